// Pixel-seeded stages and track-candidate assembly: pLS duplicate cleaning
// (SDL/Event.cc:993-1007), pixel quintuplets (SDL/Event.cc:1009-1146), pixel
// triplets (SDL/Event.cc:784-901) and track candidates (SDL/Event.cc:620-782),
// plus the flattened output lists of SDL/LST.cc:366-420.
//
// Ordering.  The reference CPU build fills pT5 and pT3 slots in the order of
// its serial loops (pLS, then connected module, then outer object) and TC
// slots in kernel order (pT5, pT3, T5, pLS), each truncated at a fixed
// capacity.  The builders here take each pLS's candidates in the same order
// (stage A one wave per pLS): a counting pass, a scan over pLS, then a
// writing pass that places every passing candidate at its serial slot.  TC
// additions rank their flags with workgroup scans.  Duplicate removal and
// cross cleaning run as all-pairs kernels in stages_clean.hip.
#include "algo_pixel.h"
#include "stages.h"
#include "wave.h"

namespace lst {
namespace {

constexpr int kWave = 64;
constexpr int kMaxConn = 64;  // >= the largest connected-module list of a superbin (40)
constexpr unsigned kPixelWaves = 4096;     // one-wave work-groups striding over the pLS
constexpr unsigned kPixWG = 256;           // stage A and the write: four waves per work-group, one pLS per wave
constexpr unsigned kPixelFullWaves = 256;  // the same for the overflow kernels, which normally return at once
constexpr unsigned kTCCap = kNMaxPixelTrackCandidates + kNMaxNonPixelTrackCandidates;

inline unsigned grid_of(uint32_t n, unsigned block) {
  unsigned g = (n + block - 1) / block;
  return g == 0 ? 1 : (g > 8192 ? 8192 : g);
}

// connected-module range of every pLS in the pixel map (SDL/Event.cc:828-856)
__global__ void k_pls_connected(LST_VIEWS) {
  LST_BIND;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    v.R.scalars[kPT5Survivors] = 0;
    v.R.scalars[kPT3Survivors] = 0;
    v.R.scalars[kPixFull] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < 2 * kPixShards) {
    v.W.pt5Shard[threadIdx.x] = 0;
    v.W.pt3Shard[threadIdx.x] = 0;
  }
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= kPT5HashBuckets; b += gridDim.x * blockDim.x)
    v.W.pt5BucketOff[b] = 0;  // hit-bucket counts of the pT5 duplicate search
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < kTCHitSet; b += gridDim.x * blockDim.x)
    v.W.tcHitSet[b] = INT32_MIN;  // the empty set of the pLS cross cleaning
  const uint32_t nPLS = v.R.scalars[kEvNPLS];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nPLS; i += gridDim.x * blockDim.x) {
    const int sb = v.P.superbin[i];
    const int type = v.P.pixelType[i];
    uint32_t base = 0, size = 0;
    if (!(sb < 0 || sb >= (int)kSizeSuperbins || type > 2 || type < 0)) {
      base = v.pixIndex[type][sb];
      size = v.pixSizes[type][sb];
    }
    v.W.connBase[i] = base;
    v.W.connSize[i] = size;
  }
}

// ------------------------------------------------------ pixel-object builders
// Pixel quintuplets (SDL/Event.cc:1009-1146) and pixel triplets
// (SDL/Event.cc:784-901) loop over the pLS, their connected modules and the
// objects of each module, and take slots in that serial order.  Candidate
// (pLS i, position L) is the L-th object over pLS i's connected modules.
//   records  per T5 / T3, the quantities of the selection that do not
//            depend on the pLS (one lane per listed object)
//   stage A  one wave per pLS: the cheap cuts; each survivor (pLS, mask
//            position, object, partial result) is appended to a list; the
//            pLS's pass-mask words are cleared
//   stage B  one lane per survivor: the costly cuts; a pass sets its bit in
//            the pLS's mask, stores its result in its survivor record, and counts the pLS;
//            pLS whose survivors did not all fit (flagged by stage A) are
//            counted instead by evaluating every candidate
//   scan     per-pLS counts to slot offsets
//   write    one lane per survivor: the passing candidates to their slots, the
//            set mask bits before each giving its rank; the flagged pLS are
//            evaluated again instead

// the connected-module prefix of pLS i in LDS: pre[k] candidates before
// module k, objBase[k] the module's first object; returns the total
__device__ __forceinline__ uint32_t conn_total(uint32_t incl, uint32_t size, unsigned lane, uint32_t* pre) {
  if (lane < size) pre[lane + 1] = incl;
  if (lane == 0) pre[0] = 0;
  __syncthreads();
  return wave_last(incl);
}
template <class K>
__device__ __forceinline__ uint32_t conn_prefix(const EventView& v, uint32_t i, uint32_t* pre, uint32_t* objBase,
                                                uint32_t& size) {
  const unsigned lane = threadIdx.x;
  size = v.W.connSize[i];
  const uint32_t base = v.W.connBase[i];
  unsigned cnt = 0;
  uint32_t ob = 0;
  if (lane < size) {
    const uint32_t m = v.M.connectedPixels[base + lane];
    if (m < v.M.nLowerModules && K::module_ok(v, m)) {
      cnt = K::n_obj(v, m);
      ob = K::obj_base(v, m);
    }
  }
  const unsigned incl = wave_inclusive_scan(cnt);
  if (lane < size) objBase[lane] = ob;
  return conn_total(incl, size, lane, pre);
}
// conn_prefix for a work-group of several waves: wave 0 fills the LDS, every
// lane reads the total back
template <class K>
__device__ __forceinline__ uint32_t conn_prefix_wg(const EventView& v, uint32_t i, uint32_t* pre, uint32_t* objBase,
                                                   uint32_t& size) {
  const unsigned lane = threadIdx.x % kWave;
  size = v.W.connSize[i];
  if (threadIdx.x < kWave) {
    const uint32_t base = v.W.connBase[i];
    unsigned cnt = 0;
    uint32_t ob = 0;
    if (lane < size) {
      const uint32_t m = v.M.connectedPixels[base + lane];
      if (m < v.M.nLowerModules && K::module_ok(v, m)) {
        cnt = K::n_obj(v, m);
        ob = K::obj_base(v, m);
      }
    }
    const unsigned incl = wave_inclusive_scan(cnt);
    if (lane < size) {
      objBase[lane] = ob;
      pre[lane + 1] = incl;
    }
    if (lane == 0) pre[0] = 0;
  }
  __syncthreads();
  return pre[size];
}
// the object of candidate L: module k is the largest with pre[k] <= L
__device__ __forceinline__ uint32_t candidate_object(const uint32_t* pre, const uint32_t* objBase, uint32_t size,
                                                     uint32_t L) {
  uint32_t lo = 0, hi = size;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= L) lo = mid;
    else hi = mid;
  }
  return objBase[lo] + (L - pre[lo]);
}

// a pLS's two lower modules and four hits (anchor and outer hit of its inner, then of its outer
// mini-doublet), from 16-byte record chunks
__device__ __forceinline__ void pls_mods_hits(const EventView& v, uint32_t pls, uint16_t mods[2], uint32_t hits[4]) {
  const uint4 s = seg_chunk(v.S, pls, 0);
  mods[0] = (uint16_t)s.z;
  mods[1] = (uint16_t)s.w;
  hits[0] = md_chunk(v.D, s.x, 0).x;
  hits[1] = md_chunk(v.D, s.x, 4).z;
  hits[2] = md_chunk(v.D, s.y, 0).x;
  hits[3] = md_chunk(v.D, s.y, 4).z;
}

struct PT5Kind {
  using Res = PT5Result;
  static constexpr uint32_t kBit = 1, kCap = kNMaxPixelQuintuplets;
  static constexpr int kSurvScalar = kPT5Survivors;
  static constexpr int kTotScalar = kTotOccPixelQuintuplets, kKeptScalar = kNPixelQuintuplets;
  static constexpr int kTnA = kTnPT5StageA, kTnB = kTnPT5StageB, kTnWrite = kTnPT5Write, kTnFull = kTnPT5FullWrite;
  __host__ __device__ static unsigned long long* mask(const EventView& v) { return v.W.pt5Mask; }
  __device__ static uint32_t* shard(const EventView& v) { return v.W.pt5Shard; }
  __host__ __device__ static uint32_t* mask_base(const EventView& v) { return v.W.pt5MaskBase; }
  __host__ __device__ static uint32_t* count(const EventView& v) { return v.W.pt5Count; }
  __device__ static bool pls_ok(const EventView& v, uint32_t i) { return v.P.isDup[i] == 0; }
  __device__ static bool module_ok(const EventView& v, uint32_t m) { return v.M.moduleType[m] != kTwoS; }
  __device__ static uint32_t n_obj(const EventView& v, uint32_t m) { return v.R.nQuintuplets[m]; }
  __device__ static uint32_t obj_base(const EventView& v, uint32_t m) { return (uint32_t)v.R.quintupletModuleIndices[m]; }
  __device__ static bool cheap(const EventView& v, const PlsCtx& p, uint32_t t, Res& r) {
    if (v.Q.isDup[t] != 0) return false;
    return pt5_cheap(v, p, t, v.W.t5Rec + (size_t)kT5RecWords * t, r);
  }
  __device__ static bool costly(const EventView& v, uint32_t pls, uint32_t i, uint32_t t, Res&) {
    return pt5_costly(v, pls, i, t);
  }
  // costly in two parts (the two pixel-tracklet tests); no result value changes
  __device__ static bool costly_half(const EventView& v, uint32_t pls, uint32_t i, uint32_t t, Res&, uint32_t part) {
    return pt3_tracklet_half(v, pls, i, v.Q.tripletIndices[2 * t], part);
  }
  __device__ static float half_out(const Res&) { return 0.f; }
  __device__ static void half_in(Res&, float) {}
  __device__ static void pack(const Res& r, uint32_t* w) {
    w[0] = __float_as_uint(r.rz); w[1] = __float_as_uint(r.rphi); w[2] = __float_as_uint(r.rphiIn);
    w[3] = __float_as_uint(r.pixR); w[4] = __float_as_uint(r.t5R); w[5] = __float_as_uint(r.cx);
    w[6] = __float_as_uint(r.cy);
  }
  __device__ static void unpack(const uint32_t* w, Res& r) {
    r.rz = __uint_as_float(w[0]); r.rphi = __uint_as_float(w[1]); r.rphiIn = __uint_as_float(w[2]);
    r.pixR = __uint_as_float(w[3]); r.t5R = __uint_as_float(w[4]); r.cx = __uint_as_float(w[5]);
    r.cy = __uint_as_float(w[6]);
  }
  // addPixelQuintupletToMemory (SDL/PixelTriplet.h:1961-2036)
  __device__ static void write(const EventView& v, uint32_t q, uint32_t pls, uint32_t i, uint32_t t5, const Res& r) {
    v.P5.pixelIndices[q] = pls;
    v.P5.T5Indices[q] = t5;
    v.P5.isDup[q] = 0;
    v.P5.score[q] = r.rphi;
    v.P5.eta[q] = v.Q.eta[t5];
    v.P5.phi[q] = v.Q.phi[t5];
    v.P5.pixelRadius[q] = r.pixR;
    v.P5.quintupletRadius[q] = r.t5R;
    v.P5.centerX[q] = r.cx;
    v.P5.centerY[q] = r.cy;
    v.P5.logicalLayers[7 * q] = 0;
    v.P5.logicalLayers[7 * q + 1] = 0;
    uint16_t pm[2];
    uint32_t ph[4];
    pls_mods_hits(v, pls, pm, ph);
    v.P5.lowerModuleIndices[7 * q] = pm[0];
    v.P5.lowerModuleIndices[7 * q + 1] = pm[1];
#pragma unroll
    for (int j = 0; j < 5; j++) {
      v.P5.logicalLayers[7 * q + 2 + j] = v.Q.logicalLayers[5 * t5 + j];
      v.P5.lowerModuleIndices[7 * q + 2 + j] = v.Q.lowerModuleIndices[5 * t5 + j];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) v.P5.hitIndices[14 * q + j] = ph[j];
    const uint2* th = reinterpret_cast<const uint2*>(&v.Q.hitIndices[10 * t5]);  // 10 words, 8-byte aligned
#pragma unroll
    for (int j = 0; j < 5; j++) {
      const uint2 h = ld_global(th + j);
      v.P5.hitIndices[14 * q + 4 + 2 * j] = h.x;
      v.P5.hitIndices[14 * q + 5 + 2 * j] = h.y;
    }
    v.P5.rzChiSquared[q] = r.rz;
    v.P5.rPhiChiSquared[q] = r.rphi;
    v.P5.rPhiChiSquaredInwards[q] = r.rphiIn;
    v.T.partOfPT5[v.Q.tripletIndices[2 * t5]] = 1;
    v.T.partOfPT5[v.Q.tripletIndices[2 * t5 + 1]] = 1;
    v.P.partOfPT5[i] = 1;
    v.Q.partOfPT5[t5] = 1;
  }
};

struct PT3Kind {
  using Res = PT3Result;
  static constexpr uint32_t kBit = 2, kCap = kNMaxPixelTriplets;
  static constexpr int kSurvScalar = kPT3Survivors;
  static constexpr int kTotScalar = kTotOccPixelTriplets, kKeptScalar = kNPixelTriplets;
  static constexpr int kTnA = kTnPT3StageA, kTnB = kTnPT3StageB, kTnWrite = kTnPT3Write, kTnFull = kTnPT3FullWrite;
  __host__ __device__ static unsigned long long* mask(const EventView& v) { return v.W.pt3Mask; }
  __device__ static uint32_t* shard(const EventView& v) { return v.W.pt3Shard; }
  __host__ __device__ static uint32_t* mask_base(const EventView& v) { return v.W.pt3MaskBase; }
  __host__ __device__ static uint32_t* count(const EventView& v) { return v.W.pt3Count; }
  __device__ static bool pls_ok(const EventView& v, uint32_t i) { return v.P.isDup[i] == 0 && v.P.partOfPT5[i] == 0; }
  __device__ static bool module_ok(const EventView& v, uint32_t m) {  // both loads issued, no branch
    const auto type = v.M.moduleType[m];
    const auto layer = v.M.layers[m];
    return (type != kTwoS) & ((layer == 1) | (layer == 2));
  }
  __device__ static uint32_t n_obj(const EventView& v, uint32_t m) { return v.R.nTriplets[m]; }
  __device__ static uint32_t obj_base(const EventView& v, uint32_t m) { return (uint32_t)v.R.tripletModuleIndices[m]; }
  // the record's last two chunks (radius, packed flags) first: most candidates end at the 2S-middle or the
  // radius test, and only the others load the points
  __device__ static bool cheap(const EventView& v, const PlsCtx& p, uint32_t t, Res& r) {
    const float4* q = reinterpret_cast<const float4*>(v.W.t3Rec + (size_t)kT3RecWords * t);
    const float4 c = ld_global(q + 2), d = ld_global(q + 3);
    const uint32_t packed = __float_as_uint(d.x);
    if ((packed >> 17) & 1) return false;  // middle module 2S
    if (!pt3_radius_ok(packed & 3, p.ptIn * kR1GeVf, p.ptErr * kR1GeVf, c.y)) return false;  // pt3_cheap's first cut
    if (v.T.partOfPT5[t] != 0) return false;
    return pt3_cheap(p, t3_rec_of(ld_global(q), ld_global(q + 1), c, d), true, r);
  }
  __device__ static bool costly(const EventView& v, uint32_t pls, uint32_t i, uint32_t t, Res& r) {
    return pt3_costly(v, pls, i, t, true, r);
  }
  // costly in two parts: the first pixel-tracklet test; the second and the
  // r-z chi2 (whose value part 1 hands over)
  __device__ static bool costly_half(const EventView& v, uint32_t pls, uint32_t i, uint32_t t, Res& r,
                                     uint32_t part) {
    return pt3_costly_half(v, pls, i, t, true, r, part);
  }
  __device__ static float half_out(const Res& r) { return r.rz; }
  __device__ static void half_in(Res& r, float rz) { r.rz = rz; }
  __device__ static void pack(const Res& r, uint32_t* w) {
    w[0] = __float_as_uint(r.pixR); w[1] = __float_as_uint(r.t3R); w[2] = __float_as_uint(r.rz);
    w[3] = __float_as_uint(r.rphi); w[4] = __float_as_uint(r.rphiIn);
  }
  __device__ static void unpack(const uint32_t* w, Res& r) {
    r.pixR = __uint_as_float(w[0]); r.t3R = __uint_as_float(w[1]); r.rz = __uint_as_float(w[2]);
    r.rphi = __uint_as_float(w[3]); r.rphiIn = __uint_as_float(w[4]);
  }
  // addPixelTripletToMemory (SDL/PixelTriplet.h:134-204)
  __device__ static void write(const EventView& v, uint32_t q, uint32_t pls, uint32_t i, uint32_t t3, const Res& r) {
    uint4 c[4];  // the T3 record's first four chunks (segments, modules, logical layers, hits)
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) c[k] = t3_chunk(v.T, t3, k);
    const uint32_t w[16] = {c[0].x, c[0].y, c[0].z, c[0].w, c[1].x, c[1].y, c[1].z, c[1].w,
                            c[2].x, c[2].y, c[2].z, c[2].w, c[3].x, c[3].y, c[3].z, c[3].w};
    constexpr uint32_t kMod = t3_word(kT3F_lowerModuleIndices), kLL = t3_word(kT3F_logicalLayers),
                       kHit = t3_word(kT3F_hitIndices);
    static_assert(kHit + 6 <= 16, "the hits are in the first four chunks");
    const uint32_t m = (uint16_t)w[kMod];
    const int adj = v.M.layers[m] == 1 ? 1 : 0;
    const uint4 sIn = seg_chunk(v.S, c[0].x, 0);
    const uint32_t mdAdj = adj ? sIn.y : sIn.x;
    const uint4 mdA = md_chunk(v.D, mdAdj, 2), mdE = md_chunk(v.D, mdAdj, 4);  // anchor phi, anchor eta
    v.P3.pixelSegmentIndices[q] = pls;
    v.P3.tripletIndices[q] = t3;
    v.P3.pixelRadius[q] = r.pixR;
    v.P3.tripletRadius[q] = r.t3R;
    v.P3.pt[q] = v.P.ptIn[i];
    v.P3.eta[q] = __uint_as_float(mdE.w);
    v.P3.phi[q] = __uint_as_float(mdA.y);
    v.P3.eta_pix[q] = v.P.eta[i];
    v.P3.phi_pix[q] = v.P.phi[i];
    v.P3.isDup[q] = 0;
    v.P3.score[q] = r.rphi + r.rphiIn;
    v.P3.centerX[q] = 0;
    v.P3.centerY[q] = 0;
    v.P3.logicalLayers[5 * q] = 0;
    v.P3.logicalLayers[5 * q + 1] = 0;
    uint16_t pm[2];
    uint32_t ph[4];
    pls_mods_hits(v, pls, pm, ph);
    v.P3.lowerModuleIndices[5 * q] = pm[0];
    v.P3.lowerModuleIndices[5 * q + 1] = pm[1];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      v.P3.logicalLayers[5 * q + 2 + j] = (uint8_t)w[kLL + j];
      v.P3.lowerModuleIndices[5 * q + 2 + j] = (uint16_t)w[kMod + j];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) v.P3.hitIndices[10 * q + j] = ph[j];
#pragma unroll
    for (int j = 0; j < 6; j++) v.P3.hitIndices[10 * q + 4 + j] = w[kHit + j];
    v.P3.rzChiSquared[q] = r.rz;
    v.P3.rPhiChiSquared[q] = r.rphi;
    v.P3.rPhiChiSquaredInwards[q] = r.rphiIn;
    v.T.partOfPT3[t3] = 1;
  }
};

// T3 slots of the modules the pixel-triplet builder connects to, module by
// module (offsets: one 1024-lane work-group; fill: one wave per module)
__global__ __launch_bounds__(1024) void k_t3_list_offsets(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[17];
  const unsigned total = scan_wide<kScanPer>(
      v.M.nLowerModules,
      [&](uint32_t m) {
        const bool ok = PT3Kind::module_ok(v, m);
        const uint32_t n = PT3Kind::n_obj(v, m);
        return ok ? n : 0u;
      },
      [&](uint32_t m, unsigned p, unsigned) { v.W.t3ListOff[m] = p; }, lds);
  if (threadIdx.x == 0) v.R.scalars[kT3Listed] = total;
}
__global__ __launch_bounds__(256) void k_t3_list_fill(LST_VIEWS) {
  LST_BIND;
  const uint32_t nL = v.M.nLowerModules;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t m = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; m < nL; m += waves) {
    if (!PT3Kind::module_ok(v, m)) continue;
    const uint32_t n = PT3Kind::n_obj(v, m);
    const uint32_t base = n ? PT3Kind::obj_base(v, m) : 0u, off = v.W.t3ListOff[m];
    for (uint32_t k = threadIdx.x % kWave; k < n; k += kWave) v.W.t3List[off + k] = base + k;
  }
}

// the records: one lane per listed T5 (those of modules that are not 2S),
// then one per listed T3
__global__ __launch_bounds__(256) void k_pix_records(LST_VIEWS) {
  LST_BIND;
  const uint32_t n5 = v.R.scalars[kT5Listed], n3 = v.R.scalars[kT3Listed];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n5 + n3; k += gridDim.x * blockDim.x) {
    if (k < n5) {
      const uint32_t t5 = v.W.t5List[k];
      if (PT5Kind::module_ok(v, v.Q.lowerModuleIndices[5 * t5]))
        make_t5_rec(v, t5, v.W.t5Rec + (size_t)kT5RecWords * t5);
    } else {
      const uint32_t t3 = v.W.t3List[k - n5];
      make_t3_rec(v, t3, v.W.t3Rec + (size_t)kT3RecWords * t3);
    }
  }
}

// exclusive scan of the per-pLS pass counts of builder K in place; the entry
// after the last pLS and the total scalar get the sum, the kept scalar the sum
// clamped to the capacity
template <class K>
__global__ __launch_bounds__(1024) void k_scan_inplace(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[17];
  uint32_t* counts = K::count(v);
  const uint32_t n = v.R.scalars[kEvNPLS];
  const unsigned sum = scan_wide<kScanPer>(
      n, [&](uint32_t i) { return counts[i]; }, [&](uint32_t i, unsigned p, unsigned) { counts[i] = p; }, lds);
  if (threadIdx.x == 0) {
    counts[n] = sum;
    v.R.scalars[K::kTotScalar] = sum;
    v.R.scalars[K::kKeptScalar] = sum < K::kCap ? sum : K::kCap;
  }
}

template <class K>
__global__ __launch_bounds__(kPixWG) void k_pix_stage_a(LST_VIEWS) {
  LST_BIND;
  constexpr uint32_t kWavesPerWG = kPixWG / kWave;
  __shared__ uint32_t preAll[kWavesPerWG][kMaxConn + 1];
  __shared__ uint32_t objBaseAll[kWavesPerWG][kMaxConn];
  const uint32_t nPLS = v.R.scalars[kEvNPLS];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  const unsigned lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  uint32_t* pre = preAll[wave];
  uint32_t* objBase = objBaseAll[wave];
  // one wave per pLS (no work-group barrier: the waves of a work-group take different pLS)
  for (uint32_t i = blockIdx.x * kWavesPerWG + wave; i < nPLS; i += gridDim.x * kWavesPerWG) {
    wave_lds_sync();  // the previous pLS's prefix in LDS is no longer read
    uint32_t total = 0, size = 0;
    const bool plsOk = K::pls_ok(v, i), connOk = v.W.connSize[i] <= (uint32_t)kMaxConn;
    if (plsOk & connOk) {  // the same branch in every lane
      size = v.W.connSize[i];
      const uint32_t base = v.W.connBase[i];
      unsigned cnt = 0;
      uint32_t ob = 0;
      if (lane < size) {
        const uint32_t m = v.M.connectedPixels[base + lane];
        if (m < v.M.nLowerModules && K::module_ok(v, m)) {
          cnt = K::n_obj(v, m);
          ob = K::obj_base(v, m);
        }
      }
      const unsigned incl = wave_inclusive_scan(cnt);
      if (lane < size) {
        objBase[lane] = ob;
        pre[lane + 1] = incl;
      }
      if (lane == 0) pre[0] = 0;
      wave_lds_sync();
      total = pre[size];
    }
    const uint32_t words = (total + kWave - 1) / kWave;
    uint32_t maskBase = ~0u;
    bool full = false;
    if (words) {
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(&K::shard(v)[i % kPixShards], words);
      b = (uint32_t)__shfl((int)b, 0, kWave);
      if (b + words <= kShardWords) maskBase = (i % kPixShards) * kShardWords + b;
      else full = true;
    }
    if (maskBase != ~0u) {
      const uint32_t pls = plsBase + i;
      const PlsCtx pc = load_pls_ctx(v, pls, i);  // once per pLS, not per candidate
      for (uint32_t w = 0; w < words; w++) {
        const uint32_t L = w * kWave + lane;
        typename K::Res r;
        uint32_t t = 0;
        bool surv = false;
        if (L < total) {
          t = candidate_object(pre, objBase, size, L);
          surv = K::cheap(v, pc, t, r);
        }
        if (lane == 0) K::mask(v)[maskBase + w] = 0ull;
        const unsigned long long sm = __ballot(surv);  // every lane of the wave takes part
        if (sm) {
          const uint32_t n = (uint32_t)__popcll(sm);
          uint32_t sb = 0;
          if (lane == 0) sb = atomicAdd(&K::shard(v)[kPixShards + i % kPixShards], n);
          sb = (uint32_t)__shfl((int)sb, 0, kWave);
          const uint32_t s = sb + lanes_below(sm);
          if (surv && s < kShardSurv) {
            uint32_t* rec = &v.W.pixSurv[(size_t)kSurvWords * ((i % kPixShards) * kShardSurv + s)];
            rec[0] = i;
            rec[1] = (maskBase + w) * kWave + lane;
            rec[2] = t;
            K::pack(r, rec + 3);
          }
          if (sb + n > kShardSurv) full = true;  // the same in every lane
        }
      }
    }
    if (lane == 0) {
      K::mask_base(v)[i] = maskBase;
      v.W.pixWords[i] = words;
      K::count(v)[i] = 0;
      v.W.pixFull[i] = full ? 1 : 0;
      if (full) atomicOr(&v.R.scalars[kPixFull], K::kBit);
    }
  }
}

// stage B, and the counting of the pLS flagged by stage A: their survivors are
// skipped here, and when any pLS is flagged the same waves then evaluate every
// candidate of each flagged pLS (its count only; the write evaluates again)
template <class K>
__global__ __launch_bounds__(kPixWG) void k_pix_stage_b(LST_VIEWS) {
  LST_BIND;
  __shared__ uint32_t pre[kMaxConn + 1];  // survivors before each shard, then the connected-module prefix
  __shared__ uint32_t objBase[kMaxConn];
  __shared__ uint32_t sPass[kPixWG / kWave / 2][kWave];
  __shared__ float sOut[kPixWG / kWave / 2][kWave];
  __shared__ uint32_t sCount[kPixWG / kWave];
  static_assert(kMaxConn >= (int)kPixShards, "the shard prefix reuses the module prefix");
  const unsigned lane = threadIdx.x % kWave, wave = threadIdx.x / kWave, part = wave & 1, pair = wave >> 1;
  if (wave == 0) {
    const uint32_t c = K::shard(v)[kPixShards + lane];
    const uint32_t incl = wave_inclusive_scan(c < kShardSurv ? c : kShardSurv);
    pre[lane + 1] = incl;
    if (lane == 0) pre[0] = 0;
  }
  __syncthreads();
  const uint32_t nS = pre[kPixShards];
  if (blockIdx.x == 0 && threadIdx.x == 0) v.R.scalars[K::kSurvScalar] = nS;
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  // waves 2p and 2p + 1 take the two parts of the costly tests of the same
  // 64 survivors; the odd wave hands its outcome over in LDS
  constexpr uint32_t kPerPass = kPixWG / 2;
  for (uint32_t g0 = blockIdx.x * kPerPass; g0 < nS; g0 += gridDim.x * kPerPass) {
    const uint32_t g = g0 + pair * kWave + lane;
    bool act = g < nS, pass = false;
    uint32_t s = 0, i = 0, slot = 0, t = 0;
    typename K::Res r;
    if (act) {
      uint32_t lo = 0, hi = kPixShards;  // the shard holding survivor g: largest lo with pre[lo] <= g
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= g) lo = mid;
        else hi = mid;
      }
      s = lo * kShardSurv + (g - pre[lo]);
      const uint32_t* rec = &v.W.pixSurv[(size_t)kSurvWords * s];
      i = rec[0]; slot = rec[1]; t = rec[2];
      act = v.W.pixFull[i] == 0;
      if (act) {
        K::unpack(rec + 3, r);
        pass = K::costly_half(v, plsBase + i, i, t, r, part);
      }
    }
    if (part == 1) {
      sPass[pair][lane] = pass ? 1u : 0u;
      sOut[pair][lane] = pass ? K::half_out(r) : 0.f;
    }
    __syncthreads();
    if (part == 0 && pass && sPass[pair][lane]) {
      K::half_in(r, sOut[pair][lane]);
      K::pack(r, &v.W.pixSurv[(size_t)kSurvWords * s + 3]);
      atomicOr(&K::mask(v)[slot / kWave], 1ull << (slot % kWave));
      atomicAdd(&K::count(v)[i], 1u);
    }
    __syncthreads();  // the LDS outcomes are read before the next survivors overwrite them
  }
  if ((v.R.scalars[kPixFull] & K::kBit) == 0) return;
  const uint32_t nPLS = v.R.scalars[kEvNPLS];
  for (uint32_t i = blockIdx.x; i < nPLS; i += gridDim.x) {
    if (!v.W.pixFull[i]) continue;  // the same in every lane
    __syncthreads();
    uint32_t size;
    const uint32_t total = conn_prefix_wg<K>(v, i, pre, objBase, size);
    unsigned count = 0;
    const PlsCtx pc = load_pls_ctx(v, plsBase + i, i);
    for (uint32_t L0 = wave * kWave; L0 < total; L0 += kPixWG) {
      const uint32_t L = L0 + lane;
      bool pass = false;
      if (L < total) {
        const uint32_t t = candidate_object(pre, objBase, size, L);
        typename K::Res r;
        pass = K::cheap(v, pc, t, r) && K::costly(v, plsBase + i, i, t, r);
      }
      count += (unsigned)__popcll(__ballot(pass));
    }
    if (lane == 0) sCount[wave] = count;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned c = 0;
      for (uint32_t w = 0; w < kPixWG / kWave; w++) c += sCount[w];
      K::count(v)[i] = c;
    }
  }
}

// One lane per survivor of stage A (the shards' lists, as stage B numbers
// them): a survivor whose mask bit stage B set is written at its slot, the
// pLS's first slot (count, scanned) plus the passing candidates before it in
// the pLS's mask words.  The pLS flagged full are written by k_pix_full_write.
template <class K>
__global__ __launch_bounds__(kPixWG) void k_pix_write(LST_VIEWS) {
  LST_BIND;
  __shared__ uint32_t pre[kPixShards + 1];  // survivors before each shard
  if (threadIdx.x < kWave) {
    const uint32_t c = K::shard(v)[kPixShards + threadIdx.x];
    const uint32_t incl = wave_inclusive_scan(c < kShardSurv ? c : kShardSurv);
    pre[threadIdx.x + 1] = incl;
    if (threadIdx.x == 0) pre[0] = 0;
  }
  __syncthreads();
  static_assert(kPixShards == kWave, "one lane per shard");
  const uint32_t nS = pre[kPixShards];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < nS; g += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = kPixShards;  // the shard holding survivor g: largest lo with pre[lo] <= g
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pre[mid] <= g) lo = mid;
      else hi = mid;
    }
    const uint32_t s = lo * kShardSurv + (g - pre[lo]);
    const uint32_t* rec = &v.W.pixSurv[(size_t)kSurvWords * s];
    const uint32_t i = rec[0], slot = rec[1], t = rec[2];
    const unsigned long long* mask = K::mask(v);
    const unsigned long long mk = mask[slot / kWave];
    if (((mk >> (slot % kWave)) & 1ull) == 0 || v.W.pixFull[i]) continue;
    uint32_t q = K::count(v)[i] + (uint32_t)__popcll(mk & ((1ull << (slot % kWave)) - 1ull));
    for (uint32_t w = K::mask_base(v)[i]; w < slot / kWave; w++) q += (uint32_t)__popcll(mask[w]);
    if (q < K::kCap) {
      typename K::Res r;
      K::unpack(rec + 3, r);
      K::write(v, q, plsBase + i, i, t, r);
    }
  }
}

// the flagged pLS: every candidate evaluated again and written to its slot
template <class K>
__global__ __launch_bounds__(kWave) void k_pix_full_write(LST_VIEWS) {
  LST_BIND;
  if ((v.R.scalars[kPixFull] & K::kBit) == 0) return;
  __shared__ uint32_t pre[kMaxConn + 1];
  __shared__ uint32_t objBase[kMaxConn];
  const uint32_t nPLS = v.R.scalars[kEvNPLS];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  const unsigned lane = threadIdx.x;
  for (uint32_t i = blockIdx.x; i < nPLS; i += gridDim.x) {
    if (!v.W.pixFull[i]) continue;
    __syncthreads();
    uint32_t size;
    const uint32_t total = conn_prefix<K>(v, i, pre, objBase, size);
    const uint32_t slotBase = K::count(v)[i], pls = plsBase + i;
    const PlsCtx pc = load_pls_ctx(v, pls, i);
    unsigned done = 0;
    for (uint32_t L0 = 0; L0 < total; L0 += kWave) {
      const uint32_t L = L0 + lane;
      bool pass = false;
      uint32_t t = 0;
      typename K::Res r;
      if (L < total) {
        t = candidate_object(pre, objBase, size, L);
        pass = K::cheap(v, pc, t, r) && K::costly(v, pls, i, t, r);
      }
      unsigned n;
      const unsigned rank = done + wave_rank(pass, n);
      if (pass && slotBase + rank < K::kCap) K::write(v, slotBase + rank, pls, i, t, r);
      done += n;
    }
  }
}

// Flag compaction over items [0, n) by 256-lane work-groups: work-group b
// takes items [256 b, 256 b + 256) and counts the flagged items before them
// itself, so the work-groups need no scan between them.  Returns this lane's
// rank among the flagged items; `last` is set in the work-group holding item
// n - 1 (work-group 0 when n is 0), whose `total` is then the flagged total.
constexpr unsigned kCompactBlock = 256;
template <class Flag>
__device__ __forceinline__ unsigned block_compact(uint32_t n, Flag flag, uint32_t item, bool& flagged, bool& last,
                                                  unsigned& total) {
  __shared__ unsigned lds[kCompactBlock / 64 + 1];
  const uint32_t b0 = blockIdx.x * kCompactBlock;
  unsigned c = 0;
  for (uint32_t q = threadIdx.x; q < b0 && q < n; q += kCompactBlock) c += flag(q) ? 1u : 0u;
  unsigned before;
  block_exclusive_scan<kCompactBlock>(c, lds, before);
  flagged = item < n && flag(item);
  unsigned tot;
  const unsigned rank = before + block_exclusive_scan<kCompactBlock>(flagged ? 1u : 0u, lds, tot);
  last = (b0 < n && n <= b0 + kCompactBlock) || (n == 0 && blockIdx.x == 0);
  total = before + tot;
  return rank;
}
inline unsigned compact_grid(uint32_t cap) { return (cap + kCompactBlock - 1) / kCompactBlock; }

// addTrackCandidateToMemory (SDL/TrackCandidate.h:126-164), with the entries
// past the object's size cleared (the reference's buffers start zeroed)
__device__ __forceinline__ void add_tc(const TCSoA& C, uint32_t idx, int16_t type, uint32_t inner, uint32_t outer,
                                       const uint8_t* ll, const uint16_t* lm, const uint32_t* hits, int seed, float cx,
                                       float cy, float radius, uint32_t direct) {
  C.trackCandidateType[idx] = type;
  C.directObjectIndices[idx] = direct;
  C.pixelSeedIndex[idx] = seed;
  C.objectIndices[2 * idx] = inner;
  C.objectIndices[2 * idx + 1] = outer;
  const int limits = type == 7 ? 7 : 5;
  for (int i = 0; i < 7; i++) {
    C.logicalLayers[7 * idx + i] = i < limits ? ll[i] : 0;
    C.lowerModuleIndices[7 * idx + i] = i < limits ? lm[i] : 0;
  }
  for (int i = 0; i < 14; i++) C.hitIndices[14 * idx + i] = i < 2 * limits ? hits[i] : 0;
  C.centerX[idx] = cx;
  C.centerY[idx] = cy;
  C.radius[idx] = radius;
}

// addpT5asTrackCandidateInGPU (SDL/TrackCandidate.h:532-580): the first TCs
__global__ __launch_bounds__(kCompactBlock) void k_tc_add_pt5(LST_VIEWS) {
  LST_BIND;
  const uint32_t n = v.R.scalars[kNPixelQuintuplets];
  if (blockIdx.x * kCompactBlock >= n && blockIdx.x != 0) return;
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  const uint32_t q = blockIdx.x * kCompactBlock + threadIdx.x;
  bool add, last;
  unsigned total;
  const unsigned slot = block_compact(n, [&](uint32_t k) { return v.P5.isDup[k] == 0; }, q, add, last, total);
  if (add && slot < kNMaxPixelTrackCandidates) {
    const uint32_t pls = v.P5.pixelIndices[q];
    const float radius = 0.5f * (v.P5.pixelRadius[q] + v.P5.quintupletRadius[q]);
    add_tc(v.C, slot, 7, pls, v.P5.T5Indices[q], &v.P5.logicalLayers[7 * q], &v.P5.lowerModuleIndices[7 * q],
           &v.P5.hitIndices[14 * q], (int)v.P.seedIdx[pls - plsBase], v.P5.centerX[q], v.P5.centerY[q], radius, q);
  }
  if (last && threadIdx.x == 0) {
    const unsigned added = total < kNMaxPixelTrackCandidates ? total : kNMaxPixelTrackCandidates;
    v.R.scalars[kNTrackCandidates] = added;
    v.R.scalars[kNTrackCandidatespT5] = added;
    v.R.scalars[kNTrackCandidatespT3] = 0;
    v.R.scalars[kNTrackCandidatespLS] = 0;
    v.R.scalars[kNTrackCandidatesT5] = 0;
  }
}

// ---------------------------------------------------------- track candidates
// addpT3asTrackCandidatesInGPU (SDL/TrackCandidate.h:381-428)
__global__ __launch_bounds__(kCompactBlock) void k_tc_add_pt3(LST_VIEWS) {
  LST_BIND;
  const uint32_t n = v.R.scalars[kNPixelTriplets];
  if (blockIdx.x * kCompactBlock >= n && blockIdx.x != 0) return;
  const uint32_t base = v.R.scalars[kNTrackCandidates];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  const uint32_t p = blockIdx.x * kCompactBlock + threadIdx.x;
  bool add, last;
  unsigned total;
  const unsigned slot = base + block_compact(n, [&](uint32_t k) { return v.P3.isDup[k] == 0; }, p, add, last, total);
  if (add && slot < kNMaxPixelTrackCandidates) {
    const float radius = 0.5f * (v.P3.pixelRadius[p] + v.P3.tripletRadius[p]);
    const uint32_t pls = v.P3.pixelSegmentIndices[p];
    add_tc(v.C, slot, 5, p, p, &v.P3.logicalLayers[5 * p], &v.P3.lowerModuleIndices[5 * p], &v.P3.hitIndices[10 * p],
           (int)v.P.seedIdx[pls - plsBase], v.P3.centerX[p], v.P3.centerY[p], radius, p);
  }
  if (last && threadIdx.x == 0) {
    const unsigned room = base < kNMaxPixelTrackCandidates ? kNMaxPixelTrackCandidates - base : 0u;
    const unsigned added = total < room ? total : room;
    v.R.scalars[kNTrackCandidates] = base + added;
    v.R.scalars[kNTrackCandidatespT3] = added;
  }
}

// addT5asTrackCandidateInGPU (SDL/TrackCandidate.h:430-482): count per module,
// scan over the eligible modules, then place in module and slot order
__device__ __forceinline__ bool t5_tc_flag(const EventView& v, uint32_t q) {
  return v.Q.isDup[q] == 0 && v.Q.partOfPT5[q] == 0 && v.Q.TightCutFlag[q] != 0;
}

// one wave per eligible module at a time, striding over the modules
__device__ __forceinline__ uint32_t wave_index() {
  return __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}
__device__ __forceinline__ uint32_t wave_count() { return (gridDim.x * blockDim.x) / kWave; }
constexpr unsigned kModuleWaveBlocks = 512;  // x 4 waves

__global__ __launch_bounds__(256) void k_tc_t5_count(LST_VIEWS) {
  LST_BIND;
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  for (uint32_t e = wave_index(); e < nElig; e += wave_count()) {
    const uint16_t m = v.R.indicesOfEligibleT5Modules[e];
    const uint32_t nq = v.R.nQuintuplets[m];
    const uint32_t base = (uint32_t)v.R.quintupletModuleIndices[m];
    unsigned c = 0;
    for (uint32_t k0 = 0; k0 < nq; k0 += kWave) {
      const uint32_t k = k0 + lane_id();
      unsigned n;
      wave_rank(k < nq && t5_tc_flag(v, base + k), n);
      c += n;
    }
    if (lane_id() == 0) v.W.t5TcCount[e] = c;
  }
}

__global__ __launch_bounds__(1024) void k_tc_t5_scan(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[17];
  const unsigned total = scan_wide<kScanPer>(
      v.R.scalars[kNEligibleT5], [&](uint32_t i) { return v.W.t5TcCount[i]; },
      [&](uint32_t i, unsigned p, unsigned) { v.W.t5TcCount[i] = p; }, lds);
  if (threadIdx.x == 0) {
    const unsigned added = total < kNMaxNonPixelTrackCandidates ? total : kNMaxNonPixelTrackCandidates;
    v.R.scalars[kNTrackCandidatesT5] = added;
    v.R.scalars[kNTrackCandidates] += added;
  }
}

__global__ __launch_bounds__(256) void k_tc_t5_write(LST_VIEWS) {
  LST_BIND;
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  const uint32_t tcBase = v.R.scalars[kNTrackCandidatespT5] + v.R.scalars[kNTrackCandidatespT3];
  for (uint32_t e = wave_index(); e < nElig; e += wave_count()) {
    const uint16_t m = v.R.indicesOfEligibleT5Modules[e];
    const uint32_t nq = v.R.nQuintuplets[m];
    const uint32_t base = (uint32_t)v.R.quintupletModuleIndices[m];
    unsigned carry = v.W.t5TcCount[e];
    for (uint32_t k0 = 0; k0 < nq; k0 += kWave) {
      const uint32_t k = k0 + lane_id();
      const bool add = k < nq && t5_tc_flag(v, base + k);
      unsigned tot;
      const unsigned rank = carry + wave_rank(add, tot);
      if (add && rank < kNMaxNonPixelTrackCandidates) {
        const uint32_t q = base + k;
        add_tc(v.C, tcBase + rank, 4, q, q, &v.Q.logicalLayers[5 * q], &v.Q.lowerModuleIndices[5 * q],
               &v.Q.hitIndices[10 * q], -1, v.Q.regressionG[q], v.Q.regressionF[q], v.Q.regressionRadius[q], q);
      }
      carry += tot;
    }
  }
}

// addpLSasTrackCandidateInGPU (SDL/TrackCandidate.h:484-530) with
// addpLSTrackCandidateToMemory (:112-124): hits stored in (x, z, y, w) order
__global__ __launch_bounds__(kCompactBlock) void k_tc_add_pls(LST_VIEWS) {
  LST_BIND;
  const uint32_t n = v.R.scalars[kEvNPLS];
  if (blockIdx.x * kCompactBlock >= n && blockIdx.x != 0) return;
  const uint32_t base = v.R.scalars[kNTrackCandidates];
  const uint32_t nT5 = v.R.scalars[kNTrackCandidatesT5];
  const uint32_t i = blockIdx.x * kCompactBlock + threadIdx.x;
  const bool triplets = v.tcPlsTriplets;
  bool add, last;
  unsigned total;
  const unsigned slot =
      base + block_compact(
                 n, [&](uint32_t k) { return triplets ? v.P.isDup[k] == 0 : (v.P.isQuad[k] && v.P.isDup[k] == 0); }, i,
                 add, last, total);
  if (add && slot - nT5 < kNMaxPixelTrackCandidates) {
    v.C.trackCandidateType[slot] = 8;
    v.C.directObjectIndices[slot] = i;
    v.C.pixelSeedIndex[slot] = (int)v.P.seedIdx[i];
    v.C.objectIndices[2 * slot] = i;
    v.C.objectIndices[2 * slot + 1] = i;
    const uint32_t* h = &v.P.pLSHitsIdxs[4 * i];
    v.C.hitIndices[14 * slot + 0] = h[0];
    v.C.hitIndices[14 * slot + 1] = h[2];
    v.C.hitIndices[14 * slot + 2] = h[1];
    v.C.hitIndices[14 * slot + 3] = h[3];
    for (int k = 4; k < 14; k++) v.C.hitIndices[14 * slot + k] = 0;
    for (int k = 0; k < 7; k++) {
      v.C.logicalLayers[7 * slot + k] = 0;
      v.C.lowerModuleIndices[7 * slot + k] = 0;
    }
    v.C.centerX[slot] = 0;  // not written by the reference for pLS candidates
    v.C.centerY[slot] = 0;
    v.C.radius[slot] = 0;
  }
  if (last && threadIdx.x == 0) {
    const unsigned limit = nT5 + kNMaxPixelTrackCandidates;
    const unsigned room = base < limit ? limit - base : 0u;
    const unsigned added = total < room ? total : room;
    v.R.scalars[kNTrackCandidates] = base + added;
    v.R.scalars[kNTrackCandidatespLS] = added;
  }
}

// LST::getHitIdxs (SDL/LST.cc:393-420): ntuple hit indices of each TC, the
// repeated fourth hit of a triplet seed dropped
// the output hit lists (outHits, outLen), and the same with the seed index and
// type as one packed block (outPacked) that a batch copies to the host at once
__global__ void k_tc_output(LST_VIEWS) {
  LST_BIND;
  const uint32_t nTC = v.R.scalars[kNTrackCandidates];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // header: the count, then which flattened builders overflowed, which capacities truncated and which
    // inputs were bad (the truncation bits are final: k_tc_prep ran before)
    const uint32_t* sc = v.R.scalars;
    uint32_t ovf = v.perModuleBuilders ? (uint32_t)kOvfPerModule : 0u;
    ovf |= sc[kMDCandTotal] > kMDCandMax ? (uint32_t)kOvfMD : 0u;
    ovf |= sc[kLSCandTotal] > kLSCandMax ? (uint32_t)kOvfLS : 0u;
    ovf |= sc[kT3CandTotal] > kT3CandMax ? (uint32_t)kOvfT3 : 0u;
    ovf |= (sc[kT5CandTotal] > kT5CandMax || sc[kT5Results] > kT5ResultMax) ? (uint32_t)kOvfT5 : 0u;
    uint32_t trunc = sc[kTruncated];
    trunc |= sc[kTotOccPixelQuintuplets] > sc[kNPixelQuintuplets] ? (uint32_t)kTruncPT5 : 0u;
    trunc |= sc[kTotOccPixelTriplets] > sc[kNPixelTriplets] ? (uint32_t)kTruncPT3 : 0u;
    v.W.outPacked[kOutNTC] = nTC;
    v.W.outPacked[kOutOverflow] = ovf;
    v.W.outPacked[kOutTruncated] = trunc;
    v.W.outPacked[kOutBadHits] = sc[kBadHits];
  }
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nTC; t += gridDim.x * blockDim.x) {
    const int16_t type = v.C.trackCandidateType[t];
    const int maxN = type == 7 ? 14 : (type == 5 || type == 4) ? 10 : type == 8 ? 4 : 0;
    // the hit slots, then their ntuple indices, loaded before any use with no
    // branch between the loads (an unused slot reads slot 0 and hit 0, both in
    // bounds, and its value is not used), then compacted
    uint32_t h[14], idx[14], hits[14];
#pragma unroll
    for (int i = 0; i < 14; i++) h[i] = v.C.hitIndices[14 * t + (i < maxN ? i : 0)];
#pragma unroll
    for (int i = 0; i < 14; i++) {
      const bool direct = type == 8 || i >= maxN;
      const uint32_t g = v.H.idxs[direct ? 0u : h[i]];
      idx[i] = direct ? h[i] : g;
    }
    uint32_t len = 0, last = 0;
#pragma unroll
    for (int i = 0; i < 14; i++) hits[i] = 0;
#pragma unroll
    for (int i = 0; i < 14; i++) {
      if (i >= maxN) break;
      if (type != 4 && len == 3 && last == idx[i]) continue;
#pragma unroll
      for (int k = 0; k < 14; k++)  // register write at a run-time position without a scratch array
        if ((uint32_t)k == len) hits[k] = idx[i];
      last = idx[i];
      len++;
    }
#pragma unroll
    for (int i = 0; i < 14; i++) v.W.outHits[14 * t + i] = hits[i];
    v.W.outLen[t] = len;
    if (t < v.outCap) {  // the packed block (possibly host memory) holds outCap records
      uint32_t* rec = v.W.outPacked + kOutHeaderWords + (size_t)kOutRecWords * t;
#pragma unroll
      for (int i = 0; i < 14; i++) rec[i] = hits[i];
      rec[14] = len;
      rec[15] = (uint32_t)v.C.pixelSeedIndex[t];
      rec[16] = (uint32_t)(int32_t)type;
    }
  }
}

// the packed output block of each event to its descriptor's device block and
// the block's header to its host block, 16 bytes per lane (the blocks are
// 256-byte aligned).  The host reads the headers and copies the records it
// needs from the device block on a copy engine, so no kernel of the launch set
// waits on writes across the host link.
__global__ __launch_bounds__(256) void k_out_copy(LST_VIEWS, const InputDesc* __restrict__ ds) {
  LST_BIND;
  const InputDesc& d = ds[blockIdx.z];
  if (!d.outHost) return;
  const uint32_t nTC = v.W.outPacked[kOutNTC];
  const uint32_t words = kOutHeaderWords + (nTC < d.outCap ? nTC : d.outCap) * kOutRecWords;
  const uint4* src4 = reinterpret_cast<const uint4*>(v.W.outPacked);
  uint4* dev4 = reinterpret_cast<uint4*>(d.outDev);
  uint4* host4 = reinterpret_cast<uint4*>(d.outHost);
  static_assert(kOutHeaderWords % 4 == 0, "the header is whole 16-byte words");
  const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (uint32_t i = t0; i < words / 4; i += stride) {
    const uint4 x = src4[i];
    dev4[i] = x;
    if (i < kOutHeaderWords / 4) host4[i] = x;
  }
  if (t0 < (words & 3u)) d.outDev[(words & ~3u) + t0] = v.W.outPacked[(words & ~3u) + t0];
}

}  // namespace

void launch_pixel_segment_cleaning(const Launch& L) { launch_pls_dup(L, false); }

template <class K>
void launch_pixel_builder(const Launch& L, int timer) {
  L.t.begin(L.s, timer);
  LST_GO_T(L, K::kTnA, k_pix_stage_a<K>, kPixelWaves, kPixWG);
  L.t.end(L.s, timer);
  LST_GO_T(L, K::kTnB, k_pix_stage_b<K>, kPixelWaves, kPixWG);
  LST_GO(L, k_scan_inplace<K>, 1, 1, 1024);
  LST_GO_T(L, K::kTnWrite, k_pix_write<K>, kPixelWaves, kPixWG);
  LST_GO_T(L, K::kTnFull, k_pix_full_write<K>, kPixelFullWaves, kWave);
}

void launch_pixel_quintuplets(const Launch& L) {
  LST_GO_T(L, kTnPlsConnected, k_pls_connected, grid_of(kNMaxPixelSegments, 256), 256);
  LST_GO(L, k_t3_list_offsets, 1, 1, 1024);
  LST_GO_T(L, kTnT3ListFill, k_t3_list_fill, kModuleWaveBlocks, 256);
  LST_GO_T(L, kTnPixRecords, k_pix_records, 256, 256);
  launch_pixel_builder<PT5Kind>(L, kTmPT5Build);
  launch_pt5_dup(L);
  LST_GO(L, k_tc_add_pt5, compact_grid(kNMaxPixelQuintuplets), 1, kCompactBlock);
}

void launch_pixel_triplets(const Launch& L) {
  launch_pixel_builder<PT3Kind>(L, kTmPT3Build);
  launch_pt3_dup(L);
}

void launch_track_candidates(const Launch& L) {
  launch_cross_pt3(L);
  LST_GO(L, k_tc_add_pt3, compact_grid(kNMaxPixelTriplets), 1, kCompactBlock);
  launch_t5_dup_before_tc(L);
  launch_cross_t5(L);
  LST_GO_T(L, kTnTcT5Count, k_tc_t5_count, kModuleWaveBlocks, 256);
  LST_GO(L, k_tc_t5_scan, 1, 1, 1024);
  LST_GO_T(L, kTnTcT5Write, k_tc_t5_write, kModuleWaveBlocks, 256);
  launch_pls_dup(L, true);
  launch_cross_pls(L);
  LST_GO(L, k_tc_add_pls, compact_grid(kNMaxPixelSegments), 1, kCompactBlock);
  LST_GO_T(L, kTnTcOutput, k_tc_output, grid_of(kTCCap, 256), 256);
}

void launch_out_copy(const Launch& L, const InputDesc* d) {
  L.tune_begin(kTnOutCopy);
  hipLaunchKernelGGL(k_out_copy, dim3(L.gx(kTnOutCopy, 64), 1, L.nEv), dim3(256), 0, L.s, L.dv, d);
  L.tune_end(kTnOutCopy);
}

}  // namespace lst
