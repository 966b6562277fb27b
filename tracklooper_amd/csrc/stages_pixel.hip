// Pixel-seeded stages and track-candidate assembly: pLS duplicate cleaning
// (SDL/Event.cc:993-1007), pixel quintuplets (SDL/Event.cc:1009-1146), pixel
// triplets (SDL/Event.cc:784-901) and track candidates (SDL/Event.cc:620-782),
// plus the flattened output lists of SDL/LST.cc:366-420.
//
// Ordering.  The reference CPU build fills pT5 and pT3 slots in the order of
// its serial loops (pLS, then connected module, then outer object) and TC
// slots in kernel order (pT5, pT3, T5, pLS), each truncated at a fixed
// capacity.  The builders here run one wave per pLS over that pLS's
// candidates in the same order: a counting pass, a scan over pLS, then a
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
constexpr unsigned kPixelWaves = 4096;  // one-wave work-groups striding over the pLS
constexpr unsigned kTCCap = kNMaxPixelTrackCandidates + kNMaxNonPixelTrackCandidates;

inline unsigned grid_of(uint32_t n, unsigned block) {
  unsigned g = (n + block - 1) / block;
  return g == 0 ? 1 : (g > 8192 ? 8192 : g);
}

// exclusive scan of counts[0, *count) in place; counts[*count] and *total get the sum
__global__ __launch_bounds__(1024) void k_scan_inplace(uint32_t* counts, const uint32_t* count, uint32_t* total) {
  __shared__ unsigned lds[17];
  const uint32_t n = *count;
  unsigned carry = 0;
  for (uint32_t b = 0; b < n; b += 1024) {
    const uint32_t i = b + threadIdx.x;
    const unsigned c = i < n ? counts[i] : 0u;
    unsigned tot;
    const unsigned ex = block_exclusive_scan<1024>(c, lds, tot);
    if (i < n) counts[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    counts[n] = carry;
    if (total) *total = carry;
  }
}

// connected-module range of every pLS in the pixel map (SDL/Event.cc:828-856)
__global__ void k_pls_connected(EventView v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    v.R.scalars[kPT5MaskWords] = 0;
    v.R.scalars[kPT3MaskWords] = 0;
    v.R.scalars[kPT5Parked] = 0;
    v.R.scalars[kPT3Parked] = 0;
  }
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

// ---------------------------------------------------------- pixel quintuplets
// one wave per pLS (striding over the pLS); candidates (connected module k, T5 b) in serial order
template <bool kWrite>
__global__ __launch_bounds__(kWave) void k_pixel_quintuplets(EventView v) {
  __shared__ uint32_t pre[kMaxConn + 1];
  const uint32_t nPLS = v.R.scalars[kEvNPLS];
  for (uint32_t i = blockIdx.x; i < nPLS; i += gridDim.x) {
    __syncthreads();  // the previous pLS's prefix in LDS is no longer read
    const unsigned lane = threadIdx.x;
    const uint32_t nL = v.M.nLowerModules;
    unsigned count = 0;
    const uint32_t size = v.W.connSize[i];
    if (v.P.isDup[i] == 0 && size <= (uint32_t)kMaxConn) {
      const uint32_t base = v.W.connBase[i];
      unsigned cnt = 0;
      if (lane < size) {
        const uint32_t m = v.M.connectedPixels[base + lane];
        if (m < nL && v.M.moduleType[m] != kTwoS) cnt = v.R.nQuintuplets[m];
      }
      const unsigned incl = wave_inclusive_scan(cnt);
      if (lane < size) pre[lane + 1] = incl;
      if (lane == 0) pre[0] = 0;
      const uint32_t total = __shfl(incl, kWave - 1, kWave);
      __syncthreads();
      const uint32_t pls = (uint32_t)v.segModuleIndices[nL] + i;
      const uint32_t slotBase = kWrite ? v.W.pt5Count[i] : 0;
      // pass masks: the counting pass records one 64-bit mask per chunk, and the
      // writing pass evaluates only the lanes those masks mark
      uint32_t maskBase = ~0u;
      if (!kWrite) {
        const uint32_t words = (total + kWave - 1) / kWave;
        uint32_t b = 0;
        if (lane == 0 && words) b = atomicAdd(&v.R.scalars[kPT5MaskWords], words);
        b = __shfl(b, 0, kWave);
        if (words && b + words <= kPixMaskWords) maskBase = b;
        if (lane == 0) v.W.pt5MaskBase[i] = maskBase;
      } else {
        maskBase = v.W.pt5MaskBase[i];
      }
      const uint32_t maskWords = (total + kWave - 1) / kWave;
      uint32_t maskHeld = ~0u, parkReg = ~0u;
      unsigned long long maskReg = 0ull;
      for (uint32_t c0 = 0; c0 < total; c0 += kWave) {
        const uint32_t L = c0 + lane, w = c0 / kWave;
        bool pass = false;
        uint32_t t5 = 0;
        PT5Result r;
        bool eval = L < total;
        if (kWrite && maskBase != ~0u) {
          // each lane holds one chunk's mask and park position; 64 chunks per refill
          const uint32_t wb = w & ~(uint32_t)(kWave - 1);
          if (wb != maskHeld) {
            maskHeld = wb;
            maskReg = wb + lane < maskWords ? v.W.pt5Mask[maskBase + wb + lane] : 0ull;
            parkReg = wb + lane < maskWords ? v.W.pt5ParkAt[maskBase + wb + lane] : ~0u;
          }
          const unsigned long long mk = __shfl(maskReg, (int)(w - wb), kWave);
          if (mk == 0ull) continue;  // no passing candidate in this chunk
          const uint32_t pb = __shfl(parkReg, (int)(w - wb), kWave);
          eval = eval && ((mk >> lane) & 1ull);
          if (pb != ~0u) {  // results parked by the counting pass: no second evaluation
            if (eval) {
              const uint32_t* rec = &v.W.pt5Park[kPT5ParkWords * (pb + lanes_below(mk))];
              t5 = rec[0];
              r.rz = __uint_as_float(rec[1]); r.rphi = __uint_as_float(rec[2]); r.rphiIn = __uint_as_float(rec[3]);
              r.pixR = __uint_as_float(rec[4]); r.t5R = __uint_as_float(rec[5]);
              r.cx = __uint_as_float(rec[6]); r.cy = __uint_as_float(rec[7]);
              pass = true;
            }
            eval = false;
          }
        }
        if (eval) {
          uint32_t lo = 0, hi = size;  // largest k with pre[k] <= L
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= L) lo = mid;
            else hi = mid;
          }
          const uint32_t m = v.M.connectedPixels[base + lo];
          t5 = (uint32_t)v.R.quintupletModuleIndices[m] + (L - pre[lo]);
          if (v.Q.isDup[t5] == 0) pass = pt5_select(v, pls, i, t5, r);
        }
        unsigned n;
        const unsigned rank = count + wave_rank(pass, n);
        const unsigned long long passMask = __ballot(pass);  // every lane of the wave takes part
        if (!kWrite && maskBase != ~0u) {
          uint32_t pb = ~0u;
          if (passMask) {
            if (lane == 0) {
              pb = atomicAdd(&v.R.scalars[kPT5Parked], (uint32_t)__popcll(passMask));
              if (pb + (uint32_t)__popcll(passMask) > kPixParkMax) pb = ~0u;
            }
            pb = __shfl(pb, 0, kWave);
          }
          if (lane == 0) {
            v.W.pt5Mask[maskBase + w] = passMask;
            v.W.pt5ParkAt[maskBase + w] = pb;
          }
          if (pass && pb != ~0u) {
            uint32_t* rec = &v.W.pt5Park[kPT5ParkWords * (pb + lanes_below(passMask))];
            rec[0] = t5;
            rec[1] = __float_as_uint(r.rz); rec[2] = __float_as_uint(r.rphi); rec[3] = __float_as_uint(r.rphiIn);
            rec[4] = __float_as_uint(r.pixR); rec[5] = __float_as_uint(r.t5R);
            rec[6] = __float_as_uint(r.cx); rec[7] = __float_as_uint(r.cy);
          }
        }
        if (kWrite && pass && slotBase + rank < kNMaxPixelQuintuplets) {
          // addPixelQuintupletToMemory (SDL/PixelTriplet.h:1961-2036)
          const uint32_t q = slotBase + rank;
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
          v.P5.lowerModuleIndices[7 * q] = v.S.innerLowerModuleIndices[pls];
          v.P5.lowerModuleIndices[7 * q + 1] = v.S.outerLowerModuleIndices[pls];
  #pragma unroll
          for (int j = 0; j < 5; j++) {
            v.P5.logicalLayers[7 * q + 2 + j] = v.Q.logicalLayers[5 * t5 + j];
            v.P5.lowerModuleIndices[7 * q + 2 + j] = v.Q.lowerModuleIndices[5 * t5 + j];
          }
          const uint32_t pi = v.S.mdIndices[2 * pls], po = v.S.mdIndices[2 * pls + 1];
          v.P5.hitIndices[14 * q] = v.D.anchorHitIndices[pi];
          v.P5.hitIndices[14 * q + 1] = v.D.outerHitIndices[pi];
          v.P5.hitIndices[14 * q + 2] = v.D.anchorHitIndices[po];
          v.P5.hitIndices[14 * q + 3] = v.D.outerHitIndices[po];
  #pragma unroll
          for (int j = 0; j < 10; j++) v.P5.hitIndices[14 * q + 4 + j] = v.Q.hitIndices[10 * t5 + j];
          v.P5.rzChiSquared[q] = r.rz;
          v.P5.rPhiChiSquared[q] = r.rphi;
          v.P5.rPhiChiSquaredInwards[q] = r.rphiIn;
          v.T.partOfPT5[v.Q.tripletIndices[2 * t5]] = 1;
          v.T.partOfPT5[v.Q.tripletIndices[2 * t5 + 1]] = 1;
          v.P.partOfPT5[i] = 1;
          v.Q.partOfPT5[t5] = 1;
        }
        count += n;
      }
    }
    if (!kWrite && lane == 0) v.W.pt5Count[i] = count;
  }
}

__global__ void k_pt5_totals(EventView v) {
  const uint32_t tot = v.R.scalars[kTotOccPixelQuintuplets];
  v.R.scalars[kNPixelQuintuplets] = tot < kNMaxPixelQuintuplets ? tot : kNMaxPixelQuintuplets;
}

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
__global__ __launch_bounds__(1024) void k_tc_add_pt5(EventView v) {
  __shared__ unsigned lds[17];
  const uint32_t n = v.R.scalars[kNPixelQuintuplets];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  unsigned carry = 0;
  for (uint32_t b = 0; b < n; b += 1024) {
    const uint32_t q = b + threadIdx.x;
    const bool add = q < n && v.P5.isDup[q] == 0;
    unsigned tot;
    const unsigned slot = carry + block_exclusive_scan<1024>(add ? 1u : 0u, lds, tot);
    if (add && slot < kNMaxPixelTrackCandidates) {
      const uint32_t pls = v.P5.pixelIndices[q];
      const float radius = 0.5f * (v.P5.pixelRadius[q] + v.P5.quintupletRadius[q]);
      add_tc(v.C, slot, 7, pls, v.P5.T5Indices[q], &v.P5.logicalLayers[7 * q], &v.P5.lowerModuleIndices[7 * q],
             &v.P5.hitIndices[14 * q], (int)v.P.seedIdx[pls - plsBase], v.P5.centerX[q], v.P5.centerY[q], radius, q);
    }
    carry += tot;
  }
  if (threadIdx.x == 0) {
    const unsigned added = carry < kNMaxPixelTrackCandidates ? carry : kNMaxPixelTrackCandidates;
    v.R.scalars[kNTrackCandidates] = added;
    v.R.scalars[kNTrackCandidatespT5] = added;
    v.R.scalars[kNTrackCandidatespT3] = 0;
    v.R.scalars[kNTrackCandidatespLS] = 0;
    v.R.scalars[kNTrackCandidatesT5] = 0;
  }
}

// ------------------------------------------------------------- pixel triplets
// one wave per pLS (striding over the pLS); candidates (connected module k, T3 b) in serial order
template <bool kWrite>
__global__ __launch_bounds__(kWave) void k_pixel_triplets(EventView v) {
  __shared__ uint32_t pre[kMaxConn + 1];
  const uint32_t nPLS = v.R.scalars[kEvNPLS];
  for (uint32_t i = blockIdx.x; i < nPLS; i += gridDim.x) {
    __syncthreads();  // the previous pLS's prefix in LDS is no longer read
    const unsigned lane = threadIdx.x;
    const uint32_t nL = v.M.nLowerModules;
    unsigned count = 0;
    const uint32_t size = v.W.connSize[i];
    if (v.P.isDup[i] == 0 && v.P.partOfPT5[i] == 0 && size <= (uint32_t)kMaxConn) {
      const uint32_t base = v.W.connBase[i];
      unsigned cnt = 0;
      if (lane < size) {
        const uint32_t m = v.M.connectedPixels[base + lane];
        if (m < nL && v.M.moduleType[m] != kTwoS && (v.M.layers[m] == 1 || v.M.layers[m] == 2))
          cnt = v.R.nTriplets[m];
      }
      const unsigned incl = wave_inclusive_scan(cnt);
      if (lane < size) pre[lane + 1] = incl;
      if (lane == 0) pre[0] = 0;
      const uint32_t total = __shfl(incl, kWave - 1, kWave);
      __syncthreads();
      const uint32_t plsBase = (uint32_t)v.segModuleIndices[nL];
      const uint32_t pls = plsBase + i;
      const uint32_t slotBase = kWrite ? v.W.pt3Count[i] : 0;
      uint32_t maskBase = ~0u;  // pass masks, as for the pixel quintuplets
      if (!kWrite) {
        const uint32_t words = (total + kWave - 1) / kWave;
        uint32_t b = 0;
        if (lane == 0 && words) b = atomicAdd(&v.R.scalars[kPT3MaskWords], words);
        b = __shfl(b, 0, kWave);
        if (words && b + words <= kPixMaskWords) maskBase = b;
        if (lane == 0) v.W.pt3MaskBase[i] = maskBase;
      } else {
        maskBase = v.W.pt3MaskBase[i];
      }
      const uint32_t maskWords = (total + kWave - 1) / kWave;
      uint32_t maskHeld = ~0u, parkReg = ~0u;
      unsigned long long maskReg = 0ull;
      for (uint32_t c0 = 0; c0 < total; c0 += kWave) {
        const uint32_t L = c0 + lane, w = c0 / kWave;
        bool pass = false;
        uint32_t t3 = 0, m = 0;
        PT3Result r;
        bool eval = L < total;
        if (kWrite && maskBase != ~0u) {
          const uint32_t wb = w & ~(uint32_t)(kWave - 1);
          if (wb != maskHeld) {
            maskHeld = wb;
            maskReg = wb + lane < maskWords ? v.W.pt3Mask[maskBase + wb + lane] : 0ull;
            parkReg = wb + lane < maskWords ? v.W.pt3ParkAt[maskBase + wb + lane] : ~0u;
          }
          const unsigned long long mk = __shfl(maskReg, (int)(w - wb), kWave);
          if (mk == 0ull) continue;
          const uint32_t pb = __shfl(parkReg, (int)(w - wb), kWave);
          eval = eval && ((mk >> lane) & 1ull);
          if (pb != ~0u) {
            if (eval) {
              const uint32_t* rec = &v.W.pt3Park[kPT3ParkWords * (pb + lanes_below(mk))];
              t3 = rec[0];
              m = v.T.lowerModuleIndices[3 * t3];
              r.pixR = __uint_as_float(rec[1]); r.t3R = __uint_as_float(rec[2]); r.rz = __uint_as_float(rec[3]);
              r.rphi = __uint_as_float(rec[4]); r.rphiIn = __uint_as_float(rec[5]);
              pass = true;
            }
            eval = false;
          }
        }
        if (eval) {
          uint32_t lo = 0, hi = size;
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= L) lo = mid;
            else hi = mid;
          }
          m = v.M.connectedPixels[base + lo];
          t3 = (uint32_t)v.R.tripletModuleIndices[m] + (L - pre[lo]);
          if (v.M.moduleType[v.T.lowerModuleIndices[3 * t3 + 1]] != kTwoS && v.T.partOfPT5[t3] == 0)
            pass = pt3_select(v, pls, i, t3, true, r);
        }
        unsigned n;
        const unsigned rank = count + wave_rank(pass, n);
        const unsigned long long passMask = __ballot(pass);  // every lane of the wave takes part
        if (!kWrite && maskBase != ~0u) {
          uint32_t pb = ~0u;
          if (passMask) {
            if (lane == 0) {
              pb = atomicAdd(&v.R.scalars[kPT3Parked], (uint32_t)__popcll(passMask));
              if (pb + (uint32_t)__popcll(passMask) > kPixParkMax) pb = ~0u;
            }
            pb = __shfl(pb, 0, kWave);
          }
          if (lane == 0) {
            v.W.pt3Mask[maskBase + w] = passMask;
            v.W.pt3ParkAt[maskBase + w] = pb;
          }
          if (pass && pb != ~0u) {
            uint32_t* rec = &v.W.pt3Park[kPT3ParkWords * (pb + lanes_below(passMask))];
            rec[0] = t3;
            rec[1] = __float_as_uint(r.pixR); rec[2] = __float_as_uint(r.t3R); rec[3] = __float_as_uint(r.rz);
            rec[4] = __float_as_uint(r.rphi); rec[5] = __float_as_uint(r.rphiIn);
          }
        }
        if (kWrite && pass && slotBase + rank < kNMaxPixelTriplets) {
          // addPixelTripletToMemory (SDL/PixelTriplet.h:134-204)
          const uint32_t q = slotBase + rank;
          const int adj = v.M.layers[m] == 1 ? 1 : 0;
          const uint32_t mdAdj = v.S.mdIndices[2 * v.T.segmentIndices[2 * t3] + adj];
          v.P3.pixelSegmentIndices[q] = pls;
          v.P3.tripletIndices[q] = t3;
          v.P3.pixelRadius[q] = r.pixR;
          v.P3.tripletRadius[q] = r.t3R;
          v.P3.pt[q] = v.P.ptIn[i];
          v.P3.eta[q] = v.D.anchorEta[mdAdj];
          v.P3.phi[q] = v.D.anchorPhi[mdAdj];
          v.P3.eta_pix[q] = v.P.eta[i];
          v.P3.phi_pix[q] = v.P.phi[i];
          v.P3.isDup[q] = 0;
          v.P3.score[q] = r.rphi + r.rphiIn;
          v.P3.centerX[q] = 0;
          v.P3.centerY[q] = 0;
          v.P3.logicalLayers[5 * q] = 0;
          v.P3.logicalLayers[5 * q + 1] = 0;
          v.P3.lowerModuleIndices[5 * q] = v.S.innerLowerModuleIndices[pls];
          v.P3.lowerModuleIndices[5 * q + 1] = v.S.outerLowerModuleIndices[pls];
  #pragma unroll
          for (int j = 0; j < 3; j++) {
            v.P3.logicalLayers[5 * q + 2 + j] = v.T.logicalLayers[3 * t3 + j];
            v.P3.lowerModuleIndices[5 * q + 2 + j] = v.T.lowerModuleIndices[3 * t3 + j];
          }
          const uint32_t pi = v.S.mdIndices[2 * pls], po = v.S.mdIndices[2 * pls + 1];
          v.P3.hitIndices[10 * q] = v.D.anchorHitIndices[pi];
          v.P3.hitIndices[10 * q + 1] = v.D.outerHitIndices[pi];
          v.P3.hitIndices[10 * q + 2] = v.D.anchorHitIndices[po];
          v.P3.hitIndices[10 * q + 3] = v.D.outerHitIndices[po];
  #pragma unroll
          for (int j = 0; j < 6; j++) v.P3.hitIndices[10 * q + 4 + j] = v.T.hitIndices[6 * t3 + j];
          v.P3.rzChiSquared[q] = r.rz;
          v.P3.rPhiChiSquared[q] = r.rphi;
          v.P3.rPhiChiSquaredInwards[q] = r.rphiIn;
          v.T.partOfPT3[t3] = 1;
        }
        count += n;
      }
    }
    if (!kWrite && lane == 0) v.W.pt3Count[i] = count;
  }
}

__global__ void k_pt3_totals(EventView v) {
  const uint32_t tot = v.R.scalars[kTotOccPixelTriplets];
  v.R.scalars[kNPixelTriplets] = tot < kNMaxPixelTriplets ? tot : kNMaxPixelTriplets;
}

// ---------------------------------------------------------- track candidates
// addpT3asTrackCandidatesInGPU (SDL/TrackCandidate.h:381-428)
__global__ __launch_bounds__(1024) void k_tc_add_pt3(EventView v) {
  __shared__ unsigned lds[17];
  const uint32_t n = v.R.scalars[kNPixelTriplets];
  const uint32_t base = v.R.scalars[kNTrackCandidates];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  unsigned carry = 0;
  for (uint32_t b = 0; b < n; b += 1024) {
    const uint32_t p = b + threadIdx.x;
    const bool add = p < n && v.P3.isDup[p] == 0;
    unsigned tot;
    const unsigned slot = base + carry + block_exclusive_scan<1024>(add ? 1u : 0u, lds, tot);
    if (add && slot < kNMaxPixelTrackCandidates) {
      const float radius = 0.5f * (v.P3.pixelRadius[p] + v.P3.tripletRadius[p]);
      const uint32_t pls = v.P3.pixelSegmentIndices[p];
      add_tc(v.C, slot, 5, p, p, &v.P3.logicalLayers[5 * p], &v.P3.lowerModuleIndices[5 * p],
             &v.P3.hitIndices[10 * p], (int)v.P.seedIdx[pls - plsBase], v.P3.centerX[p], v.P3.centerY[p], radius, p);
    }
    carry += tot;
  }
  if (threadIdx.x == 0) {
    const unsigned room = base < kNMaxPixelTrackCandidates ? kNMaxPixelTrackCandidates - base : 0u;
    const unsigned added = carry < room ? carry : room;
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

__global__ __launch_bounds__(256) void k_tc_t5_count(EventView v) {
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

__global__ __launch_bounds__(1024) void k_tc_t5_scan(EventView v) {
  __shared__ unsigned lds[17];
  const uint32_t n = v.R.scalars[kNEligibleT5];
  unsigned carry = 0;
  for (uint32_t b = 0; b < n; b += 1024) {
    const uint32_t i = b + threadIdx.x;
    const unsigned c = i < n ? v.W.t5TcCount[i] : 0u;
    unsigned tot;
    const unsigned ex = block_exclusive_scan<1024>(c, lds, tot);
    if (i < n) v.W.t5TcCount[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    const unsigned added = carry < kNMaxNonPixelTrackCandidates ? carry : kNMaxNonPixelTrackCandidates;
    v.R.scalars[kNTrackCandidatesT5] = added;
    v.R.scalars[kNTrackCandidates] += added;
  }
}

__global__ __launch_bounds__(256) void k_tc_t5_write(EventView v) {
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
__global__ __launch_bounds__(1024) void k_tc_add_pls(EventView v) {
  __shared__ unsigned lds[17];
  const uint32_t n = v.R.scalars[kEvNPLS];
  const uint32_t base = v.R.scalars[kNTrackCandidates];
  const uint32_t nT5 = v.R.scalars[kNTrackCandidatesT5];
  unsigned carry = 0;
  for (uint32_t b = 0; b < n; b += 1024) {
    const uint32_t i = b + threadIdx.x;
    bool add = false;
    if (i < n) add = v.tcPlsTriplets ? v.P.isDup[i] == 0 : (v.P.isQuad[i] && v.P.isDup[i] == 0);
    unsigned tot;
    const unsigned slot = base + carry + block_exclusive_scan<1024>(add ? 1u : 0u, lds, tot);
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
    carry += tot;
  }
  if (threadIdx.x == 0) {
    const unsigned limit = nT5 + kNMaxPixelTrackCandidates;
    const unsigned room = base < limit ? limit - base : 0u;
    const unsigned added = carry < room ? carry : room;
    v.R.scalars[kNTrackCandidates] = base + added;
    v.R.scalars[kNTrackCandidatespLS] = added;
  }
}

// LST::getHitIdxs (SDL/LST.cc:393-420): ntuple hit indices of each TC, the
// repeated fourth hit of a triplet seed dropped
__global__ void k_tc_output(EventView v) {
  const uint32_t nTC = v.R.scalars[kNTrackCandidates];
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nTC; t += gridDim.x * blockDim.x) {
    const int16_t type = v.C.trackCandidateType[t];
    const int maxN = type == 7 ? 14 : (type == 5 || type == 4) ? 10 : type == 8 ? 4 : 0;
    uint32_t len = 0, last = 0;
    for (int i = 0; i < maxN; i++) {
      const uint32_t h = v.C.hitIndices[14 * t + i];
      const uint32_t idx = type == 8 ? h : v.H.idxs[h];
      if (type != 4 && len == 3 && last == idx) continue;
      v.W.outHits[14 * t + len] = idx;
      last = idx;
      len++;
    }
    for (uint32_t i = len; i < 14; i++) v.W.outHits[14 * t + i] = 0;
    v.W.outLen[t] = len;
  }
}

}  // namespace

void launch_pixel_segment_cleaning(const EventView& v, hipStream_t s, const KernelTimers& t) {
  launch_pls_dup(v, s, false, t);
}

void launch_pixel_quintuplets(const EventView& v, hipStream_t s, const KernelTimers& t) {
  hipLaunchKernelGGL(k_pls_connected, dim3(grid_of(kNMaxPixelSegments, 256)), dim3(256), 0, s, v);
  t.begin(s, kTmPT5Build);
  hipLaunchKernelGGL(k_pixel_quintuplets<false>, dim3(kPixelWaves), dim3(kWave), 0, s, v);
  t.end(s, kTmPT5Build);
  hipLaunchKernelGGL(k_scan_inplace, dim3(1), dim3(1024), 0, s, v.W.pt5Count, v.R.scalars + kEvNPLS,
                     v.R.scalars + kTotOccPixelQuintuplets);
  hipLaunchKernelGGL(k_pt5_totals, dim3(1), dim3(1), 0, s, v);
  hipLaunchKernelGGL(k_pixel_quintuplets<true>, dim3(kPixelWaves), dim3(kWave), 0, s, v);
  launch_pt5_dup(v, s, t);
  hipLaunchKernelGGL(k_tc_add_pt5, dim3(1), dim3(1024), 0, s, v);
}

void launch_pixel_triplets(const EventView& v, hipStream_t s, const KernelTimers& t) {
  t.begin(s, kTmPT3Build);
  hipLaunchKernelGGL(k_pixel_triplets<false>, dim3(kPixelWaves), dim3(kWave), 0, s, v);
  t.end(s, kTmPT3Build);
  hipLaunchKernelGGL(k_scan_inplace, dim3(1), dim3(1024), 0, s, v.W.pt3Count, v.R.scalars + kEvNPLS,
                     v.R.scalars + kTotOccPixelTriplets);
  hipLaunchKernelGGL(k_pt3_totals, dim3(1), dim3(1), 0, s, v);
  hipLaunchKernelGGL(k_pixel_triplets<true>, dim3(kPixelWaves), dim3(kWave), 0, s, v);
  launch_pt3_dup(v, s, t);
}

void launch_track_candidates(const EventView& v, hipStream_t s, const KernelTimers& t) {
  launch_cross_pt3(v, s);
  hipLaunchKernelGGL(k_tc_add_pt3, dim3(1), dim3(1024), 0, s, v);
  launch_t5_dup_before_tc(v, s, t);
  launch_cross_t5(v, s);
  hipLaunchKernelGGL(k_tc_t5_count, dim3(kModuleWaveBlocks), dim3(256), 0, s, v);
  hipLaunchKernelGGL(k_tc_t5_scan, dim3(1), dim3(1024), 0, s, v);
  hipLaunchKernelGGL(k_tc_t5_write, dim3(kModuleWaveBlocks), dim3(256), 0, s, v);
  launch_pls_dup(v, s, true, t);
  launch_cross_pls(v, s, t);
  hipLaunchKernelGGL(k_tc_add_pls, dim3(1), dim3(1024), 0, s, v);
  hipLaunchKernelGGL(k_tc_output, dim3(grid_of(kTCCap, 256)), dim3(256), 0, s, v);
}

}  // namespace lst
