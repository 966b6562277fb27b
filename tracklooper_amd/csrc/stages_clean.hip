// Duplicate removal and cross cleaning as all-pairs kernels.
//
// Every rule of this kind in the reference sets a flag on an object x when
// some partner y satisfies a predicate on the pair (pLS cleaning,
// SDL/Kernels.h:384-473; pT5 and pT3 duplicates, :322-381; T5 duplicates
// before the track candidates, :250-320; crossCleanpT3/T5/pLS,
// SDL/TrackCandidate.h:204-379).  The predicates read only pair values and
// flags written by earlier kernels, so the flag is an OR over partners and
// the pairs can be visited in any order.  Where every x meets many y's the
// kernel tiles the (x, y) plane: a 256-lane work-group owns 256 x's, stages
// 256 y's at a time in LDS and strides over the y tiles; where an x meets a
// short range of its own (the T5 rules: its module, or its eta band) one wave
// per x runs its lanes over that range.  Flags are set with stores of a
// constant or 32-bit atomic ORs, never read-modify-write races.
#include "algo_pixel.h"
#include "stages.h"
#include "wave.h"

namespace lst {
namespace {

constexpr int kTile = 256;
constexpr int kWave = 64;

// OR `bits` into byte k of an int8/uint8 flag array (256-byte aligned base)
__device__ __forceinline__ void or_byte(void* base, uint32_t k, uint32_t bits) {
  unsigned int* word = reinterpret_cast<unsigned int*>(static_cast<char*>(base) + (k & ~3u));
  atomicOr(word, bits << (8 * (k & 3u)));
}

// The y range [0, ny) is cut into gridDim.y equal shares, one per blockIdx.y,
// so that small object counts still spread over many work-groups.
struct YShare {
  uint32_t b, e;
};
__device__ __forceinline__ YShare y_share(uint32_t ny) {
  const uint32_t chunk = (ny + gridDim.y - 1) / gridDim.y;
  const uint32_t b = blockIdx.y * chunk;
  const uint32_t e = b + chunk < ny ? b + chunk : ny;
  return {b < e ? b : e, e};
}
// Work items (x tile, y share) for a fixed 1-D grid whose blocks stride over
// the items: the share count comes from the device-side x count, so that
// about kPairItems items exist whatever the event's size.
constexpr uint32_t kPairItems = 2048, kMaxShares = 256;
struct PairItems {
  uint32_t shares, items;
};
__device__ __forceinline__ PairItems pair_items(uint32_t nx) {
  const uint32_t tiles = (nx + kTile - 1) / kTile;
  uint32_t sh = tiles ? kPairItems / tiles : 1;
  sh = sh < 1 ? 1 : (sh > kMaxShares ? kMaxShares : sh);
  return {sh, tiles * sh};
}
__device__ __forceinline__ YShare share_of(uint32_t ny, uint32_t sh, uint32_t shares) {
  const uint32_t chunk = (ny + shares - 1) / shares;
  const uint32_t b = sh * chunk;
  const uint32_t e = b + chunk < ny ? b + chunk : ny;
  return {b < e ? b : e, e};
}
__device__ __forceinline__ uint32_t tile_len(const YShare& r, uint32_t y0) {
  return r.e - y0 < (uint32_t)kTile ? r.e - y0 : (uint32_t)kTile;
}

// x tiles cover nxCap (blocks past the device-side count exit at once); the y
// shares are sized so that about 2048 work-groups hold work at nx = nxTyp
inline dim3 pair_grid(uint32_t nxCap, uint32_t nxTyp, uint32_t nyCap) {
  unsigned gx = (nxCap + kTile - 1) / kTile;
  gx = gx == 0 ? 1 : gx;
  unsigned ax = (nxTyp + kTile - 1) / kTile;
  ax = ax == 0 ? 1 : ax;
  unsigned gy = (2048 + ax - 1) / ax;
  gy = gy > 256 ? 256 : gy;
  gy = gy > nyCap ? (nyCap ? nyCap : 1) : gy;
  return dim3(gx, gy);
}

// ------------------------------------------------------------- eta bands
// The T5 and pLS rules flag a pair only when |d eta| <= 0.1 or when d eta is
// NaN (the test `> 0.1` is then false).  Objects are binned by eta (bins of
// width 1/8 > 0.1 over [-4, 4), the ends clamped, NaN in a bin of its own) and
// listed bin by bin; an x tile whose bins span [bmin, bmax] then only meets
// the y's of bins bmin-1 .. bmax+1 plus the NaN bin (every y when the tile
// holds a NaN x).  That is a superset of the pairs that can flag, so the flags
// are unchanged.
__device__ __forceinline__ int eta_bin(float eta) {
  if (eta != eta) return kNanEtaBin;
  if (!(eta > -4.f)) return 0;
  if (!(eta < 4.f)) return kEtaBins - 1;
  const int b = (int)((eta + 4.f) * 8.f);
  return b < kEtaBins ? b : kEtaBins - 1;
}

// One 1024-lane work-group: histogram, offsets, scatter (order inside a bin is
// free).  Item i is load(key(i)) (a struct with a member `eta`); store(position,
// item) writes it at its binned position.  Up to 1024 * kBinPer items are
// loaded once and kept in registers for the scatter: each lane issues all its
// key loads, then all its item loads (indices clamped to n - 1, so that no
// branch splits them); a larger count (or twoPass) loads every item twice.
struct SameKey {
  __device__ uint32_t operator()(uint32_t i) const { return i; }
};
template <int kBinPer = 16, class Load, class Store, class Key = SameKey>
__device__ void bin_by_eta(uint32_t n, bool twoPass, Load load, Store store, uint32_t* off, Key key = Key()) {
  __shared__ unsigned cnt[kEtaBinOffsets], cur[kEtaBinOffsets];
  using Item = decltype(load(0u));
  if (threadIdx.x < kEtaBinOffsets) cnt[threadIdx.x] = 0;
  __syncthreads();
  const bool fits = n <= 1024u * kBinPer && !twoPass;
  Item it[kBinPer];
  int bin[kBinPer];
  if (fits && n != 0) {
    uint32_t k[kBinPer];
#pragma unroll
    for (int j = 0; j < kBinPer; j++) {
      const uint32_t i = j * 1024u + threadIdx.x;
      k[j] = key(i < n ? i : n - 1);
    }
#pragma unroll
    for (int j = 0; j < kBinPer; j++) it[j] = load(k[j]);
#pragma unroll
    for (int j = 0; j < kBinPer; j++) {
      bin[j] = j * 1024u + threadIdx.x < n ? eta_bin(it[j].eta) : -1;
      if (bin[j] >= 0) atomicAdd(&cnt[bin[j]], 1u);
    }
  } else if (!fits) {
    for (uint32_t i = threadIdx.x; i < n; i += 1024u) atomicAdd(&cnt[eta_bin(load(key(i)).eta)], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 64) {  // exclusive scan of the kNanEtaBin + 1 counts, two bins per lane
    const uint32_t b0 = 2 * threadIdx.x, b1 = b0 + 1;
    const unsigned c0 = b0 <= (uint32_t)kNanEtaBin ? cnt[b0] : 0u, c1 = b1 <= (uint32_t)kNanEtaBin ? cnt[b1] : 0u;
    const unsigned incl = wave_inclusive_scan(c0 + c1), ex = incl - c0 - c1;
    if (b0 <= (uint32_t)kNanEtaBin) off[b0] = cur[b0] = ex;
    if (b1 <= (uint32_t)kNanEtaBin) off[b1] = cur[b1] = ex + c0;
    if (threadIdx.x == 63) off[kNanEtaBin + 1] = incl;
  }
  __syncthreads();
  if (fits && n != 0) {
#pragma unroll
    for (int j = 0; j < kBinPer; j++)
      if (bin[j] >= 0) store(atomicAdd(&cur[bin[j]], 1u), it[j]);
  } else if (!fits) {
    for (uint32_t i = threadIdx.x; i < n; i += 1024u) {
      const Item x = load(key(i));
      store(atomicAdd(&cur[eta_bin(x.eta)], 1u), x);
    }
  }
}
struct BinItem {
  uint32_t id;
  float eta;
};

struct PlsBinItem {
  float4 rec;  // (eta, phi, score, index | isQuad << 31)
  uint4 hits;
  float eta;
};
__global__ __launch_bounds__(1024) void k_bin_pls(LST_VIEWS) {
  LST_BIND;
  bin_by_eta<8>(
      v.R.scalars[kEvNPLS], v.largePaths != 0,
      [&](uint32_t i) {
        const float eta = v.P.eta[i];
        const uint32_t quad = v.P.isQuad[i] ? 1u : 0u;
        const uint4 h = reinterpret_cast<const uint4*>(v.P.pLSHitsIdxs)[i];
        return PlsBinItem{make_float4(eta, v.P.phi[i], v.P.score[i], __uint_as_float(i | quad << 31)), h, eta};
      },
      [&](uint32_t pos, const PlsBinItem& t) {
        v.W.plsBinned[pos] = __float_as_uint(t.rec.w) & 0x7fffffffu;
        v.W.plsBinRec[pos] = t.rec;
        v.W.plsBinHits[pos] = t.hits;
      },
      v.W.plsBinOff);
}

// y positions a tile meets: [b1, b1 + n1) then [b2, b2 + n2) of the binned list
struct Band {
  uint32_t b1, n1, b2, n2;
};
__device__ __forceinline__ Band tile_band(const uint32_t* off, int bmin, int bmax) {
  const uint32_t end = off[kNanEtaBin + 1];
  if (bmax == kNanEtaBin) return {0, end, 0, 0};
  const uint32_t b1 = off[bmin > 0 ? bmin - 1 : 0];
  const uint32_t e1 = off[bmax + 2 < kNanEtaBin ? bmax + 2 : kNanEtaBin];
  return {b1, e1 - b1, off[kNanEtaBin], end - off[kNanEtaBin]};
}
__device__ __forceinline__ uint32_t band_at(const Band& B, uint32_t k) {
  return k < B.n1 ? B.b1 + k : B.b2 + (k - B.n1);
}

// ------------------------------------------------------------ pLS cleaning
// checkHitspLS: for a pair (lo < hi) within 0.1 in eta the removal target is
// the triplet seed over the quadruplet, else the higher score (ties: lo); the
// first pass flags it (bit 1) when >= 3 of lo's hits are in hi, the second
// pass (quadruplets without bit 1 only) flags it (bit 2) on any shared hit or
// dR2 < 1e-5.  x is the candidate target k, y the partner j.  One wave per
// binned pLS k; its lanes run over the binned records (direction, score,
// quadruplet flag and hits, kPlsUnroll per lane per pass) of k's eta bin, the
// bins next to it and the NaN bin.
constexpr uint32_t kPlsUnroll = 4, kQuadBit = 0x80000000u;
template <bool kSecond>
__global__ __launch_bounds__(256) void k_pls_pairs(LST_VIEWS) {
  LST_BIND;
  const uint32_t n = v.R.scalars[kEvNPLS];
  const unsigned lane = threadIdx.x % kWave;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t first = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  for (uint32_t xi = first; xi < n; xi += waves) {
    const float4 ax = v.W.plsBinRec[xi];
    const uint4 hx = v.W.plsBinHits[xi];
    const uint32_t k = __float_as_uint(ax.w) & ~kQuadBit;
    const int quadK = (__float_as_uint(ax.w) & kQuadBit) ? 1 : 0;
    if (kSecond && (!quadK || (v.P.isDup[k] & 1))) continue;
    const float etaK = ax.x, phiK = ax.y, scoreK = ax.z;
    const int bx = eta_bin(etaK);
    const Band B = tile_band(v.W.plsBinOff, bx, bx);
    const uint32_t ny = B.n1 + B.n2;
    const uint32_t hk[4] = {hx.x, hx.y, hx.z, hx.w};
    bool flag = false;
    for (uint32_t k0 = 0; k0 < ny && !flag; k0 += kWave * kPlsUnroll) {
      float4 ay[kPlsUnroll];
      uint4 hy[kPlsUnroll];
#pragma unroll
      for (uint32_t u = 0; u < kPlsUnroll; u++) {
        const uint32_t kk = k0 + u * kWave + lane;
        const uint32_t at = band_at(B, kk < ny ? kk : ny - 1);
        ay[u] = v.W.plsBinRec[at];
        hy[u] = v.W.plsBinHits[at];
      }
      bool f = false;
#pragma unroll
      for (uint32_t u = 0; u < kPlsUnroll; u++) {
        const uint32_t j = __float_as_uint(ay[u].w) & ~kQuadBit;
        const int quadJ = (__float_as_uint(ay[u].w) & kQuadBit) ? 1 : 0;
        const bool kLo = k < j;
        const float etaLo = kLo ? etaK : ay[u].x, etaHi = kLo ? ay[u].x : etaK;
        const int quadDiff = kLo ? quadK - quadJ : quadJ - quadK;
        const float scoreDiff = kLo ? scoreK - ay[u].z : ay[u].z - scoreK;
        // removal target: hi when quadDiff > 0 or (== 0 and scoreDiff < 0), else lo
        const bool rmHi = quadDiff > 0 || (quadDiff == 0 && scoreDiff < 0);
        const bool cand = k0 + u * kWave + lane < ny && j != k && !(m_abs(etaHi - etaLo) > 0.1f) &&
                          (!kSecond || quadJ) && rmHi != kLo;
        if (!cand) continue;
        const uint32_t hj[4] = {hy[u].x, hy[u].y, hy[u].z, hy[u].w};
        int matched = 0;
#pragma unroll
        for (int a = 0; a < 4; a++) {
          bool hit = false;
#pragma unroll
          for (int b = 0; b < 4; b++) hit = hit || (kLo ? hk[a] == hj[b] : hj[a] == hk[b]);
          matched += hit ? 1 : 0;
        }
        if (!kSecond) {
          f = f || matched >= 3;
        } else if (!(v.P.isDup[j] & 1)) {
          const float dEta = m_abs(etaLo - etaHi);
          const float dPhi = kLo ? wrap_dphi(phiK, ay[u].y) : wrap_dphi(ay[u].y, phiK);
          f = f || matched >= 1 || dEta * dEta + dPhi * dPhi < 1e-5f;
        }
      }
      flag = __ballot(f) != 0ull;
    }
    if (flag && lane == 0) or_byte(v.P.isDup, k, kSecond ? 2u : 1u);
  }
}

// ------------------------------------------------------------ pT5 duplicates
// removeDupPixelQuintupletsInGPUFromMap: ix is dropped when a partner shares
// >= 7 of its 14 hits and scores lower (ties: the higher index is dropped).
// At least 7 of ix's 14 positions must hold a hit of the partner, so one of its
// first 8 positions does: every partner that can flag ix shares a hit with
// those 8 positions.  The (pT5, position) entries are sorted by a hash of the
// hit into contiguous buckets (count, scan, scatter); one wave per pT5 then
// checks the entries of its 8 buckets in parallel with the exact rule.
constexpr uint32_t kPT5HashMask = kPT5HashBuckets - 1;
__device__ __forceinline__ uint32_t hit_bucket(uint32_t h) { return (h * 0x9E3779B1u) >> 17 & kPT5HashMask; }

__global__ __launch_bounds__(256) void k_pt5_bucket_count(LST_VIEWS) {
  LST_BIND;
  const uint32_t n = v.R.scalars[kNPixelQuintuplets];
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < 14 * n; e += gridDim.x * blockDim.x)
    atomicAdd(&v.W.pt5BucketOff[hit_bucket(v.P5.hitIndices[e])], 1u);
}

// inclusive scan of the bucket counts in place (one 1024-lane work-group,
// kPT5HashBuckets / 1024 consecutive buckets per lane); entry kPT5HashBuckets
// gets the total
__global__ __launch_bounds__(1024) void k_pt5_bucket_scan(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[17];
  constexpr uint32_t kPer = kPT5HashBuckets / 1024;
  static_assert(kPer % 4 == 0, "buckets per lane");
  uint4* c = reinterpret_cast<uint4*>(v.W.pt5BucketOff + kPer * threadIdx.x);
  uint32_t x[kPer];
#pragma unroll
  for (uint32_t k = 0; k < kPer / 4; k++) {
    const uint4 q = c[k];
    x[4 * k] = q.x; x[4 * k + 1] = q.y; x[4 * k + 2] = q.z; x[4 * k + 3] = q.w;
  }
  unsigned sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) sum += x[k];
  unsigned tot;
  unsigned run = block_exclusive_scan<1024>(sum, lds, tot);
#pragma unroll
  for (uint32_t k = 0; k < kPer; k++) {
    run += x[k];
    x[k] = run;
  }
#pragma unroll
  for (uint32_t k = 0; k < kPer / 4; k++) c[k] = make_uint4(x[4 * k], x[4 * k + 1], x[4 * k + 2], x[4 * k + 3]);
  if (threadIdx.x == 0) v.W.pt5BucketOff[kPT5HashBuckets] = tot;
}

// entry e to its bucket; afterwards pt5BucketOff[b] is the first position of
// bucket b (the inclusive ends are counted down)
__global__ __launch_bounds__(256) void k_pt5_bucket_fill(LST_VIEWS) {
  LST_BIND;
  const uint32_t n = v.R.scalars[kNPixelQuintuplets];
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < 14 * n; e += gridDim.x * blockDim.x)
    v.W.pt5BucketEnt[atomicSub(&v.W.pt5BucketOff[hit_bucket(v.P5.hitIndices[e])], 1u) - 1] = e;
}

__global__ __launch_bounds__(256) void k_pt5_pairs(LST_VIEWS) {
  LST_BIND;
  const uint32_t n = v.R.scalars[kNPixelQuintuplets];
  const unsigned lane = threadIdx.x % kWave;
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t ix = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; ix < n; ix += waves) {
    const uint32_t mine = lane < 14 ? v.P5.hitIndices[14 * ix + lane] : 0u;
    uint32_t h[14];
#pragma unroll
    for (int a = 0; a < 14; a++) h[a] = __shfl(mine, a, kWave);
    const float s1 = v.P5.score[ix];
    // lanes 0-7: the bucket of hit position `lane`; candidates numbered over the 8 buckets
    uint32_t bStart = 0, bSize = 0;
    if (lane < 8) {
      const uint32_t b = hit_bucket(mine);
      bStart = v.W.pt5BucketOff[b];
      bSize = v.W.pt5BucketOff[b + 1] - bStart;
    }
    const uint32_t incl = wave_inclusive_scan(bSize);
    uint32_t ends[8];
#pragma unroll
    for (int j = 0; j < 8; j++) ends[j] = __shfl(incl, j, kWave);
    const uint32_t total = ends[7];
    bool dup = false;
    for (uint32_t c0 = 0; c0 < total && !dup; c0 += kWave) {
      const uint32_t c = c0 + lane;
      int k = 0;           // bucket of candidate c: the first with ends[k] > c
      uint32_t before = 0;  // candidates in the buckets before it
#pragma unroll
      for (int j = 0; j < 7; j++) {
        k += ends[j] <= c ? 1 : 0;
        before = ends[j] <= c ? ends[j] : before;
      }
      // every lane takes part in the shuffles (k < 8 for every lane)
      const uint32_t start = __shfl(bStart, k, kWave);
      const uint32_t key = __shfl(mine, k, kWave);
      bool d = false;
      if (c < total) {
        const uint32_t e = v.W.pt5BucketEnt[start + (c - before)];
        const uint32_t jx = e / 14;
        if (v.P5.hitIndices[e] == key && jx != ix) {
          const float s2 = v.P5.score[jx];
          if (s1 > s2 || (s1 == s2 && ix > jx)) {
            uint32_t g[14];
#pragma unroll
            for (int b = 0; b < 14; b++) g[b] = v.P5.hitIndices[14 * jx + b];
            int matched = 0;
#pragma unroll
            for (int a = 0; a < 14; a++) {
              bool hit = false;
#pragma unroll
              for (int b = 0; b < 14; b++) hit = hit || h[a] == g[b];
              matched += hit ? 1 : 0;
            }
            d = matched >= 7;
          }
        }
      }
      dup = __ballot(d) != 0ull;
    }
    if (dup && lane == 0) v.P5.isDup[ix] = 1;
  }
}

// ------------------------------------------------------------ pT3 duplicates
// removeDupPixelTripletsInGPUFromMap with checkHitspT3: ix is dropped when a
// partner shares >= 5 hits (pixel + outer) and reaches a lower third layer, or
// the same layer with a lower score (ties: the lower index is dropped)
__global__ __launch_bounds__(kTile) void k_pt3_pairs(LST_VIEWS) {
  LST_BIND;
  __shared__ float sScore[kTile];
  __shared__ uint8_t sLayer[kTile];
  __shared__ __align__(16) uint32_t sHits[10 * kTile];
  const uint32_t n = v.R.scalars[kNPixelTriplets];
  const YShare ys = y_share(n);
  for (uint32_t x0 = blockIdx.x * kTile; x0 < n; x0 += gridDim.x * kTile) {
    const uint32_t ix = x0 + threadIdx.x;
    const bool active = ix < n;
    uint32_t h[10];
    uint8_t li = 0;
    float si = 0;
    if (active) {
      const uint2* ph = reinterpret_cast<const uint2*>(&v.P3.hitIndices[10 * ix]);  // 10 words, 8-byte aligned
#pragma unroll
      for (int a = 0; a < 5; a++) {
        const uint2 u = ld_global(ph + a);
        h[2 * a] = u.x;
        h[2 * a + 1] = u.y;
      }
      li = v.P3.logicalLayers[5 * ix + 2];
      si = v.P3.score[ix];
    }
    bool dup = false;
    for (uint32_t y0 = ys.b; y0 < ys.e; y0 += kTile) {
      __syncthreads();
      const uint32_t jl = y0 + threadIdx.x;
      if (jl < ys.e) {
        sScore[threadIdx.x] = v.P3.score[jl];
        sLayer[threadIdx.x] = v.P3.logicalLayers[5 * jl + 2];
        const uint2* ph = reinterpret_cast<const uint2*>(&v.P3.hitIndices[10 * jl]);
        uint2* sh = reinterpret_cast<uint2*>(&sHits[threadIdx.x * 10]);
#pragma unroll
        for (int a = 0; a < 5; a++) sh[a] = ld_global(ph + a);
      }
      __syncthreads();
      const uint32_t m = tile_len(ys, y0);
      if (!active || dup) continue;
      for (uint32_t t = 0; t < m && !dup; t++) {
        const uint32_t jx = y0 + t;
        if (jx == ix) continue;
        const uint8_t lj = sLayer[t];
        const float sj = sScore[t];
        if (!(lj < li || (li == lj && si > sj) || (li == lj && si == sj && ix < jx))) continue;
        int matched = 0;
        for (int a = 0; a < 4; a++)
          for (int b = 0; b < 4; b++)
            if (h[a] == sHits[t * 10 + b]) {
              matched++;
              break;
            }
        for (int a = 4; a < 10; a++)
          for (int b = 4; b < 10; b++)
            if (h[a] == sHits[t * 10 + b]) {
              matched++;
              break;
            }
        dup = matched >= 5;
      }
    }
    if (dup) v.P3.isDup[ix] = 1;
  }
}

// ---------------------------------------------------------- crossCleanpT3
// a pT3 whose pLS direction is within dR2 < 1e-5 of a pT5's pLS is dropped
__global__ __launch_bounds__(kTile) void k_cross_pt3_pairs(LST_VIEWS) {
  LST_BIND;
  __shared__ float sEta[kTile], sPhi[kTile];
  const uint32_t n = v.R.scalars[kNPixelTriplets], n5 = v.R.scalars[kNPixelQuintuplets];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  const YShare ys = y_share(n5);
  for (uint32_t x0 = blockIdx.x * kTile; x0 < n; x0 += gridDim.x * kTile) {
    const uint32_t p = x0 + threadIdx.x;
    const bool active = p < n && v.P3.isDup[p] == 0;
    const float eta1 = active ? v.P3.eta_pix[p] : 0.f, phi1 = active ? v.P3.phi_pix[p] : 0.f;
    bool dup = false;
    for (uint32_t y0 = ys.b; y0 < ys.e; y0 += kTile) {
      __syncthreads();
      const uint32_t ql = y0 + threadIdx.x;
      if (ql < ys.e) {
        const uint32_t j = v.P5.pixelIndices[ql] - plsBase;
        sEta[threadIdx.x] = v.P.eta[j];
        sPhi[threadIdx.x] = v.P.phi[j];
      }
      __syncthreads();
      const uint32_t m = tile_len(ys, y0);
      if (!active || dup) continue;
      for (uint32_t t = 0; t < m && !dup; t++) {
        const float dEta = m_abs(eta1 - sEta[t]);
        const float dPhi = wrap_dphi(phi1, sPhi[t]);
        dup = dEta * dEta + dPhi * dPhi < 1e-5f;
      }
    }
    if (dup) v.P3.isDup[p] = 1;
  }
}

// --------------------------------------------------- T5 list in slot order
__global__ __launch_bounds__(1024) void k_t5_list_offsets(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[17];
  const uint32_t n = v.R.scalars[kNEligibleT5];
  const unsigned total = scan_wide<kScanPer>(
      n, [&](uint32_t i) { return v.R.nQuintuplets[v.R.indicesOfEligibleT5Modules[i]]; },
      [&](uint32_t i, unsigned p, unsigned) { v.W.t5ListOff[i] = p; }, lds);
  if (threadIdx.x == 0) {
    v.W.t5ListOff[n] = total;
    v.R.scalars[kT5Listed] = total;
  }
}

__global__ __launch_bounds__(256) void k_t5_list_fill(LST_VIEWS) {
  LST_BIND;
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, nWaves = (gridDim.x * blockDim.x) / 64;
  for (uint32_t e = wave; e < nElig; e += nWaves) {  // one wave per eligible module
    const uint16_t m = v.R.indicesOfEligibleT5Modules[e];
    const uint32_t nq = v.R.nQuintuplets[m];
    const uint32_t base = (uint32_t)v.R.quintupletModuleIndices[m];
    const uint32_t off = v.W.t5ListOff[e];
    for (uint32_t k = threadIdx.x & 63u; k < nq; k += 64) {
      const uint32_t q = base + k;
      v.W.t5List[off + k] = q;
      v.W.t5ListMod[off + k] = e;
      v.W.t5ListRec[off + k] = make_float4(v.Q.eta[q], v.Q.phi[q], v.Q.score_rphisum[q], __uint_as_float(q));
    }
  }
}

// ------------------------------------------------ T5 duplicates after build
// removeDupQuintupletsInGPUAfterBuild (SDL/Kernels.h:202-249): within one
// module, a close pair (i < j, |d eta| and |d phi| <= 0.1) sharing >= 7 hits
// flags i when score_i >= score_j and j otherwise.  Each 16-lane group of a
// wave takes one listed T5 q (four per wave: a module holds a few T5s, so a
// whole wave per T5 left most lanes idle) and runs its lanes over the T5s of
// q's module, 16 at a time (contiguous slots, so the loads coalesce); it
// flags q when any partner points the rule at it, with a store of 1 that does
// not depend on evaluation order.  The list records give eta, phi, score and
// slot in one 16-byte load.
__global__ __launch_bounds__(256) void k_t5_dup_after_build(LST_VIEWS) {
  LST_BIND;
  constexpr unsigned kGroup = 16;
  const uint32_t n = v.R.scalars[kT5Listed];
  const unsigned lane = threadIdx.x % kWave, g = lane / kGroup, gl = lane % kGroup;
  const uint64_t gmask = 0xFFFFull << (kGroup * g);
  const uint32_t groups = gridDim.x * (blockDim.x / kGroup);
  for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; i - g < n; i += groups) {
    // the four groups of a wave take i - g .. i - g + 3; the loop bound is the wave's first one, so
    // every lane of the wave runs the same iterations
    const bool has = i < n;
    uint32_t q = 0, yb = 0, len = 0;
    float ei = 0.f, pi = 0.f, si = 0.f;
    if (has) {
      const uint32_t mod = v.W.t5ListMod[i];
      yb = v.W.t5ListOff[mod];
      len = v.W.t5ListOff[mod + 1] - yb;
      const float4 rx = ld_global(v.W.t5ListRec + i);  // eta, phi, score, slot
      q = __float_as_uint(rx.w);
      ei = rx.x; pi = rx.y; si = rx.z;
    }
    bool dup = false;
    for (uint32_t off = 0; __ballot(has && !dup && off < len) != 0ull; off += kGroup) {
      bool d = false;
      if (has && !dup && off + gl < len) {
        const float4 ry = ld_global(v.W.t5ListRec + yb + off + gl);
        const uint32_t y = __float_as_uint(ry.w);  // the module's T5s are consecutive slots in list order
        const bool xLo = q < y;
        const float sj = ry.z, ey = ry.x, py = ry.y;
        const float dEta = m_abs(xLo ? ei - ey : ey - ei);
        const float dPhi = xLo ? wrap_dphi(pi, py) : wrap_dphi(py, pi);
        // the pair can only flag q if the score order points at it
        if (y != q && (xLo ? (si >= sj) : !(sj >= si)) && !(dEta > 0.1f) && !(m_abs(dPhi) > 0.1f)) {
          uint32_t hx[10], hy[10];
#pragma unroll
          for (int a = 0; a < 10; a++) {
            hx[a] = v.Q.hitIndices[10 * q + a];
            hy[a] = v.Q.hitIndices[10 * y + a];
          }
          int matched = 0;
#pragma unroll
          for (int a = 0; a < 10; a++) {
            bool hit = false;
#pragma unroll
            for (int b = 0; b < 10; b++) hit = hit || (xLo ? hx[a] == hy[b] : hy[a] == hx[b]);
            matched += hit ? 1 : 0;
          }
          d = matched >= 7;
        }
      }
      dup = dup || (__ballot(d) & gmask) != 0ull;
    }
    if (has && dup && gl == 0) v.Q.isDup[q] = 1;
  }
}

// ---------------------------------------------------- T5 duplicates before TC
// Among T5s not in a pT5 and not flagged after build, a close pair
// (|dEta|, |dPhi| <= 0.1, and dR2 < 0.001 or >= 5 shared hits) flags the
// higher score (ties: the lower index) with bit 2.
struct T5BinItem {
  float4 rec;  // (eta, phi, score, slot bits)
  float eta;
};
__global__ __launch_bounds__(1024) void k_bin_t5(LST_VIEWS) {
  LST_BIND;
  bin_by_eta(
      v.R.scalars[kT5Listed], v.largePaths != 0,
      [&](uint32_t i) {
        const float4 r = v.W.t5ListRec[i];
        return T5BinItem{r, r.x};
      },
      [&](uint32_t pos, const T5BinItem& t) {
        v.W.t5Binned[pos] = __float_as_uint(t.rec.w);
        v.W.t5BinRec[pos] = t.rec;
      },
      v.W.t5BinOff);
}

// One wave per binned T5 x (not in a pT5, not a duplicate after build); its
// lanes run over the binned records of x's eta bin, the bins next to it and
// the NaN bin (every record when x's eta is NaN), kT5TcUnroll records per lane
// per pass so that their loads overlap.  The partners that pass the direction
// and score cuts are queued in LDS (with a bit for dR2 < 0.001) and checked
// 64 at a time: their pT5 and duplicate flags, then the 100-compare hit count
// unless the direction alone decides, so that this work runs with every lane
// busy instead of once per pass for a few lanes.
constexpr uint32_t kT5TcUnroll = 4, kT5TcQueue = 512, kNearBit = 0x80000000u;
__global__ __launch_bounds__(256) void k_t5_tc_pairs(LST_VIEWS) {
  LST_BIND;
  __shared__ uint32_t sQueue[256 / kWave][kT5TcQueue];
  const uint32_t n = v.R.scalars[kT5Listed];
  const unsigned lane = threadIdx.x % kWave;
  uint32_t* queue = sQueue[threadIdx.x / kWave];
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t first = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  for (uint32_t xi = first; xi < n; xi += waves) {
    const float4 rx = v.W.t5BinRec[xi];
    const uint32_t x = __float_as_uint(rx.w);
    if (((unsigned)v.Q.partOfPT5[x] | ((unsigned)v.Q.isDup[x] & 1u)) != 0u) continue;
    const float etaX = rx.x, phiX = rx.y, sX = rx.z;
    const int bx = eta_bin(etaX);
    const Band B = tile_band(v.W.t5BinOff, bx, bx);
    const uint32_t ny = B.n1 + B.n2;
    bool dup = false;
    uint32_t nq = 0;  // queued partners (the same in every lane)
    for (uint32_t k0 = 0; k0 < ny && !dup; k0 += kWave * kT5TcUnroll) {
      float4 ry[kT5TcUnroll];
#pragma unroll
      for (uint32_t u = 0; u < kT5TcUnroll; u++) {
        const uint32_t k = k0 + u * kWave + lane;
        ry[u] = v.W.t5BinRec[band_at(B, k < ny ? k : ny - 1)];
      }
#pragma unroll
      for (uint32_t u = 0; u < kT5TcUnroll; u++) {
        const uint32_t y = __float_as_uint(ry[u].w);
        const float sY = ry[u].z;
        const bool xLo = x < y;
        const float dEta = m_abs(xLo ? etaX - ry[u].x : ry[u].x - etaX);
        const float dPhi = xLo ? wrap_dphi(phiX, ry[u].y) : wrap_dphi(ry[u].y, phiX);
        const bool close = k0 + u * kWave + lane < ny && y != x && (sX > sY || (!(sX < sY) && x < y)) &&
                           !(dEta > 0.1f) && !(m_abs(dPhi) > 0.1f);
        const bool near = dEta * dEta + dPhi * dPhi < 0.001f;
        const unsigned long long m = __ballot(close);
        if (close) queue[nq + lanes_below(m)] = y | (near ? kNearBit : 0u);
        nq += (uint32_t)__popcll(m);
      }
      const bool last = k0 + kWave * kT5TcUnroll >= ny;
      if (nq != 0 && (last || nq > kT5TcQueue - kWave * kT5TcUnroll)) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        bool d = false;
        for (uint32_t c0 = 0; c0 < nq; c0 += kWave) {
          if (c0 + lane >= nq) continue;
          const uint32_t e = queue[c0 + lane], y = e & ~kNearBit;
          if (((unsigned)v.Q.partOfPT5[y] | ((unsigned)v.Q.isDup[y] & 1u)) != 0u) continue;
          if (e & kNearBit) {
            d = true;
            continue;
          }
          const bool xLo = x < y;
          uint32_t hx[10], hy[10];
          const uint2* px = reinterpret_cast<const uint2*>(&v.Q.hitIndices[10 * x]);  // 10 words, 8-byte aligned
          const uint2* py = reinterpret_cast<const uint2*>(&v.Q.hitIndices[10 * y]);
#pragma unroll
          for (int a = 0; a < 5; a++) {
            const uint2 u = ld_global(px + a), w = ld_global(py + a);
            hx[2 * a] = u.x;
            hx[2 * a + 1] = u.y;
            hy[2 * a] = w.x;
            hy[2 * a + 1] = w.y;
          }
          int matched = 0;
#pragma unroll
          for (int a = 0; a < 10; a++) {
            bool hit = false;
#pragma unroll
            for (int b = 0; b < 10; b++) hit = hit || (xLo ? hx[a] == hy[b] : hy[a] == hx[b]);
            matched += hit ? 1 : 0;
          }
          d = d || matched >= 5;
        }
        dup = __ballot(d) != 0ull;
        nq = 0;
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (dup && lane == 0) or_byte(v.Q.isDup, x, 2u);
  }
}

// --------------------------------------------------------------- crossCleanT5
// a T5 (not flagged, not in a pT5) within dR2 < 1e-3 of any pT5 or pT3 is
// dropped.  dR2 < 1e-3 needs |d eta| < 0.032, and a NaN direction never flags,
// so the pT5/pT3 are binned by eta and a T5 tile (the T5s in eta-bin order)
// meets only the bins next to its own.
struct PixBinItem {
  float2 rec;  // (eta, phi)
  float eta;
};
__global__ __launch_bounds__(1024) void k_bin_pix(LST_VIEWS) {
  LST_BIND;
  const uint32_t n5 = v.R.scalars[kNPixelQuintuplets];
  bin_by_eta(
      n5 + v.R.scalars[kNPixelTriplets], v.largePaths != 0,
      [&](uint32_t j) {
        const float eta = j < n5 ? v.P5.eta[j] : v.P3.eta[j - n5], phi = j < n5 ? v.P5.phi[j] : v.P3.phi[j - n5];
        return PixBinItem{make_float2(eta, phi), eta};
      },
      [&](uint32_t pos, const PixBinItem& t) { v.W.pixBinRec[pos] = t.rec; }, v.W.pixBinOff);
}

__global__ __launch_bounds__(kTile) void k_cross_t5_pairs(LST_VIEWS) {
  LST_BIND;
  __shared__ float2 sRec[kTile];  // (eta, phi)
  const uint32_t n = v.R.scalars[kT5Listed];
  const PairItems items = pair_items(n);
  for (uint32_t it = blockIdx.x; it < items.items; it += gridDim.x) {
    const uint32_t x0 = (it / items.shares) * kTile;
    const uint32_t xl = x0 + kTile - 1 < n ? x0 + kTile - 1 : n - 1;
    const Band B = tile_band(v.W.pixBinOff, eta_bin(v.W.t5BinRec[x0].x), eta_bin(v.W.t5BinRec[xl].x));
    const YShare ys = share_of(B.n1 + B.n2, it % items.shares, items.shares);
    const uint32_t xi = x0 + threadIdx.x;
    uint32_t q = 0;
    bool active = xi < n;
    float eta1 = 0, phi1 = 0;
    if (active) {
      const float4 rx = v.W.t5BinRec[xi];  // (eta, phi, score, slot)
      q = __float_as_uint(rx.w);
      active = v.Q.isDup[q] == 0 && v.Q.partOfPT5[q] == 0;
      eta1 = rx.x;
      phi1 = rx.y;
    }
    bool dup = false;
    for (uint32_t y0 = ys.b; y0 < ys.e; y0 += kTile) {
      __syncthreads();
      if (y0 + threadIdx.x < ys.e) sRec[threadIdx.x] = v.W.pixBinRec[band_at(B, y0 + threadIdx.x)];
      __syncthreads();
      const uint32_t m = tile_len(ys, y0);
      if (!active || dup) continue;
      for (uint32_t t = 0; t < m && !dup; t++) {
        const float2 y = sRec[t];
        const float dEta = m_abs(eta1 - y.x);
        const float dPhi = wrap_dphi(phi1, y.y);
        dup = dEta * dEta + dPhi * dPhi < 1e-3f;
      }
    }
    if (dup) v.Q.isDup[q] = 1;
  }
}

// -------------------------------------------------------------- crossCleanpLS
// A quadruplet pLS not yet flagged is dropped when a T5 candidate lies within
// dR2 < 1e-3 of it, or a pT3/pT5 candidate within dR2 < 1e-6, or when one of
// its four pixel hits is a pixel hit of a pT3/pT5 candidate (checkPixelHits,
// SDL/TrackCandidate.h:166-202).  The two direction tests need |d eta| <
// 0.032 and never pass for a NaN direction, so the candidates are binned by
// eta and a pLS tile (the pLS in eta-bin order) meets only the bins next to
// its own; the hit test is a lookup in the set of the candidates' pixel hits
// (ntuple indices), done once per pLS.
__device__ __forceinline__ void pls_ntuple_hits(const EventView& v, uint32_t seg, int32_t* p) {
  const uint4 s = seg_chunk(v.S, seg, 0);  // {inner MD, outer MD, modules}
  const uint32_t a = s.x, b = s.y;
  const uint32_t h[4] = {md_chunk(v.D, a, 0).x, md_chunk(v.D, b, 0).x, md_chunk(v.D, a, 4).z, md_chunk(v.D, b, 4).z};
#pragma unroll
  for (int k = 0; k < 4; k++) p[k] = (int32_t)v.H.idxs[h[k]];  // anchor hits, then outer hits
}
constexpr int32_t kSetEmpty = INT32_MIN;  // never an ntuple index (those are >= -1); k_pls_connected clears the set
__device__ __forceinline__ uint32_t set_slot(int32_t h) { return ((uint32_t)h * 0x9E3779B1u) >> (32 - 18); }
static_assert(kTCHitSet == 1u << 18, "set_slot takes 18 bits");
__device__ __forceinline__ void set_insert(int32_t* set, int32_t h) {
  for (uint32_t s = set_slot(h);; s = (s + 1) & (kTCHitSet - 1)) {
    const int32_t old = atomicCAS(&set[s], kSetEmpty, h);
    if (old == kSetEmpty || old == h) return;
  }
}
// whether any of the four values other than -1 is in the set: the four probe
// sequences run side by side, one load of each per step
__device__ __forceinline__ bool set_has_any4(const int32_t* set, const int32_t h[4]) {
  uint32_t s[4];
  bool live[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    s[a] = set_slot(h[a]);
    live[a] = h[a] != -1;
  }
  while (live[0] || live[1] || live[2] || live[3]) {
    int32_t k[4];
#pragma unroll
    for (int a = 0; a < 4; a++) k[a] = live[a] ? set[s[a]] : kSetEmpty;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      if (!live[a]) continue;
      if (k[a] == h[a]) return true;
      if (k[a] == kSetEmpty) live[a] = false;
      else s[a] = (s[a] + 1) & (kTCHitSet - 1);
    }
  }
  return false;
}

// per track candidate: the direction a pLS is compared with and its kind (0:
// none, 1: T5, 2: pT3/pT5); the pT3/pT5 pixel hits go into the set
__global__ void k_tc_prep(LST_VIEWS) {
  LST_BIND;
  // modules whose objects a capacity truncated (the per-module counts are final here); one OR per wave
  uint32_t trunc = 0;
  const uint32_t nL = v.M.nLowerModules;
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < nL; m += gridDim.x * blockDim.x) {
    const uint32_t a = v.R.totOccupancyMDs[m], b = v.R.nMDs[m], c = v.R.totOccupancySegments[m],
                   d = v.R.nSegments[m], e = v.R.totOccupancyTriplets[m], f = v.R.nTriplets[m],
                   g = v.R.totOccupancyQuintuplets[m], h = v.R.nQuintuplets[m];
    trunc |= (a > b ? (uint32_t)kTruncMD : 0u) | (c > d ? (uint32_t)kTruncLS : 0u) |
             (e > f ? (uint32_t)kTruncT3 : 0u) | (g > h ? (uint32_t)kTruncT5 : 0u);
  }
  trunc = wave_or(trunc);
  if (trunc && (threadIdx.x & 63) == 0) atomicOr(&v.R.scalars[kTruncated], trunc);
  const uint32_t nTC = v.R.scalars[kNTrackCandidates];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nTC; t += gridDim.x * blockDim.x) {
    const int16_t type = v.C.trackCandidateType[t];
    const uint32_t inner = v.C.objectIndices[2 * t];
    float eta = 0, phi = 0;
    uint8_t kind = 0;
    int32_t p[4] = {-1, -1, -1, -1};
    if (type == 4) {
      eta = v.Q.eta[inner]; phi = v.Q.phi[inner]; kind = 1;
    } else if (type == 5) {
      eta = v.P3.eta_pix[inner]; phi = v.P3.phi_pix[inner]; kind = 2;
      pls_ntuple_hits(v, v.P3.pixelSegmentIndices[inner], p);
    } else if (type == 7) {
      eta = v.P.eta[inner - plsBase]; phi = v.P.phi[inner - plsBase]; kind = 2;
      pls_ntuple_hits(v, inner, p);
    }
    v.W.tcEta[t] = eta;
    v.W.tcPhi[t] = phi;
    v.W.tcKind[t] = kind;
    if (kind == 2)
      for (int a = 0; a < 4; a++)
        if (p[a] != -1) set_insert(v.W.tcHitSet, p[a]);
  }
}

struct TcBinItem {
  float4 rec;  // (eta, phi, dR2 bound, 0)
  float eta;
};
__global__ __launch_bounds__(1024) void k_bin_tc(LST_VIEWS) {
  LST_BIND;
  bin_by_eta(
      v.R.scalars[kNTrackCandidates], v.largePaths != 0,
      [&](uint32_t t) {
        const float eta = v.W.tcEta[t];
        const uint8_t kind = v.W.tcKind[t];  // 1: T5 (dR2 < 1e-3), 2: pT3 or pT5 (dR2 < 1e-6), 0: none
        const float bound = kind == 1 ? 1e-3f : (kind == 2 ? 0.000001f : -1.f);
        return TcBinItem{make_float4(eta, v.W.tcPhi[t], bound, 0.f), eta};
      },
      [&](uint32_t pos, const TcBinItem& t) { v.W.tcBinRec[pos] = t.rec; }, v.W.tcBinOff);
}

__global__ __launch_bounds__(kTile) void k_cross_pls_pairs(LST_VIEWS) {
  LST_BIND;
  __shared__ float4 sRec[kTile];  // (eta, phi, dR2 bound, 0)
  const uint32_t n = v.R.scalars[kEvNPLS];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  const PairItems items = pair_items(n);
  for (uint32_t it = blockIdx.x; it < items.items; it += gridDim.x) {
    const uint32_t x0 = (it / items.shares) * kTile;
    const uint32_t xl = x0 + kTile - 1 < n ? x0 + kTile - 1 : n - 1;
    const Band B = tile_band(v.W.tcBinOff, eta_bin(v.W.plsBinRec[x0].x), eta_bin(v.W.plsBinRec[xl].x));
    const YShare ys = share_of(B.n1 + B.n2, it % items.shares, items.shares);
    const uint32_t xi = x0 + threadIdx.x;
    uint32_t i = 0;
    bool active = xi < n;
    float eta1 = 0, phi1 = 0;
    bool dup = false;
    if (active) {
      const float4 rx = v.W.plsBinRec[xi];  // (eta, phi, score, index | isQuad << 31)
      i = __float_as_uint(rx.w) & 0x7fffffffu;
      active = (__float_as_uint(rx.w) & kQuadBit) != 0 && v.P.isDup[i] == 0;
      eta1 = rx.x;
      phi1 = rx.y;
      if (active && it % items.shares == 0) {  // the hit test, once per pLS
        int32_t p1[4];
        pls_ntuple_hits(v, plsBase + i, p1);
        dup = set_has_any4(v.W.tcHitSet, p1);
      }
    }
    for (uint32_t y0 = ys.b; y0 < ys.e; y0 += kTile) {
      __syncthreads();
      if (y0 + threadIdx.x < ys.e) sRec[threadIdx.x] = v.W.tcBinRec[band_at(B, y0 + threadIdx.x)];
      __syncthreads();
      const uint32_t m = tile_len(ys, y0);
      if (!active || dup) continue;
      for (uint32_t t2 = 0; t2 < m && !dup; t2++) {
        const float4 y = sRec[t2];
        const float dEta = m_abs(eta1 - y.x);
        const float dPhi = wrap_dphi(phi1, y.y);
        const float dR2 = dEta * dEta + dPhi * dPhi;
        dup = dR2 < y.z;  // never for a bound of -1 (dR2 >= 0 or NaN)
      }
    }
    if (active && dup) v.P.isDup[i] = 1;
  }
}

}  // namespace

void launch_pls_dup(const Launch& L, bool second) {
  const int tm = second ? kTmPlsDup2 : kTmPlsDup;
  if (!second) LST_GO(L, k_bin_pls, 1, 1, 1024);  // the eta order is kept for pass 2
  L.t.begin(L.s, tm);
  if (second) LST_GO_T(L, kTnPlsPairs, k_pls_pairs<true>, 1024, 256);  // 4096 waves stride over the binned pLS
  else LST_GO_T(L, kTnPlsPairs, k_pls_pairs<false>, 1024, 256);
  L.t.end(L.s, tm);
}

// the bucket counts are cleared by k_pls_connected, earlier in the chain
void launch_pt5_dup(const Launch& L) {
  LST_GO_T(L, kTnPT5BucketCount, k_pt5_bucket_count, 256, 256);
  LST_GO(L, k_pt5_bucket_scan, 1, 1, 1024);
  LST_GO_T(L, kTnPT5BucketFill, k_pt5_bucket_fill, 256, 256);
  L.t.begin(L.s, kTmPT5Dup);
  LST_GO_T(L, kTnPT5Pairs, k_pt5_pairs, 1024, 256);
  L.t.end(L.s, kTmPT5Dup);
}

void launch_pt3_dup(const Launch& L) {
  const dim3 g = pair_grid(kNMaxPixelTriplets, 512, kNMaxPixelTriplets);
  L.t.begin(L.s, kTmPT3Dup);
  L.tune_begin(kTnPT3Pairs);
  LST_GO(L, k_pt3_pairs, g.x, L.gx(kTnPT3Pairs, g.y), kTile);  // the tuned count is of y shares
  L.tune_end(kTnPT3Pairs);
  L.t.end(L.s, kTmPT3Dup);
}

void launch_cross_pt3(const Launch& L) {
  const dim3 g = pair_grid(kNMaxPixelTriplets, 512, kNMaxPixelQuintuplets);
  L.tune_begin(kTnCrossPT3);
  LST_GO(L, k_cross_pt3_pairs, g.x, L.gx(kTnCrossPT3, g.y), kTile);  // the tuned count is of y shares
  L.tune_end(kTnCrossPT3);
}

// the T5 slots of the eligible modules, module by module; built once the T5s
// are final (the pixel-quintuplet stage) and read by the later stages
void launch_t5_list(const Launch& L) {
  LST_GO(L, k_t5_list_offsets, 1, 1, 1024);
  LST_GO_T(L, kTnT5ListFill, k_t5_list_fill, 256, 256);
}

void launch_t5_dup_after_build(const Launch& L) {
  L.t.begin(L.s, kTmT5DupAfter);
  LST_GO_T(L, kTnT5DupAfter, k_t5_dup_after_build, 2048, 256);
  L.t.end(L.s, kTmT5DupAfter);
}

void launch_t5_dup_before_tc(const Launch& L) {
  LST_GO(L, k_bin_t5, 1, 1, 1024);
  L.t.begin(L.s, kTmT5DupTC);
  LST_GO_T(L, kTnT5TcPairs, k_t5_tc_pairs, 2048, 256);
  L.t.end(L.s, kTmT5DupTC);
}

void launch_cross_t5(const Launch& L) {
  LST_GO(L, k_bin_pix, 1, 1, 1024);
  LST_GO_T(L, kTnCrossT5, k_cross_t5_pairs, kPairItems, kTile);
}

void launch_cross_pls(const Launch& L) {
  LST_GO_T(L, kTnTcPrep, k_tc_prep, 128, 256);
  LST_GO(L, k_bin_tc, 1, 1, 1024);
  L.t.begin(L.s, kTmCrossPLS);
  LST_GO_T(L, kTnCrossPls, k_cross_pls_pairs, kPairItems, kTile);
  L.t.end(L.s, kTmCrossPLS);
}

}  // namespace lst
