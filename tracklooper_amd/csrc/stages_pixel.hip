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
// additions rank their flags with workgroup scans.  Every duplicate-removal
// kernel gives each object its own lane that tests all partners; the flag an
// object gets depends only on pair values and on flags written by earlier
// kernels, so the results do not depend on execution order.
#include "algo_pixel.h"
#include "stages.h"
#include "wave.h"

namespace lst {
namespace {

constexpr int kWave = 64;
constexpr int kMaxConn = 64;  // >= the largest connected-module list of a superbin (40)
constexpr unsigned kTCCap = kNMaxPixelTrackCandidates + kNMaxNonPixelTrackCandidates;

inline unsigned grid_of(uint32_t n, unsigned block) {
  unsigned g = (n + block - 1) / block;
  return g == 0 ? 1 : (g > 8192 ? 8192 : g);
}

// exclusive scan of counts[0, n) in place; counts[n] and *total get the sum
__global__ __launch_bounds__(1024) void k_scan_inplace(uint32_t* counts, uint32_t n, uint32_t* total) {
  __shared__ unsigned lds[17];
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

// --------------------------------------------------------------- pLS cleaning
// checkHitspLS (SDL/Kernels.h:384-473).  For a pair (lo < hi) within 0.1 in eta
// the reference drops the triplet seed over the quadruplet, else the higher
// score (ties: lo), when lo shares >= 3 of its hits with hi (first pass), or
// any hit or dR2 < 1e-5 (second pass, quadruplets not flagged by the first
// pass only).  Lane k collects the pairs whose removal target is k.
template <bool kSecond>
__global__ void k_pls_dup(EventView v) {
  const uint32_t n = v.nPLS;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    if (kSecond && (!v.P.isQuad[k] || (v.P.isDup[k] & 1))) continue;
    const float etaK = v.P.eta[k];
    bool flag = false;
    for (uint32_t j = 0; j < n && !flag; j++) {
      if (j == k) continue;
      const float etaJ = v.P.eta[j];
      const uint32_t lo = k < j ? k : j, hi = k < j ? j : k;
      const float etaLo = k < j ? etaK : etaJ, etaHi = k < j ? etaJ : etaK;
      if (m_abs(etaHi - etaLo) > 0.1f) continue;
      if (kSecond && (!v.P.isQuad[j] || (v.P.isDup[j] & 1))) continue;
      const int quadDiff = (int)v.P.isQuad[lo] - (int)v.P.isQuad[hi];
      const float scoreDiff = v.P.score[lo] - v.P.score[hi];
      uint32_t rm;
      if (quadDiff > 0) rm = hi;
      else if (quadDiff < 0) rm = lo;
      else if (scoreDiff < 0) rm = hi;
      else if (scoreDiff > 0) rm = lo;
      else rm = lo;
      if (rm != k) continue;
      int matched = 0;
      for (int a = 0; a < 4; a++) {
        const uint32_t h = v.P.pLSHitsIdxs[4 * lo + a];
        bool m = false;
        for (int b = 0; b < 4; b++)
          if (h == v.P.pLSHitsIdxs[4 * hi + b]) { m = true; break; }
        if (m) {
          matched++;
          if (kSecond) break;
        }
      }
      if (!kSecond) {
        flag = matched >= 3;
      } else {
        const float dEta = m_abs(etaLo - etaHi);
        const float dPhi = wrap_dphi(v.P.phi[lo], v.P.phi[hi]);
        flag = matched >= 1 || dEta * dEta + dPhi * dPhi < 1e-5f;
      }
    }
    if (flag) v.P.isDup[k] = (int8_t)(v.P.isDup[k] | (kSecond ? 2 : 1));
  }
}

// connected-module range of every pLS in the pixel map (SDL/Event.cc:828-856)
__global__ void k_pls_connected(EventView v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < v.nPLS; i += gridDim.x * blockDim.x) {
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
// one wave per pLS; candidates (connected module k, T5 b) in serial order
template <bool kWrite>
__global__ __launch_bounds__(kWave) void k_pixel_quintuplets(EventView v) {
  __shared__ uint32_t pre[kMaxConn + 1];
  const uint32_t i = blockIdx.x;
  if (i >= v.nPLS) return;
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
    for (uint32_t c0 = 0; c0 < total; c0 += kWave) {
      const uint32_t L = c0 + lane;
      bool pass = false;
      uint32_t t5 = 0;
      PT5Result r;
      if (L < total) {
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

__global__ void k_pt5_totals(EventView v) {
  const uint32_t tot = v.R.scalars[kTotOccPixelQuintuplets];
  v.R.scalars[kNPixelQuintuplets] = tot < kNMaxPixelQuintuplets ? tot : kNMaxPixelQuintuplets;
}

// removeDupPixelQuintupletsInGPUFromMap (SDL/Kernels.h:350-381): ix is dropped
// when a partner shares >= 7 of its 14 hits and scores lower (ties: the
// higher index is dropped)
__global__ void k_pt5_dup(EventView v) {
  const uint32_t n = v.R.scalars[kNPixelQuintuplets];
  for (uint32_t ix = blockIdx.x * blockDim.x + threadIdx.x; ix < n; ix += gridDim.x * blockDim.x) {
    uint32_t h[14];
#pragma unroll
    for (int a = 0; a < 14; a++) h[a] = v.P5.hitIndices[14 * ix + a];
    const float s1 = v.P5.score[ix];
    bool dup = false;
    for (uint32_t jx = 0; jx < n && !dup; jx++) {
      if (jx == ix) continue;
      const float s2 = v.P5.score[jx];
      if (!(s1 > s2 || (s1 == s2 && ix > jx))) continue;  // the flag needs this ordering
      int matched = 0;
      for (int a = 0; a < 14 && matched < 7 && matched + (14 - a) >= 7; a++) {
        for (int b = 0; b < 14; b++)
          if (h[a] == v.P5.hitIndices[14 * jx + b]) {
            matched++;
            break;
          }
      }
      dup = matched >= 7;
    }
    if (dup) v.P5.isDup[ix] = 1;
  }
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
// one wave per pLS; candidates (connected module k, T3 b) in serial order
template <bool kWrite>
__global__ __launch_bounds__(kWave) void k_pixel_triplets(EventView v) {
  __shared__ uint32_t pre[kMaxConn + 1];
  const uint32_t i = blockIdx.x;
  if (i >= v.nPLS) return;
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
    for (uint32_t c0 = 0; c0 < total; c0 += kWave) {
      const uint32_t L = c0 + lane;
      bool pass = false;
      uint32_t t3 = 0, m = 0;
      PT3Result r;
      if (L < total) {
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

__global__ void k_pt3_totals(EventView v) {
  const uint32_t tot = v.R.scalars[kTotOccPixelTriplets];
  v.R.scalars[kNPixelTriplets] = tot < kNMaxPixelTriplets ? tot : kNMaxPixelTriplets;
}

// removeDupPixelTripletsInGPUFromMap (SDL/Kernels.h:322-348) with checkHitspT3
// (:137-198): ix is dropped when a partner shares >= 5 hits and reaches a
// lower third layer, or the same layer with a lower score (ties: lower index)
__global__ void k_pt3_dup(EventView v) {
  const uint32_t n = v.R.scalars[kNPixelTriplets];
  for (uint32_t ix = blockIdx.x * blockDim.x + threadIdx.x; ix < n; ix += gridDim.x * blockDim.x) {
    uint32_t h[10];
#pragma unroll
    for (int a = 0; a < 10; a++) h[a] = v.P3.hitIndices[10 * ix + a];
    const uint8_t li = v.P3.logicalLayers[5 * ix + 2];
    const float si = v.P3.score[ix];
    bool dup = false;
    for (uint32_t jx = 0; jx < n && !dup; jx++) {
      if (jx == ix) continue;
      const uint8_t lj = v.P3.logicalLayers[5 * jx + 2];
      const float sj = v.P3.score[jx];
      if (!(lj < li || (li == lj && si > sj) || (li == lj && si == sj && ix < jx))) continue;
      int matched = 0;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
          if (h[a] == v.P3.hitIndices[10 * jx + b]) {
            matched++;
            break;
          }
      for (int a = 4; a < 10; a++)
        for (int b = 4; b < 10; b++)
          if (h[a] == v.P3.hitIndices[10 * jx + b]) {
            matched++;
            break;
          }
      dup = matched >= 5;
    }
    if (dup) v.P3.isDup[ix] = 1;
  }
}

// ---------------------------------------------------------- track candidates
// crossCleanpT3 (SDL/TrackCandidate.h:204-241)
__global__ void k_tc_cross_pt3(EventView v) {
  const uint32_t n = v.R.scalars[kNPixelTriplets], n5 = v.R.scalars[kNPixelQuintuplets];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    if (v.P3.isDup[p]) continue;
    const float eta1 = v.P3.eta_pix[p], phi1 = v.P3.phi_pix[p];
    bool dup = false;
    for (uint32_t q = 0; q < n5 && !dup; q++) {
      const uint32_t j = v.P5.pixelIndices[q] - plsBase;
      const float dEta = m_abs(eta1 - v.P.eta[j]);
      const float dPhi = wrap_dphi(phi1, v.P.phi[j]);
      dup = dEta * dEta + dPhi * dPhi < 1e-5f;
    }
    if (dup) v.P3.isDup[p] = 1;
  }
}

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

// removeDupQuintupletsInGPUBeforeTC (SDL/Kernels.h:250-320).  Among T5s not in
// a pT5 and not flagged after build, a close pair (|dEta|, |dPhi| <= 0.1 and
// dR2 < 0.001 or >= 5 shared hits) flags the higher score (ties: the lower
// index) with bit 2.  One workgroup per eligible module, one lane per T5.
__global__ __launch_bounds__(256) void k_t5_dup_before_tc(EventView v) {
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  if (blockIdx.x >= nElig) return;
  const uint16_t m1 = v.R.indicesOfEligibleT5Modules[blockIdx.x];
  const uint32_t n1 = v.R.nQuintuplets[m1];
  const uint32_t b1 = (uint32_t)v.R.quintupletModuleIndices[m1];
  for (uint32_t i1 = threadIdx.x; i1 < n1; i1 += blockDim.x) {
    const uint32_t x = b1 + i1;
    if (v.Q.partOfPT5[x] || (v.Q.isDup[x] & 1)) continue;
    const float etaX = v.Q.eta[x], phiX = v.Q.phi[x], sX = v.Q.score_rphisum[x];
    bool dup = false;
    for (uint32_t e = 0; e < nElig && !dup; e++) {
      const uint16_t m2 = v.R.indicesOfEligibleT5Modules[e];
      const uint32_t n2 = v.R.nQuintuplets[m2];
      const uint32_t b2 = (uint32_t)v.R.quintupletModuleIndices[m2];
      for (uint32_t j1 = 0; j1 < n2 && !dup; j1++) {
        const uint32_t y = b2 + j1;
        if (y == x) continue;
        if (v.Q.partOfPT5[y] || (v.Q.isDup[y] & 1)) continue;
        const float sY = v.Q.score_rphisum[y];
        const bool loses = sX > sY || (!(sX < sY) && x < y);
        if (!loses) continue;
        const float etaY = v.Q.eta[y], phiY = v.Q.phi[y];
        const uint32_t lo = x < y ? x : y, hi = x < y ? y : x;
        const float dEta = m_abs(x < y ? etaX - etaY : etaY - etaX);
        const float dPhi = x < y ? wrap_dphi(phiX, phiY) : wrap_dphi(phiY, phiX);
        if (dEta > 0.1f) continue;
        if (m_abs(dPhi) > 0.1f) continue;
        const float dR2 = dEta * dEta + dPhi * dPhi;
        int matched = 0;
        for (int a = 0; a < 10; a++) {
          const uint32_t h = v.Q.hitIndices[10 * lo + a];
          for (int b = 0; b < 10; b++)
            if (h == v.Q.hitIndices[10 * hi + b]) {
              matched++;
              break;
            }
        }
        dup = dR2 < 0.001f || matched >= 5;
      }
    }
    if (dup) v.Q.isDup[x] = (int8_t)(v.Q.isDup[x] | 2);
  }
}

// crossCleanT5 with Crossclean_T5 (SDL/TrackCandidate.h:243-294)
__global__ __launch_bounds__(256) void k_tc_cross_t5(EventView v) {
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  if (blockIdx.x >= nElig) return;
  const uint32_t n5 = v.R.scalars[kNPixelQuintuplets], n3 = v.R.scalars[kNPixelTriplets];
  const uint16_t m = v.R.indicesOfEligibleT5Modules[blockIdx.x];
  const uint32_t nq = v.R.nQuintuplets[m];
  const uint32_t base = (uint32_t)v.R.quintupletModuleIndices[m];
  for (uint32_t k = threadIdx.x; k < nq; k += blockDim.x) {
    const uint32_t q = base + k;
    if (v.Q.isDup[q] || v.Q.partOfPT5[q]) continue;
    const float eta1 = v.Q.eta[q], phi1 = v.Q.phi[q];
    bool dup = false;
    for (uint32_t j = 0; j < n5 + n3 && !dup; j++) {
      const float eta2 = j < n5 ? v.P5.eta[j] : v.P3.eta[j - n5];
      const float phi2 = j < n5 ? v.P5.phi[j] : v.P3.phi[j - n5];
      const float dEta = m_abs(eta1 - eta2);
      const float dPhi = wrap_dphi(phi1, phi2);
      dup = dEta * dEta + dPhi * dPhi < 1e-3f;
    }
    if (dup) v.Q.isDup[q] = 1;
  }
}

// addT5asTrackCandidateInGPU (SDL/TrackCandidate.h:430-482): count per module,
// scan over the eligible modules, then place in module and slot order
__device__ __forceinline__ bool t5_tc_flag(const EventView& v, uint32_t q) {
  return v.Q.isDup[q] == 0 && v.Q.partOfPT5[q] == 0 && v.Q.TightCutFlag[q] != 0;
}

__global__ __launch_bounds__(256) void k_tc_t5_count(EventView v) {
  __shared__ unsigned lds[5];
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  if (blockIdx.x >= nElig) return;
  const uint16_t m = v.R.indicesOfEligibleT5Modules[blockIdx.x];
  const uint32_t nq = v.R.nQuintuplets[m];
  const uint32_t base = (uint32_t)v.R.quintupletModuleIndices[m];
  unsigned c = 0;
  for (uint32_t k = threadIdx.x; k < nq; k += blockDim.x) c += t5_tc_flag(v, base + k) ? 1u : 0u;
  unsigned tot;
  block_exclusive_scan<256>(c, lds, tot);
  if (threadIdx.x == 0) v.W.t5TcCount[blockIdx.x] = tot;
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
  __shared__ unsigned lds[5];
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  if (blockIdx.x >= nElig) return;
  const uint16_t m = v.R.indicesOfEligibleT5Modules[blockIdx.x];
  const uint32_t nq = v.R.nQuintuplets[m];
  const uint32_t base = (uint32_t)v.R.quintupletModuleIndices[m];
  const uint32_t tcBase = v.R.scalars[kNTrackCandidatespT5] + v.R.scalars[kNTrackCandidatespT3];
  unsigned carry = v.W.t5TcCount[blockIdx.x];
  for (uint32_t k0 = 0; k0 < nq; k0 += 256) {
    const uint32_t k = k0 + threadIdx.x;
    const bool add = k < nq && t5_tc_flag(v, base + k);
    unsigned tot;
    const unsigned rank = carry + block_exclusive_scan<256>(add ? 1u : 0u, lds, tot);
    if (add && rank < kNMaxNonPixelTrackCandidates) {
      const uint32_t q = base + k;
      add_tc(v.C, tcBase + rank, 4, q, q, &v.Q.logicalLayers[5 * q], &v.Q.lowerModuleIndices[5 * q],
             &v.Q.hitIndices[10 * q], -1, v.Q.regressionG[q], v.Q.regressionF[q], v.Q.regressionRadius[q], q);
    }
    carry += tot;
  }
}

// checkPixelHits (SDL/TrackCandidate.h:166-202)
__device__ __forceinline__ int pixel_hits_matched(const EventView& v, uint32_t ix, uint32_t jx) {
  int p1[4], p2[4];
  const uint32_t a1 = v.S.mdIndices[2 * ix], b1 = v.S.mdIndices[2 * ix + 1];
  const uint32_t a2 = v.S.mdIndices[2 * jx], b2 = v.S.mdIndices[2 * jx + 1];
  p1[0] = (int)v.H.idxs[v.D.anchorHitIndices[a1]];
  p1[1] = (int)v.H.idxs[v.D.anchorHitIndices[b1]];
  p1[2] = (int)v.H.idxs[v.D.outerHitIndices[a1]];
  p1[3] = (int)v.H.idxs[v.D.outerHitIndices[b1]];
  p2[0] = (int)v.H.idxs[v.D.anchorHitIndices[a2]];
  p2[1] = (int)v.H.idxs[v.D.anchorHitIndices[b2]];
  p2[2] = (int)v.H.idxs[v.D.outerHitIndices[a2]];
  p2[3] = (int)v.H.idxs[v.D.outerHitIndices[b2]];
  int n = 0;
  for (int i = 0; i < 4; i++) {
    if (p1[i] == -1) continue;
    bool m = false;
    for (int j = 0; j < 4; j++) {
      if (p2[j] == -1) continue;
      if (p1[i] == p2[j]) { m = true; break; }
    }
    if (m) n++;
  }
  return n;
}

// crossCleanpLS (SDL/TrackCandidate.h:296-379)
__global__ void k_tc_cross_pls(EventView v) {
  const uint32_t nTC = v.R.scalars[kNTrackCandidates];
  const uint32_t plsBase = (uint32_t)v.segModuleIndices[v.M.nLowerModules];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < v.nPLS; i += gridDim.x * blockDim.x) {
    if (!v.P.isQuad[i] || v.P.isDup[i]) continue;
    const float eta1 = v.P.eta[i], phi1 = v.P.phi[i];
    bool dup = false;
    for (uint32_t t = 0; t < nTC && !dup; t++) {
      const int16_t type = v.C.trackCandidateType[t];
      const uint32_t inner = v.C.objectIndices[2 * t];
      if (type == 4) {
        const float dEta = m_abs(eta1 - v.Q.eta[inner]);
        const float dPhi = wrap_dphi(phi1, v.Q.phi[inner]);
        dup = dEta * dEta + dPhi * dPhi < 1e-3f;
      } else if (type == 5) {
        const uint32_t plsIdx = v.P3.pixelSegmentIndices[inner];
        const float dEta = m_abs(eta1 - v.P3.eta_pix[inner]);
        const float dPhi = wrap_dphi(phi1, v.P3.phi_pix[inner]);
        dup = pixel_hits_matched(v, plsBase + i, plsIdx) > 0 || dEta * dEta + dPhi * dPhi < 0.000001f;
      } else if (type == 7) {
        const uint32_t plsIdx = inner;
        const float dEta = m_abs(eta1 - v.P.eta[plsIdx - plsBase]);
        const float dPhi = wrap_dphi(phi1, v.P.phi[plsIdx - plsBase]);
        dup = pixel_hits_matched(v, plsBase + i, plsIdx) > 0 || dEta * dEta + dPhi * dPhi < 0.000001f;
      }
    }
    if (dup) v.P.isDup[i] = 1;
  }
}

// addpLSasTrackCandidateInGPU (SDL/TrackCandidate.h:484-530) with
// addpLSTrackCandidateToMemory (:112-124): hits stored in (x, z, y, w) order
__global__ __launch_bounds__(1024) void k_tc_add_pls(EventView v) {
  __shared__ unsigned lds[17];
  const uint32_t n = v.nPLS;
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

void launch_pixel_segment_cleaning(const EventView& v, hipStream_t s, const KernelTimer& t) {
  t.begin(s);
  hipLaunchKernelGGL(k_pls_dup<false>, dim3(grid_of(v.nPLS, 256)), dim3(256), 0, s, v);
  t.end(s);
}

void launch_pixel_quintuplets(const EventView& v, hipStream_t s, const KernelTimer& t) {
  const unsigned nP = v.nPLS ? v.nPLS : 1;
  hipLaunchKernelGGL(k_pls_connected, dim3(grid_of(v.nPLS, 256)), dim3(256), 0, s, v);
  t.begin(s);
  hipLaunchKernelGGL(k_pixel_quintuplets<false>, dim3(nP), dim3(kWave), 0, s, v);
  t.end(s);
  hipLaunchKernelGGL(k_scan_inplace, dim3(1), dim3(1024), 0, s, v.W.pt5Count, v.nPLS,
                     v.R.scalars + kTotOccPixelQuintuplets);
  hipLaunchKernelGGL(k_pt5_totals, dim3(1), dim3(1), 0, s, v);
  hipLaunchKernelGGL(k_pixel_quintuplets<true>, dim3(nP), dim3(kWave), 0, s, v);
  hipLaunchKernelGGL(k_pt5_dup, dim3(grid_of(kNMaxPixelQuintuplets, 256)), dim3(256), 0, s, v);
  hipLaunchKernelGGL(k_tc_add_pt5, dim3(1), dim3(1024), 0, s, v);
}

void launch_pixel_triplets(const EventView& v, hipStream_t s, const KernelTimer& t) {
  const unsigned nP = v.nPLS ? v.nPLS : 1;
  t.begin(s);
  hipLaunchKernelGGL(k_pixel_triplets<false>, dim3(nP), dim3(kWave), 0, s, v);
  t.end(s);
  hipLaunchKernelGGL(k_scan_inplace, dim3(1), dim3(1024), 0, s, v.W.pt3Count, v.nPLS,
                     v.R.scalars + kTotOccPixelTriplets);
  hipLaunchKernelGGL(k_pt3_totals, dim3(1), dim3(1), 0, s, v);
  hipLaunchKernelGGL(k_pixel_triplets<true>, dim3(nP), dim3(kWave), 0, s, v);
  hipLaunchKernelGGL(k_pt3_dup, dim3(grid_of(kNMaxPixelTriplets, 256)), dim3(256), 0, s, v);
}

void launch_track_candidates(const EventView& v, hipStream_t s, const KernelTimer& t) {
  const unsigned nElig = v.nT5SeedModules ? v.nT5SeedModules : 1;
  hipLaunchKernelGGL(k_tc_cross_pt3, dim3(grid_of(kNMaxPixelTriplets, 256)), dim3(256), 0, s, v);
  hipLaunchKernelGGL(k_tc_add_pt3, dim3(1), dim3(1024), 0, s, v);
  t.begin(s);
  hipLaunchKernelGGL(k_t5_dup_before_tc, dim3(nElig), dim3(256), 0, s, v);
  t.end(s);
  hipLaunchKernelGGL(k_tc_cross_t5, dim3(nElig), dim3(256), 0, s, v);
  hipLaunchKernelGGL(k_tc_t5_count, dim3(nElig), dim3(256), 0, s, v);
  hipLaunchKernelGGL(k_tc_t5_scan, dim3(1), dim3(1024), 0, s, v);
  hipLaunchKernelGGL(k_tc_t5_write, dim3(nElig), dim3(256), 0, s, v);
  hipLaunchKernelGGL(k_pls_dup<true>, dim3(grid_of(v.nPLS, 256)), dim3(256), 0, s, v);
  hipLaunchKernelGGL(k_tc_cross_pls, dim3(grid_of(v.nPLS, 256)), dim3(256), 0, s, v);
  hipLaunchKernelGGL(k_tc_add_pls, dim3(1), dim3(1024), 0, s, v);
  hipLaunchKernelGGL(k_tc_output, dim3(grid_of(kTCCap, 256)), dim3(256), 0, s, v);
}

}  // namespace lst
