// Stage kernels: hits, pixel line segments, mini-doublets, line segments,
// triplets.
//
// Builders run one work-group per lower module (the unit that owns an output
// range; one wave, which measured faster than four).  Candidates of the module
// are linearised in the reference's serial loop order (SDL/MiniDoublet.h:911-925,
// SDL/Segment.h:810-838, SDL/Triplet.h:886-913), evaluated one work-group width
// at a time, and the passing ones are ranked with ballot+mbcnt, so object
// slots equal the reference CPU slots and per-module capacity truncation keeps
// exactly the first `cap` objects.  Triplets use a flattened variant that
// spreads every module's candidates over the whole chip.
#include <climits>

#include "algo_md_ls_t3.h"
#include "flat_scan.h"
#include "stages.h"
#include "wave.h"

namespace lst {
namespace {

constexpr int kWave = 64;
constexpr int kBuildBlock = kWave;  // lanes per module for the MD and LS builders

// rank of a passing lane among the passing lanes of the work-group (ordered)
__device__ __forceinline__ unsigned build_rank(bool pass, unsigned* lds, unsigned& n) {
  if constexpr (kBuildBlock == kWave) {
    (void)lds;
    return wave_rank(pass, n);
  } else {
    return build_rank(pass, lds, n);
  }
}

// the detId table entry of det, ~0u when det is no module (the table is at
// most a quarter full, so a probe run ends at an empty slot)
__device__ __forceinline__ uint32_t det_lookup(const ModulesView& M, uint32_t det) {
  const uint32_t mask = (1u << M.detHashBits) - 1;
  uint32_t s = det_hash_slot(det, M.detHashBits);
  for (uint32_t n = 0; n <= mask; n++, s = (s + 1) & mask) {
    const uint2 e = M.detHash[s];
    if (e.x == det) return e.y;
    if (e.x == kNoDetId) break;
  }
  return ~0u;
}

// ------------------------------------------------------------------------ hits
__global__ void k_event_init(LST_VIEWS) {
  LST_BIND;
  const uint32_t n2 = 2 * v.M.nModules;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += gridDim.x * blockDim.x) {
    v.R.hitRanges[i] = (i & 1) ? -1 : INT_MAX;
    if (i < v.M.nModules) v.R.hitRangesUpper[i] = 0;  // per-module hit counts of k_hits
    v.R.mdRanges[i] = -1;
    v.R.segmentRanges[i] = -1;
    v.R.tripletRanges[i] = -1;
    v.R.quintupletRanges[i] = -1;
  }
  // every scalar starts at 0 except the event's hit and pLS counts, set before this kernel
  if (blockIdx.x == 0 && threadIdx.x < kNScalars && threadIdx.x != kEvNHits && threadIdx.x != kEvNPLS)
    v.R.scalars[threadIdx.x] = 0;
}

// SDL/Hit.h:217-270
__global__ __launch_bounds__(256) void k_hits(LST_VIEWS) {
  LST_BIND;
  const uint32_t nHits = v.R.scalars[kEvNHits];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nHits; i += gridDim.x * blockDim.x) {
    const float x = v.H.xs[i], y = v.H.ys[i], z = v.H.zs[i];
    const uint32_t det = v.H.detid[i];
    const float rt = m_sqrt(x * x + y * y);
    v.H.rts[i] = rt;
    v.W2.hitSlope[i] = md_hit_slope(rt, v.ptCut);
    v.H.phis[i] = phi_of(x, y);
    v.H.etas[i] = ((z > 0) - (z < 0)) * m_acosh(m_sqrt(x * x + y * y + z * z) / rt);
    float hx = 0.f, hy = 0.f, lx = 0.f, ly = 0.f;
    // module index and endcap-map entry (SDL/Hit.h:172-192 and the endcap
    // geometry map) from one table probe
    const uint32_t entry = det_lookup(v.M, det);
    const int mod = entry == ~0u ? -1 : (int)(entry & 0xFFFFu);
    if (mod < 0) {
      v.H.moduleIndices[i] = 0xFFFF;
      atomicOr(&v.R.scalars[kBadHits], (uint32_t)kBadDetId);
    } else {
      v.H.moduleIndices[i] = mod;
      {  // the strip-hit shift point, when this hit is the pixel-side hit of a pair that shifts
        const bool lower = v.M.isLower[mod] != 0;
        const uint32_t ml = lower ? (uint32_t)mod : v.M.partnerModuleIndices[mod];
        if (ml < v.M.nLowerModules) {
          const ModCtx c = load_mod(v.M, ml);
          if (md_shifts(c) && lower == md_lower_is_pixel_side(c)) v.W2.shiftPoint[i] = md_shift_point(v.M, ml, c, x, y, z, rt);
        }
      }
      if (v.M.subdets[mod] == kEndcap && v.M.moduleType[mod] == kTwoS) {
        const int e = (int)(entry >> 16) - 1;
        const float p = e >= 0 ? v.M.geoMapPhi[e] : 0.f;
        const float c = m_cos(p), s = m_sin(p);
        hx = x + 2.5f * c; lx = x - 2.5f * c;
        hy = y + 2.5f * s; ly = y - 2.5f * s;
      }
    }
    v.H.highEdgeXs[i] = hx; v.H.highEdgeYs[i] = hy; v.H.lowEdgeXs[i] = lx; v.H.lowEdgeYs[i] = ly;
    // hit range of each module: the lanes of a wave hold consecutive hits, so
    // the lanes that share a module give its smallest index at the lowest
    // lane and its largest at the highest; one atomic pair per such group
    const unsigned lane = threadIdx.x & 63u;
    unsigned long long pending = __ballot(mod >= 0);
    while (pending) {
      const int leader = __ffsll(pending) - 1;
      const int lm = __shfl(mod, leader, 64);
      const unsigned long long grp = __ballot(mod == lm);
      if ((int)lane == leader) {
        const int top = 63 - __clzll(grp);
        atomicMin(&v.R.hitRanges[2 * lm], (int)i);
        atomicMax(&v.R.hitRanges[2 * lm + 1], (int)(i - lane + (unsigned)top));
        atomicAdd(&v.R.hitRangesUpper[lm], (int)__popcll(grp));  // hit count, checked by k_module_ranges
      }
      pending &= ~grp;
    }
  }
}

// SDL/Hit.h:194-215
__global__ void k_module_ranges(LST_VIEWS) {
  LST_BIND;
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < v.M.nModules; m += gridDim.x * blockDim.x) {
    const int hi = v.R.hitRanges[2 * m + 1];
    // the hit count k_hits kept in hitRangesUpper[m] (read before this lane overwrites it; no other lane
    // reads it): the reference's ranges need at most 127 hits per outer-tracker module, all contiguous
    const int count = v.R.hitRangesUpper[m];
    if (hi != -1) {
      const int span = hi - v.R.hitRanges[2 * m] + 1;
      uint32_t bad = count > 127 && m != v.pixelModule ? (uint32_t)kBadModuleHits : 0u;  // pLS pseudo-hits
      bad |= count != span ? (uint32_t)kBadNonContiguous : 0u;
      if (bad) atomicOr(&v.R.scalars[kBadHits], bad);
    }
    int32_t lower = -1, upper = -1;
    int8_t nlo = -1, nup = -1;
    if (m < v.M.nLowerModules) {
      const uint32_t up = v.M.partnerModuleIndices[m];
      const int hiUp = v.R.hitRanges[2 * up + 1];
      if (hi != -1 && hiUp != -1) {
        lower = v.R.hitRanges[2 * m];
        upper = v.R.hitRanges[2 * up];
        nlo = hi - lower + 1;
        nup = hiUp - upper + 1;
      }
    }
    v.R.hitRangesLower[m] = lower; v.R.hitRangesUpper[m] = upper;
    v.R.hitRangesnLower[m] = nlo; v.R.hitRangesnUpper[m] = nup;
  }
}
__global__ void k_hit_ranges_fix(LST_VIEWS) {
  LST_BIND;
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < v.M.nModules; m += gridDim.x * blockDim.x)
    if (v.R.hitRanges[2 * m + 1] == -1) v.R.hitRanges[2 * m] = -1;
}

// ----------------------------------------------------------- mini-doublet store
// SDL/MiniDoublet.h:193-262
__device__ __forceinline__ void store_md(const EventView& v, uint32_t lowerHit, uint32_t upperHit, uint16_t mod,
                                         bool anchorIsUpper, const MDOut& o, uint32_t idx) {
  const HitsSoA& H = v.H;
  const uint32_t a = anchorIsUpper ? upperHit : lowerHit;
  const uint32_t b = anchorIsUpper ? lowerHit : upperHit;
  // the hot and cold records are built in registers and written with 16-byte stores: one lane writes
  // its 80 + 64 contiguous bytes in nine stores instead of 36 scattered 4-byte ones
  uint32_t hot[kMDHotBytes / 4], cold[kMDColdBytes / 4];
  const auto u = [&](MDWord k, uint32_t x) { (k < kMDHotFields ? hot : cold)[md_slot(k)] = x; };
  const auto f = [&](MDWord k, float x) { u(k, __float_as_uint(x)); };
  u(kMDw_anchorHitIndices, a);
  u(kMDw_outerHitIndices, b);
  u(kMDw_moduleIndices, mod);  // the u16 field in the low half of its slot
  f(kMDw_dphichanges, o.dphichange); f(kMDw_dphis, o.dphi); f(kMDw_dzs, o.dz);
  f(kMDw_shiftedXs, o.sx); f(kMDw_shiftedYs, o.sy); f(kMDw_shiftedZs, o.sz);
  f(kMDw_noShiftedDzs, o.nsdz); f(kMDw_noShiftedDphis, o.nsdphi); f(kMDw_noShiftedDphiChanges, o.nsdphichange);
  const float ahx = H.highEdgeXs[a], ahy = H.highEdgeYs[a], alx = H.lowEdgeXs[a], aly = H.lowEdgeYs[a];
  f(kMDw_anchorX, H.xs[a]); f(kMDw_anchorY, H.ys[a]); f(kMDw_anchorZ, H.zs[a]);
  f(kMDw_anchorRt, H.rts[a]); f(kMDw_anchorPhi, H.phis[a]); f(kMDw_anchorEta, H.etas[a]);
  f(kMDw_anchorHighEdgeX, ahx); f(kMDw_anchorHighEdgeY, ahy); f(kMDw_anchorLowEdgeX, alx); f(kMDw_anchorLowEdgeY, aly);
  f(kMDw_anchorHighEdgePhi, m_atan2(ahy, ahx));
  f(kMDw_anchorLowEdgePhi, m_atan2(aly, alx));
  f(kMDw_outerX, H.xs[b]); f(kMDw_outerY, H.ys[b]); f(kMDw_outerZ, H.zs[b]);
  f(kMDw_outerRt, H.rts[b]); f(kMDw_outerPhi, H.phis[b]); f(kMDw_outerEta, H.etas[b]);
  f(kMDw_outerHighEdgeX, H.highEdgeXs[b]); f(kMDw_outerHighEdgeY, H.highEdgeYs[b]);
  f(kMDw_outerLowEdgeX, H.lowEdgeXs[b]); f(kMDw_outerLowEdgeY, H.lowEdgeYs[b]);
  {  // the segment slope of this MD as the outer one: the anchor hit's slope (same formula) and its tan
    const float sl = v.W2.hitSlope[a];
    hot[kMDSlopeWord] = __float_as_uint(sl);
    hot[kMDSlopeWord + 1] = __float_as_uint(m_tan(sl));
  }
  uint4* h = reinterpret_cast<uint4*>(v.D.anchorHitIndices.p + (size_t)idx * kMDHotBytes);
  uint4* c = reinterpret_cast<uint4*>(v.D.shiftedXs.p + (size_t)idx * kMDColdBytes);
#pragma unroll
  for (uint32_t q = 0; q < kMDHotBytes / 16; q++) h[q] = make_uint4(hot[4 * q], hot[4 * q + 1], hot[4 * q + 2], hot[4 * q + 3]);
#pragma unroll
  for (uint32_t q = 0; q < kMDColdBytes / 16; q++)
    c[q] = make_uint4(cold[4 * q], cold[4 * q + 1], cold[4 * q + 2], cold[4 * q + 3]);
}

// One segment record built in registers and written with three 16-byte stores.
// d: dPhi, dPhiMin, dPhiMax, dPhiChange, dPhiChangeMin, dPhiChangeMax.
__device__ __forceinline__ void store_seg(const EventView& v, uint32_t s, uint32_t inMD, uint32_t outMD, uint16_t m,
                                          uint16_t om, uint32_t hitIn, uint32_t hitOut, const float (&d)[6]) {
  uint32_t w[kSegRecWords];
  w[seg_word(kSegF_mdIndices)] = inMD;
  w[seg_word(kSegF_mdIndices) + 1] = outMD;
  w[seg_word(kSegF_innerLowerModuleIndices)] = m;  // u16 fields in the low half of their word
  w[seg_word(kSegF_outerLowerModuleIndices)] = om;
  w[seg_word(kSegF_innerMiniDoubletAnchorHitIndices)] = hitIn;
  w[seg_word(kSegF_outerMiniDoubletAnchorHitIndices)] = hitOut;
  w[seg_word(kSegF_dPhis)] = __float_as_uint(d[0]);
  w[seg_word(kSegF_dPhiMins)] = __float_as_uint(d[1]);
  w[seg_word(kSegF_dPhiMaxs)] = __float_as_uint(d[2]);
  w[seg_word(kSegF_dPhiChanges)] = __float_as_uint(d[3]);
  w[seg_word(kSegF_dPhiChangeMins)] = __float_as_uint(d[4]);
  w[seg_word(kSegF_dPhiChangeMaxs)] = __float_as_uint(d[5]);
  uint4* rec = reinterpret_cast<uint4*>(v.S.mdIndices.p + (size_t)s * kSegRecBytes);
#pragma unroll
  for (uint32_t q = 0; q < kSegRecBytes / 16; q++) rec[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

// ------------------------------------------------------------- pixel segments
// SDL/Segment.h:1031-1117 and addPixelSegmentToMemory SDL/Segment.h:392-452
__global__ __launch_bounds__(256) void k_pixel_segments(LST_VIEWS) {
  LST_BIND;
  const uint32_t pix = v.pixelModule;
  const uint32_t nPLS = v.R.scalars[kEvNPLS];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    v.R.nSegments[pix] = nPLS; v.R.totOccupancySegments[pix] = nPLS;
    v.R.nMDs[pix] = 2 * nPLS; v.R.totOccupancyMDs[pix] = 2 * nPLS;
  }
  const MDOut zero{0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nPLS; t += gridDim.x * blockDim.x) {
    const uint32_t innerMD = v.mdModuleIndices[pix] + 2 * t, outerMD = innerMD + 1;
    const uint32_t seg = v.segModuleIndices[pix] + t;
    const uint32_t h0 = v.plsInHitIdx[0][t], h1 = v.plsInHitIdx[1][t], h2 = v.plsInHitIdx[2][t],
                   h3 = v.plsInHitIdx[3][t];
    store_md(v, h0, h1, pix, false, zero, innerMD);
    store_md(v, h2, h3, pix, false, zero, outerMD);
    // line through the PCA pseudo-hit with slope sinh(eta) and its residual at r3LH
    const float slope = m_sinh(v.H.ys[h1]);
    const float icpt = v.H.zs[h0] - slope * v.H.rts[h0];
    float score = (v.H.rts[h2] * slope + icpt) - v.H.zs[h2];
    score = score * score;
    {  // the reference writes no dPhi, dPhiMin/Max or dPhiChangeMin/Max for a pLS (SDL/Segment.h:392-411): 0 here
      const float d[6] = {0.f, 0.f, 0.f, v.plsInDPhiChange[t], 0.f, 0.f};
      store_seg(v, seg, innerMD, outerMD, pix, pix, h0, h2, d);
    }
    v.P.isDup[t] = 0;
    v.P.partOfPT5[t] = 0;
    v.P.score[t] = score;
    v.P.pLSHitsIdxs[4 * t + 0] = v.H.idxs[h0];
    v.P.pLSHitsIdxs[4 * t + 1] = v.H.idxs[h2];
    v.P.pLSHitsIdxs[4 * t + 2] = v.H.idxs[h1];
    v.P.pLSHitsIdxs[4 * t + 3] = v.H.idxs[h3];
    // two circle centres at distance R from the PCA point; keep the one closer to r3LH
    const float R = v.H.xs[h1] / (2 * k2Rinv1GeVf);
    const float ph = v.H.zs[h1];
    const float ax = v.H.xs[h0], ay = v.H.ys[h0];
    const float cx0 = ax + R * m_sin(ph), cx1 = ax - R * m_sin(ph);
    const float cy0 = ay - R * m_cos(ph), cy1 = ay + R * m_cos(ph);
    const float bx = v.H.xs[h2], by = v.H.ys[h2];
    const float d0 = m_abs(m_sqrt((bx - cx0) * (bx - cx0) + (by - cy0) * (by - cy0)) - R);
    const float d1 = m_abs(m_sqrt((bx - cx1) * (bx - cx1) + (by - cy1) * (by - cy1)) - R);
    // first minimum below SDL_INF wins, as the reference's scan over the two centres
    const float best0 = d0 < kSdlInf ? d0 : kSdlInf;
    const bool second = d1 < best0;
    v.P.circleCenterX[t] = second ? cx1 : cx0;
    v.P.circleCenterY[t] = second ? cy1 : cy0;
    v.P.circleRadius[t] = R;
  }
}

// module ranges (SDL/MiniDoublet.h, Segment.h, Triplet.h addRangesToEventExplicit),
// written where each module's object count is final; every range starts at -1
__device__ __forceinline__ void md_range(const EventView& v, uint32_t m, uint32_t n) {
  if (n != 0 && v.R.hitRanges[2 * m] != -1) {
    v.R.mdRanges[2 * m] = v.mdModuleIndices[m];
    v.R.mdRanges[2 * m + 1] = v.mdModuleIndices[m] + n - 1;
  }
}
__device__ __forceinline__ void seg_range(const EventView& v, uint32_t m, uint32_t n) {
  if (n != 0) {
    v.R.segmentRanges[2 * m] = v.segModuleIndices[m];
    v.R.segmentRanges[2 * m + 1] = v.segModuleIndices[m] + n - 1;
  }
}
__device__ __forceinline__ void t3_range(const EventView& v, uint32_t m, uint32_t n) {
  if (n != 0) {
    v.R.tripletRanges[2 * m] = v.R.tripletModuleIndices[m];
    v.R.tripletRanges[2 * m + 1] = v.R.tripletModuleIndices[m] + n - 1;
  }
}

// --------------------------------------------------------------- mini-doublets
// Flattened builder (the path every event takes unless it has more than
// kMDCandMax hit pairs): candidates of module m are [candOff[m], candOff[m+1]),
// pair (lower li, upper ui) at li * nUpper + ui.  A passing candidate parks its
// results at its own index L, so no slot is taken with an atomic.
__global__ __launch_bounds__(1024) void k_md_cand_scan(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[17];
  const uint32_t nL = v.M.nLowerModules;
  const unsigned total = scan_wide<kScanPer>(
      nL,
      [&](uint32_t m) {  // the counts are int8 as in the reference (<= 127 hits)
        const int lo = v.R.hitRangesLower[m], nLo = v.R.hitRangesnLower[m], nUp = v.R.hitRangesnUpper[m];
        return ((lo != -1) & (nLo > 0) & (nUp > 0)) ? (unsigned)(nLo * nUp) : 0u;
      },
      [&](uint32_t m, unsigned p, unsigned) { v.W2.candOff[m] = p; }, lds);
  if (threadIdx.x == 0) {
    v.W2.candOff[nL] = total;
    v.R.scalars[kMDCandTotal] = total;
  }
}

__device__ __forceinline__ bool md_flat_mode(const EventView& v) {
  return !v.perModuleBuilders && v.R.scalars[kMDCandTotal] <= kMDCandMax;
}
__device__ __forceinline__ bool md_flat_done(const EventView& v) {
  return md_flat_mode(v);
}

__device__ __forceinline__ uint32_t md_wave_index() {
  return __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}

// the candidates written out, L by L: one 16-lane row per module (four modules per wave), the module's candidate hit pairs in order
__global__ __launch_bounds__(256) void k_md_cand_fill(LST_VIEWS) {
  LST_BIND;
  if (!md_flat_mode(v)) return;
  const unsigned g = lane_id() / kRow, gl = lane_id() % kRow;
  const uint32_t nL = v.M.nLowerModules, groups = gridDim.x * (blockDim.x / kRow);
  for (uint32_t m = (blockIdx.x * blockDim.x + threadIdx.x) / kRow; m - g < nL; m += groups) {
    if (m >= nL) continue;
    const uint32_t n = v.W2.candOff[m + 1] - v.W2.candOff[m];
    if (n == 0) continue;
    const uint32_t off = v.W2.candOff[m], nUp = (uint32_t)v.R.hitRangesnUpper[m];
    const uint32_t lo = (uint32_t)v.R.hitRangesLower[m], up = (uint32_t)v.R.hitRangesUpper[m];
    for (uint32_t l = gl; l < n; l += kRow) v.W2.cand[off + l] = make_uint4(lo + l / nUp, up + l % nUp, m, 0u);
  }
}

// five waves per SIMD (88 VGPRs instead of the compiler's 116, no spills): with a launch set of many
// events in flight the selection is latency-bound and gains from the extra waves
__global__ __launch_bounds__(256, 5) void k_md_eval(LST_VIEWS) {
  LST_BIND;
  if (!md_flat_mode(v)) return;
  const uint32_t total = v.R.scalars[kMDCandTotal];
  for (uint32_t L0 = blockIdx.x * blockDim.x; L0 < total; L0 += gridDim.x * blockDim.x) {
    const uint32_t L = L0 + threadIdx.x;
    bool pass = false;
    MDOut o;
    if (L < total) {
      const uint4 cd = v.W2.cand[L];
      const uint32_t a = cd.x, b = cd.y, m = cd.z;
      const ModCtx c = load_mod(v.M, m);
      const float4 mt = md_terms(v.M, m);
      MDPair p{v.H.xs[a], v.H.ys[a], v.H.zs[a], v.H.rts[a], v.H.phis[a],
               v.H.xs[b], v.H.ys[b], v.H.zs[b], v.H.rts[b], v.H.phis[b], make_float4(0.f, 0.f, 0.f, 0.f)};
      if (md_shifts(c)) p.sp = v.W2.shiftPoint[md_lower_is_pixel_side(c) ? a : b];
      const float slope = v.W2.hitSlope[c.layerType == kPixelLayer ? a : b];  // the rt of the pixel-side hit
      pass = c.subdet == kBarrel ? md_barrel(c, mt, p, slope, o) : md_endcap(c, mt, p, slope, o);
    }
    if (pass) {  // parked at the candidate's own index
      float* w = reinterpret_cast<float*>(&v.W2.park[kMDParkWords * (size_t)L]);
      w[0] = o.dz; w[1] = o.dphi; w[2] = o.dphichange; w[3] = o.sx; w[4] = o.sy; w[5] = o.sz;
      w[6] = o.nsdz; w[7] = o.nsdphi; w[8] = o.nsdphichange;
    }
    store_flag_word(v.FS, L0 + (threadIdx.x & ~63u), total, pass);
  }
}


// slots: passing pairs in L order (pass-flag scan), then per-module counts
__global__ __launch_bounds__(1024) void k_md_scan_blocks(LST_VIEWS) {
  LST_BIND;
  if (!md_flat_done(v)) return;
  scan_mask_blocks(v.FS, v.R.scalars[kMDCandTotal]);
}
__global__ __launch_bounds__(1024) void k_md_scan_sums(LST_VIEWS) {
  LST_BIND;
  if (!md_flat_done(v)) return;
  scan_flag_sums(v.R.scalars[kMDCandTotal], v.FS.sums);
}

__device__ __forceinline__ uint32_t module_of(const uint32_t* off, uint32_t n, uint32_t L) {
  uint32_t lo = 0, hi = n;  // largest m with off[m] <= L
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= L) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_md_place(LST_VIEWS) {
  LST_BIND;
  if (!md_flat_done(v)) {
    // no per-module builder follows in a batch launch set: the stage is left empty (the event's overflow
    // flag sends it to the per-event path)
    if (v.noFallback)
      for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < v.M.nLowerModules; m += gridDim.x * blockDim.x) {
        v.R.totOccupancyMDs[m] = 0;
        v.R.nMDs[m] = 0;
        md_range(v, m, 0);
      }
    return;
  }
  const uint32_t total = v.R.scalars[kMDCandTotal], nL = v.M.nLowerModules;
  for_each_passing(v.FS, total, nL, [&](uint32_t L, uint32_t rankL) {
    // the loads that need L, then those that need its module, each group issued before any use
    const uint4 cd = v.W2.cand[L];
    const float* w = reinterpret_cast<const float*>(&v.W2.park[kMDParkWords * (size_t)L]);
    const MDOut o{w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8]};
    const uint32_t m = cd.z;
    const uint32_t off = v.W2.candOff[m];
    const int occ = v.mdModuleOccupancy[m];
    const ModCtx c = load_mod(v.M, m);
    const uint32_t rank = rankL - flag_rank(v.FS, off, total);
    if ((int)rank >= occ) return;
    const bool anchorIsUpper = c.type == kPS && c.layerType == kStripLayer;
    store_md(v, cd.x, cd.y, m, anchorIsUpper, o, v.mdModuleIndices[m] + rank);
  }, [&](uint32_t m) {  // then per module its counts and range
    const uint32_t count = flag_rank(v.FS, v.W2.candOff[m + 1], total) - flag_rank(v.FS, v.W2.candOff[m], total);
    v.R.totOccupancyMDs[m] = count;
    v.R.nMDs[m] = min(count, (unsigned)max(v.mdModuleOccupancy[m], 0));
    md_range(v, m, v.R.nMDs[m]);
  });
}

// Per-module builder, used when an event exceeds the flattened path's arrays:
// one wave per lower module walks its pairs in order.
__global__ __launch_bounds__(kBuildBlock) void k_mini_doublets(LST_VIEWS) {
  LST_BIND;
  __shared__ float sx[256], sy[256], sz[256], srt[256];
  __shared__ unsigned lds[kBuildBlock / 64 + 1];
  if (md_flat_done(v)) return;
  const uint32_t m = blockIdx.x;
  const unsigned lane = threadIdx.x;
  const int loBase = v.R.hitRangesLower[m];
  const int cap = v.mdModuleOccupancy[m];
  unsigned count = 0;
  const int nLo = v.R.hitRangesnLower[m], nUp = v.R.hitRangesnUpper[m];
  // counts wrapped negative (more than 127 hits, an input error) build nothing, as the reference's
  // index checks skip every pair then (SDL/MiniDoublet.h:919-923)
  if (loBase != -1 && nLo > 0 && nUp > 0) {
    const int upBase = v.R.hitRangesUpper[m];
    for (int i = lane; i < nLo; i += kBuildBlock) {
      sx[i] = v.H.xs[loBase + i]; sy[i] = v.H.ys[loBase + i]; sz[i] = v.H.zs[loBase + i]; srt[i] = v.H.rts[loBase + i];
    }
    for (int i = lane; i < nUp; i += kBuildBlock) {
      const int j = 128 + i;
      sx[j] = v.H.xs[upBase + i]; sy[j] = v.H.ys[upBase + i]; sz[j] = v.H.zs[upBase + i]; srt[j] = v.H.rts[upBase + i];
    }
    __syncthreads();
    const ModCtx c = load_mod(v.M, m);
    const float4 mt = md_terms(v.M, m);
    const bool anchorIsUpper = c.type == kPS && c.layerType == kStripLayer;
    const uint32_t base = v.mdModuleIndices[m];
    const int limit = nLo * nUp;
    for (int c0 = 0; c0 < limit; c0 += kBuildBlock) {
      const int k = c0 + lane;
      bool pass = false;
      MDOut o;
      int li = 0, ui = 0;
      if (k < limit) {
        li = k / nUp;
        ui = k - li * nUp;
        const int u = 128 + ui;
        const float slope = v.W2.hitSlope[c.layerType == kPixelLayer ? loBase + li : upBase + ui];
        MDPair p{sx[li], sy[li], sz[li], srt[li], v.H.phis[loBase + li], sx[u], sy[u], sz[u], srt[u],
                 v.H.phis[upBase + ui], make_float4(0.f, 0.f, 0.f, 0.f)};
        if (md_shifts(c)) p.sp = v.W2.shiftPoint[md_lower_is_pixel_side(c) ? loBase + li : upBase + ui];
        pass = c.subdet == kBarrel ? md_barrel(c, mt, p, slope, o) : md_endcap(c, mt, p, slope, o);
      }
      unsigned n;
      const unsigned rank = count + build_rank(pass, lds, n);
      if (pass && (int)rank < cap) store_md(v, loBase + li, upBase + ui, m, anchorIsUpper, o, base + rank);
      count += n;
    }
  }
  if (lane == 0) {
    v.R.totOccupancyMDs[m] = count;
    v.R.nMDs[m] = min(count, (unsigned)max(cap, 0));
    md_range(v, m, v.R.nMDs[m]);
  }
}


// --------------------------------------------------------------- line segments
// Flattened builder, as for the mini-doublets.  Per module, connPre[k] is the
// first candidate of connected module k (nIn * nMD(k) candidates each, inner
// MD major), and candOff the module's first global candidate.  One 16-lane
// row per module (four modules per wave), its connections 16 at a time.
__global__ __launch_bounds__(256) void k_ls_cand_count(LST_VIEWS) {
  LST_BIND;
  const unsigned g = lane_id() / kRow, gl = lane_id() % kRow;
  const uint32_t S = v.M.stride + 1, nL = v.M.nLowerModules, groups = gridDim.x * (blockDim.x / kRow);
  for (uint32_t m = (blockIdx.x * blockDim.x + threadIdx.x) / kRow; m - g < nL; m += groups) {
    if (m >= nL) continue;
    const uint32_t nIn = v.R.nMDs[m];
    const uint32_t nConn = nIn != 0 ? v.M.nConnectedModules[m] : 0;
    uint32_t run = 0;
    for (uint32_t k0 = 0; k0 < nConn; k0 += kRow) {  // the row's 16 lanes take each step together
      const uint32_t k = k0 + gl;
      const uint32_t c = k < nConn ? nIn * v.R.nMDs[v.M.moduleMap[(size_t)m * v.M.stride + k]] : 0;
      const unsigned incl = row_inclusive_scan(c);
      if (k < nConn) v.WL.connPre[(size_t)m * S + k] = run + incl - c;
      run += row_last(incl);
    }
    if (gl == 0) {
      v.WL.connPre[(size_t)m * S + nConn] = run;
      v.WL.candOff[m] = run;
    }
  }
}

__global__ __launch_bounds__(1024) void k_ls_cand_scan(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[17];
  const uint32_t nL = v.M.nLowerModules;
  const unsigned total = scan_wide<kScanPer>(
      nL, [&](uint32_t m) { return v.WL.candOff[m]; }, [&](uint32_t m, unsigned p, unsigned) { v.WL.candOff[m] = p; },
      lds);
  if (threadIdx.x == 0) {
    v.WL.candOff[nL] = total;
    v.R.scalars[kLSCandTotal] = total;
  }
}

__device__ __forceinline__ bool ls_flat_mode(const EventView& v) {
  return !v.perModuleBuilders && v.R.scalars[kLSCandTotal] <= kLSCandMax;
}
__device__ __forceinline__ bool ls_flat_done(const EventView& v) {
  return ls_flat_mode(v);
}


// the candidates written out, L by L: one 16-lane row per module (four
// modules per wave), its connected modules 16 at a time in the lanes (outer MD
// count and first MD, and the exclusive prefix of their candidate counts); the
// lanes then take the module's candidates of those connections in order, each
// finding its connection by a binary search over the prefix (row shuffles), so
// consecutive lanes write consecutive records whatever the per-connection
// counts
__global__ __launch_bounds__(256) void k_ls_cand_fill(LST_VIEWS) {
  LST_BIND;
  if (!ls_flat_mode(v)) return;
  const unsigned g = lane_id() / kRow, gl = lane_id() % kRow;
  const uint32_t S = v.M.stride + 1, nL = v.M.nLowerModules, groups = gridDim.x * (blockDim.x / kRow);
  for (uint32_t m = (blockIdx.x * blockDim.x + threadIdx.x) / kRow; m - g < nL; m += groups) {
    if (m >= nL) continue;
    const uint32_t nIn = v.R.nMDs[m];
    const uint32_t nConn = nIn != 0 ? v.M.nConnectedModules[m] : 0;
    if (nConn == 0) continue;
    const uint32_t base = v.WL.candOff[m], inBase = (uint32_t)v.R.mdRanges[2 * m];
    for (uint32_t k0 = 0; k0 < nConn; k0 += kRow) {  // the row's 16 lanes take each step together
      uint32_t nOut = 0, outBase = 0;
      if (k0 + gl < nConn) {
        const uint32_t om = v.M.moduleMap[(size_t)m * v.M.stride + k0 + gl];
        nOut = v.R.nMDs[om];
        outBase = nOut ? (uint32_t)v.R.mdRanges[2 * om] : 0u;
      }
      const uint32_t cnt = nIn * nOut;
      const uint32_t incl = row_inclusive_scan(cnt), pre = incl - cnt;  // candidates before connection k0 + gl
      const uint32_t chunk = row_last(incl);
      const uint32_t nk = nConn - k0 < kRow ? nConn - k0 : kRow;
      const uint32_t first = base + v.WL.connPre[(size_t)m * S + k0];  // L of the chunk's first candidate
      for (uint32_t l0 = 0; l0 < chunk; l0 += kRow) {
        const uint32_t l = l0 + gl;
        // the connection k of candidate l: the largest k < nk with pre_k <= l
        uint32_t k = 0;
        for (uint32_t step = kRow / 2; step >= 1; step >>= 1) {
          const uint32_t c = k + step;
          const uint32_t pc = row_shfl(pre, c < nk ? c : nk - 1);
          if (c < nk && pc <= l) k = c;
        }
        const uint32_t pk = row_shfl(pre, k), gOut = row_shfl(nOut, k), gBase = row_shfl(outBase, k);
        if (l < chunk) {
          const uint32_t r = l - pk;
          v.WL.cand[first + l] = make_uint2(inBase + r / gOut, gBase + r % gOut);
        }
      }
    }
  }
}

// four waves per SIMD (118 VGPRs with the module terms from the geometry; at most 128 kept as a bound,
// since three waves were 9 % slower in a 64-event launch set; five would spill)
__global__ __launch_bounds__(256, 4) void k_ls_eval(LST_VIEWS) {
  LST_BIND;
  if (!ls_flat_mode(v)) return;
  const uint32_t total = v.R.scalars[kLSCandTotal];
  for (uint32_t L0 = blockIdx.x * blockDim.x; L0 < total; L0 += gridDim.x * blockDim.x) {
    const uint32_t L = L0 + threadIdx.x;
    bool pass = false;
    SegOut o{0, 0, 0, 0, 0, 0};
    if (L < total) {
      const uint2 c = v.WL.cand[L];
      const uint32_t inMD = c.x, outMD = c.y;
      const MDRef a = load_md(v.D, inMD), b = load_md(v.D, outMD);
      const float2 bs = md_slope(v.D, outMD);
      const uint32_t m = a.mod, om = b.mod;
      const ModCtx ci = load_mod(v.M, m), co = load_mod(v.M, om);
      const float4 ti = seg_terms(v.M, m), to = seg_terms(v.M, om);
      pass = (ci.subdet == kBarrel && co.subdet == kBarrel) ? seg_barrel(ci, ti, co, to, v.ptCut, a, b, bs, o)
                                                            : seg_endcap(ci, ti, co, to, v.ptCut, a, b, bs, o);
    }
    if (pass) {  // parked at the candidate's own index
      float* w = reinterpret_cast<float*>(&v.WL.park[kLSParkWords * (size_t)L]);
      w[0] = o.dPhi; w[1] = o.dPhiMin; w[2] = o.dPhiMax;
      w[3] = o.dPhiChange; w[4] = o.dPhiChangeMin; w[5] = o.dPhiChangeMax;
    }
    store_flag_word(v.FS, L0 + (threadIdx.x & ~63u), total, pass);
  }
}

__global__ __launch_bounds__(1024) void k_ls_scan_blocks(LST_VIEWS) {
  LST_BIND;
  if (!ls_flat_done(v)) return;
  scan_mask_blocks(v.FS, v.R.scalars[kLSCandTotal]);
}
__global__ __launch_bounds__(1024) void k_ls_scan_sums(LST_VIEWS) {
  LST_BIND;
  if (!ls_flat_done(v)) return;
  scan_flag_sums(v.R.scalars[kLSCandTotal], v.FS.sums);
}

__global__ __launch_bounds__(256) void k_ls_place(LST_VIEWS) {
  LST_BIND;
  if (!ls_flat_done(v)) {
    // no per-module builder follows in a batch launch set: the stage is left empty (the event's overflow
    // flag sends it to the per-event path)
    if (v.noFallback)
      for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < v.M.nLowerModules; m += gridDim.x * blockDim.x) {
        v.R.totOccupancySegments[m] = 0;
        v.R.nSegments[m] = 0;
        seg_range(v, m, 0);
      }
    return;
  }
  const uint32_t total = v.R.scalars[kLSCandTotal], nL = v.M.nLowerModules;
  for_each_passing(v.FS, total, nL, [&](uint32_t L, uint32_t rankL) {
    // the loads that need L, then those that need its mini-doublets and module, each group issued before any use
    const uint2 c = v.WL.cand[L];
    static_assert(kLSParkWords % 2 == 0, "parked results in 8-byte pairs");
    const float2* w = reinterpret_cast<const float2*>(&v.WL.park[kLSParkWords * (size_t)L]);
    const float2 p0 = ld_global(w), p1 = ld_global(w + 1), p2 = ld_global(w + 2);
    const uint32_t inMD = c.x, outMD = c.y;
    const uint4 a = md_chunk(v.D, inMD, 0), b = md_chunk(v.D, outMD, 0);  // {anchor hit, module, ...}
    const uint32_t m = (uint16_t)a.y, om = (uint16_t)b.y, hitIn = a.x, hitOut = b.x;
    const uint32_t off = v.WL.candOff[m];
    const int occ = v.segModuleOccupancy[m];
    const uint32_t segBase = v.segModuleIndices[m];
    const uint32_t rank = rankL - flag_rank(v.FS, off, total);
    if ((int)rank >= occ) return;
    const uint32_t s = segBase + rank;
    const float d[6] = {p0.x, p0.y, p1.x, p1.y, p2.x, p2.y};
    store_seg(v, s, inMD, outMD, m, om, hitIn, hitOut, d);
  }, [&](uint32_t m) {  // then per module its counts and range
    const uint32_t count = flag_rank(v.FS, v.WL.candOff[m + 1], total) - flag_rank(v.FS, v.WL.candOff[m], total);
    v.R.totOccupancySegments[m] = count;
    v.R.nSegments[m] = min(count, (unsigned)max(v.segModuleOccupancy[m], 0));
    seg_range(v, m, v.R.nSegments[m]);
  });
}

// Per-module builder, used when an event exceeds the flattened path's arrays.
__global__ __launch_bounds__(kBuildBlock) void k_segments(LST_VIEWS) {
  LST_BIND;
  __shared__ uint32_t pre[160];
  __shared__ uint16_t outMod[160];
  __shared__ uint32_t outCnt[160];
  __shared__ unsigned lds[kBuildBlock / 64 + 1];
  const uint32_t m = blockIdx.x;
  const unsigned lane = threadIdx.x;
  const uint32_t nIn = v.R.nMDs[m];
  const int cap = v.segModuleOccupancy[m];
  unsigned count = 0;
  if (ls_flat_done(v)) return;
  if (nIn != 0) {
    const uint32_t nConn = v.M.nConnectedModules[m];
    for (uint32_t k = lane; k < nConn; k += kBuildBlock) {
      const uint16_t om = v.M.moduleMap[(size_t)m * v.M.stride + k];
      outMod[k] = om;
      outCnt[k] = v.R.nMDs[om];
    }
    __syncthreads();
    if (lane == 0) {
      uint32_t run = 0;
      for (uint32_t k = 0; k < nConn; k++) { pre[k] = run; run += nIn * outCnt[k]; }
      pre[nConn] = run;
    }
    __syncthreads();
    const uint32_t total = pre[nConn];
    const ModCtx ci = load_mod(v.M, m);
    const float4 ti = seg_terms(v.M, m);
    const uint32_t inBase = v.R.mdRanges[2 * m];
    const uint32_t segBase = v.segModuleIndices[m];
    for (uint32_t c0 = 0; c0 < total; c0 += kBuildBlock) {
      const uint32_t L = c0 + lane;
      bool pass = false;
      SegOut o{0, 0, 0, 0, 0, 0};
      uint32_t inMD = 0, outMD = 0;
      uint16_t om = 0;
      if (L < total) {
        uint32_t k = 0, hiK = nConn;  // largest k with pre[k] <= L
        while (hiK - k > 1) {
          const uint32_t mid = (k + hiK) >> 1;
          if (pre[mid] <= L) k = mid;
          else hiK = mid;
        }
        om = outMod[k];
        const uint32_t local = L - pre[k];
        const uint32_t nOut = outCnt[k];
        inMD = inBase + local / nOut;
        outMD = v.R.mdRanges[2 * om] + local % nOut;
        const ModCtx co = load_mod(v.M, om);
        const float4 to = seg_terms(v.M, om);
        const MDRef a = load_md(v.D, inMD), b = load_md(v.D, outMD);
        const float2 bs = md_slope(v.D, outMD);
        pass = (ci.subdet == kBarrel && co.subdet == kBarrel) ? seg_barrel(ci, ti, co, to, v.ptCut, a, b, bs, o)
                                                              : seg_endcap(ci, ti, co, to, v.ptCut, a, b, bs, o);
      }
      unsigned n;
      const unsigned rank = count + build_rank(pass, lds, n);
      if (pass && (int)rank < cap) {
        const uint32_t s = segBase + rank;
        const float d[6] = {o.dPhi, o.dPhiMin, o.dPhiMax, o.dPhiChange, o.dPhiChangeMin, o.dPhiChangeMax};
        store_seg(v, s, inMD, outMD, m, om, v.D.anchorHitIndices[inMD], v.D.anchorHitIndices[outMD], d);
      }
      count += n;
    }
  }
  if (lane == 0) {
    v.R.totOccupancySegments[m] = count;
    v.R.nSegments[m] = min(count, (unsigned)max(cap, 0));
    seg_range(v, m, v.R.nSegments[m]);
  }
}


// -------------------------------------------------------------------- triplets
// capacity offsets: a module with segments gets its tabulated capacity, others 0
// (SDL/Triplet.h:1007-1102); exclusive scan in module order.
__global__ __launch_bounds__(1024) void k_triplet_offsets(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[17];
  const unsigned total = scan_wide<kScanPer>(
      v.M.nLowerModules,
      [&](uint32_t m) {
        const unsigned ns = v.R.nSegments[m], occ = (unsigned)v.t3OccIfAny[m];
        return ns != 0 ? occ : 0u;
      },
      [&](uint32_t m, unsigned p, unsigned occ) {
        v.R.tripletModuleOccupancy[m] = (int)occ;
        v.R.tripletModuleIndices[m] = (int)p;
      },
      lds);
  if (threadIdx.x == 0) v.R.scalars[kNTotalTrips] = total;
}

// runTripletConstraintsAndAlgo (SDL/Triplet.h:796-872) for the segment pair
// (inSeg on m1, outSeg), including the shared middle mini-doublet check
__device__ __forceinline__ bool t3_select(const EventView& v, uint32_t m1, uint32_t inSeg, uint32_t outSeg,
                                          float& betaIn, float& radius, float& gx, float& gy) {
  const uint4 si = seg_chunk(v.S, inSeg, 0), so = seg_chunk(v.S, outSeg, 0);
  const uint32_t md1 = si.x, md2 = so.x, md3 = so.y;
  if (si.y != md2) return false;  // SDL/Triplet.h:823-825
  const uint16_t m2 = (uint16_t)si.w, m3 = (uint16_t)so.w;
  T3Geo p;  // the anchor points: chunk 1 of each record, and the first one's phi from its chunk 2
  const uint4 a1 = md_chunk(v.D, md1, 1), a2 = md_chunk(v.D, md2, 1), a3 = md_chunk(v.D, md3, 1);
  p.x1 = __uint_as_float(a1.x); p.y1 = __uint_as_float(a1.y); p.z1 = __uint_as_float(a1.z); p.r1 = __uint_as_float(a1.w);
  p.phi1 = v.D.anchorPhi[md1];
  p.x2 = __uint_as_float(a2.x); p.y2 = __uint_as_float(a2.y); p.z2 = __uint_as_float(a2.z); p.r2 = __uint_as_float(a2.w);
  p.x3 = __uint_as_float(a3.x); p.y3 = __uint_as_float(a3.y); p.z3 = __uint_as_float(a3.z); p.r3 = __uint_as_float(a3.w);
  p.slope3 = md_slope(v.D, md3);
  const uint32_t w1 = mod_word(v.M, m1), w2 = mod_word(v.M, m2), w3 = mod_word(v.M, m3);
  const uint4 sa = seg_chunk(v.S, inSeg, 2);  // dPhiMax, dPhiChange, dPhiChangeMin, dPhiChangeMax
  ModCtx c1, c2, c3;
  unpack_mod(w1, c1); unpack_mod(w2, c2); unpack_mod(w3, c3);
  if (!t3_rz(unpack_sdl_layer(w1), unpack_sdl_layer(w2), unpack_sdl_layer(w3), p)) return false;
  const int16_t s1 = c1.subdet, s2 = c2.subdet, s3 = c3.subdet;
  const bool psIn = c1.type == kPS, psOut = c3.type == kPS;
  const float alphaIn = __uint_as_float(sa.y);
  bool pass = false;
  if (s1 == kBarrel && s2 == kBarrel && s3 == kBarrel) pass = t3_bbb(psIn, psOut, v.ptCut, p, alphaIn, betaIn);
  else if (s1 == kBarrel && s3 == kEndcap) pass = t3_bbe(psIn, psOut, v.ptCut, p, alphaIn, betaIn);
  else if (s1 == kEndcap && s2 == kEndcap && s3 == kEndcap)
    pass = t3_eee(psIn, psOut, v.ptCut, p, alphaIn, __uint_as_float(sa.z), __uint_as_float(sa.w), betaIn);
  if (pass) radius = circle_radius3(p.x1, p.y1, p.x2, p.y2, p.x3, p.y3, gx, gy);
  return pass;
}

// addTripletToMemory (SDL/Triplet.h:153-239)
__device__ __forceinline__ void t3_store(const EventView& v, uint32_t t, uint32_t m1, uint32_t inSeg, uint32_t outSeg,
                                         float betaIn, float radius, float gx, float gy) {
  const uint4 sIn = seg_chunk(v.S, inSeg, 0), sOut = seg_chunk(v.S, outSeg, 0);  // {MD in, MD out, modules}
  const uint16_t m2 = (uint16_t)sIn.w, m3 = (uint16_t)sOut.w;
  const uint32_t md1 = sIn.x, mdB = sIn.y, md3 = sOut.y;
  // the record built in registers (u16 and u8 fields in the low bytes of their words, flags 0) and
  // written with 16-byte stores
  uint32_t w[kT3ObjWords] = {};
  w[t3_word(kT3F_segmentIndices)] = inSeg;
  w[t3_word(kT3F_segmentIndices) + 1] = outSeg;
  const uint16_t mods[3] = {(uint16_t)m1, m2, m3};
  const uint32_t mds[3] = {md1, mdB, md3};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    w[t3_word(kT3F_lowerModuleIndices) + k] = mods[k];
    ModCtx c;
    unpack_mod(mod_word(v.M, mods[k]), c);
    w[t3_word(kT3F_logicalLayers) + k] = (uint8_t)(c.layer + (c.subdet == kEndcap) * 6);
    w[t3_word(kT3F_hitIndices) + 2 * k] = md_chunk(v.D, mds[k], 0).x;      // anchor hit
    w[t3_word(kT3F_hitIndices) + 2 * k + 1] = md_chunk(v.D, mds[k], 4).z;  // outer hit
  }
  w[t3_word(kT3F_betaIn)] = __float_as_uint(betaIn);
  w[t3_word(kT3F_circleRadius)] = __float_as_uint(radius);
  w[t3_word(kT3F_circleCenterX)] = __float_as_uint(gx);
  w[t3_word(kT3F_circleCenterY)] = __float_as_uint(gy);
  uint4* rec = reinterpret_cast<uint4*>(v.T.segmentIndices.p + (size_t)t * kT3ObjBytes);
#pragma unroll
  for (uint32_t q = 0; q < kT3ObjBytes / 16; q++) rec[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

// ----------------------------------------------------- flattened T3 builder
// Candidates are numbered in the reference's serial order (module m1, inner
// segment a, outer segment b on the inner segment's outer module): module m1
// owns [modOff[m1], modOff[m1+1]); the shared-MD candidates that the list
// keeps are [survOff[m1], survOff[m1+1]) in the same order.
__device__ __forceinline__ bool t3_flat_mode(const EventView& v) {
  return !v.perModuleBuilders && v.R.scalars[kT3CandTotal] <= kT3CandMax;
}
__device__ __forceinline__ bool t3_flat_done(const EventView& v) {
  return t3_flat_mode(v);
}

// The flattened T3 kernels below run one wave per lower module at a time and
// stride over the modules, so a fixed grid that fills the chip covers every
// module without dispatching a work-group per (mostly small) module.
__device__ __forceinline__ uint32_t wave_index() {
  return __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}
__device__ __forceinline__ uint32_t wave_count() { return (gridDim.x * blockDim.x) / kWave; }
constexpr unsigned kModuleWaveBlocks = 1024;  // x 4 waves

// per module, the candidates of its inner segments (the outer modules'
// segments) and their shared-MD candidates (the segments listed under each
// inner segment's outer mini-doublet).  One 16-lane group per module (four
// modules per wave: most hold a few segments); the lanes stride over the inner
// segments and the group sums their counts.
__global__ __launch_bounds__(256) void k_t3_cand_count(LST_VIEWS) {
  LST_BIND;
  const unsigned g = lane_id() / kRow, gl = lane_id() % kRow;
  const uint32_t nL = v.M.nLowerModules, groups = gridDim.x * (blockDim.x / kRow);
  // the wave's four groups take consecutive modules; the loop runs on the first one's index, so every lane
  // of the wave takes the same iterations
  for (uint32_t m1 = (blockIdx.x * blockDim.x + threadIdx.x) / kRow; m1 - g < nL; m1 += groups) {
    const bool has = m1 < nL;
    const uint32_t nIn = has && v.M.nConnectedModules[m1] != 0 ? v.R.nSegments[m1] : 0;
    const uint32_t segBase = has ? v.segModuleIndices[m1] : 0;
    uint32_t cnt = 0, sc = 0;
    for (uint32_t a = gl; a < nIn; a += kRow) {
      const uint4 s0 = seg_chunk(v.S, segBase + a, 0);  // outer MD X = s0.y, outer module = s0.w
      cnt += v.R.nSegments[(uint16_t)s0.w];
      sc += v.W3.mdSegEnd[s0.y] - v.W3.mdSegOff[s0.y];
    }
    cnt = row_sum(cnt);
    sc = row_sum(sc);
    if (has && gl == 0) {
      v.W3.modOff[m1] = cnt;
      v.W3.survOff[m1] = sc;
    }
  }
}

__global__ __launch_bounds__(1024) void k_t3_cand_scan(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[34];
  const uint32_t n = v.M.nLowerModules;
  const UVec<2> tot = scan_wide_n<kScanPer, 2>(
      n, [&](uint32_t m) { return UVec<2>{{v.W3.modOff[m], v.W3.survOff[m]}}; },
      [&](uint32_t m, const UVec<2>& p, const UVec<2>&) {
        v.W3.modOff[m] = p.v[0];
        v.W3.survOff[m] = p.v[1];
      },
      lds);
  const unsigned total = tot.v[0], surv = tot.v[1];
  if (threadIdx.x == 0) {
    v.W3.modOff[n] = total;
    v.W3.survOff[n] = surv;
    v.R.scalars[kT3CandTotal] = total;
    v.R.scalars[kT3List] = surv;
  }
}

// stage 1: the candidates whose inner segment's outer mini-doublet X is the
// outer segment's inner one.  Per mini-doublet, the segments starting at it
// are listed first, in slot order (four lower modules per wave, one per row
// where they are small: counts, offsets, a scatter inside the module's own segment range, then each
// segment placed at its rank within its short list); the candidate count
// pass also counts, per inner segment,
// the segments listed under X; then one wave per module m1 walks its inner
// segments and writes, for each segment listed under X, the candidate at the
// position those counts give.  The list is thereby in the reference's serial
// order (module, inner segment, outer segment).
__global__ __launch_bounds__(256) void k_md_seg_lists(LST_VIEWS) {
  LST_BIND;
  __shared__ __align__(16) WaveLists sLists[4];
  md_lists_all(
      v.M.nLowerModules,
      [&](uint32_t m, uint32_t& nMD, uint32_t& mdBase, uint32_t& n, uint32_t& base) {
        nMD = v.R.nMDs[m];
        n = v.R.nSegments[m];
        mdBase = (uint32_t)v.mdModuleIndices[m];
        base = (uint32_t)v.segModuleIndices[m];
      },
      [&](uint32_t s) { return seg_chunk(v.S, s, 0).x; }, v.W3.mdSegOff, v.W3.mdSegEnd, v.W3.segByMD, v.W3.segTmp,
      sLists[threadIdx.x / kWave], v.largePaths != 0);
}

// One 16-lane group per module m1 (four modules per wave), its inner segments
// 16 at a time in the lanes (the segments listed under each one's outer
// mini-doublet: range and count, and the exclusive prefix of the counts); the
// lanes then take the chunk's candidates in order, each finding its inner
// segment by a binary search over the prefix (row shuffles), so consecutive
// lanes write consecutive entries.  Every group-level step runs with the
// group's 16 lanes active (its loop bounds are the group's own).
__global__ __launch_bounds__(256) void k_t3_enum(LST_VIEWS) {
  LST_BIND;
  if (!t3_flat_mode(v)) return;
  const unsigned g = lane_id() / kRow, gl = lane_id() % kRow;
  uint2* list = reinterpret_cast<uint2*>(v.W3.list);
  const uint32_t nL = v.M.nLowerModules, groups = gridDim.x * (blockDim.x / kRow);
  for (uint32_t m1 = (blockIdx.x * blockDim.x + threadIdx.x) / kRow; m1 - g < nL; m1 += groups) {
    const bool has = m1 < nL;
    const uint32_t nIn = has && v.M.nConnectedModules[m1] != 0 ? v.R.nSegments[m1] : 0;
    if (nIn == 0) continue;
    const uint32_t segBase = v.segModuleIndices[m1];
    uint32_t run = v.W3.survOff[m1];  // list position of the chunk's first candidate
    for (uint32_t a0 = 0; a0 < nIn; a0 += kRow) {
      const uint32_t a = a0 + gl;
      uint32_t cnt = 0, b = 0, seg2Base = 0;
      if (a < nIn) {
        const uint4 s0 = seg_chunk(v.S, segBase + a, 0);  // outer MD X = s0.y, outer module = s0.w
        b = v.W3.mdSegOff[s0.y];
        cnt = v.W3.mdSegEnd[s0.y] - b;
        seg2Base = v.segModuleIndices[(uint16_t)s0.w];
      }
      const uint32_t incl = row_inclusive_scan(cnt), pre = incl - cnt;
      const uint32_t chunk = row_last(incl);
      const uint32_t nk = nIn - a0 < kRow ? nIn - a0 : kRow;
      for (uint32_t l0 = 0; l0 < chunk; l0 += kRow) {
        const uint32_t l = l0 + gl;
        uint32_t k = 0;  // the largest k < nk with pre_k <= l
        for (uint32_t step = kRow / 2; step >= 1; step >>= 1) {
          const uint32_t c = k + step;
          const uint32_t pc = row_shfl(pre, c < nk ? c : nk - 1);
          if (c < nk && pc <= l) k = c;
        }
        const uint32_t pk = row_shfl(pre, k), bk = row_shfl(b, k), sk = row_shfl(seg2Base, k);
        if (l < chunk) {
          const uint32_t j = v.W3.segByMD[bk + (l - pk)] - sk;
          list[run + l] = make_uint2(m1, (a0 + k) | (j << 16));
        }
      }
      run += chunk;
    }
  }
}

// stage 2: the triplet selection; results parked at the list position
// four waves per SIMD: at most 128 VGPRs (no spills; the compiler's own allocation can exceed 128)
__global__ __launch_bounds__(256, 4) void k_t3_select(LST_VIEWS) {
  LST_BIND;
  if (!t3_flat_mode(v)) return;
  const uint32_t n = v.R.scalars[kT3List];
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    bool pass = false;
    if (i < n) {
      const uint32_t m1 = v.W3.list[2 * i], ab = v.W3.list[2 * i + 1];
      const uint32_t inSeg = v.segModuleIndices[m1] + (ab & 0xFFFFu);
      const uint32_t outSeg = v.segModuleIndices[v.S.outerLowerModuleIndices[inSeg]] + (ab >> 16);
      float betaIn = 0, radius = 0, gx = 0, gy = 0;
      pass = t3_select(v, m1, inSeg, outSeg, betaIn, radius, gx, gy);
      if (pass) {  // results at the list position: no slot taken with an atomic
        uint32_t* w = &v.W3.results[kT3ResultWords * (size_t)i];
        w[0] = inSeg; w[1] = outSeg;
        w[2] = __float_as_uint(betaIn); w[3] = __float_as_uint(radius);
        w[4] = __float_as_uint(gx); w[5] = __float_as_uint(gy);
      }
    }
    store_flag_word(v.FS, i0 + (threadIdx.x & ~63u), n, pass);
  }
}

// stage 3: the list is in serial order (modules, inner segments and each
// mini-doublet's segments in order), so per module the passing entries take
// the slots in list order: a pass-flag scan over the list positions
__global__ __launch_bounds__(1024) void k_t3_scan_blocks(LST_VIEWS) {
  LST_BIND;
  if (!t3_flat_done(v)) return;
  scan_mask_blocks(v.FS, v.R.scalars[kT3List]);
}
__global__ __launch_bounds__(1024) void k_t3_scan_sums(LST_VIEWS) {
  LST_BIND;
  if (!t3_flat_done(v)) return;
  scan_flag_sums(v.R.scalars[kT3List], v.FS.sums);
}

// passing entries take their slots in list order; then per-module counts
__global__ __launch_bounds__(256) void k_t3_place(LST_VIEWS) {
  LST_BIND;
  if (!t3_flat_done(v)) {
    // no per-module builder follows in a batch launch set: the stage is left empty (the event's overflow
    // flag sends it to the per-event path)
    if (v.noFallback)
      for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < v.M.nLowerModules; m += gridDim.x * blockDim.x) {
        v.R.totOccupancyTriplets[m] = 0;
        v.R.nTriplets[m] = 0;
        t3_range(v, m, 0);
      }
    return;
  }
  const uint32_t total = v.R.scalars[kT3List], nL = v.M.nLowerModules;
  for_each_passing(v.FS, total, nL, [&](uint32_t k, uint32_t rankK) {
    const uint32_t m1 = v.W3.list[2 * k];
    const uint32_t* w = &v.W3.results[kT3ResultWords * (size_t)k];
    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5];
    const uint32_t rank = rankK - flag_rank(v.FS, v.W3.survOff[m1], total);
    if ((int)rank >= v.R.tripletModuleOccupancy[m1]) return;
    t3_store(v, v.R.tripletModuleIndices[m1] + rank, m1, w0, w1, __uint_as_float(w2), __uint_as_float(w3),
             __uint_as_float(w4), __uint_as_float(w5));
  }, [&](uint32_t m) {  // then per module its counts and range
    const uint32_t count = flag_rank(v.FS, v.W3.survOff[m + 1], total) - flag_rank(v.FS, v.W3.survOff[m], total);
    v.R.totOccupancyTriplets[m] = count;
    v.R.nTriplets[m] = min(count, (unsigned)max(v.R.tripletModuleOccupancy[m], 0));
    t3_range(v, m, v.R.nTriplets[m]);
  });
}

// Per-module builder, used when an event exceeds the flattened path's lists:
// one wave per lower module walks its candidates in order.
__global__ __launch_bounds__(kWave) void k_triplets(LST_VIEWS) {
  LST_BIND;
  __shared__ uint32_t pre[1025];
  if (t3_flat_done(v)) return;
  const uint32_t m1 = blockIdx.x;
  const unsigned lane = threadIdx.x;
  const uint32_t nIn = v.R.nSegments[m1];
  const int cap = v.R.tripletModuleOccupancy[m1];
  unsigned count = 0;
  if (nIn != 0 && v.M.nConnectedModules[m1] != 0 && nIn <= 1024) {
    const uint32_t segBase = v.segModuleIndices[m1];
    uint32_t run = 0;
    for (uint32_t a0 = 0; a0 < nIn; a0 += kWave) {
      const uint32_t a = a0 + lane;
      const uint32_t cnt = a < nIn ? v.R.nSegments[v.S.outerLowerModuleIndices[segBase + a]] : 0;
      const uint32_t incl = wave_inclusive_scan(cnt);
      if (a < nIn) pre[a + 1] = run + incl;
      run += wave_last(incl);
    }
    if (lane == 0) pre[0] = 0;
    __syncthreads();
    const uint32_t total = run;
    const uint32_t t3Base = v.R.tripletModuleIndices[m1];
    for (uint32_t c0 = 0; c0 < total; c0 += kWave) {
      const uint32_t L = c0 + lane;
      bool pass = false;
      float betaIn = 0, radius = 0, gx = 0, gy = 0;
      uint32_t inSeg = 0, outSeg = 0;
      if (L < total) {
        uint32_t lo = 0, hi = nIn;  // largest a with pre[a] <= L
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (pre[mid] <= L) lo = mid;
          else hi = mid;
        }
        inSeg = segBase + lo;
        outSeg = v.segModuleIndices[v.S.outerLowerModuleIndices[inSeg]] + (L - pre[lo]);
        pass = t3_select(v, m1, inSeg, outSeg, betaIn, radius, gx, gy);
      }
      unsigned n;
      const unsigned rank = count + wave_rank(pass, n);
      if (pass && (int)rank < cap) t3_store(v, t3Base + rank, m1, inSeg, outSeg, betaIn, radius, gx, gy);
      count += n;
    }
  }
  if (lane == 0) {
    v.R.totOccupancyTriplets[m1] = count;
    v.R.nTriplets[m1] = min(count, (unsigned)max(cap, 0));
    t3_range(v, m1, v.R.nTriplets[m1]);
  }
}


inline unsigned grid_for(uint32_t n, unsigned block) {
  unsigned g = (n + block - 1) / block;
  return g == 0 ? 1 : (g > 4096 ? 4096 : g);
}

// one row of blocks per buffer; 4-byte words, then the 0-3 tail bytes
__global__ __launch_bounds__(256) void k_copy_multi(CopyList c) {
  const int e = blockIdx.y;
  if (e >= c.n) return;
  const uint32_t words = c.bytes[e] >> 2;
  const uint32_t* src = static_cast<const uint32_t*>(c.src[e]);
  uint32_t* dst = static_cast<uint32_t*>(c.dst[e]);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) dst[i] = src[i];
  if (blockIdx.x == 0 && threadIdx.x < (c.bytes[e] & 3u))
    static_cast<char*>(c.dst[e])[4 * words + threadIdx.x] = static_cast<const char*>(c.src[e])[4 * words + threadIdx.x];
}

}  // namespace

__global__ void k_set_scalar(uint32_t* p, uint32_t value) { *p = value; }

__global__ __launch_bounds__(256) void k_stage_in(LST_VIEWS, const InputDesc* __restrict__ ds) {
  LST_BIND;
  const InputDesc* d = ds + blockIdx.z;
  const uint32_t nh = d->nHits, np = d->nPLS;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    v.R.scalars[kEvNHits] = nh;
    v.R.scalars[kEvNPLS] = np;
  }
  const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (uint32_t i = t0; i < nh; i += stride) {
    v.H.xs[i] = d->x[i]; v.H.ys[i] = d->y[i]; v.H.zs[i] = d->z[i];
    v.H.detid[i] = d->det[i]; v.H.idxs[i] = d->idx[i];
  }
  for (uint32_t i = t0; i < np; i += stride) {
#pragma unroll
    for (int k = 0; k < 4; k++) v.plsInHitIdx[k][i] = d->hitIdx[k][i];
    v.plsInDPhiChange[i] = d->dPhiChange[i];
    v.P.ptIn[i] = d->ptIn[i]; v.P.ptErr[i] = d->ptErr[i];
    v.P.px[i] = d->px[i]; v.P.py[i] = d->py[i]; v.P.pz[i] = d->pz[i];
    v.P.eta[i] = d->eta[i]; v.P.etaErr[i] = d->etaErr[i]; v.P.phi[i] = d->phi[i];
    v.P.charge[i] = d->charge[i]; v.P.seedIdx[i] = d->seedIdx[i]; v.P.superbin[i] = d->superbin[i];
    v.P.pixelType[i] = d->pixelType[i]; v.P.isQuad[i] = d->isQuad[i];
  }
}

void launch_stage_in(const Launch& L, const InputDesc* d) {
  L.tune_begin(kTnStageIn);
  hipLaunchKernelGGL(k_stage_in, dim3(L.gx(kTnStageIn, 512), 1, L.nEv), dim3(256), 0, L.s, L.dv, d);
  L.tune_end(kTnStageIn);
}

void launch_set_scalar(uint32_t* p, uint32_t value, hipStream_t s) {
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, s, p, value);
}

void launch_copy_multi(const CopyList& c, hipStream_t s) {
  if (c.n <= 0) return;
  uint32_t maxBytes = 0;
  for (int e = 0; e < c.n; e++) maxBytes = c.bytes[e] > maxBytes ? c.bytes[e] : maxBytes;
  unsigned gx = (maxBytes / 4 + 255) / 256;
  gx = gx == 0 ? 1 : (gx > 64 ? 64 : gx);
  hipLaunchKernelGGL(k_copy_multi, dim3(gx, c.n), dim3(256), 0, s, c);
}

void launch_hits(const Launch& L) {
  const EventView& v = L.v;
  LST_GO_T(L, kTnEventInit, k_event_init, grid_for(2 * v.M.nModules, 256), 256);
  L.t.begin(L.s, kTmHits);
  LST_GO_T(L, kTnHits, k_hits, 1024, 256);
  L.t.end(L.s, kTmHits);
  LST_GO_T(L, kTnModuleRanges, k_module_ranges, grid_for(v.M.nModules, 256), 256);
  LST_GO_T(L, kTnHitRangesFix, k_hit_ranges_fix, grid_for(v.M.nModules, 256), 256);
}

void launch_pixel_segments(const Launch& L) {
  L.t.begin(L.s, kTmPixelSegments);
  LST_GO_T(L, kTnPixelSegments, k_pixel_segments, 196, 256);
  L.t.end(L.s, kTmPixelSegments);
}

void launch_mini_doublets(const Launch& L) {
  const EventView& v = L.v;
  LST_GO(L, k_md_cand_scan, 1, 1, 1024);
  LST_GO_T(L, kTnMdCandFill, k_md_cand_fill, 1024, 256);
  L.t.begin(L.s, kTmMiniDoublets);
  LST_GO_T(L, kTnMdEval, k_md_eval, 2048, 256);
  L.t.end(L.s, kTmMiniDoublets);
  LST_GO_T(L, kTnMdScanBlocks, k_md_scan_blocks, kScanGrid, 1024);
  LST_GO(L, k_md_scan_sums, 1, 1, 1024);
  LST_GO_T(L, kTnMdPlace, k_md_place, 2048, 256);
  if (L.fallbacks) LST_GO(L, k_mini_doublets, v.M.nLowerModules, 1, kBuildBlock);
}

void launch_segments(const Launch& L) {
  const EventView& v = L.v;
  LST_GO_T(L, kTnLsCandCount, k_ls_cand_count, 1024, 256);
  LST_GO(L, k_ls_cand_scan, 1, 1, 1024);
  LST_GO_T(L, kTnLsCandFill, k_ls_cand_fill, kModuleWaveBlocks, 256);
  L.t.begin(L.s, kTmSegments);
  LST_GO_T(L, kTnLsEval, k_ls_eval, 4096, 256);
  L.t.end(L.s, kTmSegments);
  LST_GO_T(L, kTnLsScanBlocks, k_ls_scan_blocks, kScanGrid, 1024);
  LST_GO(L, k_ls_scan_sums, 1, 1, 1024);
  LST_GO_T(L, kTnLsPlace, k_ls_place, 2048, 256);
  if (L.fallbacks) LST_GO(L, k_segments, v.M.nLowerModules, 1, kBuildBlock);
}

void launch_triplets(const Launch& L) {
  const EventView& v = L.v;
  LST_GO(L, k_triplet_offsets, 1, 1, 1024);
  LST_GO_T(L, kTnMdSegLists, k_md_seg_lists, (v.M.nLowerModules + 15) / 16 ? (v.M.nLowerModules + 15) / 16 : 1, 256);
  LST_GO_T(L, kTnT3CandCount, k_t3_cand_count, kModuleWaveBlocks, 256);
  LST_GO(L, k_t3_cand_scan, 1, 1, 1024);
  L.t.begin(L.s, kTmT3Filter);
  LST_GO_T(L, kTnT3Enum, k_t3_enum, kModuleWaveBlocks, 256);
  L.t.end(L.s, kTmT3Filter);
  L.t.begin(L.s, kTmT3Select);
  LST_GO_T(L, kTnT3Select, k_t3_select, 2048, 256);
  L.t.end(L.s, kTmT3Select);
  LST_GO_T(L, kTnT3ScanBlocks, k_t3_scan_blocks, kScanGrid, 1024);
  LST_GO(L, k_t3_scan_sums, 1, 1, 1024);
  LST_GO_T(L, kTnT3Place, k_t3_place, 2048, 256);
  if (L.fallbacks) LST_GO(L, k_triplets, v.M.nLowerModules, 1, kWave);
}

}  // namespace lst
