// Quintuplet stage: eligible-module scan, builder, after-build duplicate
// removal and ranges (SDL/Event.cc:903-991).
//
// Candidates (inner T3 a of an eligible module, outer T3 b on the inner T3's
// third module) are numbered in the reference's serial order (module, then a,
// then b; SDL/Quintuplet.h:2988-3115).  The flattened builder spreads them over
// the whole chip: a shared-MD filter and the two T4-style tests compact the
// survivors into lists (order within a list does not matter), the final
// selection parks each passing candidate's results, and one work-group per
// module then gives passing candidates their slots in candidate order, so T5
// slots equal the reference CPU slots and the per-module capacity keeps exactly
// the first `cap` passing candidates.  Events with more than kT5CandMax
// candidates take the per-module builder instead (same order, same results).
#include "../../include/lst_amd.h"
#include "algo_t5.h"
#include "flat_scan.h"
#include "stages.h"
#include "wave.h"

namespace lst {
namespace {

constexpr int kT5Block = 256;
constexpr int kMaxInnerT3 = 1024;   // > the largest T3 module capacity (755)
constexpr int kMaxT5PerModule = 416;  // >= the largest T5 module capacity (414)

// addQuintupletRangesToEventExplicit (SDL/Quintuplet.h:3215-3235), written
// where a module's T5 count is final; every range starts at -1
__device__ __forceinline__ void t5_range(const EventView& v, uint32_t m, uint32_t n) {
  const int base = v.R.quintupletModuleIndices[m];
  if (n != 0 && base != -1) {
    v.R.quintupletRanges[2 * m] = base;
    v.R.quintupletRanges[2 * m + 1] = base + (int)n - 1;
  }
}

// createEligibleModulesListForQuintupletsGPU (SDL/Quintuplet.h:3117-3213):
// seed modules holding a triplet, their capacities and the exclusive scan of
// the capacities in module order.
__global__ __launch_bounds__(1024) void k_t5_eligible(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[34];
  const uint32_t nL = v.M.nLowerModules;
  // the capacities and the eligibility flags scanned together; both loads
  // issued with no branch, and the puts use the scanned values only
  const UVec<2> tot = scan_wide_n<kScanPer, 2>(
      nL,
      [&](uint32_t m) {
        const int occ = v.t5OccIfAny[m];
        const auto nt = v.R.nTriplets[m];
        const bool e = (occ >= 0) & (nt != 0);
        return UVec<2>{{e ? (unsigned)occ : 0u, e ? 1u : 0u}};
      },
      [&](uint32_t m, const UVec<2>& p, const UVec<2>& x) {
        v.R.quintupletModuleIndices[m] = x.v[1] ? (int)p.v[0] : -1;
        v.R.quintupletModuleOccupancy[m] = (int)x.v[0];
        v.R.nQuintuplets[m] = 0;
        v.R.totOccupancyQuintuplets[m] = 0;
        if (x.v[1]) v.R.indicesOfEligibleT5Modules[p.v[1]] = (uint16_t)m;
      },
      lds);
  const unsigned totOcc = tot.v[0], totElig = tot.v[1];
  if (threadIdx.x == 0) {
    v.R.scalars[kNEligibleT5] = totElig;
    v.R.scalars[kNTotalQuints] = totOcc;
  }
}

// brute-force radius range when any of the three points is an endcap 2S
// mini-doublet; with three plain points the 27 evaluations coincide
__device__ __forceinline__ void radius_range_of(const ModLite& ma, const MDPoint& a, const ModLite& mb,
                                                const MDPoint& b, const ModLite& mc, const MDPoint& c, float& rmin,
                                                float& rmax) {
  const bool ea = is_endcap_2s(ma), eb = is_endcap_2s(mb), ec = is_endcap_2s(mc);
  if (!(ea || eb || ec)) {
    float g, f;
    const float r = circle_radius3(a.x, a.y, b.x, b.y, c.x, c.y, g, f);
    rmin = m_min(r, kSdlInf);
    rmax = m_max(r, 0.f);
    return;
  }
  const float xa[3] = {a.x, ea ? a.lx : a.x, ea ? a.hx : a.x}, ya[3] = {a.y, ea ? a.ly : a.y, ea ? a.hy : a.y};
  const float xb[3] = {b.x, eb ? b.lx : b.x, eb ? b.hx : b.x}, yb[3] = {b.y, eb ? b.ly : b.y, eb ? b.hy : b.y};
  const float xc[3] = {c.x, ec ? c.lx : c.x, ec ? c.hx : c.x}, yc[3] = {c.y, ec ? c.ly : c.y, ec ? c.hy : c.y};
  radius_range(xa, ya, xb, yb, xc, yc, rmin, rmax);
}

// subdetector pattern of the five modules -> radius-matching category
__device__ __forceinline__ int radii_category(const ModLite* m) {
  const bool B0 = m[0].subdet == kBarrel, B1 = m[1].subdet == kBarrel, B2 = m[2].subdet == kBarrel,
             B3 = m[3].subdet == kBarrel, B4 = m[4].subdet == kBarrel;
  const bool E1 = m[1].subdet == kEndcap, E2 = m[2].subdet == kEndcap, E3 = m[3].subdet == kEndcap,
             E4 = m[4].subdet == kEndcap;
  if (B0 && B1 && B2 && B3 && B4) return 0;
  if (B0 && B1 && B2 && B3 && E4) return 1;
  if (B0 && B1 && B2 && E3 && E4) return m[0].layer == 1 ? 2 : (m[0].layer == 2 ? 3 : 4);
  if (B0 && B1 && E2 && E3 && E4) return 5;
  if (B0 && E1 && E2 && E3 && E4) return 6;
  return 7;
}

// runQuintupletDefaultAlgo, plain chi2 mode (SDL/Quintuplet.h:2508-2986), in
// three parts so that the flattened builder can compact survivors between them:
// the shared mini-doublet check, the two T4-style tests, and the rest.
__device__ __forceinline__ bool t5_shared_md(const EventView& v, uint32_t innerT3, uint32_t outerT3) {
  const uint32_t seg2 = v.T.segmentIndices[2 * innerT3 + 1], seg3 = v.T.segmentIndices[2 * outerT3];
  return v.S.mdIndices[2 * seg2 + 1] == v.S.mdIndices[2 * seg3];
}

struct T5Cand {  // the five modules and mini-doublets of a candidate
  uint16_t mods[5];
  uint32_t md[5];
  uint32_t seg1, seg3, seg4;
};
__device__ __forceinline__ T5Cand t5_cand(const EventView& v, uint32_t innerT3, uint32_t outerT3) {
  T5Cand c;  // the first two chunks of both triplet records, then the first chunk of the four segments
  const uint4 i0 = t3_chunk(v.T, innerT3, 0), i1 = t3_chunk(v.T, innerT3, 1);
  const uint4 o0 = t3_chunk(v.T, outerT3, 0), o1 = t3_chunk(v.T, outerT3, 1);
  c.seg1 = i0.x;
  const uint32_t seg2 = i0.y;
  c.seg3 = o0.x;
  c.seg4 = o0.y;
  c.mods[0] = (uint16_t)i0.z;
  c.mods[1] = (uint16_t)i0.w;
  c.mods[2] = (uint16_t)i1.x;
  c.mods[3] = (uint16_t)o0.w;
  c.mods[4] = (uint16_t)o1.x;
  const uint4 s1 = seg_chunk(v.S, c.seg1, 0), s2 = seg_chunk(v.S, seg2, 0);
  const uint4 s3 = seg_chunk(v.S, c.seg3, 0), s4 = seg_chunk(v.S, c.seg4, 0);
  c.md[0] = s1.x;
  c.md[1] = s2.x;
  c.md[2] = s2.y;
  c.md[3] = s3.y;
  c.md[4] = s4.y;
  return c;
}

__device__ __forceinline__ bool t5_t4_tests(const EventView& v, const T5Cand& c) {
  ModLite m[5];
  MDPoint p[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    m[k] = load_lite(v.M, c.mods[k]);
    p[k] = load_point(v.D, c.md[k]);
  }
  const SegAlpha a1 = load_alpha(v.S, c.seg1);
  {
    T4Args q{m[0], m[1], m[2], m[3], p[0], p[1], p[2], p[3], a1, load_alpha(v.S, c.seg3), md_slope(v.D, c.md[2])};
    if (!t4_select(q, v.ptCut)) return false;
  }
  T4Args q{m[0], m[1], m[3], m[4], p[0], p[1], p[3], p[4], a1, load_alpha(v.S, c.seg4), md_slope(v.D, c.md[3])};
  return t4_select(q, v.ptCut);
}

__device__ __forceinline__ bool t5_rest(const EventView& v, const T5Cand& c, uint32_t innerT3, uint32_t outerT3,
                                        T5Result& r) {
  ModLite m[5];
  MDPoint p[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    m[k] = load_lite(v.M, c.mods[k]);
    p[k] = load_point(v.D, c.md[k]);
  }
  float g, f;
  r.outerR = __uint_as_float(t3_chunk(v.T, outerT3, 3).w);  // circle radius
  r.bridgeR = circle_radius3(p[1].x, p[1].y, p[2].x, p[2].y, p[3].x, p[3].y, g, f);
  const uint4 i3 = t3_chunk(v.T, innerT3, 3), i4 = t3_chunk(v.T, innerT3, 4);  // radius; centre x, y
  r.innerR = __uint_as_float(i3.w);
  g = __uint_as_float(i4.x);
  f = __uint_as_float(i4.y);
  const float innerPt = 2 * k2Rinv1GeVf * r.innerR;
  r.tight = false;
  if (!t5_rz_chi2(m, p, innerPt, r.innerR, g, f, r.rzChi2, r.tight)) return false;
  if (!(r.innerR >= 0.95f * v.ptCut / (2.f * k2Rinv1GeVf))) return false;
  const int cat = radii_category(m);
  float iMin = 0, iMax = 0, bMin = 0, bMax = 0;
  if (cat >= 6) radius_range_of(m[0], p[0], m[1], p[1], m[2], p[2], iMin, iMax);
  if (cat >= 2) radius_range_of(m[1], p[1], m[2], p[2], m[3], p[3], bMin, bMax);
  if (!radii_compatible(cat, r.innerR, r.bridgeR, iMin, iMax, bMin, bMax)) return false;
  float x[5], y[5], d1[5], d2[5];
  bool flat[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    x[k] = p[k].x;
    y[k] = p[k].y;
  }
  regression_sigmas(m, true, d1, d2, flat);
  r.regR = regression_fit(v.M, c.mods, x, y, d1, d2, flat, r.regG, r.regF, r.chi2);
  // r-phi chi2 cut of the plain mode (USE_RPHICHI2, SDL/Quintuplet.h:2943-2956);
  // the DNN mode applies the T5 DNN instead (k_t5_dnn / t5dnn_score)
  if (!v.t5Dnn && r.regR < 5.0f / (2.f * k2Rinv1GeVf) &&
      !t5_rphi_chi2_ok(m[0].sdl, m[1].sdl, m[2].sdl, m[3].sdl, m[4].sdl, r.chi2))
    return false;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const float2 o = md_outer_xy(v.D, c.md[k]);
    x[k] = o.x;
    y[k] = o.y;
  }
  regression_sigmas(m, false, d1, d2, flat);
  r.nonAnchorChi2 = circle_chi2(v.M, c.mods, x, y, d1, d2, flat, r.regG, r.regF, r.regR);
  return true;
}

// The flattened builder gives each candidate two lanes (an even lane and the
// odd one after it), which evaluate the two halves of the tests side by side;
// the candidate passes when both halves pass.  The per-module builder runs
// the same tests in one lane (t5_t4_tests, t5_rest).

// half `part` of t5_t4_tests: part 0 the T4 of modules 0-3, part 1 the T4 of
// modules 0, 1, 3, 4
__device__ __forceinline__ bool t5_t4_half(const EventView& v, const T5Cand& c, uint32_t part) {
  const uint16_t mc = part ? c.mods[3] : c.mods[2], md = part ? c.mods[4] : c.mods[3];
  const uint32_t pc = part ? c.md[3] : c.md[2], pd = part ? c.md[4] : c.md[3];
  const T4Args q{load_lite(v.M, c.mods[0]), load_lite(v.M, c.mods[1]), load_lite(v.M, mc), load_lite(v.M, md),
                 load_point(v.D, c.md[0]),  load_point(v.D, c.md[1]),  load_point(v.D, pc),  load_point(v.D, pd),
                 load_alpha(v.S, c.seg1),   load_alpha(v.S, part ? c.seg4 : c.seg3), md_slope(v.D, pc)};
  return t4_select(q, v.ptCut);
}

// half `part` of t5_rest: part 0 the radii, the r-z chi2 and the radius
// compatibility (filling innerR, bridgeR, outerR, rzChi2, tight); part 1 the
// regression, the r-phi chi2 cut and the non-anchor chi2 (filling regR, regG,
// regF, chi2, nonAnchorChi2)
__device__ __forceinline__ bool t5_rest_half(const EventView& v, const T5Cand& c, uint32_t innerT3,
                                             uint32_t outerT3, uint32_t part, T5Result& r) {
  ModLite m[5];
  MDPoint p[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    m[k] = load_lite(v.M, c.mods[k]);
    p[k] = load_point(v.D, c.md[k]);
  }
  if (part == 0) {
    float g, f;
    r.outerR = __uint_as_float(t3_chunk(v.T, outerT3, 3).w);  // circle radius
    r.bridgeR = circle_radius3(p[1].x, p[1].y, p[2].x, p[2].y, p[3].x, p[3].y, g, f);
    const uint4 i3 = t3_chunk(v.T, innerT3, 3), i4 = t3_chunk(v.T, innerT3, 4);  // radius; centre x, y
    r.innerR = __uint_as_float(i3.w);
    g = __uint_as_float(i4.x);
    f = __uint_as_float(i4.y);
    const float innerPt = 2 * k2Rinv1GeVf * r.innerR;
    r.tight = false;
    if (!t5_rz_chi2(m, p, innerPt, r.innerR, g, f, r.rzChi2, r.tight)) return false;
    if (!(r.innerR >= 0.95f * v.ptCut / (2.f * k2Rinv1GeVf))) return false;
    const int cat = radii_category(m);
    float iMin = 0, iMax = 0, bMin = 0, bMax = 0;
    if (cat >= 6) radius_range_of(m[0], p[0], m[1], p[1], m[2], p[2], iMin, iMax);
    if (cat >= 2) radius_range_of(m[1], p[1], m[2], p[2], m[3], p[3], bMin, bMax);
    return radii_compatible(cat, r.innerR, r.bridgeR, iMin, iMax, bMin, bMax);
  }
  float x[5], y[5], d1[5], d2[5];
  bool flat[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    x[k] = p[k].x;
    y[k] = p[k].y;
  }
  regression_sigmas(m, true, d1, d2, flat);
  r.regR = regression_fit(v.M, c.mods, x, y, d1, d2, flat, r.regG, r.regF, r.chi2);
  if (!v.t5Dnn && r.regR < 5.0f / (2.f * k2Rinv1GeVf) &&
      !t5_rphi_chi2_ok(m[0].sdl, m[1].sdl, m[2].sdl, m[3].sdl, m[4].sdl, r.chi2))
    return false;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const float2 o = md_outer_xy(v.D, c.md[k]);
    x[k] = o.x;
    y[k] = o.y;
  }
  regression_sigmas(m, false, d1, d2, flat);
  r.nonAnchorChi2 = circle_chi2(v.M, c.mods, x, y, d1, d2, flat, r.regG, r.regF, r.regR);
  return true;
}

// ------------------------------------------------------------------ T5 DNN
// T5DNN::runInference (SDL/NeuralNetwork.h:21-165), the cut of the DNN mode
// (USE_T5_DNN, SDL/Quintuplet.h:2921-2941).  Parameters: the float tables of
// SDL/NeuralNetworkWeights.h, kept as data (row-major wgtT_k[in][out]).  Every
// lane evaluates its own candidate; the loops unroll, so the weights are
// uniform scalar loads and each output accumulates its inputs in index order
// (no FMA contraction), as the serial reference does.
__constant__ float cDnnBias0[32] = {
#include "../data/t5dnn/bias_0.inc"
};
__constant__ float cDnnW0[38 * 32] = {
#include "../data/t5dnn/wgtT_0.inc"
};
__constant__ float cDnnBias2[32] = {
#include "../data/t5dnn/bias_2.inc"
};
__constant__ float cDnnW2[32 * 32] = {
#include "../data/t5dnn/wgtT_2.inc"
};
__constant__ float cDnnBias4[1] = {
#include "../data/t5dnn/bias_4.inc"
};
__constant__ float cDnnW4[32] = {
#include "../data/t5dnn/wgtT_4.inc"
};
__constant__ float cDnnWP = 0.6177366f;    // working point, by default T5DNN::LSTWP2 (SDL/Constants.h:145)
constexpr float kMagneticField = 3.8112f;  // SDL/Constants.h:135

__device__ __forceinline__ float dnn_log10(float x) { return m_log(x) / 2.302585093f; }

// the 38 inputs (SDL/NeuralNetwork.h:36-114)
__device__ __forceinline__ void t5dnn_features(const EventView& v, const T5Cand& c, uint32_t innerT3, float innerR,
                                               float bridgeR, float outerR, float* in) {
  const int hitOf[6] = {0, 1, 2, 2, 3, 4};  // inner T3 hits 1-3, outer T3 hits 3-5
#pragma unroll
  for (int b = 0; b < 6; b++) {
    const uint32_t md = c.md[hitOf[b]];
    const uint16_t mod = c.mods[hitOf[b]];
    const float x = v.D.anchorX[md], y = v.D.anchorY[md];
    float* o = in + (b < 3 ? 1 + 5 * b : 17 + 5 * (b - 3));
    o[0] = v.D.anchorEta[md];
    o[1] = v.D.anchorPhi[md];
    o[2] = v.D.anchorZ[md];
    o[3] = m_sqrt(x * x + y * y);
    o[4] = float(v.M.layers[mod] + 6 * (v.M.subdets[mod] == kEndcap));
  }
  const int adj = v.M.layers[c.mods[0]] == 1 ? 1 : 0;
  const uint32_t mdEtaPhi = v.S.mdIndices[2 * v.T.segmentIndices[2 * innerT3 + adj]];
  in[0] = dnn_log10(2 * k2Rinv1GeVf * innerR);
  in[16] = dnn_log10(2 * k2Rinv1GeVf * outerR);
  in[32] = dnn_log10((innerR + outerR) * kMagneticField * 1.602f / (2 * 100 * 5.39f));
  in[33] = v.D.anchorEta[mdEtaPhi];
  in[34] = v.D.anchorPhi[mdEtaPhi];
  in[35] = dnn_log10(innerR);
  in[36] = dnn_log10(bridgeR);
  in[37] = dnn_log10(outerR);
}

// Linear(38,32) ReLU Linear(32,32) ReLU Linear(32,1) sigmoid (SDL/NeuralNetwork.h:116-162)
__device__ __forceinline__ float t5dnn_forward(const float* in) {
  float h1[32];
#pragma unroll
  for (int o = 0; o < 32; o++) {
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < 38; i++) acc += in[i] * cDnnW0[32 * i + o];
    acc += cDnnBias0[o];
    h1[o] = acc > 0.f ? acc : 0.f;
  }
  float out = 0.f;
#pragma unroll
  for (int o = 0; o < 32; o++) {
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < 32; i++) acc += h1[i] * cDnnW2[32 * i + o];
    acc += cDnnBias2[o];
    const float h2 = acc > 0.f ? acc : 0.f;
    out += h2 * cDnnW4[o];
  }
  out += cDnnBias4[0];
  return m_exp(out) / (m_exp(out) + 1);  // NaN once exp overflows, as in the reference
}

__device__ __forceinline__ float t5dnn_score(const EventView& v, const T5Cand& c, uint32_t innerT3, float innerR,
                                             float bridgeR, float outerR) {
  float in[38];
  t5dnn_features(v, c, innerT3, innerR, bridgeR, outerR, in);
  return t5dnn_forward(in);
}

__device__ bool t5_candidate(const EventView& v, const uint16_t* mods, uint32_t innerT3, uint32_t outerT3,
                             uint32_t* md, T5Result& r) {
  if (!t5_shared_md(v, innerT3, outerT3)) return false;
  const T5Cand c = t5_cand(v, innerT3, outerT3);
#pragma unroll
  for (int k = 0; k < 5; k++) md[k] = c.md[k];
  if (!t5_t4_tests(v, c)) return false;
  return t5_rest(v, c, innerT3, outerT3, r);
}

// addQuintupletToMemory (SDL/Quintuplet.h:150-215) plus the T3 flags
__device__ __forceinline__ void t5_store(const EventView& v, uint32_t q, uint32_t innerT3, uint32_t outerT3,
                                         const T5Result& r, int layer) {
  // the two T3 records' first four chunks (segments, modules, logical layers, hits), one load each
  uint4 a[4], b[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    a[k] = t3_chunk(v.T, innerT3, k);
    b[k] = t3_chunk(v.T, outerT3, k);
  }
  const uint32_t ai[16] = {a[0].x, a[0].y, a[0].z, a[0].w, a[1].x, a[1].y, a[1].z, a[1].w,
                           a[2].x, a[2].y, a[2].z, a[2].w, a[3].x, a[3].y, a[3].z, a[3].w};
  const uint32_t bi[16] = {b[0].x, b[0].y, b[0].z, b[0].w, b[1].x, b[1].y, b[1].z, b[1].w,
                           b[2].x, b[2].y, b[2].z, b[2].w, b[3].x, b[3].y, b[3].z, b[3].w};
  constexpr uint32_t kMod = t3_word(kT3F_lowerModuleIndices), kLL = t3_word(kT3F_logicalLayers),
                     kHit = t3_word(kT3F_hitIndices);
  static_assert(kHit + 6 <= 16, "the hits are in the first four chunks");
  const int adj = layer == 1 ? 1 : 0;
  const uint32_t mdAdj = seg_chunk(v.S, adj ? a[0].y : a[0].x, 0).x;  // first MD of the inner or middle segment
  const uint4 mdA = md_chunk(v.D, mdAdj, 2), mdE = md_chunk(v.D, mdAdj, 4);  // anchor phi, anchor eta
  v.Q.tripletIndices[2 * q] = innerT3;
  v.Q.tripletIndices[2 * q + 1] = outerT3;
  v.Q.lowerModuleIndices[5 * q + 0] = (uint16_t)ai[kMod];
  v.Q.lowerModuleIndices[5 * q + 1] = (uint16_t)ai[kMod + 1];
  v.Q.lowerModuleIndices[5 * q + 2] = (uint16_t)ai[kMod + 2];
  v.Q.lowerModuleIndices[5 * q + 3] = (uint16_t)bi[kMod + 1];
  v.Q.lowerModuleIndices[5 * q + 4] = (uint16_t)bi[kMod + 2];
  v.Q.innerRadius[q] = r.innerR;
  v.Q.outerRadius[q] = r.outerR;
  v.Q.bridgeRadius[q] = r.bridgeR;
  v.Q.pt[q] = (r.innerR + r.outerR) * 3.8f * 1.602f / (2 * 100 * 5.39f);
  v.Q.eta[q] = __uint_as_float(mdE.w);
  v.Q.phi[q] = __uint_as_float(mdA.y);
  v.Q.score_rphisum[q] = r.chi2 + r.nonAnchorChi2;
  v.Q.layer[q] = (uint8_t)layer;
  v.Q.isDup[q] = 0;
  v.Q.TightCutFlag[q] = r.tight;
  v.Q.partOfPT5[q] = 0;
  v.Q.regressionRadius[q] = r.regR;
  v.Q.regressionG[q] = r.regG;
  v.Q.regressionF[q] = r.regF;
  v.Q.rzChiSquared[q] = r.rzChi2;
  v.Q.chiSquared[q] = r.chi2;
  v.Q.nonAnchorChiSquared[q] = r.nonAnchorChi2;
  v.Q.logicalLayers[5 * q + 0] = (uint8_t)ai[kLL];
  v.Q.logicalLayers[5 * q + 1] = (uint8_t)ai[kLL + 1];
  v.Q.logicalLayers[5 * q + 2] = (uint8_t)ai[kLL + 2];
  v.Q.logicalLayers[5 * q + 3] = (uint8_t)bi[kLL + 1];
  v.Q.logicalLayers[5 * q + 4] = (uint8_t)bi[kLL + 2];
#pragma unroll
  for (int j = 0; j < 6; j++) v.Q.hitIndices[10 * q + j] = ai[kHit + j];
#pragma unroll
  for (int j = 0; j < 4; j++) v.Q.hitIndices[10 * q + 6 + j] = bi[kHit + 2 + j];
  v.T.partOfT5[innerT3] = 1;
  v.T.partOfT5[outerT3] = 1;
}

// ------------------------------------------------------ flattened T5 builder
// Candidates of module b (eligible-list position) are [modOff[b], modOff[b+1]),
// inner T3 by inner T3; the shared-MD candidates that list A keeps are
// [survOff[b], survOff[b+1]) in the same order.

// The per-module kernels of the flattened builder run one wave per eligible
// module at a time and stride over the modules (a fixed grid of
// kModuleWaveBlocks work-groups), so small modules do not each dispatch one.
constexpr unsigned kModuleWaveBlocks = 512;  // x 4 waves
constexpr int kWave = 64;
__device__ __forceinline__ uint32_t wave_index() {
  return __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / 64);
}
__device__ __forceinline__ uint32_t wave_count() { return (gridDim.x * blockDim.x) / 64; }

// per eligible module: candidate count and per-inner-T3 exclusive prefix
// Per mini-doublet, the T3s whose first mini-doublet it is, in slot order (four
// lower modules per wave, one per row where they are small: counts, offsets, a scatter inside the module's own T3
// range, then each T3 placed at its rank within its short list).
__global__ __launch_bounds__(256) void k_md_t3_lists(LST_VIEWS) {
  LST_BIND;
  __shared__ __align__(16) WaveLists sLists[4];
  md_lists_all(
      v.M.nLowerModules,
      [&](uint32_t m, uint32_t& nMD, uint32_t& mdBase, uint32_t& n, uint32_t& base) {
        nMD = v.R.nMDs[m];
        n = v.R.nTriplets[m];
        mdBase = (uint32_t)v.mdModuleIndices[m];
        base = n ? (uint32_t)v.R.tripletModuleIndices[m] : 0u;
      },
      [&](uint32_t t) { return seg_chunk(v.S, t3_chunk(v.T, t, 0).x, 0).x; },  // the inner segment's inner MD
      v.W5.mdT3Off, v.W5.mdT3End, v.W5.t3ByMD, v.W5.t3Tmp, sLists[threadIdx.x / kWave], v.largePaths != 0);
}

// the outer mini-doublet of a T3, which an outer T3 must start at
__device__ __forceinline__ uint32_t t3_last_md(const EventView& v, uint32_t t3) {
  return seg_chunk(v.S, t3_chunk(v.T, t3, 0).y, 0).y;  // the outer segment's outer mini-doublet
}

// per eligible module, the candidates of its inner T3s (the T3s of their
// third modules) and their shared-MD candidates (the T3s listed under each
// inner T3's last mini-doublet).  One 16-lane group per eligible module (four
// per wave); the lanes stride over the inner T3s and the group sums.
__global__ __launch_bounds__(256) void k_t5_cand_count(LST_VIEWS) {
  LST_BIND;
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  const unsigned g = lane_id() / kRow, gl = lane_id() % kRow;
  const uint32_t groups = gridDim.x * (blockDim.x / kRow);
  // the loop runs on the wave's first group's index, so every lane of the wave takes the same iterations
  for (uint32_t e = (blockIdx.x * blockDim.x + threadIdx.x) / kRow; e - g < nElig; e += groups) {
    const bool has = e < nElig;
    uint32_t nIn = 0, t3Base = 0;
    if (has) {
      const uint16_t m1 = v.R.indicesOfEligibleT5Modules[e];
      const int layer = v.M.layers[m1];
      nIn = (layer == 1 || layer == 2) ? v.R.nTriplets[m1] : 0;
      t3Base = (uint32_t)v.R.tripletModuleIndices[m1];
    }
    uint32_t cnt = 0, sc = 0;
    for (uint32_t a = gl; a < nIn; a += kRow) {
      const uint32_t inner = t3Base + a, X = t3_last_md(v, inner);
      cnt += v.R.nTriplets[(uint16_t)t3_chunk(v.T, inner, 1).x];  // the inner T3's third module
      sc += v.W5.mdT3End[X] - v.W5.mdT3Off[X];
    }
    cnt = row_sum(cnt);
    sc = row_sum(sc);
    if (has && gl == 0) {
      v.W5.modOff[e] = cnt;
      v.W5.survOff[e] = sc;
    }
  }
}

// exclusive scan of the per-module counts (count = nEligible, read on the device)
__global__ __launch_bounds__(1024) void k_t5_cand_scan(LST_VIEWS) {
  LST_BIND;
  __shared__ unsigned lds[34];
  const uint32_t n = v.R.scalars[kNEligibleT5];
  const UVec<2> tot = scan_wide_n<kScanPer, 2>(
      n, [&](uint32_t i) { return UVec<2>{{v.W5.modOff[i], v.W5.survOff[i]}}; },
      [&](uint32_t i, const UVec<2>& p, const UVec<2>&) {
        v.W5.modOff[i] = p.v[0];
        v.W5.survOff[i] = p.v[1];
      },
      lds);
  const unsigned total = tot.v[0], surv = tot.v[1];
  if (threadIdx.x == 0) {
    v.W5.modOff[n] = total;
    v.W5.survOff[n] = surv;
    v.R.scalars[kT5CandTotal] = total;
    v.R.scalars[kT5ListA] = surv;
    v.R.scalars[kT5ListB] = 0;
    v.R.scalars[kT5Results] = 0;
  }
}

__device__ __forceinline__ bool t5_flat_mode(const EventView& v) {
  return !v.perModuleBuilders && v.R.scalars[kT5CandTotal] <= kT5CandMax;
}
// the flattened path completed: every candidate fitted and every result was parked
__device__ __forceinline__ bool t5_flat_done(const EventView& v) {
  return t5_flat_mode(v) && v.R.scalars[kT5Results] <= kT5ResultMax;
}


__device__ __forceinline__ void list_push(uint32_t* list, uint32_t* counter, uint32_t L, uint32_t a, uint32_t b) {
  // wave-aggregated append: one atomic per wave
  const uint64_t mask = __ballot(1);
  const unsigned leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader, 64);
  const uint32_t k = base + lanes_below(mask);
  list[3 * k] = L;
  list[3 * k + 1] = a;
  list[3 * k + 2] = b;
}

// stage 1: the candidates whose outer T3 starts at the inner T3's last
// mini-doublet, written at the positions the count pass gives (one wave per
// eligible module, its inner T3s 64 at a time in the lanes; the lanes then
// take the chunk's candidates in order, each finding its inner T3 by a binary
// search over the prefix of the counts, so consecutive lanes write
// consecutive entries).  With each mini-doublet's T3s in slot order, list A is
// in the reference's serial order (module, inner T3, outer T3).
__global__ __launch_bounds__(256) void k_t5_enum(LST_VIEWS) {
  LST_BIND;
  if (!t5_flat_mode(v)) return;
  const uint32_t nElig = v.R.scalars[kNEligibleT5];
  const unsigned lane = lane_id();
  for (uint32_t e = wave_index(); e < nElig; e += wave_count()) {
    const uint16_t m1 = v.R.indicesOfEligibleT5Modules[e];
    const int layer = v.M.layers[m1];
    const uint32_t nIn = (layer == 1 || layer == 2) ? v.R.nTriplets[m1] : 0;
    if (nIn == 0) continue;
    const uint32_t t3Base = (uint32_t)v.R.tripletModuleIndices[m1];
    uint32_t run = v.W5.survOff[e];  // list position of the chunk's first candidate
    for (uint32_t a0 = 0; a0 < nIn; a0 += kWave) {
      const uint32_t a = a0 + lane;
      uint32_t cnt = 0, b = 0;
      if (a < nIn) {
        const uint32_t X = t3_last_md(v, t3Base + a);
        b = v.W5.mdT3Off[X];
        cnt = v.W5.mdT3End[X] - b;
      }
      const uint32_t incl = wave_inclusive_scan(cnt), pre = incl - cnt;
      const uint32_t chunk = wave_last(incl);
      const uint32_t nk = nIn - a0 < (uint32_t)kWave ? nIn - a0 : (uint32_t)kWave;
      for (uint32_t l0 = 0; l0 < chunk; l0 += kWave) {
        const uint32_t l = l0 + lane;
        uint32_t k = 0;  // the largest k < nk with pre_k <= l (every lane takes part)
        for (uint32_t step = kWave / 2; step >= 1; step >>= 1) {
          const uint32_t c = k + step;
          const uint32_t pc = (uint32_t)__shfl((int)pre, (int)(c < nk ? c : nk - 1), kWave);
          if (c < nk && pc <= l) k = c;
        }
        const uint32_t pk = (uint32_t)__shfl((int)pre, (int)k, kWave);
        const uint32_t bk = (uint32_t)__shfl((int)b, (int)k, kWave);
        if (l < chunk) {
          const uint32_t kk = run + l;
          v.W5.listA[3 * kk] = e;
          v.W5.listA[3 * kk + 1] = t3Base + a0 + k;
          v.W5.listA[3 * kk + 2] = v.W5.t3ByMD[bk + (l - pk)];
        }
      }
      run += chunk;
    }
  }
}

// stage 2: the two T4-style tests
// three waves per SIMD (168 VGPRs instead of the compiler's 184, which allowed two; 24 bytes of
// spills per lane): 12 % faster in a 64-event launch set
__global__ __launch_bounds__(256, 3) void k_t5_filter_t4(LST_VIEWS) {
  LST_BIND;
  if (!t5_flat_mode(v)) return;
  const uint32_t n = v.R.scalars[kT5ListA];
  for (uint32_t g0 = blockIdx.x * blockDim.x; g0 < 2 * n; g0 += gridDim.x * blockDim.x) {
    const uint32_t g = g0 + threadIdx.x, i = g >> 1, part = g & 1;
    uint32_t inner = 0, outer = 0;
    bool pass = false;
    if (i < n) {
      inner = v.W5.listA[3 * i + 1];
      outer = v.W5.listA[3 * i + 2];
      pass = t5_t4_half(v, t5_cand(v, inner, outer), part);
    }
    const int other = __shfl_xor((int)pass, 1, kWave);  // every lane of the wave takes part
    const bool both = pass && other != 0;
    if (i < n && part == 0) {
      if (both) list_push(v.W5.listB, &v.R.scalars[kT5ListB], i, inner, outer);
      else v.W5.resIdx[i] = ~0u;
    }
  }
}

// stage 3: radii, chi2 cuts and the regression; results parked per candidate.
// The two halves run in different waves (so that neither waits for the
// other's branch): waves 2p and 2p + 1 of a work-group take halves 0 and 1 of
// the same 64 candidates, and the odd wave hands its results over in LDS.
__global__ __launch_bounds__(256) void k_t5_final(LST_VIEWS) {
  LST_BIND;
  __shared__ float sHalf[2][6][kWave];  // per wave pair: pass, regR, regG, regF, chi2, nonAnchorChi2
  if (!t5_flat_mode(v)) return;
  const uint32_t n = v.R.scalars[kT5ListB];
  const unsigned lane = threadIdx.x % kWave, wave = threadIdx.x / kWave, part = wave & 1, pair = wave >> 1;
  for (uint32_t c0 = blockIdx.x * 2 * kWave; c0 < n; c0 += gridDim.x * 2 * kWave) {
    const uint32_t i = c0 + pair * kWave + lane;
    uint32_t L = 0, inner = 0, outer = 0;
    bool pass = false;
    T5Result r{};
    if (i < n) {
      L = v.W5.listB[3 * i];
      inner = v.W5.listB[3 * i + 1];
      outer = v.W5.listB[3 * i + 2];
      pass = t5_rest_half(v, t5_cand(v, inner, outer), inner, outer, part, r);
    }
    if (part == 1) {
      float* h = &sHalf[pair][0][lane];
      h[0] = pass ? 1.f : 0.f;
      h[kWave] = r.regR; h[2 * kWave] = r.regG; h[3 * kWave] = r.regF;
      h[4 * kWave] = r.chi2; h[5 * kWave] = r.nonAnchorChi2;
    }
    __syncthreads();
    if (part == 0 && i < n) {
      const float* h = &sHalf[pair][0][lane];
      v.W5.resIdx[L] = ~0u;  // L: the list-A position; set below when the candidate passes
      if (pass && h[0] != 0.f) {
        r.regR = h[kWave]; r.regG = h[2 * kWave]; r.regF = h[3 * kWave];
        r.chi2 = h[4 * kWave]; r.nonAnchorChi2 = h[5 * kWave];
        // wave-aggregated slot: one atomic per wave
        const uint64_t mask = __ballot(1);
        const unsigned leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(&v.R.scalars[kT5Results], (uint32_t)__popcll(mask));
        const uint32_t k = __shfl(base, leader, 64) + lanes_below(mask);
        if (k < kT5ResultMax) {  // overflow: the per-module builder redoes the stage
          uint32_t* w = &v.W5.results[kT5ResultWords * k];
          w[0] = inner; w[1] = outer;
          w[2] = __float_as_uint(r.innerR); w[3] = __float_as_uint(r.bridgeR); w[4] = __float_as_uint(r.outerR);
          w[5] = __float_as_uint(r.regR); w[6] = __float_as_uint(r.regG); w[7] = __float_as_uint(r.regF);
          w[8] = __float_as_uint(r.rzChi2); w[9] = __float_as_uint(r.chi2);
          w[10] = __float_as_uint(r.nonAnchorChi2);
          w[11] = r.tight ? 1u : 0u;
          v.W5.resIdx[L] = k;
        }
      }
    }
    __syncthreads();  // the LDS halves are read before the next candidates overwrite them
  }
}

// stage 3b (DNN mode): the T5 DNN over the parked results; a result that
// fails gets its inner T3 index set to ~0 and is not placed.  The cut cannot
// change the tight flag of a candidate that passes it.
__global__ __launch_bounds__(256) void k_t5_dnn(LST_VIEWS) {
  LST_BIND;
  if (!v.t5Dnn || !t5_flat_mode(v)) return;
  const uint32_t nRes = v.R.scalars[kT5Results];
  const uint32_t n = nRes < kT5ResultMax ? nRes : kT5ResultMax;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    uint32_t* w = &v.W5.results[kT5ResultWords * k];
    const uint32_t inner = w[0], outer = w[1];
    float in[38];
    t5dnn_features(v, t5_cand(v, inner, outer), inner, __uint_as_float(w[2]), __uint_as_float(w[3]),
                   __uint_as_float(w[4]), in);
    const float score = t5dnn_forward(in);
    if (v.dnnDump.features && k < v.dnnDump.capacity) {  // training rows: every candidate at the DNN call site
      for (int a = 0; a < 38; a++) v.dnnDump.features[38 * (size_t)k + a] = in[a];
      v.dnnDump.score[k] = score;
      v.dnnDump.t3Pair[2 * k] = inner;
      v.dnnDump.t3Pair[2 * k + 1] = outer;
      for (int a = 0; a < 6; a++) v.dnnDump.hitIdxs[10 * (size_t)k + a] = v.H.idxs[v.T.hitIndices[6 * inner + a]];
      for (int a = 0; a < 4; a++)
        v.dnnDump.hitIdxs[10 * (size_t)k + 6 + a] = v.H.idxs[v.T.hitIndices[6 * outer + 2 + a]];
    }
    if (!(score > cDnnWP)) w[0] = ~0u;
  }
}

// stage 4: passing candidates take their slots in candidate order (pass-flag
// scan; a result the DNN rejected does not pass), then per-module counts
__device__ __forceinline__ bool t5_passed(const EventView& v, uint32_t k) {
  const uint32_t r = v.W5.resIdx[k];
  return r != ~0u && v.W5.results[kT5ResultWords * r] != ~0u;
}
__global__ __launch_bounds__(1024) void k_t5_scan_blocks(LST_VIEWS) {
  LST_BIND;
  if (!t5_flat_done(v)) return;
  scan_flag_blocks(v.FS, v.R.scalars[kT5ListA], [&](uint32_t k) { return t5_passed(v, k); });
}
__global__ __launch_bounds__(1024) void k_t5_scan_sums(LST_VIEWS) {
  LST_BIND;
  if (!t5_flat_done(v)) return;
  scan_flag_sums(v.R.scalars[kT5ListA], v.FS.sums);
}

// list A is in serial order (eligible modules, inner T3s and each
// mini-doublet's T3s in order), so per module the passing entries take the
// slots in list order
__global__ __launch_bounds__(256) void k_t5_place(LST_VIEWS) {
  LST_BIND;
  if (!t5_flat_done(v)) return;
  const uint32_t total = v.R.scalars[kT5ListA], nElig = v.R.scalars[kNEligibleT5];
  for_each_passing(v.FS, total, nElig, [&](uint32_t k, uint32_t rankK) {
    const uint32_t e = v.W5.listA[3 * k];
    const uint16_t m1 = v.R.indicesOfEligibleT5Modules[e];
    const uint32_t rank = rankK - flag_rank(v.FS, v.W5.survOff[e], total);
    if ((int)rank >= v.R.quintupletModuleOccupancy[m1]) return;
    const uint32_t* w = &v.W5.results[kT5ResultWords * v.W5.resIdx[k]];
    T5Result r;
    r.innerR = __uint_as_float(w[2]); r.bridgeR = __uint_as_float(w[3]); r.outerR = __uint_as_float(w[4]);
    r.regR = __uint_as_float(w[5]); r.regG = __uint_as_float(w[6]); r.regF = __uint_as_float(w[7]);
    r.rzChi2 = __uint_as_float(w[8]); r.chi2 = __uint_as_float(w[9]); r.nonAnchorChi2 = __uint_as_float(w[10]);
    r.tight = w[11] != 0;
    t5_store(v, (uint32_t)v.R.quintupletModuleIndices[m1] + rank, w[0], w[1], r, v.M.layers[m1]);
  }, [&](uint32_t e) {  // then per eligible module its counts and range
    const uint16_t m1 = v.R.indicesOfEligibleT5Modules[e];
    const uint32_t count = flag_rank(v.FS, v.W5.survOff[e + 1], total) - flag_rank(v.FS, v.W5.survOff[e], total);
    v.R.totOccupancyQuintuplets[m1] = count;
    v.R.nQuintuplets[m1] = min(count, (unsigned)max(v.R.quintupletModuleOccupancy[m1], 0));
    t5_range(v, m1, v.R.nQuintuplets[m1]);
  });
}

// Per-module builder, used when an event has more than kT5CandMax candidates
// or more than kT5ResultMax passing ones:
// one 256-lane work-group per eligible module walks its candidates in order.
__global__ __launch_bounds__(kT5Block) void k_quintuplets(LST_VIEWS) {
  LST_BIND;
  __shared__ uint32_t pre[kMaxInnerT3 + 1];
  __shared__ unsigned lds[kT5Block / 64 + 1];
  if (t5_flat_done(v)) return;
  if (blockIdx.x >= v.R.scalars[kNEligibleT5]) return;
  const uint16_t m1 = v.R.indicesOfEligibleT5Modules[blockIdx.x];
  const int layer = v.M.layers[m1];
  const int cap = v.R.quintupletModuleOccupancy[m1];
  const uint32_t nIn = v.R.nTriplets[m1];
  unsigned count = 0;
  if ((layer == 1 || layer == 2) && nIn <= (uint32_t)kMaxInnerT3) {
    const uint32_t t3Base = v.R.tripletModuleIndices[m1];
    uint32_t run = 0;
    for (uint32_t a0 = 0; a0 < nIn; a0 += kT5Block) {
      const uint32_t a = a0 + threadIdx.x;
      const uint32_t cnt = a < nIn ? v.R.nTriplets[v.T.lowerModuleIndices[3 * (t3Base + a) + 2]] : 0;
      unsigned tot;
      const unsigned ex = block_exclusive_scan<kT5Block>(cnt, lds, tot);
      if (a < nIn) pre[a + 1] = run + ex + cnt;
      run += tot;
    }
    if (threadIdx.x == 0) pre[0] = 0;
    __syncthreads();
    const uint32_t total = run;
    const uint32_t t5Base = (uint32_t)v.R.quintupletModuleIndices[m1];
    for (uint32_t c0 = 0; c0 < total; c0 += kT5Block) {
      const uint32_t L = c0 + threadIdx.x;
      bool pass = false;
      uint32_t innerT3 = 0, outerT3 = 0, md[5];
      uint16_t mods[5];
      T5Result r;
      if (L < total) {
        uint32_t lo = 0, hi = nIn;  // largest a with pre[a] <= L
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (pre[mid] <= L) lo = mid;
          else hi = mid;
        }
        innerT3 = t3Base + lo;
        mods[0] = m1;
        mods[1] = v.T.lowerModuleIndices[3 * innerT3 + 1];
        mods[2] = v.T.lowerModuleIndices[3 * innerT3 + 2];
        outerT3 = (uint32_t)v.R.tripletModuleIndices[mods[2]] + (L - pre[lo]);
        mods[3] = v.T.lowerModuleIndices[3 * outerT3 + 1];
        mods[4] = v.T.lowerModuleIndices[3 * outerT3 + 2];
        pass = t5_candidate(v, mods, innerT3, outerT3, md, r);
        if (pass && v.t5Dnn)
          pass = t5dnn_score(v, t5_cand(v, innerT3, outerT3), innerT3, r.innerR, r.bridgeR, r.outerR) > cDnnWP;
      }
      unsigned n;
      const unsigned rank = count + block_exclusive_scan<kT5Block>(pass ? 1u : 0u, lds, n);
      if (pass && (int)rank < cap) t5_store(v, t5Base + rank, innerT3, outerT3, r, layer);
      count += n;
    }
  }
  if (threadIdx.x == 0) {
    v.R.totOccupancyQuintuplets[m1] = count;
    v.R.nQuintuplets[m1] = min(count, (unsigned)max(cap, 0));
    t5_range(v, m1, v.R.nQuintuplets[m1]);
  }
}


}  // namespace

void launch_quintuplets(const Launch& L) {
  const EventView& v = L.v;
  const unsigned nBlocks = v.nT5SeedModules ? v.nT5SeedModules : 1;
  const unsigned flatGrid = 2048;  // grid-stride over candidates / survivors
  LST_GO(L, k_t5_eligible, 1, 1, 1024);
  LST_GO_T(L, kTnMdT3Lists, k_md_t3_lists, (v.M.nLowerModules + 15) / 16 ? (v.M.nLowerModules + 15) / 16 : 1, 256);
  LST_GO_T(L, kTnT5CandCount, k_t5_cand_count, kModuleWaveBlocks, 256);
  LST_GO(L, k_t5_cand_scan, 1, 1, 1024);
  LST_GO_T(L, kTnT5Enum, k_t5_enum, kModuleWaveBlocks, 256);
  LST_GO_T(L, kTnT5FilterT4, k_t5_filter_t4, flatGrid, 256);
  L.t.begin(L.s, kTmT5Final);
  LST_GO_T(L, kTnT5Final, k_t5_final, flatGrid, 256);
  L.t.end(L.s, kTmT5Final);
  if (v.t5Dnn) {
    L.t.begin(L.s, kTmT5Dnn);
    LST_GO_T(L, kTnT5Dnn, k_t5_dnn, flatGrid, 256);
    L.t.end(L.s, kTmT5Dnn);
  }
  LST_GO_T(L, kTnT5ScanBlocks, k_t5_scan_blocks, kScanGrid, 1024);
  LST_GO(L, k_t5_scan_sums, 1, 1, 1024);
  LST_GO_T(L, kTnT5Place, k_t5_place, 2048, 256);
  if (L.fallbacks) LST_GO(L, k_quintuplets, nBlocks, 1, kT5Block);
  launch_t5_list(L);
  launch_t5_dup_after_build(L);
}

}  // namespace lst
// process-wide T5 DNN parameters (packed: bias_0, wgtT_0, bias_2, wgtT_2, bias_4, wgtT_4)
extern "C" lst_status lst_t5dnn_set_weights(const float* params) {
  if (!params) return LST_E_INPUT;
  const float* p = params;
  if (hipMemcpyToSymbol(HIP_SYMBOL(lst::cDnnBias0), p, 32 * 4) != hipSuccess) return LST_E_DEVICE;
  p += 32;
  if (hipMemcpyToSymbol(HIP_SYMBOL(lst::cDnnW0), p, 38 * 32 * 4) != hipSuccess) return LST_E_DEVICE;
  p += 38 * 32;
  if (hipMemcpyToSymbol(HIP_SYMBOL(lst::cDnnBias2), p, 32 * 4) != hipSuccess) return LST_E_DEVICE;
  p += 32;
  if (hipMemcpyToSymbol(HIP_SYMBOL(lst::cDnnW2), p, 32 * 32 * 4) != hipSuccess) return LST_E_DEVICE;
  p += 32 * 32;
  if (hipMemcpyToSymbol(HIP_SYMBOL(lst::cDnnBias4), p, 4) != hipSuccess) return LST_E_DEVICE;
  p += 1;
  if (hipMemcpyToSymbol(HIP_SYMBOL(lst::cDnnW4), p, 32 * 4) != hipSuccess) return LST_E_DEVICE;
  return hipDeviceSynchronize() == hipSuccess ? LST_OK : LST_E_DEVICE;
}

extern "C" lst_status lst_t5dnn_set_working_point(float wp) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(lst::cDnnWP), &wp, sizeof(float)) != hipSuccess) return LST_E_DEVICE;
  return hipDeviceSynchronize() == hipSuccess ? LST_OK : LST_E_DEVICE;
}

extern "C" float lst_t5dnn_get_working_point(void) {
  float wp = 0.f;
  return hipMemcpyFromSymbol(&wp, HIP_SYMBOL(lst::cDnnWP), sizeof(float)) == hipSuccess ? wp : -1.f;
}

extern "C" lst_status lst_t5dnn_get_weights(float* params) {
  if (!params) return LST_E_INPUT;
  float* p = params;
  if (hipMemcpyFromSymbol(p, HIP_SYMBOL(lst::cDnnBias0), 32 * 4) != hipSuccess) return LST_E_DEVICE;
  p += 32;
  if (hipMemcpyFromSymbol(p, HIP_SYMBOL(lst::cDnnW0), 38 * 32 * 4) != hipSuccess) return LST_E_DEVICE;
  p += 38 * 32;
  if (hipMemcpyFromSymbol(p, HIP_SYMBOL(lst::cDnnBias2), 32 * 4) != hipSuccess) return LST_E_DEVICE;
  p += 32;
  if (hipMemcpyFromSymbol(p, HIP_SYMBOL(lst::cDnnW2), 32 * 32 * 4) != hipSuccess) return LST_E_DEVICE;
  p += 32 * 32;
  if (hipMemcpyFromSymbol(p, HIP_SYMBOL(lst::cDnnBias4), 4) != hipSuccess) return LST_E_DEVICE;
  p += 1;
  if (hipMemcpyFromSymbol(p, HIP_SYMBOL(lst::cDnnW4), 32 * 4) != hipSuccess) return LST_E_DEVICE;
  return LST_OK;
}
