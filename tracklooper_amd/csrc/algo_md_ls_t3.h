// Per-candidate selections of the first three LST stages, evaluated by one
// lane each.  Module attributes are wave-uniform and gathered once per wave
// into a ModCtx (scalar registers) instead of being re-read per candidate.
//
// Selections restate: mini-doublets SDL/MiniDoublet.h:264-894, line segments
// SDL/Segment.h:177-796, triplets SDL/Triplet.h:241-872.
#pragma once
#include "lst_types.h"
#include "module_terms.h"
#include "soa.h"

namespace lst {

// A module's attributes and the module-only terms of the cuts come from the
// geometry's per-module records (module_terms.h): one 16-byte load each.
__device__ __forceinline__ ModCtx load_mod(const ModulesView& M, unsigned m) {
  const float4 r = ld_global(M.modCtx + m);
  ModCtx c;
  unpack_mod(__float_as_uint(r.x), c);
  c.drdz = r.y; c.dxdy = r.z; c.partnerDrdz = r.w;
  return c;
}
// the packed attribute word of module m (module_terms.h layout)
__device__ __forceinline__ uint32_t mod_word(const ModulesView& M, unsigned m) {
  return ld_global(reinterpret_cast<const uint32_t*>(M.modCtx + m));
}
// {muls, segTilt, segRes, -} of module m
__device__ __forceinline__ float4 seg_terms(const ModulesView& M, unsigned m) { return ld_global(M.modTerms + 2 * m); }
// {mdBase, mdTilt2, mdRoot, -} of module m
__device__ __forceinline__ float4 md_terms(const ModulesView& M, unsigned m) {
  return ld_global(M.modTerms + 2 * m + 1);
}

// ----------------------------------------------------------------- mini-doublets
// dphi threshold: SDL/MiniDoublet.h:339-401.  slope = asin(min(rt * k2Rinv1GeVf /
// ptCut, kSinAlphaMax)) of the hit on the pixel side, computed once per hit
__device__ __forceinline__ float md_hit_slope(float rt, float ptCut) {
  return m_asin(m_min(rt * k2Rinv1GeVf / ptCut, kSinAlphaMax));
}
// mt = md_terms of the module: mdBase = muls^2 + pvOff^2, mdTilt2 = tilt^2, mdRoot = sqrt(mdBase)
__device__ __forceinline__ float md_dphi_cut(const ModCtx& c, float4 mt, float slope, float dPhi, float dz) {
  if (c.subdet == kBarrel && c.side == kCenter) return slope + mt.z;
  if (c.subdet == kBarrel) return slope + m_sqrt(mt.x + mt.y * slope * slope);
  const float lum = m_abs(dPhi * kDeltaZLum / dz);
  return slope + m_sqrt(mt.x + lum * lum);
}

// strip-hit shift along the line of sight of the PS pixel hit: SDL/MiniDoublet.h:403-567.
// Everything but the last step depends only on the module pair and the hit on
// its pixel side, so it is computed once per such hit (md_shift_point, in
// k_hits) instead of once per hit pair; md_shift_pair finishes it for a pair.
// The same expressions on the same floats: the results are those of the
// reference's per-pair function.
__device__ __forceinline__ bool md_lower_is_pixel_side(const ModCtx& c) {
  return c.type == kPS ? (c.subdet == kBarrel ? c.layerType != kPixelLayer : c.layerType == kPixelLayer) : true;
}
__device__ __forceinline__ bool md_shifts(const ModCtx& c) { return c.subdet == kEndcap || c.side != kCenter; }
// c: the lower module m of the pair; (xp, yp, zp, rtp): its pixel-side hit.  The module-only terms (angle
// B, and sin / cos of the slope angle per quadrant of the hit) come from the geometry (es.cpp).
__device__ __forceinline__ float4 md_shift_point(const ModulesView& M, uint32_t m, const ModCtx& c, float xp, float yp,
                                                 float zp, float rtp) {
  const float angleA = m_abs(m_atan(rtp / zp));
  const float angleB = M.shiftAngleB[m];
  float sep = gap_size(c);
  if (c.type == kPS && c.layerType != kPixelLayer) sep *= -1;
  const float drprime = (sep / m_sin(angleA + angleB)) * m_sin(angleA);
  const int q = (xp > 0 && yp > 0) ? 0 : (xp > 0 && yp < 0) ? 1 : (xp < 0 && yp < 0) ? 2 : 3;
  const float xa = xp + drprime * M.shiftSinM[4 * m + q];
  const float ya = yp + drprime * M.shiftCosM[4 * m + q];
  const float absdz = m_abs(sep / m_sin(angleA + angleB) * m_cos(angleA));
  const float abszn = c.layerType == kPixelLayer ? m_abs(zp) + absdz : m_abs(zp) - absdz;
  return make_float4(xa, ya, abszn * ((zp > 0) ? 1 : -1), 0.f);
}
// the shifted strip hit from the pixel-side shift point sp and the other hit (xo, yo)
__device__ __forceinline__ void md_shift_pair(const ModCtx& c, float4 sp, float xo, float yo, float& xn, float& yn) {
  const float slope = c.dxdy, xa = sp.x, ya = sp.y;
  if (slope == kSdlInf) { xn = xa; yn = yo; }
  else if (slope == 0) { xn = xo; yn = ya; }
  else {
    xn = (slope * xa + (1.f / slope) * xo - ya + yo) / (slope + (1.f / slope));
    yn = (xn - xa) * slope + ya;
  }
}

struct MDOut {
  float dz, dphi, dphichange, sx, sy, sz, nsdz, nsdphi, nsdphichange;
};

// A candidate hit pair: coordinates, the hits' phi (phi_of of their x, y, kept per hit) and, for a module
// pair that shifts its strip hits, the shift point of the pixel-side hit.  A deltaPhi of two points whose
// phi is known is the wrapped difference of those phis (the same floats the reference computes).
struct MDPair {
  float xL, yL, zL, rtL, phiL, xU, yU, zU, rtU, phiU;
  float4 sp;
};
// deltaPhiChange(x1, y1, x2, y2) with phi1 = phi_of(x1, y1) known (SDL/Hit.h:140-149)
__device__ __forceinline__ float md_dphi_change(float x1, float y1, float phi1, float x2, float y2) {
  return phi_mpi_pi(phi_of(x2 - x1, y2 - y1) - phi1);
}

// SDL/MiniDoublet.h:645-772
__device__ __forceinline__ bool md_barrel(const ModCtx& c, float4 mt, const MDPair& p, float slope, MDOut& o) {
  const float xL = p.xL, yL = p.yL, zL = p.zL, rtL = p.rtL, xU = p.xU, yU = p.yU, zU = p.zU, rtU = p.rtU;
  o.dz = zL - zU;
  const float dzCut = c.type == kPS ? 2.f : 10.f;
  const float sign = ((o.dz > 0) - (o.dz < 0)) * ((zL > 0) - (zL < 0));
  const float crosser = (m_abs(o.dz) > 2) * sign;
  if (!(m_abs(o.dz) < dzCut && crosser <= 0)) return false;
  const float cut = md_dphi_cut(c, mt, slope, 0.f, 0.f);
  float shiftedRt = 0.f, phiS = 0.f;
  if (c.side != kCenter) {
    float xn, yn;
    const bool lowerPixelSide = md_lower_is_pixel_side(c);
    md_shift_pair(c, p.sp, lowerPixelSide ? xU : xL, lowerPixelSide ? yU : yL, xn, yn);
    o.sx = xn; o.sy = yn;
    shiftedRt = m_sqrt(xn * xn + yn * yn);
    phiS = phi_of(xn, yn);
    if (c.layerType == kPixelLayer) {
      o.sz = zU;
      o.dphi = phi_mpi_pi(phiS - p.phiL);  // deltaPhi(lower, shifted)
    } else {
      o.sz = zL;
      o.dphi = phi_mpi_pi(p.phiU - phiS);  // deltaPhi(shifted, upper)
    }
    o.nsdphi = phi_mpi_pi(p.phiU - p.phiL);
  } else {
    o.sx = 0; o.sy = 0; o.sz = 0;
    o.dphi = phi_mpi_pi(p.phiU - p.phiL);
    o.nsdphi = o.dphi;
  }
  if (!(m_abs(o.dphi) < cut)) return false;
  if (c.side != kCenter) {
    if (c.layerType != kPixelLayer)
      o.dphichange = rtL < shiftedRt ? md_dphi_change(xL, yL, p.phiL, o.sx, o.sy)
                                     : md_dphi_change(o.sx, o.sy, phiS, xL, yL);
    else
      o.dphichange = shiftedRt < rtU ? md_dphi_change(o.sx, o.sy, phiS, xU, yU)
                                     : md_dphi_change(xU, yU, p.phiU, o.sx, o.sy);
    o.nsdphichange = rtL < rtU ? md_dphi_change(xL, yL, p.phiL, xU, yU) : md_dphi_change(xU, yU, p.phiU, xL, yL);
  } else {
    o.dphichange = md_dphi_change(xL, yL, p.phiL, xU, yU);
    o.nsdphichange = o.dphichange;
  }
  o.nsdz = 0;
  return m_abs(o.dphichange) < cut;
}

// SDL/MiniDoublet.h:774-894 (o.dz holds drt, as the reference stores it)
__device__ __forceinline__ bool md_endcap(const ModCtx& c, float4 mt, const MDPair& p, float slope, MDOut& o) {
  const float xL = p.xL, yL = p.yL, zL = p.zL, rtL = p.rtL, xU = p.xU, yU = p.yU, zU = p.zU, rtU = p.rtU;
  float dz = zL - zU;
  if (!(m_abs(dz) < 1.f)) return false;
  o.dz = rtL - rtU;
  if (!(m_abs(o.dz) < (c.type == kPS ? 2.f : 10.f))) return false;
  float xn, yn;
  const bool lowerPixelSide = md_lower_is_pixel_side(c);
  md_shift_pair(c, p.sp, lowerPixelSide ? xU : xL, lowerPixelSide ? yU : yL, xn, yn);
  const float zn = p.sp.z;
  o.sx = xn; o.sy = yn;
  const float phiS = phi_of(xn, yn);
  if (c.type == kPS && c.layerType != kPixelLayer) {
    o.sz = zL;
    o.dphi = phi_mpi_pi(p.phiU - phiS);  // deltaPhi(shifted, upper)
  } else {
    o.sz = zU;
    o.dphi = phi_mpi_pi(phiS - p.phiL);  // deltaPhi(lower, shifted)
  }
  o.nsdphi = phi_mpi_pi(p.phiU - p.phiL);
  if (c.type == kPS) dz = c.layerType == kPixelLayer ? zL - zn : zU - zn;
  const float cut = md_dphi_cut(c, mt, slope, o.dphi, dz);
  if (!(m_abs(o.dphi) < cut)) return false;
  const float frac = m_abs(dz) / m_abs(zL);
  o.dphichange = o.dphi / frac * (1.f + frac);
  o.nsdphichange = o.nsdphi / frac * (1.f + frac);
  o.nsdz = 0;
  return m_abs(o.dphichange) < cut;
}

// ------------------------------------------------------------------ line segments
struct MDRef {  // the MD fields a segment selection reads
  float x, y, z, rt, phi, dphichange, dphi, dz, hiEdgePhi, loEdgePhi;
  uint32_t anchorHit;
  uint16_t mod;
};
// the first three 16-byte chunks of the record
__device__ __forceinline__ MDRef load_md(const MDsSoA& D, unsigned i) {
  const uint4 q0 = md_chunk(D, i, 0), q1 = md_chunk(D, i, 1), q2 = md_chunk(D, i, 2);
  MDRef r;
  r.anchorHit = q0.x; r.mod = (uint16_t)q0.y; r.dphichange = __uint_as_float(q0.z); r.dz = __uint_as_float(q0.w);
  r.x = __uint_as_float(q1.x); r.y = __uint_as_float(q1.y); r.z = __uint_as_float(q1.z); r.rt = __uint_as_float(q1.w);
  r.dphi = __uint_as_float(q2.x); r.phi = __uint_as_float(q2.y);
  r.hiEdgePhi = __uint_as_float(q2.z); r.loEdgePhi = __uint_as_float(q2.w);
  return r;
}

struct SegOut {
  float dPhi, dPhiMin, dPhiMax, dPhiChange, dPhiChangeMin, dPhiChangeMax;
};

// dalpha thresholds: SDL/Segment.h:264-356.  ti, to: seg_terms of the inner
// and outer module (muls, the tilt term that is 0 unless a tilted barrel
// module, 0.02 / gap); the resolution of both ends uses the inner module's gap,
// as the reference's does
__device__ __forceinline__ void seg_dalpha_cuts(const ModCtx& ci, float4 ti, const ModCtx& co, float4 to, float ptCut,
                                                const MDRef& a, const MDRef& b, float th[3]) {
  const float muls = ti.x;
  const float dr = m_sqrt((b.y - a.y) * (b.y - a.y) + (b.x - a.x) * (b.x - a.x));
  const float bfield = m_asin(m_min(dr * k2Rinv1GeVf / ptCut, kSinAlphaMax));
  const float tI = ti.y, tO = to.y;
  const float lumI = ci.subdet == kBarrel ? tI * bfield : a.dphi * 15.0f / a.dz;
  const float lumO = co.subdet == kBarrel ? tO * bfield : b.dphi * 15.0f / b.dz;
  const float resI = ti.z * (ci.subdet == kBarrel ? 1.0f : m_abs(a.z) / a.rt);
  const float resO = ti.z * (co.subdet == kBarrel ? 1.0f : m_abs(b.z) / b.rt);
  const float res = resI + resO;
  const float base2 = res * res + muls * muls;
  th[0] = (ci.subdet == kBarrel && ci.side == kCenter) ? bfield + m_sqrt(base2) : bfield + m_sqrt(base2 + lumI * lumI);
  th[1] = (co.subdet == kBarrel && co.side == kCenter) ? bfield + m_sqrt(base2) : bfield + m_sqrt(base2 + lumO * lumO);
  th[2] = bfield + m_sqrt(base2);
}

__device__ __forceinline__ bool seg_dalpha_pass(const ModCtx& ci, float4 ti, const ModCtx& co, float4 to, float ptCut,
                                                const MDRef& a, const MDRef& b, float dPhiChange) {
  float th[3];
  seg_dalpha_cuts(ci, ti, co, to, ptCut, a, b, th);
  if (!(m_abs(a.dphichange - dPhiChange) < th[0])) return false;
  if (!(m_abs(b.dphichange - dPhiChange) < th[1])) return false;
  return m_abs(a.dphichange - b.dphichange) < th[2];
}

// SDL/Segment.h:454-558
// bSlope: (asin(min(b.rt * k2Rinv1GeVf / ptCut, kSinAlphaMax)), the tan of
// it), a function of the outer mini-doublet's anchor hit only, computed once
// per MD from the hit's slope (md_hit_slope, the same formula)
__device__ __forceinline__ bool seg_barrel(const ModCtx& ci, float4 ti, const ModCtx& co, float4 to, float ptCut,
                                           const MDRef& a, const MDRef& b, float2 bSlope, SegOut& o) {
  const float muls = ti.x;
  const float slope = bSlope.x;
  const float pvOff = 0.1f / b.rt;
  const float scale = bSlope.y / slope;
  const float zGeom = ci.layer <= 2 ? 2.f * kPixelPSZpitch : 2.f * kStrip2SZpitch;
  const float zLo = a.z + (a.z - kDeltaZLum) * (b.rt / a.rt - 1.f) * (a.z > 0.f ? 1.f : scale) - zGeom;
  const float zHi = a.z + (a.z + kDeltaZLum) * (b.rt / a.rt - 1.f) * (a.z < 0.f ? 1.f : scale) + zGeom;
  if (!(b.z >= zLo && b.z <= zHi)) return false;
  const float cut = slope + m_sqrt(muls * muls + pvOff * pvOff);
  o.dPhi = phi_mpi_pi(b.phi - a.phi);
  if (!(m_abs(o.dPhi) <= cut)) return false;
  o.dPhiChange = phi_mpi_pi(phi_of(b.x - a.x, b.y - a.y) - a.phi);
  if (!(m_abs(o.dPhiChange) <= cut)) return false;
  return seg_dalpha_pass(ci, ti, co, to, ptCut, a, b, o.dPhiChange);
}

// SDL/Segment.h:560-704
__device__ __forceinline__ bool seg_endcap(const ModCtx& ci, float4 ti, const ModCtx& co, float4 to, float ptCut,
                                           const MDRef& a, const MDRef& b, float2 bSlope, SegOut& o) {
  const float slope = bSlope.x;
  const float r2S = 60.f;
  const float rtGeom = (a.rt < r2S && b.rt < r2S)   ? (2.f * kPixelPSZpitch)
                       : (a.rt < r2S || b.rt < r2S) ? (kPixelPSZpitch + kStrip2SZpitch)
                                                    : (2.f * kStrip2SZpitch);
  if (!(a.z * b.z >= 0)) return false;
  const float dz = b.z - a.z;
  const float dLum = sign_copy(kDeltaZLum, a.z);
  const float scale = slope / bSlope.y;
  const float rtLo = m_max(a.rt * (1.f + dz / (a.z + dLum) * scale) - rtGeom, a.rt - 0.5f * rtGeom);
  const float rtHi = a.rt * (b.z - dLum) / (a.z - dLum) + rtGeom;
  if (!(b.rt >= rtLo && b.rt <= rtHi)) return false;
  o.dPhi = phi_mpi_pi(b.phi - a.phi);
  if (co.subdet == kEndcap && co.type == kTwoS) {
    const float hi = phi_mpi_pi(b.hiEdgePhi - a.phi);
    const float lo = phi_mpi_pi(b.loEdgePhi - a.phi);
    const bool hiWider = m_abs(hi) > m_abs(lo);
    o.dPhiMax = hiWider ? hi : lo;
    o.dPhiMin = hiWider ? lo : hi;
  } else {
    o.dPhiMax = o.dPhi;
    o.dPhiMin = o.dPhi;
  }
  if (!(m_abs(o.dPhi) <= slope)) return false;
  const float frac = dz / a.z;
  o.dPhiChange = o.dPhi / frac * (1.f + frac);
  o.dPhiChangeMin = o.dPhiMin / frac * (1.f + frac);
  o.dPhiChangeMax = o.dPhiMax / frac * (1.f + frac);
  if (!(m_abs(o.dPhiChange) <= slope)) return false;
  return seg_dalpha_pass(ci, ti, co, to, ptCut, a, b, o.dPhiChange);
}

// ------------------------------------------------------------------------ triplets
struct T3Geo {  // anchor coordinates of the three MDs
  float x1, y1, z1, r1, phi1, x2, y2, z2, r2, x3, y3, z3, r3;
  float2 slope3;  // asin(min(r3 * k2Rinv1GeVf / ptCut, kSinAlphaMax)) and its tan, kept per MD
};

// SDL/Triplet.h:241-298
__device__ __forceinline__ bool t3_rz(int l1, int l2, int l3, const T3Geo& p) {
  const float res = m_abs(p.z2 - ((p.z3 - p.z1) / (p.r3 - p.r1) * (p.r2 - p.r1) + p.z1));
  const int key = l1 * 10000 + l2 * 100 + l3;
  switch (key) {
    case 121314: case 131415: case 141516: return false;
    case 10203: return res < 0.53f;
    case 20304: return res < 1.21f;
    case 30405: return res < 2.7f;
    case 40506: return res < 3.06f;
    case 10207: case 10708: case 20307: case 20708: case 70809: case 80910: case 91011: return res < 1.f;
    default: return res < 5;
  }
}

// SDL/Triplet.h:300-402 (barrel-barrel-barrel)
__device__ __forceinline__ bool t3_bbb(bool psIn, bool psOut, float ptCut, const T3Geo& p, float alphaIn,
                                       float& betaIn) {
  const float a1 = p.slope3.x;
  const float ratio = p.r3 / p.r1;
  const float scale = p.slope3.y / a1;
  const float zpIn = psIn ? kPixelPSZpitch : kStrip2SZpitch;
  const float zpOut = psOut ? kPixelPSZpitch : kStrip2SZpitch;
  const float zHi = p.z1 + (p.z1 + kDeltaZLum) * (ratio - 1.f) * (p.z1 < 0.f ? 1.f : scale) + (zpIn + zpOut);
  const float zLo = p.z1 + (p.z1 - kDeltaZLum) * (ratio - 1.f) * (p.z1 > 0.f ? 1.f : scale) - (zpIn + zpOut);
  if (!(p.z3 >= zLo && p.z3 <= zHi)) return false;
  const float drtOutIn = p.r3 - p.r1;
  const float r3In = m_sqrt(p.z1 * p.z1 + p.r1 * p.r1);
  const float drtIn = p.r2 - p.r1;
  const float dzIn = p.z2 - p.z1;
  const float dr3In = m_sqrt(p.r2 * p.r2 + p.z2 * p.z2) - m_sqrt(p.r1 * p.r1 + p.z1 * p.z1);
  const float coshEta = dr3In / drtIn;
  float dzErr = (zpIn + zpOut) * (zpIn + zpOut) * 2.f;
  const float thetaMuls = 0.015f * m_sqrt(0.1f + 0.2f * (p.r3 - p.r1) / 50.f) * m_sqrt(r3In / p.r1);
  const float muls = thetaMuls * 3.f / ptCut * 4.f;
  dzErr += muls * muls * drtOutIn * drtOutIn / 3.f * coshEta * coshEta;
  dzErr = m_sqrt(dzErr);
  const float dzMean = dzIn / drtIn * drtOutIn;
  const float zWin = dzErr / drtIn * drtOutIn + (zpIn + zpOut);
  const float zLoP = p.z1 + dzMean * (p.z1 > 0.f ? 1.f : scale) - zWin;
  const float zHiP = p.z1 + dzMean * (p.z1 < 0.f ? 1.f : scale) + zWin;
  const bool pointed = p.z3 >= zLoP && p.z3 <= zHiP;
  const float tx = p.x3 - p.x1, ty = p.y3 - p.y1;
  betaIn = alphaIn - phi_mpi_pi(phi_of(tx, ty) - p.phi1);
  const float dtl = m_sqrt(tx * tx + ty * ty);
  const float rtIn = m_sqrt((p.x2 - p.x1) * (p.x2 - p.x1) + (p.y2 - p.y1) * (p.y2 - p.y1));
  const float cut = m_asin(m_min((-rtIn + dtl) * k2Rinv1GeVf / ptCut, kSinAlphaMax)) + (0.02f / drtIn);
  return pointed && m_abs(betaIn) < cut;
}

// SDL/Triplet.h:404-526 (barrel-barrel-endcap and barrel-endcap-endcap)
__device__ __forceinline__ bool t3_bbe(bool psIn, bool psOut, float ptCut, const T3Geo& p, float alphaIn,
                                       float& betaIn) {
  const float a1 = p.slope3.x;
  const float scale = p.slope3.y / a1;
  const float zGeom = (psIn ? kPixelPSZpitch : kStrip2SZpitch) + (psOut ? kPixelPSZpitch : kStrip2SZpitch);
  if (!(p.z1 * p.z3 > 0)) return false;
  const float dLum = sign_copy(kDeltaZLum, p.z1);
  const float rtGeom1 = psOut ? kPixelPSZpitch : kStrip2SZpitch;
  const float zGeom1 = sign_copy(zGeom, p.z1);
  const float rtLo = p.r1 * (1.f + (p.z3 - p.z1 - zGeom1) / (p.z1 + zGeom1 + dLum) / scale) - rtGeom1;
  float zInForHi = p.z1 - zGeom1 - dLum;
  if (zInForHi * p.z1 < 0) zInForHi = sign_copy(0.1f, p.z1);
  const float rtHi = p.r1 * (1.f + (p.z3 - p.z1 + zGeom1) / zInForHi) + rtGeom1;
  if (!(p.r3 >= rtLo && p.r3 <= rtHi)) return false;
  const float kZ = (p.z3 - p.z1) / (p.z2 - p.z1);
  const float tx = p.x3 - p.x1, ty = p.y3 - p.y1;
  betaIn = alphaIn - phi_mpi_pi(phi_of(tx, ty) - p.phi1);
  const float dIn = m_sqrt((p.x2 - p.x1) * (p.x2 - p.x1) + (p.y2 - p.y1) * (p.y2 - p.y1));
  const float dtl = m_sqrt(tx * tx + ty * ty);
  const float cut = m_asin(m_min((-dIn + dtl) * k2Rinv1GeVf / ptCut, kSinAlphaMax)) + (0.02f / (p.r2 - p.r1));
  // the reference's rt-z error estimate here (SDL/Triplet.h:474-488) feeds no cut
  return kZ >= 0 && m_abs(betaIn) < cut;
}

// SDL/Triplet.h:528-663 (endcap-endcap-endcap)
__device__ __forceinline__ bool t3_eee(bool psIn, bool psOut, float ptCut, const T3Geo& p, float alphaIn,
                                       float alphaMin, float alphaMax, float& betaIn) {
  const float a1 = p.slope3.x;
  const float scale = p.slope3.y / a1;
  if (!(p.z1 * p.z3 > 0)) return false;
  const float dLum = sign_copy(kDeltaZLum, p.z1);
  const float rtGeom = (psIn && psOut) ? 2.f * kPixelPSZpitch
                       : (psIn || psOut) ? kPixelPSZpitch + kStrip2SZpitch
                                         : 2.f * kStrip2SZpitch;
  const float dz = p.z3 - p.z1;
  const float rtLo = p.r1 * (1.f + dz / (p.z1 + dLum) / scale) - rtGeom;
  const float rtHi = p.r1 * (1.f + dz / (p.z1 - dLum)) + rtGeom;
  if (!(p.r3 >= rtLo && p.r3 <= rtHi)) return false;
  bool pass = true;
  if (psIn && psOut) {
    const float drtIn = p.r2 - p.r1, dzIn = p.z2 - p.z1;
    const float dr3In = m_sqrt(p.r2 * p.r2 + p.z2 * p.z2) - m_sqrt(p.r1 * p.r1 + p.z1 * p.z1);
    const float coshEta = dr3In / drtIn;
    const float dzAbs = m_abs(p.z3 - p.z1);
    const float multDzDr = dzAbs * coshEta / (coshEta * coshEta - 1.f);
    const float kZ = (p.z3 - p.z1) / dzIn;
    const float thetaMuls = 0.015f * m_sqrt(0.1f + 0.2f * (p.r3 - p.r1) / 50.f);
    const float muls = thetaMuls * 3.f / ptCut * 4.f;
    const float drtErr = m_sqrt(kPixelPSZpitch * kPixelPSZpitch * 2.f / (dzIn * dzIn) * (dzAbs * dzAbs) +
                                muls * muls * multDzDr * multDzDr / 3.f * coshEta * coshEta);
    const float drtMean = drtIn * dzAbs / m_abs(dzIn);
    const float win = drtErr + rtGeom;
    pass = kZ >= 0 && p.r3 >= p.r1 + drtMean / scale - win && p.r3 <= p.r1 + drtMean + win;
  }
  const float tx = p.x3 - p.x1, ty = p.y3 - p.y1;
  betaIn = alphaIn - phi_mpi_pi(phi_of(tx, ty) - p.phi1);
  float bMin = betaIn + alphaMin - alphaIn;
  const float bMax = betaIn + alphaMax - alphaIn;
  if (m_abs(bMin) > m_abs(bMax)) bMin = bMax;
  const float dIn = m_sqrt((p.x2 - p.x1) * (p.x2 - p.x1) + (p.y2 - p.y1) * (p.y2 - p.y1));
  const float dtl = m_sqrt(tx * tx + ty * ty);
  const float cut = m_asin(m_min((-dIn + dtl) * k2Rinv1GeVf / ptCut, kSinAlphaMax)) + (0.02f / (p.r2 - p.r1));
  return pass && m_abs(bMin) < cut;
}

// circumcircle of three points (SDL/Triplet.h:762-794); radius -1 if degenerate
__device__ __forceinline__ float circle_radius3(float x1, float y1, float x2, float y2, float x3, float y3, float& g,
                                                float& f) {
  const float det = (y1 - y3) * (x2 - x3) - (x1 - x3) * (y2 - y3);
  const float inv = 1.0f / det;
  const float s1 = x1 * x1 + y1 * y1, s2 = x2 * x2 + y2 * y2, s3 = x3 * x3 + y3 * y3;
  g = 0.5f * ((y3 - y2) * s1 + (y1 - y3) * s2 + (y2 - y1) * s3) * inv;
  f = 0.5f * ((x2 - x3) * s1 + (x3 - x1) * s2 + (x1 - x2) * s3) * inv;
  const float c = ((x2 * y3 - x3 * y2) * s1 + (x3 * y1 - x1 * y3) * s2 + (x1 * y2 - x2 * y1) * s3) * inv;
  if (det == 0 || (g * g + f * f - c < 0)) return -1.f;
  return m_sqrt(g * g + f * f - c);
}

}  // namespace lst
