// Quintuplet (T5) selection evaluated by one lane per (inner T3, outer T3)
// candidate.  Restates SDL/Quintuplet.h:143-2986 in plain-chi2 mode
// (-DUSE_RZCHI2 -DUSE_RPHICHI2).  The mini-doublet coordinates a candidate needs
// are gathered once into registers (MDPoint) instead of re-read per formula.
#pragma once
#include "algo_md_ls_t3.h"

namespace lst {

struct MDPoint {  // anchor hit of a mini-doublet, plus its strip edges
  float x, y, z, rt, phi, hx, hy, lx, ly, hphi, lphi;
};
// chunks 1-3 of the record
__device__ __forceinline__ MDPoint load_point(const MDsSoA& D, unsigned i) {
  const uint4 q1 = md_chunk(D, i, 1), q2 = md_chunk(D, i, 2), q3 = md_chunk(D, i, 3);
  MDPoint p;
  p.x = __uint_as_float(q1.x); p.y = __uint_as_float(q1.y); p.z = __uint_as_float(q1.z); p.rt = __uint_as_float(q1.w);
  p.phi = __uint_as_float(q2.y); p.hphi = __uint_as_float(q2.z); p.lphi = __uint_as_float(q2.w);
  p.hx = __uint_as_float(q3.x); p.hy = __uint_as_float(q3.y); p.lx = __uint_as_float(q3.z); p.ly = __uint_as_float(q3.w);
  return p;
}

struct ModLite {  // the module attributes the T5 selection reads
  int16_t subdet, side, layer;
  uint8_t type;
  int sdl;
  float drdz, dxdy;
};
__device__ __forceinline__ ModLite load_lite(const ModulesView& M, unsigned m) {
  const float4 r = ld_global(M.modCtx + m);  // one 16-byte load (module_terms.h)
  const uint32_t w = __float_as_uint(r.x);
  ModCtx full;
  unpack_mod(w, full);
  ModLite c;
  c.subdet = full.subdet; c.side = full.side; c.layer = full.layer; c.type = full.type;
  c.sdl = unpack_sdl_layer(w); c.drdz = r.y; c.dxdy = r.z;
  return c;
}
__device__ __forceinline__ bool is_endcap_2s(const ModLite& c) { return c.subdet == kEndcap && c.type == kTwoS; }

// Segment quantities of the inner (seg1) and outer (seg3) segment of a T4-style test.
struct SegAlpha {
  float dPhiChange, dPhiChangeMin, dPhiChangeMax, dPhiMin, dPhiMax;
};
__device__ __forceinline__ SegAlpha load_alpha(const SegsSoA& S, unsigned s) {
  const uint4 c1 = seg_chunk(S, s, 1), c2 = seg_chunk(S, s, 2);
  return SegAlpha{__uint_as_float(c2.y), __uint_as_float(c2.z), __uint_as_float(c2.w), __uint_as_float(c1.w),
                  __uint_as_float(c2.x)};
}

// beta refinement by the pT estimate: SDL/Quintuplet.h:1370-1464
__device__ __forceinline__ float dbeta_step(float dr, float absPt) {
  return m_asin(m_min(dr * k2Rinv1GeVf / absPt, kSinAlphaMax));
}
__device__ __forceinline__ void delta_beta_iterations(float& betaIn, float& betaOut, float& betaAv, float& ptBeta,
                                                      float drIn, float drOut, float dr, float lIn) {
  if (lIn == 0) {
    betaOut += sign_copy(dbeta_step(drOut, m_abs(ptBeta)), betaOut);
    return;
  }
  if (betaIn * betaOut > 0.f &&
      (m_abs(ptBeta) < 4.f * kPtBetaMax || (lIn >= 11 && m_abs(ptBeta) < 8.f * kPtBetaMax))) {
    const float inUpd = betaIn + sign_copy(dbeta_step(drIn, m_abs(ptBeta)), betaIn);
    const float outUpd = betaOut + sign_copy(dbeta_step(drOut, m_abs(ptBeta)), betaOut);
    betaAv = 0.5f * (inUpd + outUpd);
    const float invPt = 1.f / m_abs(dr * k2Rinv1GeVf / m_sin(betaAv));
    betaIn += sign_copy(m_asin(m_min(drIn * k2Rinv1GeVf * invPt, kSinAlphaMax)), betaIn);
    betaOut += sign_copy(m_asin(m_min(drOut * k2Rinv1GeVf * invPt, kSinAlphaMax)), betaOut);
    betaAv = 0.5f * (betaIn + betaOut);
    ptBeta = dr * k2Rinv1GeVf / m_sin(betaAv);
  } else if (lIn < 11 && m_abs(betaOut) < 0.2f * m_abs(betaIn) && m_abs(ptBeta) < 12.f * kPtBetaMax) {
    const float ptIn = dr * k2Rinv1GeVf / m_sin(betaIn);
    const float inUpd = betaIn + sign_copy(dbeta_step(drIn, m_abs(ptIn)), betaIn);
    const float outUpd = betaOut + sign_copy(dbeta_step(drOut, m_abs(ptIn)), betaIn);
    betaAv = m_abs(betaOut) > 0.2f * m_abs(betaIn) ? 0.5f * (inUpd + outUpd) : inUpd;
    ptBeta = dr * k2Rinv1GeVf / m_sin(betaAv);
    betaIn += sign_copy(dbeta_step(drIn, m_abs(ptBeta)), betaIn);
    betaOut += sign_copy(dbeta_step(drOut, m_abs(ptBeta)), betaIn);
    betaAv = 0.5f * (betaIn + betaOut);
    ptBeta = dr * k2Rinv1GeVf / m_sin(betaAv);
  }
}

struct T4Args {
  ModLite m1, m2, m3, m4;  // modules of the four mini-doublets
  MDPoint a, b, c, d;      // the four mini-doublets (inner segment a-b, outer segment c-d)
  SegAlpha s1, s3;         // inner and outer segment
  float2 cSlope;           // asin(min(c.rt * k2Rinv1GeVf / ptCut, kSinAlphaMax)) and its tan, kept per MD
};

__device__ __forceinline__ float planar_dist(float x1, float y1, float x2, float y2) {
  return m_sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
}

// Common tail of the three T4 flavours: beta cuts after the delta-beta iterations.
struct BetaTail {
  float betaIn, betaOut, betaInMin, betaInMax, betaOutMin, betaOutMax;
  float thetaMuls, drIn, drOut, dOut, dInCut, dr, dPhi;
  float alphaInAbs, alphaOutAbs;  // |alpha| used for the luminous-region term
  float lIn, lOut, dBetaROut;
};
__device__ __forceinline__ bool beta_tail(const T4Args& q, float ptCut, BetaTail t) {
  float betaAv = 0.5f * (t.betaIn + t.betaOut);
  float ptBeta = t.dr * k2Rinv1GeVf / m_sin(betaAv);
  delta_beta_iterations(t.betaIn, t.betaOut, betaAv, ptBeta, t.drIn, t.drOut, t.dr, t.lIn);
  const float sIn = m_abs(t.betaInMin + t.betaInMax) > 0 ? (2.f * t.betaIn / m_abs(t.betaInMin + t.betaInMax)) : 0.f;
  const float sOut =
      m_abs(t.betaOutMin + t.betaOutMax) > 0 ? (2.f * t.betaOut / m_abs(t.betaOutMin + t.betaOutMax)) : 0.f;
  t.betaInMin *= sIn; t.betaInMax *= sIn; t.betaOutMin *= sOut; t.betaOutMax *= sOut;
  const float mulsB = t.thetaMuls * 4.f / m_min(m_abs(ptBeta), kPtBetaMax);
  const float regIn = m_max(t.alphaInAbs, m_asin(m_min(q.a.rt * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
  const float regOut = m_max(t.alphaOutAbs, m_asin(m_min(q.c.rt * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
  const float lumIn = t.lIn < 11 ? 0.0f : m_abs(regIn * kDeltaZLum / q.a.z);
  const float lumOut = t.lOut < 11 ? 0.0f : m_abs(regOut * kDeltaZLum / q.c.z);
  const float lum2 = (lumIn + lumOut) * (lumIn + lumOut);
  const float outCut = m_asin(m_min(t.dr * k2Rinv1GeVf / ptCut, kSinAlphaMax)) + (0.02f / t.dOut) +
                       m_sqrt(lum2 + mulsB * mulsB);
  if (!(m_abs(t.betaOut) < outCut)) return false;
  const float res = 0.02f / m_min(t.dOut, t.dInCut);
  const float spread = m_abs(t.betaInMin - t.betaInMax) + m_abs(t.betaOutMin - t.betaOutMax);
  const float rOut2 = t.dBetaROut * t.dBetaROut;
  const float cut2 = (res * res * 2.0f + mulsB * mulsB + lum2 + 0.f + rOut2 + 0.25f * spread * spread);
  const float dBeta = t.betaIn - t.betaOut;
  return dBeta * dBeta <= cut2;
}

__device__ __forceinline__ float edge_span(const MDPoint& d) {
  return m_sqrt(d.hx * d.hx + d.hy * d.hy) - m_sqrt(d.lx * d.lx + d.ly * d.ly);
}

// phi window checks shared by the three flavours (SDL/Quintuplet.h:1559-1577 and analogues)
__device__ __forceinline__ bool phi_windows(const T4Args& q, float cut, float& dPhi) {
  if (!(m_abs(phi_mpi_pi(q.d.phi - q.b.phi)) <= cut)) return false;
  dPhi = delta_phi(0.5f * (q.a.x + q.c.x), 0.5f * (q.a.y + q.c.y), q.c.x - q.a.x, q.c.y - q.a.y);
  return m_abs(dPhi) <= cut;
}

// SDL/Quintuplet.h:1467-1742 (barrel inner segment, also BBBE)
__device__ __forceinline__ bool t4_bbbb(const T4Args& q, float ptCut) {
  const float zpIn = q.m1.type == kPS ? kPixelPSZpitch : kStrip2SZpitch;
  const float zpOut = q.m3.type == kPS ? kPixelPSZpitch : kStrip2SZpitch;
  const float a1 = q.cSlope.x;
  const float ratio = q.c.rt / q.a.rt;
  const float scale = q.cSlope.y / a1;
  const float zHi = q.a.z + (q.a.z + kDeltaZLum) * (ratio - 1.f) * (q.a.z < 0.f ? 1.f : scale) + (zpIn + zpOut);
  const float zLo = q.a.z + (q.a.z - kDeltaZLum) * (ratio - 1.f) * (q.a.z > 0.f ? 1.f : scale) - (zpIn + zpOut);
  if (!(q.c.z >= zLo && q.c.z <= zHi)) return false;
  const float drtOutIn = q.c.rt - q.a.rt;
  const float r3In = m_sqrt(q.a.z * q.a.z + q.a.rt * q.a.rt);
  const float drtIn = q.b.rt - q.a.rt;
  const float dzIn = q.b.z - q.a.z;
  const float dr3In = m_sqrt(q.b.rt * q.b.rt + q.b.z * q.b.z) - m_sqrt(q.a.rt * q.a.rt + q.a.z * q.a.z);
  const float coshEta = dr3In / drtIn;
  float dzErr = (zpIn + zpOut) * (zpIn + zpOut) * 2.f;
  const float thetaMuls = 0.015f * m_sqrt(0.1f + 0.2f * (q.c.rt - q.a.rt) / 50.f) * m_sqrt(r3In / q.a.rt);
  const float muls = thetaMuls * 3.f / ptCut * 4.f;
  dzErr += muls * muls * drtOutIn * drtOutIn / 3.f * coshEta * coshEta;
  dzErr = m_sqrt(dzErr);
  const float dzMean = dzIn / drtIn * drtOutIn;
  const float zWin = dzErr / drtIn * drtOutIn + (zpIn + zpOut);
  if (!(q.c.z >= q.a.z + dzMean * (q.a.z > 0.f ? 1.f : scale) - zWin &&
        q.c.z <= q.a.z + dzMean * (q.a.z < 0.f ? 1.f : scale) + zWin))
    return false;
  const float pvOff = 0.1f / q.c.rt;
  const float cut = a1 + m_sqrt(muls * muls + pvOff * pvOff);
  float dPhi;
  if (!phi_windows(q, cut, dPhi)) return false;
  const bool ecLast = is_endcap_2s(q.m4);
  const float alphaOutUp = phi_mpi_pi(phi_of(q.d.x - q.c.x, q.d.y - q.c.y) - q.d.phi);
  const float tx = q.d.x - q.a.x, ty = q.d.y - q.a.y;
  BetaTail t;
  t.betaIn = q.s1.dPhiChange - phi_mpi_pi(phi_of(tx, ty) - q.a.phi);
  t.betaInMin = t.betaIn;
  t.betaInMax = t.betaIn;
  t.betaOut = -alphaOutUp + phi_mpi_pi(phi_of(tx, ty) - q.d.phi);
  t.betaOutMin = t.betaOut;
  t.betaOutMax = t.betaOut;
  if (ecLast) {
    const float hiAlpha = phi_mpi_pi(phi_of(q.d.hx - q.c.x, q.d.hy - q.c.y) - q.d.hphi);
    const float loAlpha = phi_mpi_pi(phi_of(q.d.lx - q.c.x, q.d.ly - q.c.y) - q.d.lphi);
    t.betaOutMin = -hiAlpha + phi_mpi_pi(phi_of(q.d.hx - q.a.x, q.d.hy - q.a.y) - q.d.hphi);
    t.betaOutMax = -loAlpha + phi_mpi_pi(phi_of(q.d.lx - q.a.x, q.d.ly - q.a.y) - q.d.lphi);
  }
  t.dr = m_sqrt(tx * tx + ty * ty);
  t.drIn = planar_dist(q.a.x, q.a.y, q.b.x, q.b.y);
  const float inCut = m_asin(m_min((-t.drIn * 1.f + t.dr) * k2Rinv1GeVf / ptCut, kSinAlphaMax)) + (0.02f / drtIn);
  if (!(m_abs(t.betaInMin) < inCut)) return false;
  t.thetaMuls = thetaMuls;
  t.drOut = planar_dist(q.c.x, q.c.y, q.d.x, q.d.y);
  t.dOut = q.d.rt - q.c.rt;
  t.dInCut = drtIn;
  t.dPhi = dPhi;
  t.alphaInAbs = m_abs(q.s1.dPhiChange);
  t.alphaOutAbs = m_abs(q.s3.dPhiChange);
  t.lIn = 5;
  t.lOut = ecLast ? 11 : 5;
  t.dBetaROut = ecLast ? edge_span(q.d) * m_sin(dPhi) / t.dr : 0.f;
  return beta_tail(q, ptCut, t);
}

// SDL/Quintuplet.h:1744-2028 (barrel to endcap, also BEEE)
__device__ __forceinline__ bool t4_bbee(const T4Args& q, float ptCut) {
  const float zpIn = q.m1.type == kPS ? kPixelPSZpitch : kStrip2SZpitch;
  const float zpOut = q.m3.type == kPS ? kPixelPSZpitch : kStrip2SZpitch;
  const float a1 = q.cSlope.x;
  const float scale = q.cSlope.y / a1;
  if (!(q.a.z * q.c.z > 0)) return false;
  const float dLum = sign_copy(kDeltaZLum, q.a.z);
  const float rtGeom1 = q.m3.type == kPS ? kPixelPSZpitch : kStrip2SZpitch;
  const float zGeom1 = sign_copy(zpIn + zpOut, q.a.z);
  const float rtLo = q.a.rt * (1.f + (q.c.z - q.a.z - zGeom1) / (q.a.z + zGeom1 + dLum) / scale) - rtGeom1;
  if (!(q.c.rt >= rtLo)) return false;
  float zInForHi = q.a.z - zGeom1 - dLum;
  if (zInForHi * q.a.z < 0) zInForHi = sign_copy(0.1f, q.a.z);
  const float rtHi = q.a.rt * (1.f + (q.c.z - q.a.z + zGeom1) / zInForHi) + rtGeom1;
  if (!(q.c.rt >= rtLo && q.c.rt <= rtHi)) return false;
  const float rIn = m_sqrt(q.a.z * q.a.z + q.a.rt * q.a.rt);
  const float kZ = (q.c.z - q.a.z) / (q.b.z - q.a.z);
  const float thetaMuls = 0.015f * m_sqrt(0.1f + 0.2f * (q.c.rt - q.a.rt) / 50.f) * m_sqrt(rIn / q.a.rt);
  const float muls = thetaMuls * 3.f / ptCut * 4.f;
  if (!(kZ >= 0)) return false;
  const float pvOff = 0.1f / q.c.rt;
  const float cut = a1 + m_sqrt(muls * muls + pvOff * pvOff);
  float dPhi;
  if (!phi_windows(q, cut, dPhi)) return false;
  const float alphaIn = q.s1.dPhiChange;
  const float alphaOut = phi_mpi_pi(phi_of(q.d.x - q.c.x, q.d.y - q.c.y) - q.d.phi);
  const float alphaOutMin = phi_mpi_pi(q.s3.dPhiChangeMin - q.s3.dPhiMin);
  const float alphaOutMax = phi_mpi_pi(q.s3.dPhiChangeMax - q.s3.dPhiMax);
  const float tx = q.d.x - q.a.x, ty = q.d.y - q.a.y;
  BetaTail t;
  t.betaIn = alphaIn - phi_mpi_pi(phi_of(tx, ty) - q.a.phi);
  t.betaInMin = t.betaIn;
  t.betaInMax = t.betaIn;
  t.betaOut = -alphaOut + phi_mpi_pi(phi_of(tx, ty) - q.d.phi);
  if (is_endcap_2s(q.m2)) {
    t.betaInMin = t.betaIn - q.s1.dPhiChangeMin + alphaIn;
    t.betaInMax = t.betaIn - q.s1.dPhiChangeMax + alphaIn;
  }
  t.betaOutMin = t.betaOut - alphaOutMin + alphaOut;
  t.betaOutMax = t.betaOut - alphaOutMax + alphaOut;
  if (m_abs(t.betaOutMin) > m_abs(t.betaOutMax)) { const float s = t.betaOutMin; t.betaOutMin = t.betaOutMax; t.betaOutMax = s; }
  if (m_abs(t.betaInMin) > m_abs(t.betaInMax)) { const float s = t.betaInMin; t.betaInMin = t.betaInMax; t.betaInMax = s; }
  t.drIn = planar_dist(q.a.x, q.a.y, q.b.x, q.b.y);
  const float dIn = q.b.rt - q.a.rt;
  t.dr = m_sqrt(tx * tx + ty * ty);
  const float inCut = m_asin(m_min((-t.drIn * 1.f + t.dr) * k2Rinv1GeVf / ptCut, kSinAlphaMax)) + (0.02f / dIn);
  if (!(m_abs(t.betaInMin) < inCut)) return false;
  t.thetaMuls = thetaMuls;
  t.drOut = planar_dist(q.c.x, q.c.y, q.d.x, q.d.y);
  t.dOut = q.d.rt - q.c.rt;
  t.dInCut = dIn;
  t.dPhi = dPhi;
  t.alphaInAbs = m_abs(alphaIn);
  t.alphaOutAbs = m_abs(alphaIn);  // the reference reuses the inner alpha here
  t.lIn = 5;
  t.lOut = 11;
  t.dBetaROut = q.m4.type == kTwoS ? edge_span(q.d) * m_sin(dPhi) / t.dr : 0.f;
  return beta_tail(q, ptCut, t);
}

// SDL/Quintuplet.h:2030-2298 (endcap only)
__device__ __forceinline__ bool t4_eeee(const T4Args& q, float ptCut) {
  const float a1 = q.cSlope.x;
  const float scale = q.cSlope.y / a1;
  if (!((q.a.z * q.c.z) > 0)) return false;
  const float dLum = sign_copy(kDeltaZLum, q.a.z);
  const bool psOutLo = q.m3.type == kPS, psInLo = q.m1.type == kPS;
  const float rtGeom = (psInLo && psOutLo) ? 2.f * kPixelPSZpitch
                       : (psInLo || psOutLo) ? kPixelPSZpitch + kStrip2SZpitch
                                             : 2.f * kStrip2SZpitch;
  const float dz = q.c.z - q.a.z;
  const float rtLo = q.a.rt * (1.f + dz / (q.a.z + dLum) / scale) - rtGeom;
  const float rtHi = q.a.rt * (1.f + dz / (q.a.z - dLum)) + rtGeom;
  if (!(q.c.rt >= rtLo && q.c.rt <= rtHi)) return false;
  const float drtIn = q.b.rt - q.a.rt, dzIn = q.b.z - q.a.z;
  const float thetaMuls = 0.015f * m_sqrt(0.1f + 0.2f * (q.c.rt - q.a.rt) / 50.f);
  const float muls = thetaMuls * 3.f / ptCut * 4.f;
  if (psInLo && q.m2.type == kPS) {
    const float dr3In = m_sqrt(q.b.rt * q.b.rt + q.b.z * q.b.z) - m_sqrt(q.a.rt * q.a.rt + q.a.z * q.a.z);
    const float coshEta = dr3In / drtIn;
    const float dzAbs = m_abs(q.c.z - q.a.z);
    const float multDzDr = dzAbs * coshEta / (coshEta * coshEta - 1.f);
    const float kZ = (q.c.z - q.a.z) / dzIn;
    const float drtErr = m_sqrt(kPixelPSZpitch * kPixelPSZpitch * 2.f / (dzIn * dzIn) * (dzAbs * dzAbs) +
                                muls * muls * multDzDr * multDzDr / 3.f * coshEta * coshEta);
    const float drtMean = drtIn * dzAbs / m_abs(dzIn);
    const float win = drtErr + rtGeom;
    if (!(kZ >= 0 && q.c.rt >= q.a.rt + drtMean / scale - win && q.c.rt <= q.a.rt + drtMean + win)) return false;
  }
  const float pvOff = 0.1f / q.c.rt;
  const float cut = a1 + m_sqrt(muls * muls + pvOff * pvOff);
  float dPhi;
  if (!phi_windows(q, cut, dPhi)) return false;
  const float alphaIn = q.s1.dPhiChange;
  const float dPhiPosOut = phi_mpi_pi(q.d.phi - q.c.phi);
  const float alphaOutMin = phi_mpi_pi(q.s3.dPhiChangeMin - dPhiPosOut);
  const float alphaOutMax = phi_mpi_pi(q.s3.dPhiChangeMax - dPhiPosOut);
  const float alphaOut = phi_mpi_pi(q.s3.dPhiChange - dPhiPosOut);
  const float tx = q.d.x - q.a.x, ty = q.d.y - q.a.y;
  BetaTail t;
  t.betaIn = alphaIn - phi_mpi_pi(phi_of(tx, ty) - q.a.phi);
  t.betaInMin = t.betaIn + q.s1.dPhiChangeMin - alphaIn;
  t.betaInMax = t.betaIn + q.s1.dPhiChangeMax - alphaIn;
  t.betaOut = -alphaOut + phi_mpi_pi(phi_of(tx, ty) - q.d.phi);
  t.betaOutMin = t.betaOut - alphaOutMin + alphaOut;
  t.betaOutMax = t.betaOut - alphaOutMax + alphaOut;
  if (m_abs(t.betaOutMin) > m_abs(t.betaOutMax)) { const float s = t.betaOutMin; t.betaOutMin = t.betaOutMax; t.betaOutMax = s; }
  if (m_abs(t.betaInMin) > m_abs(t.betaInMax)) { const float s = t.betaInMin; t.betaInMin = t.betaInMax; t.betaInMax = s; }
  t.drIn = planar_dist(q.a.x, q.a.y, q.b.x, q.b.y);
  t.dr = m_sqrt(tx * tx + ty * ty);
  const float inCut = m_asin(m_min((-t.drIn * 1.f + t.dr) * k2Rinv1GeVf / ptCut, kSinAlphaMax)) + (0.02f / drtIn);
  if (!(m_abs(t.betaInMin) < inCut)) return false;
  t.thetaMuls = thetaMuls;
  t.drOut = planar_dist(q.c.x, q.c.y, q.d.x, q.d.y);
  t.dOut = q.d.rt - q.c.rt;
  t.dInCut = drtIn;
  t.dPhi = dPhi;
  t.alphaInAbs = m_abs(alphaIn);
  t.alphaOutAbs = m_abs(alphaIn);  // the reference reuses the inner alpha here
  t.lIn = 11;
  t.lOut = 13;
  t.dBetaROut = 0.f;
  return beta_tail(q, ptCut, t);
}

// dispatch by subdetector pattern: SDL/Quintuplet.h:2299-2505
__device__ __forceinline__ bool t4_select(const T4Args& q, float ptCut) {
  const bool b1 = q.m1.subdet == kBarrel, b2 = q.m2.subdet == kBarrel, b3 = q.m3.subdet == kBarrel,
             b4 = q.m4.subdet == kBarrel;
  const bool e1 = q.m1.subdet == kEndcap, e2 = q.m2.subdet == kEndcap, e3 = q.m3.subdet == kEndcap,
             e4 = q.m4.subdet == kEndcap;
  if (b1 && b2 && b3 && (b4 || e4)) return t4_bbbb(q, ptCut);
  if (b1 && ((b2 && e3) || (e2 && e3)) && e4) return t4_bbee(q, ptCut);
  if (e1 && e2 && e3 && e4) return t4_eeee(q, ptCut);
  return false;
}

// ---------------------------------------------------------------- radius checks
// brute-force radius range over the strip-edge positions (SDL/Quintuplet.h:753-782)
__device__ __forceinline__ void radius_range(const float* xa, const float* ya, const float* xb, const float* yb,
                                             const float* xc, const float* yc, float& rmin, float& rmax) {
  rmin = kSdlInf;
  rmax = 0.f;
  float g, f;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        const float r = circle_radius3(xa[i], ya[i], xb[j], yb[j], xc[k], yc[k], g, f);
        rmax = m_max(r, rmax);
        rmin = m_min(r, rmin);
      }
}

// inverse-radius window overlap of the inner and bridge circles
// (SDL/Quintuplet.h:785-1134; nine categories share this shape)
__device__ __forceinline__ bool radii_compatible(int category, float innerR, float bridgeR, float iMin2S, float iMax2S,
                                                 float bMin2S, float bMax2S) {
  // category: 0 BBBBB, 1 BBBBE, 2 BBBEE12378, 3 BBBEE23478, 4 BBBEE34578, 5 BBEEE, 6 BEEEE, 7 EEEEE
  float ib, bb;
  bool high = innerR > 2.0f / (2.f * k2Rinv1GeVf);
  switch (category) {
    case 0: ib = high ? 0.4449f : 0.1512f; bb = high ? 0.4033f : 0.1781f; break;
    case 1: ib = high ? 0.4750f : 0.1781f; bb = high ? 0.3903f : 0.2167f; break;
    case 2: ib = 0.178f; bb = 0.507f; break;
    case 3: ib = 0.2097f; bb = 0.8557f; break;
    case 4: ib = 0.066f; bb = 0.617f; break;
    case 5: ib = high ? 12.9173f : 0.6376f; bb = high ? 5.1700f : 2.1381f; break;
    case 6: ib = high ? 23.2713f : 1.9382f; bb = high ? 21.7980f : 3.7280f; break;
    default: ib = high ? 22.5226f : 1.9382f; bb = high ? 21.0966f : 2.2091f; break;
  }
  float iLo = m_max(0.f, (1.f - ib) / innerR), iHi = (1.f + ib) / innerR;
  float bLo = m_max(0.f, (1.f - bb) / bridgeR), bHi = (1.f + bb) / bridgeR;
  if (category >= 2) {  // widen the bridge window by the 2S brute-force range
    bLo = m_min(bLo, 1.0f / bMax2S);
    bHi = m_max(bHi, 1.0f / bMin2S);
  }
  if (category >= 6) {  // ... and the inner window (double literal 1.0 there; the float result is identical)
    iLo = m_min(iLo, 1.0f / iMax2S);
    iHi = m_max(iHi, 1.0f / iMin2S);
  }
  return (iLo <= bLo && bLo < iHi) || (bLo < iLo && iLo < bHi);
}

// ------------------------------------------------------------- rz chi-squared
// SDL/Quintuplet.h:302-735; returns pass, sets chi2 and the tight flag
__device__ __forceinline__ bool t5_rz_chi2(const ModLite* mod, const MDPoint* p, float innerPt, float innerRadius,
                                           float g, float f, float& chi2, bool& tight) {
  float rt[5], z[5];
  for (int k = 0; k < 5; k++) { rt[k] = p[k].rt / 100; z[k] = p[k].z / 100; }
  const float x1 = p[0].x / 100, x2 = p[1].x / 100, x3 = p[2].x / 100, x4 = p[3].x / 100;
  const float y1 = p[0].y / 100, y2 = p[1].y / 100, y3 = p[2].y / 100, y4 = p[3].y / 100;
  const int t3 = mod[2].type;
  const float xc = g / 100, yc = f / 100;
  const int init = t3 == 1 ? 1 : 2;  // 2S third module: start from the second MD
  const float xi = p[init].x / 100, yi = p[init].y / 100, zi = p[init].z / 100, rti = p[init].rt / 100;
  int charge = 0;
  const float s3 = (y3 - yc) / (x3 - xc), s1 = (y1 - yc) / (x1 - xc);
  if ((y3 - yc) > 0 && (y1 - yc) > 0) {
    if (s1 > 0 && s3 < 0) charge = -1;
    else if (s1 < 0 && s3 > 0) charge = 1;
    else if (s3 > s1) charge = -1;
    else if (s3 < s1) charge = 1;
  } else if ((y3 - yc) < 0 && (y1 - yc) < 0) {
    if (s1 < 0 && s3 > 0) charge = 1;
    else if (s1 > 0 && s3 < 0) charge = -1;
    else if (s3 > s1) charge = -1;
    else if (s3 < s1) charge = 1;
  } else if ((y3 - yc) < 0 && (y1 - yc) > 0) {
    if ((x3 - xc) > 0 && (x1 - xc) > 0) charge = 1;
    else if ((x3 - xc) < 0 && (x1 - xc) < 0) charge = -1;
  } else if ((y3 - yc) > 0 && (y1 - yc) < 0) {
    if ((x3 - xc) > 0 && (x1 - xc) > 0) charge = -1;
    else if ((x3 - xc) < 0 && (x1 - xc) < 0) charge = 1;
  }
  const float psi = m_atan((yi - yc) / (xi - xc));
  const float Pt = innerPt;
  float Px = Pt * m_abs(m_sin(psi));
  // the reference's unqualified cos() resolves to double cos on the CPU build
  float Py = (float)((double)Pt * ::fabs(::cos((double)psi)));
  if (xi > xc && yi > yc) { if (charge == 1) Py = -Py; if (charge == -1) Px = -Px; }
  if (xi < xc && yi > yc) { if (charge == -1) { Px = -Px; Py = -Py; } }
  if (xi < xc && yi < yc) { if (charge == 1) Px = -Px; if (charge == -1) Py = -Py; }
  if (xi > xc && yi < yc) { if (charge == 1) { Px = -Px; Py = -Py; } }
  if (t3 == 0) {
    if (x4 < x3 && x3 < x2) Px = -m_abs(Px);
    if (x4 > x3 && x3 > x2) Px = m_abs(Px);
    if (y4 < y3 && y3 < y2) Py = -m_abs(Py);
    if (y4 > y3 && y3 > y2) Py = m_abs(Py);
  } else if (t3 == 1) {
    if (x3 < x2 && x2 < x1) Px = -m_abs(Px);
    if (x3 > x2 && x2 > x1) Px = m_abs(Px);
    if (y3 < y2 && y2 < y1) Py = -m_abs(Py);
    if (y3 > y2 && y2 > y1) Py = m_abs(Py);
  }
  const float AO = m_sqrt((x1 - xc) * (x1 - xc) + (y1 - yc) * (y1 - yc));
  const float BO = m_sqrt((xi - xc) * (xi - xc) + (yi - yc) * (yi - yc));
  const float AB = m_sqrt((x1 - xi) * (x1 - xi) + (y1 - yi) * (y1 - yi));
  const float dPhi = m_acos((AO * AO + BO * BO - AB * AB) / (2 * AO * BO));
  const float ds = innerRadius / 100 * dPhi;
  const float Pz = (zi - z[0]) / ds * Pt;
  const float pp = m_sqrt(Px * Px + Py * Py + Pz * Pz);
  const float a = -0.299792 * kMagneticField * charge;
  chi2 = 0;
  for (int i = 2; i < 6; i++) {
    if (t3 == 0 ? i == 3 : i == 2) continue;
    const int k = i - 1;
    const float zs = z[k], rts = rt[k];
    const int layer = mod[k].sdl, type = mod[k].type;
    const float rou = a / pp;
    const float s = (zs - zi) * pp / Pz;
    const float x = xi + Px / a * m_sin(rou * s) - Py / a * (1 - m_cos(rou * s));
    const float y = yi + Py / a * m_sin(rou * s) + Px / a * (1 - m_cos(rou * s));
    const float diffr = (rts - m_sqrt(x * x + y * y)) * 100;
    float diffz = 0;
    if (layer <= 6) {
      const float pA = rti * rti + 2 * (Px * Px + Py * Py) / (a * a) + 2 * (yi * Px - xi * Py) / a - rts * rts;
      const float pB = 2 * (xi * Px + yi * Py) / a;
      const float pC = 2 * (yi * Px - xi * Py) / a + 2 * (Px * Px + Py * Py) / (a * a);
      const float A = pB * pB + pC * pC, B = 2 * pA * pB, C = pA * pA - pC * pC;
      const float sol1 = (-B + m_sqrt(B * B - 4 * A * C)) / (2 * A);
      const float sol2 = (-B - m_sqrt(B * B - 4 * A * C)) / (2 * A);
      const float dz1 = (m_asin(sol1) / rou * Pz / pp + zi - zs) * 100;
      const float dz2 = (m_asin(sol2) / rou * Pz / pp + zi - zs) * 100;
      if (isnan(dz1)) diffz = dz2;
      else if (isnan(dz2)) diffz = dz1;
      else diffz = m_abs(dz1) < m_abs(dz2) ? dz1 : dz2;
    }
    float residual = layer > 6 ? diffr : diffz;
    float error = type == 0 ? 0.15f : 5.0f;
    if (i == 2 || i == 3) {
      const float drdz = m_abs(mod[k].drdz);
      const bool flat = mod[k].subdet == kEndcap || mod[k].side == kCenter;
      residual = (layer <= 6 && (mod[k].side == kCenter || drdz < 1)) ? diffz : diffr;
      float proj = 1;
      if (drdz < 1) proj = flat ? 1.f : 1 / m_sqrt(1 + drdz * drdz);
      if (drdz > 1) proj = flat ? 1.f : drdz / m_sqrt(1 + drdz * drdz);
      error = error * proj;
    }
    chi2 += 12 * (residual * residual) / (error * error);
  }
  const int l1 = mod[0].sdl, l2 = mod[1].sdl, l3 = mod[2].sdl, l4 = mod[3].sdl, l5 = mod[4].sdl;
  if (innerPt > 100 || isnan(chi2)) {  // straight-line fallback
    const float slope = (mod[0].type == 0 && mod[1].type == 0 && t3 == 1) ? (z[1] - z[0]) / (rt[1] - rt[0])
                                                                          : (z[2] - z[0]) / (rt[2] - rt[0]);
    float r4 = l4 <= 6 ? ((z[3] - z[0]) - slope * (rt[3] - rt[0])) : ((rt[3] - rt[0]) - (z[3] - z[0]) / slope);
    float r5 = l4 <= 6 ? ((z[4] - z[0]) - slope * (rt[4] - rt[0])) : ((rt[4] - rt[0]) - (z[4] - z[0]) / slope);
    r4 = mod[3].type == 0 ? r4 / 0.15f : r4 / 5.0f;
    r5 = mod[4].type == 0 ? r5 / 0.15f : r5 / 5.0f;
    r4 = r4 * 100;
    r5 = r5 * 100;
    chi2 = 12 * (r4 * r4 + r5 * r5);
    return chi2 < 4.677f;
  }
  tight = false;
  // (tight, loose) thresholds per layer pattern; loose < 0 means no loose cut
  float tc = -1, lc = -1;
  const int key = ((l1 * 20 + l2) * 20 + l3) * 400 + l4 * 20 + l5;
  auto K = [](int a, int b, int c, int d, int e) { return ((a * 20 + b) * 20 + c) * 400 + d * 20 + e; };
  if (key == K(7, 8, 9, 10, 11)) { tc = 94.470f; }
  else if (key == K(7, 8, 9, 10, 16)) { tc = 22.099f; lc = 37.956f; }
  else if (key == K(7, 8, 9, 15, 16)) { tc = 7.992f; lc = 11.622f; }
  else if (key == K(1, 7, 8, 9, 10)) { tc = 111.390f; }
  else if (key == K(1, 7, 8, 9, 15)) { tc = 18.351f; lc = 37.941f; }
  else if (key == K(1, 2, 7, 8, 9)) { tc = 116.148f; }
  else if (key == K(1, 2, 7, 8, 14)) { tc = 19.352f; lc = 52.561f; }
  else if (key == K(1, 2, 7, 13, 14)) { tc = 10.392f; lc = 13.76f; }
  else if (key == K(1, 2, 3, 7, 8)) { tc = 27.824f; lc = 44.247f; }
  else if (key == K(1, 2, 3, 7, 13)) { tc = 18.145f; lc = 33.752f; }
  else if (key == K(1, 2, 3, 12, 13)) { tc = 13.308f; lc = 21.213f; }
  else if (key == K(1, 2, 3, 4, 5)) { tc = 15.627f; lc = 29.035f; }
  else if (key == K(1, 2, 3, 4, 12)) { tc = 14.64f; lc = 23.037f; }
  else if (key == K(2, 7, 8, 9, 15)) { tc = 24.662f; lc = 41.036f; }
  else if (key == K(2, 7, 8, 14, 15)) { tc = 8.866f; lc = 14.092f; }
  else if (key == K(2, 3, 7, 8, 14)) { tc = 23.730f; lc = 23.748f; }
  else if (key == K(2, 3, 7, 13, 14)) { tc = 10.772f; lc = 17.945f; }
  else if (key == K(2, 3, 4, 5, 6)) { tc = 6.065f; lc = 8.803f; }
  else if (key == K(2, 3, 4, 5, 12)) { tc = 5.693f; lc = 7.930f; }
  else if (key == K(2, 3, 4, 12, 13)) { tc = 5.473f; lc = 7.626f; }
  if (tc >= 0 && chi2 < tc) tight = true;
  return lc < 0 ? true : chi2 < lc;
}

// r-phi chi2 thresholds: SDL/Quintuplet.h:218-298 (returns true where the reference applies no cut)
__device__ __forceinline__ bool t5_rphi_chi2_ok(int l1, int l2, int l3, int l4, int l5, float chi2) {
  auto K = [](int a, int b, int c, int d, int e) { return ((a * 20 + b) * 20 + c) * 400 + d * 20 + e; };
  const int key = K(l1, l2, l3, l4, l5);
  float cut = -1;
  if (key == K(7, 8, 9, 10, 11)) cut = 0.01788f;
  else if (key == K(7, 8, 9, 10, 16) || key == K(7, 8, 9, 15, 16)) cut = 0.04725f;
  else if (key == K(1, 7, 8, 9, 10)) cut = 0.01788f;
  else if (key == K(1, 7, 8, 9, 15)) cut = 0.08234f;
  else if (key == K(1, 2, 7, 8, 9)) cut = 0.02360f;
  else if (key == K(1, 2, 7, 8, 14)) cut = 0.07167f;
  else if (key == K(1, 2, 7, 13, 14)) cut = 0.08234f;
  else if (key == K(1, 2, 3, 7, 8)) cut = 0.01026f;
  else if (key == K(1, 2, 3, 7, 13) || key == K(1, 2, 3, 12, 13)) cut = 0.06238f;
  else if (key == K(2, 7, 8, 9, 10)) cut = 0.00512f;
  else if (key == K(2, 7, 8, 9, 15)) cut = 0.04112f;
  else if (key == K(2, 7, 8, 14, 15)) cut = 0.06238f;
  else if (key == K(2, 3, 7, 8, 14)) cut = 0.07167f;
  else if (key == K(2, 3, 7, 13, 14)) cut = 0.06238f;
  else if (key == K(2, 3, 4, 5, 6)) cut = 0.08234f;
  else if (key == K(2, 3, 4, 5, 12) || key == K(2, 3, 4, 12, 13)) cut = 0.10870f;
  else if (key == K(3, 7, 8, 14, 15) || key == K(3, 4, 5, 12, 13)) cut = 0.09461f;
  return cut < 0 ? true : chi2 < cut;
}

// ---------------------------------------------------------- circle regression
// per-point uncertainties (SDL/Quintuplet.h:1136-1221)
// (the slope of a non-flat module enters through its rotation table, point_sigma)
__device__ __forceinline__ void regression_sigmas(const ModLite* mod, bool anchor, float* d1, float* d2, bool* flat) {
  const float inv1 = 0.01f / 0.009f, inv2 = 0.15f / 0.009f, inv3 = 2.4f / 0.009f;
  for (int i = 0; i < 5; i++) {
    const ModLite& m = mod[i];
    if (m.subdet == kBarrel && m.type == kPS && m.side == kCenter) {
      d1[i] = inv1; d2[i] = inv1; flat[i] = true;
    } else if (m.subdet == kBarrel && m.type == kTwoS) {
      d1[i] = 1.f; d2[i] = 1.f; flat[i] = true;
    } else if (m.subdet == kBarrel && m.type == kPS) {
      d1[i] = inv1; flat[i] = false;
      d2[i] = (anchor ? inv2 : inv3) * m.drdz / m_sqrt(1 + m.drdz * m.drdz);
    } else if (m.subdet == kEndcap && m.type == kPS) {
      d1[i] = inv1; flat[i] = false;
      d2[i] = anchor ? inv2 : inv3;
    } else if (m.subdet == kEndcap && m.type == kTwoS) {
      d1[i] = 1.f; d2[i] = 500.f * inv1; flat[i] = false;
    }
  }
}

// the r-phi uncertainty of point (x, y) on module m: the point is rotated by an angle of the module's
// slope and the point's quadrant, whose cosine and sine the geometry holds (sigRot); a point on an axis
// keeps angle 0 (cos 1, sin 0)
__device__ __forceinline__ float point_sigma(const ModulesView& M, unsigned m, float x, float y, float d1, float d2,
                                             bool flat) {
  float xp = x, yp = y;
  if (!flat) {
    const int q = (x > 0 && y > 0) ? 0 : (x < 0 && y > 0) ? 1 : (x < 0 && y < 0) ? 2 : (x > 0 && y < 0) ? 3 : 4;
    float c = 1.f, s = 0.f;
    if (q < 4) {
      const float2 r = ld_global(M.sigRot + 4 * m + q);
      c = r.x;
      s = r.y;
    }
    xp = x * c + y * s;
    yp = y * c - x * s;
  }
  return 2 * m_sqrt((xp * d1) * (xp * d1) + (yp * d2) * (yp * d2));
}

// weighted algebraic circle fit and its chi2 (SDL/Quintuplet.h:1223-1320)
__device__ __forceinline__ float regression_fit(const ModulesView& M, const uint16_t* mods, const float* x,
                                                const float* y, const float* d1, const float* d2, const bool* flat,
                                                float& g, float& f, float& chi2) {
  float sxx = 0.f, syy = 0.f, sxy = 0.f, sxr = 0.f, syr = 0.f, sr = 0.f, sx = 0.f, sy = 0.f, s1 = 0.f;
  float sig[5];
  for (int i = 0; i < 5; i++) {
    sig[i] = point_sigma(M, mods[i], x[i], y[i], d1[i], d2[i], flat[i]);
    const float w2 = sig[i] * sig[i];
    sxx += (x[i] * x[i]) / w2;
    syy += (y[i] * y[i]) / w2;
    sxy += (x[i] * y[i]) / w2;
    sxr += (x[i] * (x[i] * x[i] + y[i] * y[i])) / w2;
    syr += (y[i] * (x[i] * x[i] + y[i] * y[i])) / w2;
    sr += (x[i] * x[i] + y[i] * y[i]) / w2;
    sx += x[i] / w2;
    sy += y[i] / w2;
    s1 += 1.0f / w2;
  }
  const float den = (sxy - sx * sy) * (sxy - sx * sy) - (sxx - sx * sx) * (syy - sy * sy);
  const float twoG = ((syr - sy * sr) * (sxy - sx * sy) - (sxr - sx * sr) * (syy - sy * sy)) / den;
  const float twoF = ((sxr - sx * sr) * (sxy - sx * sy) - (syr - sy * sr) * (sxx - sx * sx)) / den;
  const float c = -(sr - twoG * sx - twoF * sy) / s1;
  g = 0.5f * twoG;
  f = 0.5f * twoF;
  if (g * g + f * f - c < 0) {
    chi2 = -1;
    return -1;
  }
  const float r = m_sqrt(g * g + f * f - c);
  chi2 = 0.f;
  for (int i = 0; i < 5; i++) {
    const float t = x[i] * x[i] + y[i] * y[i] - twoG * x[i] - twoF * y[i] + c;
    chi2 += t * t / (sig[i] * sig[i]);
  }
  return r;
}

// chi2 of points against a given circle (SDL/Quintuplet.h:1322-1367)
__device__ __forceinline__ float circle_chi2(const ModulesView& M, const uint16_t* mods, const float* x, const float* y,
                                             const float* d1, const float* d2, const bool* flat, float g, float f,
                                             float r) {
  const float c = g * g + f * f - r * r;
  float chi2 = 0.f;
  for (int i = 0; i < 5; i++) {
    const float s = point_sigma(M, mods[i], x[i], y[i], d1[i], d2[i], flat[i]);
    const float t = x[i] * x[i] + y[i] * y[i] - 2 * g * x[i] - 2 * f * y[i] + c;
    chi2 += t * t / (s * s);
  }
  return chi2;
}

struct T5Result {
  float innerR, bridgeR, outerR, regR, regG, regF, rzChi2, chi2, nonAnchorChi2;
  bool tight;
};

}  // namespace lst
