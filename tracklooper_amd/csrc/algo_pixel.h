// Selections of the pixel-seeded objects, evaluated by one lane per
// candidate: the pixel-segment/outer-segment compatibility test
// (SDL/PixelTriplet.h:1167-1869), the pixel-triplet selection
// (SDL/PixelTriplet.h:316-1024) and the pixel-quintuplet selection
// (SDL/PixelTriplet.h:2038-2677).
#pragma once
#include "algo_t5.h"
#include "stages.h"

namespace lst {

// sdl layer numbering of the pixel-object cut tables (SDL/PixelTriplet.h:321-333)
__device__ __forceinline__ int pix_layer(const ModulesView& M, unsigned m) {
  const bool ec = M.subdets[m] == kEndcap;
  return M.layers[m] + 6 * ec + 5 * (ec && M.moduleType[m] == kTwoS);
}

struct PixelSeed {  // the pLS quantities the selections read
  float ptIn, ptErr, px, py, pz, etaErr, g, f, radius;
  int charge;
};
__device__ __forceinline__ PixelSeed load_seed(const PLSSoA& P, unsigned i) {
  return PixelSeed{P.ptIn[i], P.ptErr[i], P.px[i], P.py[i], P.pz[i], P.etaErr[i],
                   P.circleCenterX[i], P.circleCenterY[i], P.circleRadius[i], P.charge[i]};
}

// pixel segment (anchor points a, b) against an outer-tracker segment (c, d)
// on modules mIn (lower) and mOut (upper); alphaIn/alphaOut are the two
// segments' dPhiChange; slopeC is c's slope asin(min(rt * k2Rinv1GeVf / ptCut,
// kSinAlphaMax)) and its tan, kept per mini-doublet
__device__ __forceinline__ bool pixel_tracklet(const ModulesView& M, const MDsSoA& D, const PixelSeed& s,
                                               uint32_t mdA, uint32_t mdB, uint32_t mdC, uint32_t mdD, unsigned mIn,
                                               unsigned mOut, float alphaIn, float alphaOut, float ptCut,
                                               float2 slopeC) {
  const int16_t subIn = M.subdets[mIn], subOut = M.subdets[mOut];
  const bool bb = subIn == kBarrel && (subOut == kBarrel || subOut == kEndcap);
  const bool ee = subIn == kEndcap && subOut == kEndcap;
  if (!bb && !ee) return false;
  const float x1 = D.anchorX[mdA], y1 = D.anchorY[mdA], rt1 = D.anchorRt[mdA];
  const float x2 = D.anchorX[mdB], y2 = D.anchorY[mdB], z2 = D.anchorZ[mdB], rt2 = D.anchorRt[mdB];
  const float x3 = D.anchorX[mdC], y3 = D.anchorY[mdC], z3 = D.anchorZ[mdC], rt3 = D.anchorRt[mdC];
  const float x4 = D.anchorX[mdD], y4 = D.anchorY[mdD], rt4 = D.anchorRt[mdD];
  const bool psOut = M.moduleType[mIn] == kPS;
  const float zpitchOut = psOut ? kPixelPSZpitch : kStrip2SZpitch;
  const float ptIn = s.ptIn, pz = s.pz;
  float thetaMuls, dPhi, drtInSeg;
  if (bb) {  // runTripletDefaultAlgoPPBB (:1264-1553)
    if (!(m_abs(delta_phi(x2, y2, x3, y3)) <= 0.5f * kPi)) return false;
    const float a1 = slopeC.x;
    const float ratio = rt3 / rt2;
    const float scale = slopeC.y / a1;
    const float zpIn = 0.05f, zpInOut = 0.05f;
    const float zGeom = zpIn + zpitchOut;
    const float zHi = z2 + (z2 + kDeltaZLum) * (ratio - 1.f) * (z2 < 0.f ? 1.f : scale) + (zpInOut + zpitchOut);
    const float zLo = z2 + (z2 - kDeltaZLum) * (ratio - 1.f) * (z2 > 0.f ? 1.f : scale) - (zpInOut + zpitchOut);
    if (!((z3 >= zLo) && (z3 <= zHi))) return false;
    const float coshEta = m_sqrt(ptIn * ptIn + pz * pz) / ptIn;
    const float drOutIn = rt3 - rt2;
    const float r3In = m_sqrt(z2 * z2 + rt2 * rt2);
    drtInSeg = rt2 - rt1;
    thetaMuls = 0.015f * m_sqrt(0.1f + 0.2f * (rt3 - rt2) / 50.f) * m_sqrt(r3In / rt2);
    const float muls = thetaMuls * 3.f / ptCut * 4.f;
    float dzErr = drOutIn * s.etaErr * coshEta;
    dzErr *= dzErr;
    dzErr += 0.03f * 0.03f;
    dzErr *= 9.f;
    dzErr += muls * muls * drOutIn * drOutIn / 3.f * coshEta * coshEta;
    dzErr += zGeom * zGeom;
    dzErr = m_sqrt(dzErr);
    const float dzDrIn = pz / ptIn;
    const float zWin = dzErr / drtInSeg * drOutIn + zGeom;
    const float dzMean =
        dzDrIn * drOutIn * (1.f + drOutIn * drOutIn * 4 * k2Rinv1GeVf * k2Rinv1GeVf / ptIn / ptIn / 24.f);
    if (!((z3 >= z2 + dzMean - zWin) && (z3 <= z2 + dzMean + zWin))) return false;
    const float pvOff = 0.1f / rt3;
    const float cut = a1 + m_sqrt(muls * muls + pvOff * pvOff);
    dPhi = delta_phi(0.5f * (x1 + x3), 0.5f * (y1 + y3), x3 - x1, y3 - y1);
    if (!(m_abs(dPhi) <= cut)) return false;
  } else {  // runTripletDefaultAlgoPPEE (:1555-1869)
    if (!(z2 * z3 > 0)) return false;
    const float zpIn = 0.05f;
    const float zGeom = zpIn + zpitchOut;
    const float slope = slopeC.x;
    const float scale = slopeC.y / slope;
    const float dLum = sign_copy(kDeltaZLum, z2);
    const float rtGeom1 = psOut ? kPixelPSZpitch : kStrip2SZpitch;
    const float zGeom1 = sign_copy(zGeom, z2);
    const float rtLo = rt2 * (1.f + (z3 - z2 - zGeom1) / (z2 + zGeom1 + dLum) / scale) - rtGeom1;
    float zInForHi = z2 - zGeom1 - dLum;
    if (zInForHi * z2 < 0) zInForHi = sign_copy(0.1f, z2);
    const float rtHi = rt2 * (1.f + (z3 - z2 + zGeom1) / zInForHi) + rtGeom1;
    if (!((rt3 >= rtLo) && (rt3 <= rtHi))) return false;
    const float dzAbs = m_abs(z3 - z2);
    const float coshEta = m_sqrt(ptIn * ptIn + pz * pz) / ptIn;
    const float multDzDr = dzAbs * coshEta / (coshEta * coshEta - 1.f);
    const float r3In = m_sqrt(z2 * z2 + rt2 * rt2);
    thetaMuls = 0.015f * m_sqrt(0.1f + 0.2f * (rt3 - rt2) / 50.f) * m_sqrt(r3In / rt2);
    const float muls = thetaMuls * 3.f / ptCut * 4.f;
    float drtErr = s.etaErr * multDzDr;
    drtErr *= drtErr;
    drtErr += 0.03f * 0.03f;
    drtErr *= 9.f;
    drtErr += muls * muls * multDzDr * multDzDr / 3.f * coshEta * coshEta;
    drtErr = m_sqrt(drtErr);
    const float drtDzIn = m_abs(ptIn / pz);
    const float drOutIn = rt3 - rt2;
    const float rtWin = drtErr + rtGeom1;
    const float drtMean =
        drtDzIn * dzAbs * (1.f - drOutIn * drOutIn * 4 * k2Rinv1GeVf * k2Rinv1GeVf / ptIn / ptIn / 24.f);
    if (!((rt3 >= rt2 + drtMean - rtWin) && (rt3 <= rt2 + drtMean + rtWin))) return false;
    const float a1 = slopeC.x;
    const float pvOff = 0.1f / rt3;
    const float cut = a1 + m_sqrt(muls * muls + pvOff * pvOff);
    dPhi = delta_phi(0.5f * (x1 + x3), 0.5f * (y1 + y3), x3 - x1, y3 - y1);
    if (!(m_abs(dPhi) <= cut)) return false;
    drtInSeg = rt2 - rt1;
  }
  // beta cuts shared by both flavours
  const bool ecLast = M.subdets[mOut] == kEndcap && M.moduleType[mOut] == kTwoS;
  const float alphaOutUp = delta_phi(x4, y4, x4 - x3, y4 - y3);
  const float tx = x4 - x2, ty = y4 - y2;
  float betaIn = -delta_phi(s.px, s.py, tx, ty);
  float betaInMin = betaIn, betaInMax = betaIn;
  float betaOut = -alphaOutUp + delta_phi(x4, y4, tx, ty);
  float betaOutMin = betaOut, betaOutMax = betaOut;
  const float hx = D.anchorHighEdgeX[mdD], hy = D.anchorHighEdgeY[mdD];
  const float lx = D.anchorLowEdgeX[mdD], ly = D.anchorLowEdgeY[mdD];
  if (ecLast) {
    betaOutMin = -delta_phi(hx, hy, hx - x3, hy - y3) + delta_phi(hx, hy, hx - x2, hy - y2);
    betaOutMax = -delta_phi(lx, ly, lx - x3, ly - y3) + delta_phi(lx, ly, lx - x2, ly - y2);
  }
  const float dr = m_sqrt(tx * tx + ty * ty);
  const float drIn = m_sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
  float betaAv = 0.5f * (betaIn + betaOut);
  float ptBeta = ptIn;
  const float lOut = ecLast ? 11.f : 5.f;
  const float drOut = m_sqrt((x4 - x3) * (x4 - x3) + (y4 - y3) * (y4 - y3));
  const float dOut = rt4 - rt3;
  delta_beta_iterations(betaIn, betaOut, betaAv, ptBeta, drIn, drOut, dr, 0.f);
  const float sIn = m_abs(betaInMin + betaInMax) > 0 ? (2.f * betaIn / m_abs(betaInMin + betaInMax)) : 0.f;
  const float sOut = m_abs(betaOutMin + betaOutMax) > 0 ? (2.f * betaOut / m_abs(betaOutMin + betaOutMax)) : 0.f;
  betaInMin *= sIn; betaInMax *= sIn; betaOutMin *= sOut; betaOutMax *= sOut;
  const float mulsB = thetaMuls * 4.f / m_min(m_abs(ptBeta), kPtBetaMax);
  const float regOut = m_max(m_abs(alphaOut), m_asin(m_min(rt3 * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
  const float lumOut = lOut < 11 ? 0.0f : m_abs(regOut * kDeltaZLum / z3);
  const float lum2 = (0.0f + lumOut) * (0.0f + lumOut);  // the inner term is 0 for a pixel segment
  float dBetaROut = 0;
  if (ecLast) dBetaROut = (m_sqrt(hx * hx + hy * hy) - m_sqrt(lx * lx + ly * ly)) * m_sin(dPhi) / dr;
  const float outCut =
      m_asin(m_min(dr * k2Rinv1GeVf / ptCut, kSinAlphaMax)) + (0.02f / dOut) + m_sqrt(lum2 + mulsB * mulsB);
  if (!(m_abs(betaOut) < outCut)) return false;
  const float res = 0.02f / m_min(dOut, drtInSeg);
  const float spread = m_abs(betaInMin - betaInMax) + m_abs(betaOutMin - betaOutMax);
  const float cut2 = (res * res * 2.0f + mulsB * mulsB + lum2 + 0.f + dBetaROut * dBetaROut + 0.25f * spread * spread);
  const float dBeta = betaIn - betaOut;
  (void)alphaIn;
  return dBeta * dBeta <= cut2;
}

// radius-compatibility class of a T3 (SDL/PixelTriplet.h:615-734): 3 when its
// lower module is in the endcap, else 2 when its middle one is, else 1 when
// its upper one is, else 0
__device__ __forceinline__ int pt3_radius_kind(const ModulesView& M, unsigned lo, unsigned mid, unsigned up) {
  if (M.subdets[lo] == kEndcap) return 3;
  if (M.subdets[mid] == kEndcap) return 2;
  if (M.subdets[up] == kEndcap) return 1;
  return 0;
}

// SDL/PixelTriplet.h:615-734
__device__ __forceinline__ bool pt3_radius_ok(int kind, float pixR, float pixRErr, float t3R) {
  float tb, pb, tbHi, pbHi;
  if (kind == 3) { tb = 1.7006f; pb = 0.26367f; pbHi = 2.286f; tbHi = 2.436f; }
  else if (kind == 2) { tb = 1.59294f; pb = 0.255181f; pbHi = 2.2091f; tbHi = 2.3548f; }
  else if (kind == 1) { tb = 0.45972f; pb = 0.19644f; pbHi = 0.6805f; tbHi = 0.8557f; }
  else { tb = 0.15624f; pb = 0.17235f; pbHi = 0.6375f; tbHi = 0.6588f; }
  if (pixR > 2.0f * kR1GeVf) { pb = pbHi; tb = tbHi; }
  const float tMax = (1 + tb) / t3R;
  const float tMin = m_max((1 - tb) / t3R, 0.0f);
  const float pMax = m_max((1 + pb) / pixR, 1.f / (pixR - pixRErr));
  float pMin = m_min((1 - pb) / pixR, 1.f / (pixR + pixRErr));
  if (kind >= 2) pMin = m_max(pMin, 0.0f);
  return ((tMin <= pMin) && (pMin < tMax)) || ((pMin < tMin) && (tMin < pMax));
}

__device__ __forceinline__ bool pt3_rz_ok(int l1, int l2, int l3, float c) {  // :316-369
  if (l1 == 8 && l2 == 9 && l3 == 10) return c < 13.6067f;
  if (l1 == 8 && l2 == 9 && l3 == 15) return c < 5.5953f;
  if (l1 == 7 && l2 == 8 && l3 == 9) return c < 3.9263f;
  if (l1 == 1 && l2 == 2 && l3 == 3) return c < 9.4377f;
  if (l1 == 1 && l2 == 2 && l3 == 7) return c < 9.9975f;
  if (l1 == 1 && l2 == 7 && l3 == 8) return c < 8.6369f;
  if (l1 == 2 && l2 == 3 && l3 == 7) return c < 37.945f;
  if (l1 == 2 && l2 == 3 && l3 == 12) return c < 43.0167f;
  if (l1 == 2 && l2 == 3 && l3 == 4) return c < 8.6923f;
  if (l1 == 2 && l2 == 7 && l3 == 8) return c < 11.9672f;
  if (l1 == 2 && l2 == 7 && l3 == 13) return c < 16.2133f;
  return true;
}
__device__ __forceinline__ bool pt3_rphi_ok(int l1, int l2, int l3, float c) {  // :501-546
  if (l1 == 8 && l2 == 9 && l3 == 10) return c < 7.003f;
  if (l1 == 8 && l2 == 9 && l3 == 15) return c < 0.5f;
  if (l1 == 7 && l2 == 8 && l3 == 9) return c < 8.046f;
  if (l1 == 7 && l2 == 8 && l3 == 14) return c < 0.575f;
  if (l1 == 1 && l2 == 2 && l3 == 7) return c < 5.304f;
  if (l1 == 1 && l2 == 2 && l3 == 3) return c < 10.6211f;
  if (l1 == 1 && l2 == 7 && l3 == 8) return c < 4.617f;
  if (l1 == 2 && l2 == 7 && l3 == 8) return c < 8.046f;
  if (l1 == 2 && l2 == 7 && l3 == 13) return c < 0.435f;
  if (l1 == 2 && l2 == 3 && l3 == 7) return c < 9.244f;
  if (l1 == 2 && l2 == 3 && l3 == 12) return c < 0.287f;
  if (l1 == 2 && l2 == 3 && l3 == 4) return c < 18.509f;
  return true;
}
__device__ __forceinline__ bool pt3_rphi_in_ok(int l1, int l2, int l3, float c) {  // :548-605
  if (l1 == 7 && l2 == 8 && l3 == 9) return c < 22016.8055f;
  if (l1 == 7 && l2 == 8 && l3 == 14) return c < 935179.56807f;
  if (l1 == 8 && l2 == 9 && l3 == 10) return c < 29064.12959f;
  if (l1 == 8 && l2 == 9 && l3 == 15) return c < 935179.5681f;
  if (l1 == 1 && l2 == 2 && l3 == 3) return c < 1370.0113195101474f;
  if (l1 == 1 && l2 == 2 && l3 == 7) return c < 5492.110048314815f;
  if (l1 == 2 && l2 == 3 && l3 == 4) return c < 4160.410806470067f;
  if (l1 == 1 && l2 == 7 && l3 == 8) return c < 29064.129591225726f;
  if (l1 == 2 && l2 == 3 && l3 == 7) return c < 12634.215376250893f;
  if (l1 == 2 && l2 == 3 && l3 == 12) return c < 353821.69361145404f;
  if (l1 == 2 && l2 == 7 && l3 == 8) return c < 33393.26076341235f;
  if (l1 == 2 && l2 == 7 && l3 == 13) return c < 935179.5680742573f;
  return true;
}

// helix r-z residuals of the three triplet points (SDL/PixelTriplet.h:736-824)
__device__ __forceinline__ float pt3_rz_chi2(const ModulesView& M, const uint16_t* mods, float xPix1, float yPix1,
                                             float zPix1, float rtPix1, const float* rts, const float* zs, float Px,
                                             float Py, float Pz, int charge) {
  float rmse = 0;
  const float x1 = xPix1 / 100, y1 = yPix1 / 100, z1 = zPix1 / 100, r1 = rtPix1 / 100;
  const float a = -0.299792 * kMagneticField * charge;
  for (int i = 0; i < 3; i++) {
    const float zsi = zs[i] / 100, rtsi = rts[i] / 100;
    const unsigned m = mods[i];
    const int type = M.moduleType[m], side = M.sides[m], subdet = M.subdets[m];
    float diffr = 0, diffz = 0;
    const float p = m_sqrt(Px * Px + Py * Py + Pz * Pz);
    const float rou = a / p;
    if (subdet == kEndcap) {
      const float sArc = (zsi - z1) * p / Pz;
      const float x = x1 + Px / a * m_sin(rou * sArc) - Py / a * (1 - m_cos(rou * sArc));
      const float y = y1 + Py / a * m_sin(rou * sArc) + Px / a * (1 - m_cos(rou * sArc));
      diffr = m_abs(rtsi - m_sqrt(x * x + y * y)) * 100;
    }
    if (subdet == kBarrel) {
      const float pA = r1 * r1 + 2 * (Px * Px + Py * Py) / (a * a) + 2 * (y1 * Px - x1 * Py) / a - rtsi * rtsi;
      const float pB = 2 * (x1 * Px + y1 * Py) / a;
      const float pC = 2 * (y1 * Px - x1 * Py) / a + 2 * (Px * Px + Py * Py) / (a * a);
      const float A = pB * pB + pC * pC, B = 2 * pA * pB, C = pA * pA - pC * pC;
      const float sol1 = (-B + m_sqrt(B * B - 4 * A * C)) / (2 * A);
      const float sol2 = (-B - m_sqrt(B * B - 4 * A * C)) / (2 * A);
      const float dz1 = m_abs(m_asin(sol1) / rou * Pz / p + z1 - zsi) * 100;
      const float dz2 = m_abs(m_asin(sol2) / rou * Pz / p + z1 - zsi) * 100;
      diffz = m_min(dz1, dz2);
    }
    const float residual = subdet == kBarrel ? diffz : diffr;
    float error = type == 0 ? 0.15f : 5.0f;
    if (type == 0 && subdet == kBarrel && side != kCenter) {
      const float drdz = M.drdzs[m];
      error /= m_sqrt(1 + drdz * drdz);
    }
    rmse += (residual * residual) / (error * error);
  }
  return m_sqrt(0.2f * rmse);
}

// per-point uncertainty table shared by the pT3/pT5 r-phi chi2 (anchor hits)
__device__ __forceinline__ void pixel_sigmas(const ModulesView& M, const uint16_t* mods, int n, float* d1, float* d2,
                                             bool* flat) {
  const float inv1 = 0.01f / 0.009f, inv2 = 0.15f / 0.009f;
  for (int i = 0; i < n; i++) {
    const unsigned m = mods[i];
    const int16_t sub = M.subdets[m], side = M.sides[m];
    const uint8_t type = M.moduleType[m];
    const float drdz = M.drdzs[m];
    d1[i] = 0; d2[i] = 0; flat[i] = false;
    if (sub == kBarrel && type == kPS && side == kCenter) {
      d1[i] = inv1; d2[i] = inv1; flat[i] = true;
    } else if (sub == kBarrel && type == kTwoS) {
      d1[i] = 1.f; d2[i] = 1.f; flat[i] = true;
    } else if (sub == kBarrel && type == kPS) {
      d1[i] = inv1; d2[i] = inv2 * drdz / m_sqrt(1 + drdz * drdz);
    } else if (sub == kEndcap && type == kPS) {
      d1[i] = inv1; d2[i] = inv2;
    } else if (sub == kEndcap && type == kTwoS) {
      d1[i] = 1.f; d2[i] = 500.f * inv1;
    }
  }
}

// r-phi uncertainty of each point (SDL/PixelTriplet.h:419-487, 2218-2387)
__device__ __forceinline__ void point_sigmas(const ModulesView& M, const uint16_t* mods, int n, const float* x,
                                             const float* y, float* sig) {
  float d1[5], d2[5];
  bool flat[5];
  pixel_sigmas(M, mods, n, d1, d2, flat);
  for (int i = 0; i < n; i++) sig[i] = point_sigma(M, mods[i], x[i], y[i], d1[i], d2[i], flat[i]);
}

// chi2 of points against the circle (g, f, r), sigma per point given
__device__ __forceinline__ float points_chi2(int n, const float* x, const float* y, const float* sig, float g,
                                             float f, float r) {
  const float c = g * g + f * f - r * r;
  float chi2 = 0.f;
  for (int i = 0; i < n; i++) {
    const float t = x[i] * x[i] + y[i] * y[i] - 2 * g * x[i] - 2 * f * y[i] + c;
    chi2 += t * t / (sig[i] * sig[i]);
  }
  return chi2;
}

struct PT3Result {
  float pixR, t3R, rz, rphi, rphiIn;
};

// Per-T3 quantities of the pixel-triplet selection that do not depend on the
// pLS, gathered once per T3 of a pixel-connected module (kT3RecWords words):
//   [0,3) x and [3,6) y of the anchor points of its three mini-doublets,
//   [6,9) their r-phi sigmas, [9] circle radius, [10] circle centre x,
//   [11] centre y, [12] packed: bits 0-1 radius class, bits 2-16 the pixel-
//   object layer numbers of its three modules (5 bits each), bit 17 middle
//   module is 2S
struct T3Rec {
  float x[3], y[3], sig[3], r, g, f;
  uint32_t packed;
};
__device__ __forceinline__ int rec_layer(uint32_t packed, int shift, int k) { return (packed >> (shift + 5 * k)) & 31; }

__device__ __forceinline__ void make_t3_rec(const EventView& v, uint32_t t3, float* out) {
  const ModulesView& M = v.M;
  // the record's chunks and the segments' and mini-doublets' first chunks, each one 16-byte load
  const uint4 c0 = t3_chunk(v.T, t3, 0), c1 = t3_chunk(v.T, t3, 1), c3 = t3_chunk(v.T, t3, 3),
              c4 = t3_chunk(v.T, t3, 4);
  const uint16_t mods[3] = {(uint16_t)c0.z, (uint16_t)c0.w, (uint16_t)c1.x};
  const uint4 sIn = seg_chunk(v.S, c0.x, 0), sOut = seg_chunk(v.S, c0.y, 0);
  const uint32_t md[3] = {sIn.x, sIn.y, sOut.y};
  float x[3], y[3], sig[3];
  for (int k = 0; k < 3; k++) {
    const uint4 p = md_chunk(v.D, md[k], 1);
    x[k] = __uint_as_float(p.x);
    y[k] = __uint_as_float(p.y);
  }
  point_sigmas(M, mods, 3, x, y, sig);
  uint32_t packed = (uint32_t)pt3_radius_kind(M, mods[0], mods[1], mods[2]);
  for (int k = 0; k < 3; k++) packed |= (uint32_t)pix_layer(M, mods[k]) << (2 + 5 * k);
  if (M.moduleType[mods[1]] == kTwoS) packed |= 1u << 17;
  float4* o = reinterpret_cast<float4*>(out);
  o[0] = make_float4(x[0], x[1], x[2], y[0]);
  o[1] = make_float4(y[1], y[2], sig[0], sig[1]);
  o[2] = make_float4(sig[2], __uint_as_float(c3.w), __uint_as_float(c4.x), __uint_as_float(c4.y));
  o[3] = make_float4(__uint_as_float(packed), 0.f, 0.f, 0.f);
}

__device__ __forceinline__ T3Rec t3_rec_of(float4 a, float4 b, float4 c, float4 d) {
  T3Rec r;
  r.x[0] = a.x; r.x[1] = a.y; r.x[2] = a.z;
  r.y[0] = a.w; r.y[1] = b.x; r.y[2] = b.y;
  r.sig[0] = b.z; r.sig[1] = b.w; r.sig[2] = c.x;
  r.r = c.y; r.g = c.z; r.f = c.w;
  r.packed = __float_as_uint(d.x);
  return r;
}
__device__ __forceinline__ T3Rec load_t3_rec(const float* rec) {
  const float4* q = reinterpret_cast<const float4*>(rec);
  return t3_rec_of(ld_global(q), ld_global(q + 1), ld_global(q + 2), ld_global(q + 3));
}

// the two pixel-tracklet tests of runPixelTripletDefaultAlgo (SDL/PixelTriplet.h:
// 826-1024): the pLS against the inner and against the outer segment of T3 t3
__device__ __forceinline__ bool pt3_tracklets(const EventView& v, uint32_t pls, uint32_t arr, uint32_t t3) {
  const ModulesView& M = v.M;
  // record chunks: T3 {segments, modules 0-1}, {module 2, ...}; segment {MDs, modules}, {..., dPhiChange, ...}
  const uint4 t0 = t3_chunk(v.T, t3, 0);
  const uint16_t lo = (uint16_t)t0.z, mid = (uint16_t)t0.w, up = (uint16_t)t3_chunk(v.T, t3, 1).x;
  const uint4 sp = seg_chunk(v.S, pls, 0), sIn = seg_chunk(v.S, t0.x, 0), sOut = seg_chunk(v.S, t0.y, 0);
  const PixelSeed s = load_seed(v.P, arr);
  const float alphaPix = __uint_as_float(seg_chunk(v.S, pls, 2).y);
  if (!pixel_tracklet(M, v.D, s, sp.x, sp.y, sIn.x, sIn.y, lo, mid, alphaPix,
                      __uint_as_float(seg_chunk(v.S, t0.x, 2).y), v.ptCut, md_slope(v.D, sIn.x)))
    return false;
  return pixel_tracklet(M, v.D, s, sp.x, sp.y, sOut.x, sOut.y, mid, up, alphaPix,
                        __uint_as_float(seg_chunk(v.S, t0.y, 2).y), v.ptCut, md_slope(v.D, sOut.x));
}

// one of the two pixel-tracklet tests of pt3_tracklets: part 0 the inner
// segment of the T3, part 1 the outer one
__device__ __forceinline__ bool pt3_tracklet_half(const EventView& v, uint32_t pls, uint32_t arr, uint32_t t3,
                                                  uint32_t part) {
  const ModulesView& M = v.M;
  const uint4 t0 = t3_chunk(v.T, t3, 0);  // {segments, modules 0-1}
  const uint16_t mA = (uint16_t)(part ? t0.w : t0.z);
  const uint16_t mB = (uint16_t)(part ? t3_chunk(v.T, t3, 1).x : t0.w);
  const uint32_t seg = part ? t0.y : t0.x;
  const uint4 sp = seg_chunk(v.S, pls, 0), sg = seg_chunk(v.S, seg, 0);  // {MDs, modules}
  const PixelSeed s = load_seed(v.P, arr);
  const float alphaPix = __uint_as_float(seg_chunk(v.S, pls, 2).y);  // dPhiChange
  return pixel_tracklet(M, v.D, s, sp.x, sp.y, sg.x, sg.y, mA, mB, alphaPix,
                        __uint_as_float(seg_chunk(v.S, seg, 2).y), v.ptCut, md_slope(v.D, sg.x));
}

// runPixelTripletDefaultAlgo (SDL/PixelTriplet.h:826-1024) is the AND of pure
// cuts, evaluated here in two parts: pt3_cheap (radius compatibility, the
// r-phi and inward r-phi chi2 cuts) and pt3_costly (the two pixel-tracklet
// tests and the r-z chi2 cut), which need double-precision transcendentals.
// Either order gives the same pass/fail and the same values.  `pls` is the
// global segment index of the pLS, `arr` its pLS array index.
// the pLS values the cheap cuts read, loaded once per pLS by the builders
// that evaluate many candidates of one pLS
struct PlsCtx {
  float ptIn, ptErr, cx, cy, pixR;        // seed pT and its error, the pLS circle
  float xa, ya, za, ra, xb, yb, zb, rb;   // anchor x, y, z, rt of its inner and outer mini-doublet
};
__device__ __forceinline__ PlsCtx load_pls_ctx(const EventView& v, uint32_t pls, uint32_t arr) {
  PlsCtx p;
  const uint32_t pixIn = v.S.mdIndices[2 * pls], pixOut = v.S.mdIndices[2 * pls + 1];
  p.ptIn = v.P.ptIn[arr]; p.ptErr = v.P.ptErr[arr];
  p.cx = v.P.circleCenterX[arr]; p.cy = v.P.circleCenterY[arr]; p.pixR = v.P.circleRadius[arr];
  p.xa = v.D.anchorX[pixIn]; p.ya = v.D.anchorY[pixIn]; p.za = v.D.anchorZ[pixIn]; p.ra = v.D.anchorRt[pixIn];
  p.xb = v.D.anchorX[pixOut]; p.yb = v.D.anchorY[pixOut]; p.zb = v.D.anchorZ[pixOut]; p.rb = v.D.anchorRt[pixOut];
  return p;
}

__device__ __forceinline__ bool pt3_cheap(const PlsCtx& p, const T3Rec& t, bool chi2Cuts, PT3Result& r) {
  const float ptIn = p.ptIn;
  r.pixR = ptIn * kR1GeVf;
  const float pixRErr = p.ptErr * kR1GeVf;
  r.t3R = t.r;
  r.rz = -1;
  if (!pt3_radius_ok(t.packed & 3, r.pixR, pixRErr, r.t3R)) return false;
  const int l1 = rec_layer(t.packed, 2, 0), l2 = rec_layer(t.packed, 2, 1), l3 = rec_layer(t.packed, 2, 2);
  const bool cuts = chi2Cuts && ptIn < 5.0f;
  r.rphi = points_chi2(3, t.x, t.y, t.sig, p.cx, p.cy, p.pixR);
  if (cuts && !pt3_rphi_ok(l1, l2, l3, r.rphi)) return false;
  {
    const float xa = p.xa, ya = p.ya, xb = p.xb, yb = p.yb;
    const float g = t.g, f = t.f, rr = r.t3R;
    float res = (xa - g) * (xa - g) + (ya - f) * (ya - f) - rr * rr;
    float c2 = res * res;
    res = (xb - g) * (xb - g) + (yb - f) * (yb - f) - rr * rr;
    c2 += res * res;
    c2 *= 0.5f;
    r.rphiIn = c2;
  }
  return !(cuts && !pt3_rphi_in_ok(l1, l2, l3, r.rphiIn));
}

__device__ __forceinline__ bool pt3_costly(const EventView& v, uint32_t pls, uint32_t arr, uint32_t t3, bool chi2Cuts,
                                           PT3Result& r) {
  if (!pt3_tracklets(v, pls, arr, t3)) return false;
  const PixelSeed s = load_seed(v.P, arr);
  if (!(chi2Cuts && s.ptIn < 5.0f)) return true;
  const ModulesView& M = v.M;
  const uint16_t mods[3] = {v.T.lowerModuleIndices[3 * t3], v.T.lowerModuleIndices[3 * t3 + 1],
                            v.T.lowerModuleIndices[3 * t3 + 2]};
  const uint32_t segIn = v.T.segmentIndices[2 * t3], segOut = v.T.segmentIndices[2 * t3 + 1];
  const uint32_t md1 = v.S.mdIndices[2 * segIn], md2 = v.S.mdIndices[2 * segIn + 1],
                 mdC = v.S.mdIndices[2 * segOut + 1];
  const uint32_t pixOut = v.S.mdIndices[2 * pls + 1];
  const float rts[3] = {v.D.anchorRt[md1], v.D.anchorRt[md2], v.D.anchorRt[mdC]};
  const float zs[3] = {v.D.anchorZ[md1], v.D.anchorZ[md2], v.D.anchorZ[mdC]};
  r.rz = pt3_rz_chi2(M, mods, v.D.anchorX[pixOut], v.D.anchorY[pixOut], v.D.anchorZ[pixOut], v.D.anchorRt[pixOut], rts,
                     zs, s.px, s.py, s.pz, s.charge);
  return pt3_rz_ok(pix_layer(M, mods[0]), pix_layer(M, mods[1]), pix_layer(M, mods[2]), r.rz);
}

// pt3_costly in two parts that can run side by side: part 0 the first
// pixel-tracklet test, part 1 the second and the r-z chi2 cut (filling r.rz)
__device__ __forceinline__ bool pt3_costly_half(const EventView& v, uint32_t pls, uint32_t arr, uint32_t t3,
                                                bool chi2Cuts, PT3Result& r, uint32_t part) {
  if (!pt3_tracklet_half(v, pls, arr, t3, part)) return false;
  if (part == 0) return true;
  const PixelSeed s = load_seed(v.P, arr);
  if (!(chi2Cuts && s.ptIn < 5.0f)) return true;
  const ModulesView& M = v.M;
  const uint16_t mods[3] = {v.T.lowerModuleIndices[3 * t3], v.T.lowerModuleIndices[3 * t3 + 1],
                            v.T.lowerModuleIndices[3 * t3 + 2]};
  const uint32_t segIn = v.T.segmentIndices[2 * t3], segOut = v.T.segmentIndices[2 * t3 + 1];
  const uint32_t md1 = v.S.mdIndices[2 * segIn], md2 = v.S.mdIndices[2 * segIn + 1],
                 mdC = v.S.mdIndices[2 * segOut + 1];
  const uint32_t pixOut = v.S.mdIndices[2 * pls + 1];
  const float rts[3] = {v.D.anchorRt[md1], v.D.anchorRt[md2], v.D.anchorRt[mdC]};
  const float zs[3] = {v.D.anchorZ[md1], v.D.anchorZ[md2], v.D.anchorZ[mdC]};
  r.rz = pt3_rz_chi2(M, mods, v.D.anchorX[pixOut], v.D.anchorY[pixOut], v.D.anchorZ[pixOut], v.D.anchorRt[pixOut], rts,
                     zs, s.px, s.py, s.pz, s.charge);
  return pt3_rz_ok(pix_layer(M, mods[0]), pix_layer(M, mods[1]), pix_layer(M, mods[2]), r.rz);
}

__device__ __forceinline__ bool pt5_rz_ok(int l1, int l2, int l3, int l4, int l5, float c) {  // :2038-2126
  if (l1 == 1 && l2 == 2 && l3 == 3) {
    if (l4 == 12 && l5 == 13) return c < 451.141f;
    if (l4 == 4 && l5 == 12) return c < 392.654f;
    if (l4 == 4 && l5 == 5) return c < 225.322f;
    if (l4 == 7 && l5 == 13) return c < 595.546f;
    if (l4 == 7 && l5 == 8) return c < 196.111f;
  } else if (l1 == 1 && l2 == 2 && l3 == 7) {
    if (l4 == 13 && l5 == 14) return c < 297.446f;
    if (l4 == 8 && l5 == 14) return c < 451.141f;
    if (l4 == 8 && l5 == 9) return c < 518.339f;
  } else if (l1 == 1 && l2 == 7 && l3 == 8) {
    if (l4 == 9 && l5 == 10) return c < 341.75f;
    if (l4 == 9 && l5 == 15) return c < 341.75f;
  } else if (l1 == 2 && l2 == 3 && l3 == 4) {
    if (l4 == 12 && l5 == 13) return c < 392.655f;
    if (l4 == 5 && l5 == 12) return c < 341.75f;
    if (l4 == 5 && l5 == 6) return c < 112.537f;
  } else if (l1 == 2 && l2 == 3 && l4 == 7) {
    // the reference tests l4 == 7 here, so neither inner case can match
  } else if (l1 == 2 && l2 == 7 && l3 == 8) {
    if (l4 == 14 && l5 == 15) return c < 518.339f;
    if (l4 == 9 && l5 == 10) return c < 8.046f;
    if (l4 == 9 && l5 == 15) return c < 451.141f;
  } else if (l1 == 3 && l2 == 7 && l3 == 8 && l4 == 14 && l5 == 15) {
    return c < 56.207f;
  } else if (l1 == 7 && l2 == 8 && l3 == 9) {
    if (l4 == 10 && l5 == 11) return c < 64.578f;
    if (l4 == 10 && l5 == 16) return c < 85.250f;
    if (l4 == 15 && l5 == 16) return c < 85.250f;
  }
  return true;
}
__device__ __forceinline__ bool pt5_rphi_ok(int l1, int l2, int l3, int l4, int l5, float c) {  // :2128-2217
  if (l1 == 1 && l2 == 2 && l3 == 3) {
    if (l4 == 12 && l5 == 13) return c < 48.921f;
    if (l4 == 4 && l5 == 12) return c < 97.948f;
    if (l4 == 4 && l5 == 5) return c < 129.3f;
    if (l4 == 7 && l5 == 13) return c < 56.21f;
    if (l4 == 7 && l5 == 8) return c < 74.198f;
  } else if (l1 == 1 && l2 == 2 && l3 == 7) {
    if (l4 == 13 && l5 == 14) return c < 21.265f;
    if (l4 == 8 && l5 == 14) return c < 37.058f;
    if (l4 == 8 && l5 == 9) return c < 42.578f;
  } else if (l1 == 1 && l2 == 7 && l3 == 8) {
    if (l4 == 9 && l5 == 10) return c < 32.253f;
    if (l4 == 9 && l5 == 15) return c < 37.058f;
  } else if (l1 == 2 && l2 == 3 && l3 == 4) {
    if (l4 == 12 && l5 == 13) return c < 97.947f;
    if (l4 == 5 && l5 == 12) return c < 129.3f;
    if (l4 == 5 && l5 == 6) return c < 170.68f;
  } else if (l1 == 2 && l2 == 3 && l3 == 7) {
    if (l4 == 13 && l5 == 14) return c < 48.92f;
    if (l4 == 8 && l5 == 14) return c < 74.2f;
  } else if (l1 == 2 && l2 == 7 && l3 == 8) {
    if (l4 == 14 && l5 == 15) return c < 42.58f;
    if (l4 == 9 && l5 == 10) return c < 37.06f;
    if (l4 == 9 && l5 == 15) return c < 48.92f;
  } else if (l1 == 3 && l2 == 7 && l3 == 8 && l4 == 14 && l5 == 15) {
    return c < 85.25f;
  } else if (l1 == 7 && l2 == 8 && l3 == 9) {
    if (l4 == 10 && l5 == 11) return c < 42.58f;
    if (l4 == 10 && l5 == 16) return c < 37.06f;
    if (l4 == 15 && l5 == 16) return c < 37.06f;
  }
  return true;
}
__device__ __forceinline__ bool pt5_rphi_in_ok(int l1, int l2, int l3, int l4, int l5, float c) {  // :2389-2478
  if (l1 == 1 && l2 == 2 && l3 == 3) {
    if (l4 == 12 && l5 == 13) return c < 451.141f;
    if (l4 == 4 && l5 == 12) return c < 786.173f;
    if (l4 == 4 && l5 == 5) return c < 595.545f;
    if (l4 == 7 && l5 == 13) return c < 581.339f;
    if (l4 == 7 && l5 == 8) return c < 112.537f;
  } else if (l1 == 1 && l2 == 2 && l3 == 7) {
    if (l4 == 13 && l5 == 14) return c < 225.322f;
    if (l4 == 8 && l5 == 14) return c < 1192.402f;
    if (l4 == 8 && l5 == 9) return c < 786.173f;
  } else if (l1 == 1 && l2 == 7 && l3 == 8) {
    if (l4 == 9 && l5 == 10) return c < 1037.817f;
    if (l4 == 9 && l5 == 15) return c < 1808.536f;
  } else if (l1 == 2 && l2 == 3 && l3 == 4) {
    if (l4 == 12 && l5 == 13) return c < 684.253f;
    if (l4 == 5 && l5 == 12) return c < 684.253f;
    if (l4 == 5 && l5 == 6) return c < 684.253f;
  } else if (l1 == 2 && l2 == 3 && l3 == 7) {
    if (l4 == 13 && l5 == 14) return c < 451.141f;
    if (l4 == 8 && l5 == 14) return c < 518.34f;
  } else if (l1 == 2 && l2 == 7 && l3 == 8) {
    if (l4 == 14 && l5 == 15) return c < 2077.92f;
    if (l4 == 9 && l5 == 10) return c < 74.20f;
    if (l4 == 9 && l5 == 15) return c < 1808.536f;
  } else if (l1 == 3 && l2 == 7 && l3 == 8 && l4 == 14 && l5 == 15) {
    return c < 786.173f;
  } else if (l1 == 7 && l2 == 8 && l3 == 9) {
    if (l4 == 10 && l5 == 11) return c < 1574.076f;
    if (l4 == 10 && l5 == 16) return c < 5492.11f;
    if (l4 == 15 && l5 == 16) return c < 2743.037f;
  }
  return true;
}

struct PT5Result {
  float rz, rphi, rphiIn, pixR, t5R, cx, cy;
};

// Per-T5 quantities of the pixel-quintuplet selection that do not depend on
// the pLS, gathered once per T5 of a pixel-connected module (kT5RecWords words):
//   [0,5) x, [5,10) y, [10,15) z, [15,20) rt of the anchor points of its five
//   mini-doublets, [20,25) their r-phi sigmas, [25,30) their r-z errors,
//   [30] circle radius of the inner T3, [31] packed: bits 0-4 barrel flag of
//   each module, bits 5-6 radius class of the inner T3, bits 7-31 the pixel-
//   object layer numbers of the five modules (5 bits each)

__device__ __forceinline__ void make_t5_rec(const EventView& v, uint32_t t5, float* out) {
  const ModulesView& M = v.M;
  const uint32_t innerT3 = v.Q.tripletIndices[2 * t5], outerT3 = v.Q.tripletIndices[2 * t5 + 1];
  const uint4 ti = t3_chunk(v.T, innerT3, 0), to = t3_chunk(v.T, outerT3, 0);
  const uint4 g1 = seg_chunk(v.S, ti.x, 0), g2 = seg_chunk(v.S, ti.y, 0), g3 = seg_chunk(v.S, to.x, 0),
              g4 = seg_chunk(v.S, to.y, 0);
  const uint32_t md[5] = {g1.x, g2.x, g2.y, g3.y, g4.y};
  uint16_t mods[5];
  float w[kT5RecWords];
  uint32_t packed = 0;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    mods[k] = v.Q.lowerModuleIndices[5 * t5 + k];
    const unsigned m = mods[k];
    const uint4 p = md_chunk(v.D, md[k], 1);  // x, y, z, rt
    w[k] = __uint_as_float(p.x);
    w[5 + k] = __uint_as_float(p.y);
    w[10 + k] = __uint_as_float(p.z);
    w[15 + k] = __uint_as_float(p.w);
    // r-z error of the point (computePT5RZChiSquared, SDL/PixelTriplet.h:2636-2677)
    const int type = M.moduleType[m], side = M.sides[m], sub = M.subdets[m];
    const float drdz = M.drdzs[m];
    float error = type == 0 ? 0.15f : 5.0f;
    if (type == 0 && sub == kBarrel && side != kCenter) error /= m_sqrt(1.f + drdz * drdz);
    w[25 + k] = error;
    if (sub == kBarrel) packed |= 1u << k;
    packed |= (uint32_t)pix_layer(M, m) << (7 + 5 * k);
  }
  packed |= (uint32_t)pt3_radius_kind(M, mods[0], mods[1], mods[2]) << 5;
  point_sigmas(M, mods, 5, w, w + 5, w + 20);
  w[30] = __uint_as_float(t3_chunk(v.T, innerT3, 3).w);  // circle radius
  w[31] = __uint_as_float(packed);
  float4* o = reinterpret_cast<float4*>(out);
#pragma unroll
  for (int k = 0; k < kT5RecWords / 4; k++) o[k] = make_float4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}

// runPixelQuintupletDefaultAlgo (SDL/PixelTriplet.h:2480-2634), the AND of pure
// cuts, in two parts as for the pixel triplets: pt5_cheap (the inner T3's
// radius compatibility, the r-z, r-phi and inward r-phi chi2 cuts, which set
// every value of the result) and pt5_costly (the pixel-tracklet tests of the
// inner T3).  `rec` is the T5's kT5RecWords record.
__device__ __forceinline__ bool pt5_cheap(const EventView& v, const PlsCtx& p, uint32_t t5, const float* rec,
                                          PT5Result& r) {
  // the last chunk (inner radius, packed flags) first: the radius test ends most candidates, and only the
  // others load the rest of the record
  const float4* q4 = reinterpret_cast<const float4*>(rec);
  const float4 last = ld_global(q4 + kT5RecWords / 4 - 1);
  const uint32_t packed = __float_as_uint(last.w);
  // inner T3 (runPixelTripletDefaultAlgo without chi2 cuts): radius compatibility
  if (!pt3_radius_ok((packed >> 5) & 3, p.ptIn * kR1GeVf, p.ptErr * kR1GeVf, last.z)) return false;
  float w[kT5RecWords];
#pragma unroll
  for (int k = 0; k < kT5RecWords - 4; k += 4) {
    const float4 q = ld_global(q4 + k / 4);
    w[k] = q.x; w[k + 1] = q.y; w[k + 2] = q.z; w[k + 3] = q.w;
  }
  w[kT5RecWords - 4] = last.x; w[kT5RecWords - 3] = last.y; w[kT5RecWords - 2] = last.z; w[kT5RecWords - 1] = last.w;
  int l[5];
#pragma unroll
  for (int k = 0; k < 5; k++) l[k] = rec_layer(packed, 7, k);
  const float zP0 = p.za, zP1 = p.zb;
  const float rP0 = p.ra, rP1 = p.rb;
  r.rz = 0;
  r.cx = p.cx;
  r.cy = p.cy;
  r.pixR = p.pixR;
  if (r.pixR < 5.0f * kR1GeVf) {  // computePT5RZChiSquared (:2636-2677)
    const float slope = (zP1 - zP0) / (rP1 - rP0);
    float rmse = 0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const float zs = w[10 + i], rts = w[15 + i], error = w[25 + i];
      const float residual = ((packed >> i) & 1) ? (zs - zP0) - slope * (rts - rP0) : (rts - rP0) - (zs - zP0) / slope;
      rmse += (residual * residual) / (error * error);
    }
    r.rz = m_sqrt(0.2f * rmse);
    if (!pt5_rz_ok(l[0], l[1], l[2], l[3], l[4], r.rz)) return false;
  }
  const float t5g = v.Q.regressionG[t5], t5f = v.Q.regressionF[t5];
  r.t5R = v.Q.regressionRadius[t5];
  r.rphi = points_chi2(5, w, w + 5, w + 20, r.cx, r.cy, r.pixR);
  if (r.pixR < 5.0f * kR1GeVf && !pt5_rphi_ok(l[0], l[1], l[2], l[3], l[4], r.rphi)) return false;
  {
    const float xa = p.xa, ya = p.ya, xb = p.xb, yb = p.yb;
    const float rr = r.t5R;
    float c2 = 0;
    float res = (xa - t5g) * (xa - t5g) + (ya - t5f) * (ya - t5f) - rr * rr;
    c2 += res * res;
    res = (xb - t5g) * (xb - t5g) + (yb - t5f) * (yb - t5f) - rr * rr;
    c2 += res * res;
    c2 *= 0.5f;
    r.rphiIn = c2;
  }
  if (r.t5R < 5.0f * kR1GeVf && !pt5_rphi_in_ok(l[0], l[1], l[2], l[3], l[4], r.rphiIn)) return false;
  r.cx = (r.cx + t5g) / 2;
  r.cy = (r.cy + t5f) / 2;
  return true;
}

__device__ __forceinline__ bool pt5_costly(const EventView& v, uint32_t pls, uint32_t arr, uint32_t t5) {
  return pt3_tracklets(v, pls, arr, v.Q.tripletIndices[2 * t5]);
}

}  // namespace lst
