// ORACLE — test infrastructure only (see orc_common.h).
// pLS duplicate cleaning (SDL/Kernels.h:384-473), pixel quintuplets
// (SDL/PixelTriplet.h:1871-2790, SDL/Event.cc:1009-1146), pixel triplets
// (SDL/PixelTriplet.h:20-1870, SDL/Event.cc:784-901), track candidates
// (SDL/TrackCandidate.h, SDL/Kernels.h:250-383, SDL/Event.cc:620-782) and the
// LST output lists (SDL/LST.cc:366-420), in the reference's serial CPU order.
#include <algorithm>

#include "orc_event.h"
#include "orc_t5.h"

namespace orc {
namespace {

constexpr unsigned kTCCapacity = kNMaxNonPixelTrackCandidates + kNMaxPixelTrackCandidates;

// sdl layer number as the pixel-object cuts compute it (SDL/PixelTriplet.h:321-333)
int sdlLayerOf(const Geometry& G, uint16_t m) {
  return G.layers[m] + 6 * (G.subdets[m] == Endcap) + 5 * (G.subdets[m] == Endcap && G.moduleType[m] == TwoS);
}

// connected-module range of a pLS in the pixel map (SDL/Event.cc:828-856)
void connectedRange(const Event& ev, unsigned i, unsigned& base, unsigned& size) {
  const Geometry& G = *ev.g;
  const int8_t type = ev.plsPixelType[i];
  const int sb = ev.plsSuperbin[i];
  base = size = 0;
  if (sb < 0 || sb >= (int)kSizeSuperbins || type > 2 || type < 0) return;
  const unsigned offPos = G.cpIndex[kSizeSuperbins - 1] + G.cpSizes[kSizeSuperbins - 1];
  const unsigned offNeg = G.cpIndexPos[kSizeSuperbins - 1] + G.cpSizesPos[kSizeSuperbins - 1] + offPos;
  if (type == 0) { size = G.cpSizes[sb]; base = G.cpIndex[sb]; }
  else if (type == 1) { size = G.cpSizesPos[sb]; base = G.cpIndexPos[sb] + offPos; }
  else { size = G.cpSizesNeg[sb]; base = G.cpIndexNeg[sb] + offNeg; }
}

// SDL/Kernels.h:384-473
void checkHitspLS(Event& ev, bool secondpass) {
  const unsigned n = std::min(ev.nPLS, kNMaxPixelSegments);
  for (unsigned ix = 0; ix < n; ix++) {
    if (secondpass && (!ev.plsIsQuad[ix] || (ev.plsIsDup[ix] & 1))) continue;
    const unsigned* h1 = &ev.plsHitsIdxs[4 * ix];
    const float eta1 = ev.plsEta[ix], phi1 = ev.plsPhi[ix];
    for (unsigned jx = ix + 1; jx < n; jx++) {
      const float eta2 = ev.plsEta[jx], phi2 = ev.plsPhi[jx];
      if (std::fabs(eta2 - eta1) > 0.1f) continue;
      if (secondpass && (!ev.plsIsQuad[jx] || (ev.plsIsDup[jx] & 1))) continue;
      const int8_t quadDiff = ev.plsIsQuad[ix] - ev.plsIsQuad[jx];
      const float scoreDiff = ev.plsScore[ix] - ev.plsScore[jx];
      unsigned rm;
      if (quadDiff > 0) rm = jx;
      else if (quadDiff < 0) rm = ix;
      else if (scoreDiff < 0) rm = jx;
      else if (scoreDiff > 0) rm = ix;
      else rm = ix;
      const unsigned* h2 = &ev.plsHitsIdxs[4 * jx];
      int npMatched = 0;
      for (int a = 0; a < 4; a++) {
        bool m = false;
        for (int b = 0; b < 4; b++)
          if (h1[a] == h2[b]) { m = true; break; }
        if (m) {
          npMatched++;
          if (secondpass) break;
        }
      }
      if (npMatched >= 3) ev.plsIsDup[rm] |= 1 + secondpass;
      if (secondpass) {
        const float dEta = std::fabs(eta1 - eta2);
        const float dPhi = calculate_dPhi(phi1, phi2);
        const float dR2 = dEta * dEta + dPhi * dPhi;
        if (npMatched >= 1 || dR2 < 1e-5f) ev.plsIsDup[rm] |= 1 + secondpass;
      }
    }
  }
}

// pixel segment against an outer-tracker segment: runTripletDefaultAlgoPPBB
// (SDL/PixelTriplet.h:1264-1553) and runTripletDefaultAlgoPPEE (:1555-1869),
// dispatched as runPixelTrackletDefaultAlgopT3 (:206-314)
struct PixelTracklet {
  const Event& ev;
  unsigned pls, plsArr;   // global segment index and pLS array index
  uint16_t mIn, mOut;     // lower modules of the outer segment
  unsigned outerSeg;
  float x1, y1, rt1, x2, y2, z2, rt2, x3, y3, z3, rt3, x4, y4, rt4;
  unsigned md4;

  bool run() {
    const Geometry& G = *ev.g;
    const unsigned md1 = ev.segMdIndices[2 * pls], md2 = ev.segMdIndices[2 * pls + 1];
    const unsigned md3 = ev.segMdIndices[2 * outerSeg];
    md4 = ev.segMdIndices[2 * outerSeg + 1];
    x1 = ev.anchorX[md1]; y1 = ev.anchorY[md1]; rt1 = ev.anchorRt[md1];
    x2 = ev.anchorX[md2]; y2 = ev.anchorY[md2]; z2 = ev.anchorZ[md2]; rt2 = ev.anchorRt[md2];
    x3 = ev.anchorX[md3]; y3 = ev.anchorY[md3]; z3 = ev.anchorZ[md3]; rt3 = ev.anchorRt[md3];
    x4 = ev.anchorX[md4]; y4 = ev.anchorY[md4]; rt4 = ev.anchorRt[md4];
    const short sIn = G.subdets[mIn], sOut = G.subdets[mOut];
    if (sIn == Barrel && (sOut == Barrel || sOut == Endcap)) return ppbb();
    if (sIn == Endcap && sOut == Endcap) return ppee();
    return false;
  }

  // common beta part of both flavours (from the EC_lastLayer flag onwards)
  bool betaCuts(float dPhi, float sdlThetaMulsF, float drtInSeg) {
    const Geometry& G = *ev.g;
    const float ptCut = ev.cfg.ptCut;
    const float ptIn = ev.plsPtIn[plsArr], px = ev.plsPx[plsArr], py = ev.plsPy[plsArr];
    const float alpha_InLo = ev.segDPhiChanges[pls];
    const float alpha_OutLo = ev.segDPhiChanges[outerSeg];
    const bool isEC_lastLayer = G.subdets[mOut] == Endcap && G.moduleType[mOut] == TwoS;
    const float alpha_OutUp = deltaPhi(x4, y4, x4 - x3, y4 - y3);
    const float tl_axis_x = x4 - x2, tl_axis_y = y4 - y2;
    float betaIn = -deltaPhi(px, py, tl_axis_x, tl_axis_y);
    float betaInRHmin = betaIn, betaInRHmax = betaIn;
    float betaOut = -alpha_OutUp + deltaPhi(x4, y4, tl_axis_x, tl_axis_y);
    float betaOutRHmin = betaOut, betaOutRHmax = betaOut;
    const float hx = ev.anchorHighEdgeX[md4], hy = ev.anchorHighEdgeY[md4];
    const float lx = ev.anchorLowEdgeX[md4], ly = ev.anchorLowEdgeY[md4];
    if (isEC_lastLayer) {
      const float aHigh = deltaPhi(hx, hy, hx - x3, hy - y3);
      const float aLow = deltaPhi(lx, ly, lx - x3, ly - y3);
      betaOutRHmin = -aHigh + deltaPhi(hx, hy, hx - x2, hy - y2);
      betaOutRHmax = -aLow + deltaPhi(lx, ly, lx - x2, ly - y2);
    }
    const float drt_tl_axis = fsqrt(tl_axis_x * tl_axis_x + tl_axis_y * tl_axis_y);
    const float rt_InSeg = fsqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
    float betaAv = 0.5f * (betaIn + betaOut);
    float pt_beta = ptIn;
    const int lIn = 0;
    const int lOut = isEC_lastLayer ? 11 : 5;
    const float sdOut_dr = fsqrt((x4 - x3) * (x4 - x3) + (y4 - y3) * (y4 - y3));
    const float sdOut_d = rt4 - rt3;
    runDeltaBetaIterationsT5(betaIn, betaOut, betaAv, pt_beta, rt_InSeg, sdOut_dr, drt_tl_axis, lIn);
    const float betaInMMSF =
        (std::fabs(betaInRHmin + betaInRHmax) > 0) ? (2.f * betaIn / std::fabs(betaInRHmin + betaInRHmax)) : 0.;
    const float betaOutMMSF =
        (std::fabs(betaOutRHmin + betaOutRHmax) > 0) ? (2.f * betaOut / std::fabs(betaOutRHmin + betaOutRHmax)) : 0.;
    betaInRHmin *= betaInMMSF; betaInRHmax *= betaInMMSF;
    betaOutRHmin *= betaOutMMSF; betaOutRHmax *= betaOutMMSF;
    const float dBetaMuls = sdlThetaMulsF * 4.f / fminf_(std::fabs(pt_beta), pt_betaMax);
    const float alphaInAbsReg = fmaxf_(std::fabs(alpha_InLo), fasin(fminf_(rt2 * k2Rinv1GeVf / 3.0f, sinAlphaMax)));
    const float alphaOutAbsReg = fmaxf_(std::fabs(alpha_OutLo), fasin(fminf_(rt3 * k2Rinv1GeVf / 3.0f, sinAlphaMax)));
    const float dBetaInLum = lIn < 11 ? 0.0f : std::fabs(alphaInAbsReg * deltaZLum / z2);
    const float dBetaOutLum = lOut < 11 ? 0.0f : std::fabs(alphaOutAbsReg * deltaZLum / z3);
    const float dBetaLum2 = (dBetaInLum + dBetaOutLum) * (dBetaInLum + dBetaOutLum);
    const float sinDPhi = fsin(dPhi);
    const float dBetaRIn2 = 0;
    float dBetaROut = 0;
    if (isEC_lastLayer)
      dBetaROut = (fsqrt(hx * hx + hy * hy) - fsqrt(lx * lx + ly * ly)) * sinDPhi / drt_tl_axis;
    const float dBetaROut2 = dBetaROut * dBetaROut;
    const float betaOutCut = fasin(fminf_(drt_tl_axis * k2Rinv1GeVf / ptCut, sinAlphaMax)) + (0.02f / sdOut_d) +
                             fsqrt(dBetaLum2 + dBetaMuls * dBetaMuls);
    if (!(std::fabs(betaOut) < betaOutCut)) return false;
    const float dBetaRes = 0.02f / fminf_(sdOut_d, drtInSeg);
    const float spread = std::fabs(betaInRHmin - betaInRHmax) + std::fabs(betaOutRHmin - betaOutRHmax);
    const float dBetaCut2 = (dBetaRes * dBetaRes * 2.0f + dBetaMuls * dBetaMuls + dBetaLum2 + dBetaRIn2 + dBetaROut2 +
                             0.25f * spread * spread);
    const float dBeta = betaIn - betaOut;
    return dBeta * dBeta <= dBetaCut2;
  }

  bool ppbb() {
    const Geometry& G = *ev.g;
    const float ptCut = ev.cfg.ptCut;
    const bool isPS_OutLo = G.moduleType[mIn] == PS;
    if (!(std::fabs(deltaPhi(x2, y2, x3, y3)) <= 0.5f * float(M_PI))) return false;
    const float ptIn = ev.plsPtIn[plsArr], pz = ev.plsPz[plsArr], etaErr = ev.plsEtaErr[plsArr];
    const float alpha1GeV_OutLo = fasin(fminf_(rt3 * k2Rinv1GeVf / ptCut, sinAlphaMax));
    const float rtRatio = rt3 / rt2;
    const float dzDrtScale = ftan(alpha1GeV_OutLo) / alpha1GeV_OutLo;
    const float zpitch_InLo = 0.05f, zpitch_InOut = 0.05f;
    const float zpitch_OutLo = isPS_OutLo ? pixelPSZpitch : strip2SZpitch;
    const float zGeom = zpitch_InLo + zpitch_OutLo;
    const float zHi = z2 + (z2 + deltaZLum) * (rtRatio - 1.f) * (z2 < 0.f ? 1.f : dzDrtScale) + (zpitch_InOut + zpitch_OutLo);
    const float zLo = z2 + (z2 - deltaZLum) * (rtRatio - 1.f) * (z2 > 0.f ? 1.f : dzDrtScale) - (zpitch_InOut + zpitch_OutLo);
    if (!((z3 >= zLo) && (z3 <= zHi))) return false;
    const float coshEta = fsqrt(ptIn * ptIn + pz * pz) / ptIn;
    const float drt_OutLo_InUp = rt3 - rt2;
    const float r3_InUp = fsqrt(z2 * z2 + rt2 * rt2);
    const float drt_InSeg = rt2 - rt1;
    const float sdlThetaMulsF = 0.015f * fsqrt(0.1f + 0.2f * (rt3 - rt2) / 50.f) * fsqrt(r3_InUp / rt2);
    const float sdlMuls = sdlThetaMulsF * 3.f / ptCut * 4.f;
    float dzErr = drt_OutLo_InUp * etaErr * coshEta;
    dzErr *= dzErr;
    dzErr += 0.03f * 0.03f;
    dzErr *= 9.f;
    dzErr += sdlMuls * sdlMuls * drt_OutLo_InUp * drt_OutLo_InUp / 3.f * coshEta * coshEta;
    dzErr += zGeom * zGeom;
    dzErr = fsqrt(dzErr);
    const float dzDrIn = pz / ptIn;
    const float zWindow = dzErr / drt_InSeg * drt_OutLo_InUp + zGeom;
    const float dzMean = dzDrIn * drt_OutLo_InUp *
                         (1.f + drt_OutLo_InUp * drt_OutLo_InUp * 4 * k2Rinv1GeVf * k2Rinv1GeVf / ptIn / ptIn / 24.f);
    const float zLoPointed = z2 + dzMean - zWindow;
    const float zHiPointed = z2 + dzMean + zWindow;
    if (!((z3 >= zLoPointed) && (z3 <= zHiPointed))) return false;
    const float sdlPVoff = 0.1f / rt3;
    const float sdlCut = alpha1GeV_OutLo + fsqrt(sdlMuls * sdlMuls + sdlPVoff * sdlPVoff);
    const float dPhi = deltaPhi(0.5f * (x1 + x3), 0.5f * (y1 + y3), x3 - x1, y3 - y1);
    if (!(std::fabs(dPhi) <= sdlCut)) return false;
    return betaCuts(dPhi, sdlThetaMulsF, drt_InSeg);
  }

  bool ppee() {
    const Geometry& G = *ev.g;
    const float ptCut = ev.cfg.ptCut;
    const bool isPS_OutLo = G.moduleType[mIn] == PS;
    if (!(z2 * z3 > 0)) return false;
    const float ptIn = ev.plsPtIn[plsArr], pz = ev.plsPz[plsArr], etaErr = ev.plsEtaErr[plsArr];
    const float zpitch_InLo = 0.05f;
    const float zpitch_OutLo = isPS_OutLo ? pixelPSZpitch : strip2SZpitch;
    const float zGeom = zpitch_InLo + zpitch_OutLo;
    const float sdlSlope = fasin(fminf_(rt3 * k2Rinv1GeVf / ptCut, sinAlphaMax));
    const float dzDrtScale = ftan(sdlSlope) / sdlSlope;
    const float dLum = copysignf_(deltaZLum, z2);
    const float rtGeom1 = isPS_OutLo ? pixelPSZpitch : strip2SZpitch;
    const float zGeom1 = copysignf_(zGeom, z2);
    const float rtLo = rt2 * (1.f + (z3 - z2 - zGeom1) / (z2 + zGeom1 + dLum) / dzDrtScale) - rtGeom1;
    float zInForHi = z2 - zGeom1 - dLum;
    if (zInForHi * z2 < 0) zInForHi = copysignf_(0.1f, z2);
    const float rtHi = rt2 * (1.f + (z3 - z2 + zGeom1) / zInForHi) + rtGeom1;
    if (!((rt3 >= rtLo) && (rt3 <= rtHi))) return false;
    const float dzOutInAbs = std::fabs(z3 - z2);
    const float coshEta = fsqrt(ptIn * ptIn + pz * pz) / ptIn;
    const float multDzDr = dzOutInAbs * coshEta / (coshEta * coshEta - 1.f);
    const float r3_InUp = fsqrt(z2 * z2 + rt2 * rt2);
    const float sdlThetaMulsF = 0.015f * fsqrt(0.1f + 0.2f * (rt3 - rt2) / 50.f) * fsqrt(r3_InUp / rt2);
    const float sdlMuls = sdlThetaMulsF * 3.f / ptCut * 4.f;
    float drtErr = etaErr * multDzDr;
    drtErr *= drtErr;
    drtErr += 0.03f * 0.03f;
    drtErr *= 9.f;
    drtErr += sdlMuls * sdlMuls * multDzDr * multDzDr / 3.f * coshEta * coshEta;
    drtErr = fsqrt(drtErr);
    const float drtDzIn = std::fabs(ptIn / pz);
    const float drt_OutLo_InUp = rt3 - rt2;
    const float rtWindow = drtErr + rtGeom1;
    const float drtMean = drtDzIn * dzOutInAbs *
                          (1.f - drt_OutLo_InUp * drt_OutLo_InUp * 4 * k2Rinv1GeVf * k2Rinv1GeVf / ptIn / ptIn / 24.f);
    const float rtLo_point = rt2 + drtMean - rtWindow;
    const float rtHi_point = rt2 + drtMean + rtWindow;
    if (!((rt3 >= rtLo_point) && (rt3 <= rtHi_point))) return false;
    const float alpha1GeV_OutLo = fasin(fminf_(rt3 * k2Rinv1GeVf / ptCut, sinAlphaMax));
    const float sdlPVoff = 0.1f / rt3;
    const float sdlCut = alpha1GeV_OutLo + fsqrt(sdlMuls * sdlMuls + sdlPVoff * sdlPVoff);
    const float dPhi = deltaPhi(0.5f * (x1 + x3), 0.5f * (y1 + y3), x3 - x1, y3 - y1);
    if (!(std::fabs(dPhi) <= sdlCut)) return false;
    return betaCuts(dPhi, sdlThetaMulsF, rt2 - rt1);
  }
};

bool pixelTracklet(const Event& ev, unsigned pls, uint16_t mIn, uint16_t mOut, unsigned outerSeg) {
  const unsigned arr = pls - ev.segmentModuleIndices[ev.g->nLowerModules];
  PixelTracklet t{ev, pls, arr, mIn, mOut, outerSeg};
  return t.run();
}

bool checkIntervalOverlap(float aMin, float aMax, float bMin, float bMax) {  // PixelTriplet.h:607-612
  return ((aMin <= bMin) && (bMin < aMax)) || ((bMin < aMin) && (aMin < bMax));
}

// SDL/PixelTriplet.h:615-734
bool passRadiusCriterion(const Geometry& G, float pixelRadius, float pixelRadiusError, float tripletRadius,
                         uint16_t lower, uint16_t middle, uint16_t upper) {
  float tb, pb, tbHi, pbHi;
  int kind;  // 0 BBB, 1 BBE, 2 BEE, 3 EEE
  if (G.subdets[lower] == Endcap) { kind = 3; tb = 1.7006f; pb = 0.26367f; pbHi = 2.286f; tbHi = 2.436f; }
  else if (G.subdets[middle] == Endcap) { kind = 2; tb = 1.59294f; pb = 0.255181f; pbHi = 2.2091f; tbHi = 2.3548f; }
  else if (G.subdets[upper] == Endcap) { kind = 1; tb = 0.45972f; pb = 0.19644f; pbHi = 0.6805f; tbHi = 0.8557f; }
  else { kind = 0; tb = 0.15624f; pb = 0.17235f; pbHi = 0.6375f; tbHi = 0.6588f; }
  if (pixelRadius > 2.0f * kR1GeVf) { pb = pbHi; tb = tbHi; }
  const float tInvMax = (1 + tb) / tripletRadius;
  const float tInvMin = fmaxf_((1 - tb) / tripletRadius, 0.0f);
  const float pInvMax = fmaxf_((1 + pb) / pixelRadius, 1.f / (pixelRadius - pixelRadiusError));
  float pInvMin = fminf_((1 - pb) / pixelRadius, 1.f / (pixelRadius + pixelRadiusError));
  if (kind == 2) pInvMin = fmaxf_(pInvMin, 0.0f);
  if (kind == 3) pInvMin = fmaxf_(0.0f, pInvMin);
  return checkIntervalOverlap(tInvMin, tInvMax, pInvMin, pInvMax);
}

// SDL/PixelTriplet.h:316-369
bool passPT3RZChiSquaredCuts(int l1, int l2, int l3, float c) {
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

// SDL/PixelTriplet.h:501-546
bool passPT3RPhiChiSquaredCuts(int l1, int l2, int l3, float c) {
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

// SDL/PixelTriplet.h:548-605
bool passPT3RPhiChiSquaredInwardsCuts(int l1, int l2, int l3, float c) {
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

// SDL/PixelTriplet.h:736-824
float computePT3RZChiSquared(const Geometry& G, const uint16_t* mods, const float* rtPix, const float* xPix,
                             const float* yPix, const float* zPix, const float* rts, const float* zs, float Px,
                             float Py, float Pz, int charge) {
  float RMSE = 0;
  const float x1 = xPix[1] / 100, y1 = yPix[1] / 100, z1 = zPix[1] / 100, r1 = rtPix[1] / 100;
  const float a = -0.299792 * magnetic_field * charge;
  for (int i = 0; i < 3; i++) {
    const float zsi = zs[i] / 100, rtsi = rts[i] / 100;
    const int type = G.moduleType[mods[i]], side = G.sides[mods[i]], subdet = G.subdets[mods[i]];
    float diffr = 0, diffz = 0;  // the reference leaves the unused one uninitialised
    const float p = fsqrt(Px * Px + Py * Py + Pz * Pz);
    const float rou = a / p;
    if (subdet == Endcap) {
      const float s = (zsi - z1) * p / Pz;
      const float x = x1 + Px / a * fsin(rou * s) - Py / a * (1 - fcos(rou * s));
      const float y = y1 + Py / a * fsin(rou * s) + Px / a * (1 - fcos(rou * s));
      diffr = std::fabs(rtsi - fsqrt(x * x + y * y)) * 100;
    }
    if (subdet == Barrel) {
      const float paraA = r1 * r1 + 2 * (Px * Px + Py * Py) / (a * a) + 2 * (y1 * Px - x1 * Py) / a - rtsi * rtsi;
      const float paraB = 2 * (x1 * Px + y1 * Py) / a;
      const float paraC = 2 * (y1 * Px - x1 * Py) / a + 2 * (Px * Px + Py * Py) / (a * a);
      const float A = paraB * paraB + paraC * paraC;
      const float B = 2 * paraA * paraB;
      const float C = paraA * paraA - paraC * paraC;
      const float sol1 = (-B + fsqrt(B * B - 4 * A * C)) / (2 * A);
      const float sol2 = (-B - fsqrt(B * B - 4 * A * C)) / (2 * A);
      const float solz1 = fasin(sol1) / rou * Pz / p + z1;
      const float solz2 = fasin(sol2) / rou * Pz / p + z1;
      const float diffz1 = std::fabs(solz1 - zsi) * 100;
      const float diffz2 = std::fabs(solz2 - zsi) * 100;
      diffz = fminf_(diffz1, diffz2);
    }
    const float residual = subdet == Barrel ? diffz : diffr;
    float error = type == 0 ? 0.15f : 5.0f;
    if (type == 0 && subdet == Barrel && side != Center) {
      const float drdz = G.drdzs[mods[i]];
      error /= fsqrt(1 + drdz * drdz);
    }
    RMSE += (residual * residual) / (error * error);
  }
  return fsqrt(0.2f * RMSE);
}

struct PT3Out {
  float pixelRadius = 0, pixelRadiusError = 0, tripletRadius = 0, centerX = 0, centerY = 0, rzChiSquared = 0,
        rPhiChiSquared = 0, rPhiChiSquaredInwards = 0;
};

// SDL/PixelTriplet.h:826-1024
bool runPixelTripletDefaultAlgo(const Event& ev, unsigned pls, unsigned t3, bool runChiSquaredCuts, PT3Out& o) {
  const Geometry& G = *ev.g;
  const uint16_t pixelModule = ev.segInnerLowerModule[pls];
  const uint16_t lower = ev.t3LowerModuleIndices[3 * t3], middle = ev.t3LowerModuleIndices[3 * t3 + 1],
                 upper = ev.t3LowerModuleIndices[3 * t3 + 2];
  const unsigned segIn = ev.t3SegmentIndices[2 * t3], segOut = ev.t3SegmentIndices[2 * t3 + 1];
  if (!pixelTracklet(ev, pls, lower, middle, segIn)) return false;
  if (!pixelTracklet(ev, pls, middle, upper, segOut)) return false;
  const unsigned arr = pls - ev.segmentModuleIndices[pixelModule];
  const float pt = ev.plsPtIn[arr], ptErr = ev.plsPtErr[arr];
  const float pixelG = ev.plsCircleCenterX[arr], pixelF = ev.plsCircleCenterY[arr],
              pixelRadiusPCA = ev.plsCircleRadius[arr];
  const unsigned pixIn = ev.segMdIndices[2 * pls], pixOut = ev.segMdIndices[2 * pls + 1];
  o.pixelRadius = pt * kR1GeVf;
  o.pixelRadiusError = ptErr * kR1GeVf;
  const unsigned md1 = ev.segMdIndices[2 * segIn], md2 = ev.segMdIndices[2 * segIn + 1],
                 md3 = ev.segMdIndices[2 * segOut + 1];
  float xs[3] = {ev.anchorX[md1], ev.anchorX[md2], ev.anchorX[md3]};
  float ys[3] = {ev.anchorY[md1], ev.anchorY[md2], ev.anchorY[md3]};
  o.tripletRadius = ev.t3CircleRadius[t3];
  const float g = ev.t3CircleCenterX[t3], f = ev.t3CircleCenterY[t3];
  if (!passRadiusCriterion(G, o.pixelRadius, o.pixelRadiusError, o.tripletRadius, lower, middle, upper)) return false;
  const uint16_t mods[3] = {lower, middle, upper};
  const int l1 = sdlLayerOf(G, lower), l2 = sdlLayerOf(G, middle), l3 = sdlLayerOf(G, upper);
  const bool cuts = runChiSquaredCuts && pt < 5.0f;
  if (cuts) {
    const float rts[3] = {ev.anchorRt[md1], ev.anchorRt[md2], ev.anchorRt[md3]};
    const float zs[3] = {ev.anchorZ[md1], ev.anchorZ[md2], ev.anchorZ[md3]};
    const float rtPix[2] = {ev.anchorRt[pixIn], ev.anchorRt[pixOut]};
    const float xPix[2] = {ev.anchorX[pixIn], ev.anchorX[pixOut]};
    const float yPix[2] = {ev.anchorY[pixIn], ev.anchorY[pixOut]};
    const float zPix[2] = {ev.anchorZ[pixIn], ev.anchorZ[pixOut]};
    o.rzChiSquared = computePT3RZChiSquared(G, mods, rtPix, xPix, yPix, zPix, rts, zs, ev.plsPx[arr], ev.plsPy[arr],
                                            ev.plsPz[arr], ev.plsCharge[arr]);
    if (!passPT3RZChiSquaredCuts(l1, l2, l3, o.rzChiSquared)) return false;
  } else {
    o.rzChiSquared = -1;
  }
  {  // computePT3RPhiChiSquared (:419-487)
    float d1[3] = {0, 0, 0}, d2[3] = {0, 0, 0}, sl[3];
    bool flat[3] = {false, false, false};
    sigmasForRegression(G, mods, d1, d2, sl, flat, 3, true);
    o.rPhiChiSquared = chiSquaredOfCircle(3, xs, ys, d1, d2, sl, flat, pixelG, pixelF, pixelRadiusPCA);
  }
  if (cuts && !passPT3RPhiChiSquaredCuts(l1, l2, l3, o.rPhiChiSquared)) return false;
  {  // computePT3RPhiChiSquaredInwards (:489-499)
    const float xPix[2] = {ev.anchorX[pixIn], ev.anchorX[pixOut]};
    const float yPix[2] = {ev.anchorY[pixIn], ev.anchorY[pixOut]};
    const float r = o.tripletRadius;
    float residual = (xPix[0] - g) * (xPix[0] - g) + (yPix[0] - f) * (yPix[0] - f) - r * r;
    float c2 = residual * residual;
    residual = (xPix[1] - g) * (xPix[1] - g) + (yPix[1] - f) * (yPix[1] - f) - r * r;
    c2 += residual * residual;
    c2 *= 0.5f;
    o.rPhiChiSquaredInwards = c2;
  }
  if (cuts && !passPT3RPhiChiSquaredInwardsCuts(l1, l2, l3, o.rPhiChiSquaredInwards)) return false;
  o.centerX = 0;
  o.centerY = 0;
  return true;
}

// SDL/PixelTriplet.h:2038-2126
bool passPT5RZChiSquaredCuts(int l1, int l2, int l3, int l4, int l5, float c) {
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
    // as in the reference: this branch tests l4 == 7, so its inner cases never match
    if (l4 == 13 && l5 == 14) return c < 595.545f;
    if (l4 == 8 && l5 == 14) return c < 74.198f;
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

// SDL/PixelTriplet.h:2128-2217
bool passPT5RPhiChiSquaredCuts(int l1, int l2, int l3, int l4, int l5, float c) {
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

// SDL/PixelTriplet.h:2389-2478
bool passPT5RPhiChiSquaredInwardsCuts(int l1, int l2, int l3, int l4, int l5, float c) {
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

struct PT5Out {
  float rzChiSquared = 0, rPhiChiSquared = 0, rPhiChiSquaredInwards = 0, pixelRadius = 0, quintupletRadius = 0,
        centerX = 0, centerY = 0;
};

// SDL/PixelTriplet.h:2480-2634 (with computePT5RZChiSquared, :2636-2677)
bool runPixelQuintupletDefaultAlgo(const Event& ev, unsigned pls, unsigned t5, unsigned plsArr, PT5Out& o) {
  const Geometry& G = *ev.g;
  const unsigned innerT3 = ev.t5TripletIndices[2 * t5], outerT3 = ev.t5TripletIndices[2 * t5 + 1];
  {
    PT3Out tmp;
    if (!runPixelTripletDefaultAlgo(ev, pls, innerT3, false, tmp)) return false;
  }
  const unsigned s1 = ev.t3SegmentIndices[2 * innerT3], s2 = ev.t3SegmentIndices[2 * innerT3 + 1];
  const unsigned s3 = ev.t3SegmentIndices[2 * outerT3], s4 = ev.t3SegmentIndices[2 * outerT3 + 1];
  const unsigned pixIn = ev.segMdIndices[2 * pls], pixOut = ev.segMdIndices[2 * pls + 1];
  const unsigned md[5] = {ev.segMdIndices[2 * s1], ev.segMdIndices[2 * s2], ev.segMdIndices[2 * s2 + 1],
                          ev.segMdIndices[2 * s3 + 1], ev.segMdIndices[2 * s4 + 1]};
  uint16_t mods[5];
  for (int k = 0; k < 5; k++) mods[k] = ev.t5LowerModuleIndices[5 * t5 + k];
  int l[5];
  for (int k = 0; k < 5; k++) l[k] = sdlLayerOf(G, mods[k]);
  const float zPix[2] = {ev.anchorZ[pixIn], ev.anchorZ[pixOut]};
  const float rtPix[2] = {ev.anchorRt[pixIn], ev.anchorRt[pixOut]};
  o.rzChiSquared = 0;
  o.centerX = ev.plsCircleCenterX[plsArr];
  o.centerY = ev.plsCircleCenterY[plsArr];
  o.pixelRadius = ev.plsCircleRadius[plsArr];
  if (o.pixelRadius < 5.0f * kR1GeVf) {
    const float slope = (zPix[1] - zPix[0]) / (rtPix[1] - rtPix[0]);
    float RMSE = 0;
    for (int i = 0; i < 5; i++) {
      const int type = G.moduleType[mods[i]], side = G.sides[mods[i]], subdet = G.subdets[mods[i]];
      const float zs = ev.anchorZ[md[i]], rts = ev.anchorRt[md[i]];
      const float residual = (subdet == Barrel) ? (zs - zPix[0]) - slope * (rts - rtPix[0])
                                                : (rts - rtPix[0]) - (zs - zPix[0]) / slope;
      const float drdz = G.drdzs[mods[i]];
      float error = type == 0 ? 0.15f : 5.0f;
      if (type == 0 && subdet == Barrel && side != Center) error /= fsqrt(1.f + drdz * drdz);
      RMSE += (residual * residual) / (error * error);
    }
    o.rzChiSquared = fsqrt(0.2f * RMSE);
    if (!passPT5RZChiSquaredCuts(l[0], l[1], l[2], l[3], l[4], o.rzChiSquared)) return false;
  }
  float xs[5], ys[5];
  for (int k = 0; k < 5; k++) { xs[k] = ev.anchorX[md[k]]; ys[k] = ev.anchorY[md[k]]; }
  const float T5CenterX = ev.t5RegressionG[t5], T5CenterY = ev.t5RegressionF[t5];
  o.quintupletRadius = ev.t5RegressionRadius[t5];
  {
    float d1[5], d2[5], sl[5];
    bool flat[5];
    sigmasForRegression(G, mods, d1, d2, sl, flat, 5, true);
    o.rPhiChiSquared = chiSquaredOfCircle(5, xs, ys, d1, d2, sl, flat, o.centerX, o.centerY, o.pixelRadius);
  }
  if (o.pixelRadius < 5.0f * kR1GeVf && !passPT5RPhiChiSquaredCuts(l[0], l[1], l[2], l[3], l[4], o.rPhiChiSquared))
    return false;
  {  // computePT5RPhiChiSquaredInwards (SDL/PixelTriplet.h:2374-2387)
    const float xPix[2] = {ev.anchorX[pixIn], ev.anchorX[pixOut]};
    const float yPix[2] = {ev.anchorY[pixIn], ev.anchorY[pixOut]};
    const float r = o.quintupletRadius;
    float c2 = 0;
    for (int i = 0; i < 2; i++) {
      const float residual = (xPix[i] - T5CenterX) * (xPix[i] - T5CenterX) +
                             (yPix[i] - T5CenterY) * (yPix[i] - T5CenterY) - r * r;
      c2 += residual * residual;
    }
    c2 *= 0.5f;
    o.rPhiChiSquaredInwards = c2;
  }
  if (ev.t5RegressionRadius[t5] < 5.0f * kR1GeVf &&
      !passPT5RPhiChiSquaredInwardsCuts(l[0], l[1], l[2], l[3], l[4], o.rPhiChiSquaredInwards))
    return false;
  o.centerX = (o.centerX + T5CenterX) / 2;
  o.centerY = (o.centerY + T5CenterY) / 2;
  return true;
}

void tcInit(Event& ev) {
  const unsigned n = kTCCapacity;
  ev.nTrackCandidates = ev.nTrackCandidatespT3 = ev.nTrackCandidatespT5 = ev.nTrackCandidatespLS =
      ev.nTrackCandidatesT5 = 0;
  ev.tcType.assign(n, 0); ev.tcDirectObjectIndices.assign(n, 0); ev.tcObjectIndices.assign(2 * n, 0);
  ev.tcHitIndices.assign(14 * n, 0); ev.tcPixelSeedIndex.assign(n, 0); ev.tcLogicalLayers.assign(7 * n, 0);
  ev.tcLowerModuleIndices.assign(7 * n, 0); ev.tcCenterX.assign(n, 0.f); ev.tcCenterY.assign(n, 0.f);
  ev.tcRadius.assign(n, 0.f);
}

// addTrackCandidateToMemory (SDL/TrackCandidate.h:126-164)
void addTC(Event& ev, int16_t type, unsigned inner, unsigned outer, const uint8_t* ll, const uint16_t* lm,
           const unsigned* hits, int seed, float cx, float cy, float r, unsigned idx, unsigned direct) {
  ev.tcType[idx] = type;
  ev.tcDirectObjectIndices[idx] = direct;
  ev.tcPixelSeedIndex[idx] = seed;
  ev.tcObjectIndices[2 * idx] = inner;
  ev.tcObjectIndices[2 * idx + 1] = outer;
  const unsigned limits = type == 7 ? 7 : 5;
  for (unsigned i = 0; i < limits; i++) {
    ev.tcLogicalLayers[7 * idx + i] = ll[i];
    ev.tcLowerModuleIndices[7 * idx + i] = lm[i];
  }
  for (unsigned i = 0; i < 2 * limits; i++) ev.tcHitIndices[14 * idx + i] = hits[i];
  ev.tcCenterX[idx] = cx;
  ev.tcCenterY[idx] = cy;
  ev.tcRadius[idx] = r;
}

// checkPixelHits (SDL/TrackCandidate.h:166-202): ntuple hit indices of the
// four pLS hits, matched position by position of the first segment
int checkPixelHits(const Event& ev, unsigned ix, unsigned jx) {
  auto hitsOf = [&](unsigned s, int* h) {
    const unsigned a = ev.segMdIndices[2 * s], b = ev.segMdIndices[2 * s + 1];
    h[0] = ev.idxs[ev.mdAnchorHitIndices[a]];
    h[1] = ev.idxs[ev.mdAnchorHitIndices[b]];
    h[2] = ev.idxs[ev.mdOuterHitIndices[a]];
    h[3] = ev.idxs[ev.mdOuterHitIndices[b]];
  };
  int p1[4], p2[4];
  hitsOf(ix, p1);
  hitsOf(jx, p2);
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

}  // namespace

// SDL/Event.cc:993-1007 (first pass of checkHitspLS)
void event_pixel_segment_cleaning(Event& ev) {
  if (!ev.cfg.noPlsDupClean) checkHitspLS(ev, false);
}

// SDL/Event.cc:1009-1146: createPixelQuintupletsInGPUFromMapv2 (SDL/PixelTriplet.h:2679-2786),
// removeDupPixelQuintupletsInGPUFromMap (SDL/Kernels.h:350-381), addpT5asTrackCandidateInGPU
// (SDL/TrackCandidate.h:532-580)
void event_create_pixel_quintuplets(Event& ev) {
  const Geometry& G = *ev.g;
  const unsigned nL = G.nLowerModules, cap = kNMaxPixelQuintuplets;
  ev.nPixelQuintuplets = ev.totOccupancyPixelQuintuplets = 0;
  ev.pt5PixelIndices.assign(cap, 0); ev.pt5T5Indices.assign(cap, 0); ev.pt5HitIndices.assign(14 * cap, 0);
  ev.pt5IsDup.assign(cap, 0); ev.pt5LogicalLayers.assign(7 * cap, 0); ev.pt5LowerModuleIndices.assign(7 * cap, 0);
  for (auto* v : {&ev.pt5Score, &ev.pt5Eta, &ev.pt5Phi, &ev.pt5PixelRadius, &ev.pt5QuintupletRadius, &ev.pt5CenterX,
                  &ev.pt5CenterY, &ev.pt5RzChiSquared, &ev.pt5RPhiChiSquared, &ev.pt5RPhiChiSquaredInwards})
    v->assign(cap, 0.f);
  tcInit(ev);
  const unsigned nPLS = ev.nPLS;
  const unsigned plsBase = ev.segmentModuleIndices[nL];
  for (unsigned i = 0; i < nPLS; i++) {
    unsigned base, size;
    connectedRange(ev, i, base, size);
    for (unsigned k = base; k < base + size; k++) {
      const uint16_t m = G.connectedPixels[k];
      if (m >= nL) continue;
      if (G.moduleType[m] == TwoS) continue;
      if (ev.plsIsDup[i]) continue;
      const unsigned nOuter = ev.nQuintuplets[m];
      if (nOuter == 0) continue;
      const unsigned pls = plsBase + i;
      for (unsigned b = 0; b < nOuter; b++) {
        const unsigned t5 = ev.quintupletModuleIndices[m] + b;
        if (ev.t5IsDup[t5]) continue;
        PT5Out o;
        if (!runPixelQuintupletDefaultAlgo(ev, pls, t5, i, o)) continue;
        if (ev.totOccupancyPixelQuintuplets++ >= cap) continue;
        const unsigned q = ev.nPixelQuintuplets++;
        // addPixelQuintupletToMemory (SDL/PixelTriplet.h:1961-2036)
        ev.pt5PixelIndices[q] = pls;
        ev.pt5T5Indices[q] = t5;
        ev.pt5IsDup[q] = 0;
        ev.pt5Score[q] = o.rPhiChiSquared;
        ev.pt5Eta[q] = ev.t5Eta[t5];
        ev.pt5Phi[q] = ev.t5Phi[t5];
        ev.pt5PixelRadius[q] = o.pixelRadius;
        ev.pt5QuintupletRadius[q] = o.quintupletRadius;
        ev.pt5CenterX[q] = o.centerX;
        ev.pt5CenterY[q] = o.centerY;
        ev.pt5LogicalLayers[7 * q] = 0;
        ev.pt5LogicalLayers[7 * q + 1] = 0;
        for (int j = 0; j < 5; j++) ev.pt5LogicalLayers[7 * q + 2 + j] = ev.t5LogicalLayers[5 * t5 + j];
        ev.pt5LowerModuleIndices[7 * q] = ev.segInnerLowerModule[pls];
        ev.pt5LowerModuleIndices[7 * q + 1] = ev.segOuterLowerModule[pls];
        for (int j = 0; j < 5; j++) ev.pt5LowerModuleIndices[7 * q + 2 + j] = ev.t5LowerModuleIndices[5 * t5 + j];
        const unsigned pi = ev.segMdIndices[2 * pls], po = ev.segMdIndices[2 * pls + 1];
        ev.pt5HitIndices[14 * q] = ev.mdAnchorHitIndices[pi];
        ev.pt5HitIndices[14 * q + 1] = ev.mdOuterHitIndices[pi];
        ev.pt5HitIndices[14 * q + 2] = ev.mdAnchorHitIndices[po];
        ev.pt5HitIndices[14 * q + 3] = ev.mdOuterHitIndices[po];
        for (int j = 0; j < 10; j++) ev.pt5HitIndices[14 * q + 4 + j] = ev.t5HitIndices[10 * t5 + j];
        ev.pt5RzChiSquared[q] = o.rzChiSquared;
        ev.pt5RPhiChiSquared[q] = o.rPhiChiSquared;
        ev.pt5RPhiChiSquaredInwards[q] = o.rPhiChiSquaredInwards;
        ev.t3PartOfPT5[ev.t5TripletIndices[2 * t5]] = 1;
        ev.t3PartOfPT5[ev.t5TripletIndices[2 * t5 + 1]] = 1;
        ev.plsPartOfPT5[i] = 1;
        ev.t5PartOfPT5[t5] = 1;
      }
    }
  }
  // removeDupPixelQuintupletsInGPUFromMap: ix is dropped if a partner shares
  // >= 7 of its 14 hits and has a lower score (ties: the higher index is dropped)
  const unsigned np = ev.nPixelQuintuplets;
  for (unsigned ix = 0; ix < np; ix++) {
    const float score1 = ev.pt5Score[ix];
    for (unsigned jx = 0; jx < np; jx++) {
      if (ix == jx) continue;
      int nMatched = 0;
      for (int a = 0; a < 14; a++)
        for (int b = 0; b < 14; b++)
          if (ev.pt5HitIndices[14 * ix + a] == ev.pt5HitIndices[14 * jx + b]) { nMatched++; break; }
      const float score2 = ev.pt5Score[jx];
      if (nMatched >= 7 && (score1 > score2 || (score1 == score2 && ix > jx))) {
        ev.pt5IsDup[ix] = 1;
        break;
      }
    }
  }
  // addpT5asTrackCandidateInGPU (SDL/TrackCandidate.h:531-581)
  for (unsigned q = 0; q < np; q++) {
    if (ev.pt5IsDup[q]) continue;
    const unsigned idx = ev.nTrackCandidates++;
    if (idx >= kNMaxPixelTrackCandidates) {
      ev.nTrackCandidates--;
      break;
    }
    ev.nTrackCandidatespT5++;
    const float radius = 0.5f * (ev.pt5PixelRadius[q] + ev.pt5QuintupletRadius[q]);
    const unsigned pls = ev.pt5PixelIndices[q];
    addTC(ev, 7, pls, ev.pt5T5Indices[q], &ev.pt5LogicalLayers[7 * q], &ev.pt5LowerModuleIndices[7 * q],
          &ev.pt5HitIndices[14 * q], (int)ev.plsSeedIdx[pls - plsBase], ev.pt5CenterX[q], ev.pt5CenterY[q], radius,
          idx, q);
  }
}

// SDL/Event.cc:784-901: createPixelTripletsInGPUFromMapv2 (SDL/PixelTriplet.h:1026-1164),
// removeDupPixelTripletsInGPUFromMap (SDL/Kernels.h:322-348)
void event_create_pixel_triplets(Event& ev) {
  const Geometry& G = *ev.g;
  const unsigned nL = G.nLowerModules, cap = kNMaxPixelTriplets;
  ev.nPixelTriplets = ev.totOccupancyPixelTriplets = 0;
  ev.pt3PixelSegmentIndices.assign(cap, 0); ev.pt3TripletIndices.assign(cap, 0); ev.pt3HitIndices.assign(10 * cap, 0);
  ev.pt3IsDup.assign(cap, 0); ev.pt3LogicalLayers.assign(5 * cap, 0); ev.pt3LowerModuleIndices.assign(5 * cap, 0);
  for (auto* v : {&ev.pt3PixelRadius, &ev.pt3TripletRadius, &ev.pt3Pt, &ev.pt3Eta, &ev.pt3Phi, &ev.pt3EtaPix,
                  &ev.pt3PhiPix, &ev.pt3Score, &ev.pt3CenterX, &ev.pt3CenterY, &ev.pt3RzChiSquared,
                  &ev.pt3RPhiChiSquared, &ev.pt3RPhiChiSquaredInwards})
    v->assign(cap, 0.f);
  const unsigned nPLS = ev.nPLS;
  const unsigned plsBase = ev.segmentModuleIndices[nL];
  for (unsigned i = 0; i < nPLS; i++) {
    unsigned base, size;
    connectedRange(ev, i, base, size);
    for (unsigned k = base; k < base + size; k++) {
      const uint16_t m = G.connectedPixels[k];
      if (G.moduleType[m] == TwoS) continue;
      const unsigned nOuter = ev.nTriplets[m];
      if (nOuter == 0) continue;
      const unsigned pls = plsBase + i;
      if (ev.plsIsDup[i]) continue;
      if (ev.plsPartOfPT5[i]) continue;
      short adj;
      if (G.layers[m] == 1) adj = 1;
      else if (G.layers[m] == 2) adj = 0;
      else continue;
      for (unsigned b = 0; b < nOuter; b++) {
        const unsigned t3 = ev.tripletModuleIndices[m] + b;
        if (G.moduleType[ev.t3LowerModuleIndices[3 * t3 + 1]] == TwoS) continue;
        if (ev.t3PartOfPT5[t3]) continue;
        PT3Out o;
        if (!runPixelTripletDefaultAlgo(ev, pls, t3, true, o)) continue;
        const unsigned mdAdj = ev.segMdIndices[2 * ev.t3SegmentIndices[2 * t3] + adj];
        const float phi = ev.anchorPhi[mdAdj], eta = ev.anchorEta[mdAdj];
        const float score = o.rPhiChiSquared + o.rPhiChiSquaredInwards;
        if (ev.totOccupancyPixelTriplets++ >= cap) continue;
        const unsigned q = ev.nPixelTriplets++;
        // addPixelTripletToMemory (SDL/PixelTriplet.h:134-204)
        ev.pt3PixelSegmentIndices[q] = pls;
        ev.pt3TripletIndices[q] = t3;
        ev.pt3PixelRadius[q] = o.pixelRadius;
        ev.pt3TripletRadius[q] = o.tripletRadius;
        ev.pt3Pt[q] = ev.plsPtIn[i];
        ev.pt3Eta[q] = eta;
        ev.pt3Phi[q] = phi;
        ev.pt3EtaPix[q] = ev.plsEta[i];
        ev.pt3PhiPix[q] = ev.plsPhi[i];
        ev.pt3IsDup[q] = 0;
        ev.pt3Score[q] = score;
        ev.pt3CenterX[q] = o.centerX;
        ev.pt3CenterY[q] = o.centerY;
        ev.pt3LogicalLayers[5 * q] = 0;
        ev.pt3LogicalLayers[5 * q + 1] = 0;
        for (int j = 0; j < 3; j++) ev.pt3LogicalLayers[5 * q + 2 + j] = ev.t3LogicalLayers[3 * t3 + j];
        ev.pt3LowerModuleIndices[5 * q] = ev.segInnerLowerModule[pls];
        ev.pt3LowerModuleIndices[5 * q + 1] = ev.segOuterLowerModule[pls];
        for (int j = 0; j < 3; j++) ev.pt3LowerModuleIndices[5 * q + 2 + j] = ev.t3LowerModuleIndices[3 * t3 + j];
        const unsigned pi = ev.segMdIndices[2 * pls], po = ev.segMdIndices[2 * pls + 1];
        ev.pt3HitIndices[10 * q] = ev.mdAnchorHitIndices[pi];
        ev.pt3HitIndices[10 * q + 1] = ev.mdOuterHitIndices[pi];
        ev.pt3HitIndices[10 * q + 2] = ev.mdAnchorHitIndices[po];
        ev.pt3HitIndices[10 * q + 3] = ev.mdOuterHitIndices[po];
        for (int j = 0; j < 6; j++) ev.pt3HitIndices[10 * q + 4 + j] = ev.t3HitIndices[6 * t3 + j];
        ev.pt3RPhiChiSquared[q] = o.rPhiChiSquared;
        ev.pt3RPhiChiSquaredInwards[q] = o.rPhiChiSquaredInwards;
        ev.pt3RzChiSquared[q] = o.rzChiSquared;
        ev.t3PartOfPT3[t3] = 1;
      }
    }
  }
  // removeDupPixelTripletsInGPUFromMap: ix is dropped if a partner shares >= 5
  // hits (pixel + outer, counted by checkHitspT3, SDL/Kernels.h:137-198) and
  // reaches a lower third logical layer, or the same layer with a lower score
  // (ties: the lower index is dropped)
  const unsigned np = ev.nPixelTriplets;
  for (unsigned ix = 0; ix < np; ix++) {
    for (unsigned jx = 0; jx < np; jx++) {
      if (ix == jx) continue;
      int np1 = 0, nt = 0;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
          if ((int)ev.pt3HitIndices[10 * ix + a] == (int)ev.pt3HitIndices[10 * jx + b]) { np1++; break; }
      for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++)
          if ((int)ev.pt3HitIndices[10 * ix + 4 + a] == (int)ev.pt3HitIndices[10 * jx + 4 + b]) { nt++; break; }
      if (np1 + nt >= 5) {
        const uint8_t li = ev.pt3LogicalLayers[5 * ix + 2], lj = ev.pt3LogicalLayers[5 * jx + 2];
        const float si = ev.pt3Score[ix], sj = ev.pt3Score[jx];
        if (lj < li || (li == lj && si > sj) || (li == lj && si == sj && ix < jx)) {
          ev.pt3IsDup[ix] = 1;
          break;
        }
      }
    }
  }
}

// SDL/Event.cc:620-782
void event_create_track_candidates(Event& ev) {
  const Geometry& G = *ev.g;
  const unsigned nL = G.nLowerModules;
  const unsigned plsBase = ev.segmentModuleIndices[nL];
  const unsigned nPT3 = ev.nPixelTriplets, nPT5 = ev.nPixelQuintuplets;
  // crossCleanpT3 (SDL/TrackCandidate.h:204-241)
  for (unsigned p = 0; p < nPT3; p++) {
    if (ev.pt3IsDup[p]) continue;
    const float eta1 = ev.pt3EtaPix[p], phi1 = ev.pt3PhiPix[p];
    for (unsigned q = 0; q < nPT5; q++) {
      const unsigned j = ev.pt5PixelIndices[q] - plsBase;
      const float dEta = std::fabs(eta1 - ev.plsEta[j]);
      const float dPhi = calculate_dPhi(phi1, ev.plsPhi[j]);
      if (dEta * dEta + dPhi * dPhi < 1e-5f) ev.pt3IsDup[p] = 1;
    }
  }
  // addpT3asTrackCandidatesInGPU (SDL/TrackCandidate.h:381-428)
  for (unsigned p = 0; p < nPT3; p++) {
    if (ev.pt3IsDup[p]) continue;
    const unsigned idx = ev.nTrackCandidates++;
    if (idx >= kNMaxPixelTrackCandidates) {
      ev.nTrackCandidates--;
      break;
    }
    ev.nTrackCandidatespT3++;
    const float radius = 0.5f * (ev.pt3PixelRadius[p] + ev.pt3TripletRadius[p]);
    const unsigned pls = ev.pt3PixelSegmentIndices[p];
    addTC(ev, 5, p, p, &ev.pt3LogicalLayers[5 * p], &ev.pt3LowerModuleIndices[5 * p], &ev.pt3HitIndices[10 * p],
          (int)ev.plsSeedIdx[pls - plsBase], ev.pt3CenterX[p], ev.pt3CenterY[p], radius, idx, p);
  }
  // removeDupQuintupletsInGPUBeforeTC (SDL/Kernels.h:250-320): over pairs of
  // eligible modules (i <= j); skips read only the after-build bit, so the
  // flags do not depend on the pair order
  const unsigned nElig = ev.nEligibleT5Modules;
  for (unsigned a = 0; a < nElig; a++) {
    const uint16_t m1 = ev.indicesOfEligibleT5Modules[a];
    const unsigned n1 = ev.nQuintuplets[m1];
    if (n1 == 0) continue;
    const unsigned b1 = ev.quintupletModuleIndices[m1];
    for (unsigned b = a; b < nElig; b++) {
      const uint16_t m2 = ev.indicesOfEligibleT5Modules[b];
      const unsigned n2 = ev.nQuintuplets[m2];
      if (n2 == 0) continue;
      const unsigned b2 = ev.quintupletModuleIndices[m2];
      for (unsigned i1 = 0; i1 < n1; i1++) {
        const unsigned ix = b1 + i1;
        if (ev.t5PartOfPT5[ix] || (ev.t5IsDup[ix] & 1)) continue;
        for (unsigned j1 = 0; j1 < n2; j1++) {
          const unsigned jx = b2 + j1;
          if (ix == jx) continue;
          if (ev.t5PartOfPT5[jx] || (ev.t5IsDup[jx] & 1)) continue;
          const float dEta = std::fabs(ev.t5Eta[ix] - ev.t5Eta[jx]);
          const float dPhi = calculate_dPhi(ev.t5Phi[ix], ev.t5Phi[jx]);
          if (dEta > 0.1f) continue;
          if (std::fabs(dPhi) > 0.1f) continue;
          const float dR2 = dEta * dEta + dPhi * dPhi;
          int nMatched = 0;
          for (int u = 0; u < 10; u++)
            for (int w = 0; w < 10; w++)
              if (ev.t5HitIndices[10 * ix + u] == ev.t5HitIndices[10 * jx + w]) { nMatched++; break; }
          if (dR2 < 0.001f || nMatched >= 5) {
            const float s1 = ev.t5ScoreRphiSum[ix], s2 = ev.t5ScoreRphiSum[jx];
            if (s1 > s2) ev.t5IsDup[ix] |= 2;
            else if (s1 < s2) ev.t5IsDup[jx] |= 2;
            else ev.t5IsDup[ix < jx ? ix : jx] |= 2;
          }
        }
      }
    }
  }
  // crossCleanT5 with Crossclean_T5 (SDL/TrackCandidate.h:243-294)
  for (unsigned m = 0; m < nL; m++) {
    if (ev.quintupletModuleIndices[m] == -1) continue;
    const unsigned nq = ev.nQuintuplets[m];
    for (unsigned k = 0; k < nq; k++) {
      const unsigned q = ev.quintupletModuleIndices[m] + k;
      if (ev.t5IsDup[q] || ev.t5PartOfPT5[q]) continue;
      const float eta1 = ev.t5Eta[q], phi1 = ev.t5Phi[q];
      for (unsigned jx = 0; jx < nPT5 + nPT3; jx++) {
        float eta2, phi2;
        if (jx < nPT5) { eta2 = ev.pt5Eta[jx]; phi2 = ev.pt5Phi[jx]; }
        else { eta2 = ev.pt3Eta[jx - nPT5]; phi2 = ev.pt3Phi[jx - nPT5]; }
        const float dEta = std::fabs(eta1 - eta2);
        const float dPhi = calculate_dPhi(phi1, phi2);
        if (dEta * dEta + dPhi * dPhi < 1e-3f) ev.t5IsDup[q] = 1;
      }
    }
  }
  // addT5asTrackCandidateInGPU (SDL/TrackCandidate.h:430-482)
  for (unsigned m = 0; m < nL; m++) {
    if (ev.quintupletModuleIndices[m] == -1) continue;
    const unsigned nq = ev.nQuintuplets[m];
    for (unsigned k = 0; k < nq; k++) {
      const unsigned q = ev.quintupletModuleIndices[m] + k;
      if (ev.t5IsDup[q] || ev.t5PartOfPT5[q]) continue;
      if (!ev.t5TightCutFlag[q]) continue;
      const unsigned idx = ev.nTrackCandidates++;
      if (idx - ev.nTrackCandidatespT5 - ev.nTrackCandidatespT3 >= kNMaxNonPixelTrackCandidates) {
        ev.nTrackCandidates--;
        break;
      }
      ev.nTrackCandidatesT5++;
      addTC(ev, 4, q, q, &ev.t5LogicalLayers[5 * q], &ev.t5LowerModuleIndices[5 * q], &ev.t5HitIndices[10 * q], -1,
            ev.t5RegressionG[q], ev.t5RegressionF[q], ev.t5RegressionRadius[q], idx, q);
    }
  }
  // second pass of checkHitspLS
  if (!ev.cfg.noPlsDupClean) checkHitspLS(ev, true);
  // crossCleanpLS (SDL/TrackCandidate.h:296-379)
  const unsigned nPix = ev.nPLS;
  const unsigned nTC = ev.nTrackCandidates;
  for (unsigned i = 0; i < nPix; i++) {
    if (!ev.plsIsQuad[i] || ev.plsIsDup[i]) continue;
    const float eta1 = ev.plsEta[i], phi1 = ev.plsPhi[i];
    for (unsigned t = 0; t < nTC; t++) {
      const int16_t type = ev.tcType[t];
      const unsigned inner = ev.tcObjectIndices[2 * t];
      if (type == 4) {
        const float dEta = std::fabs(eta1 - ev.t5Eta[inner]);
        const float dPhi = calculate_dPhi(phi1, ev.t5Phi[inner]);
        if (dEta * dEta + dPhi * dPhi < 1e-3f) ev.plsIsDup[i] = 1;
      }
      if (type == 5) {
        const int plsIdx = ev.pt3PixelSegmentIndices[inner];
        if (checkPixelHits(ev, plsBase + i, plsIdx) > 0) ev.plsIsDup[i] = 1;
        const float dEta = std::fabs(eta1 - ev.pt3EtaPix[inner]);
        const float dPhi = calculate_dPhi(phi1, ev.pt3PhiPix[inner]);
        if (dEta * dEta + dPhi * dPhi < 0.000001f) ev.plsIsDup[i] = 1;
      }
      if (type == 7) {
        const unsigned plsIdx = inner;
        if (checkPixelHits(ev, plsBase + i, plsIdx) > 0) ev.plsIsDup[i] = 1;
        const float dEta = std::fabs(eta1 - ev.plsEta[plsIdx - plsBase]);
        const float dPhi = calculate_dPhi(phi1, ev.plsPhi[plsIdx - plsBase]);
        if (dEta * dEta + dPhi * dPhi < 0.000001f) ev.plsIsDup[i] = 1;
      }
    }
  }
  // addpLSasTrackCandidateInGPU (SDL/TrackCandidate.h:484-530)
  for (unsigned i = 0; i < nPix; i++) {
    if (ev.cfg.tcPlsTriplets ? (bool)ev.plsIsDup[i] : (!ev.plsIsQuad[i] || ev.plsIsDup[i])) continue;
    const unsigned idx = ev.nTrackCandidates++;
    if (idx - ev.nTrackCandidatesT5 >= kNMaxPixelTrackCandidates) {
      ev.nTrackCandidates--;
      break;
    }
    ev.nTrackCandidatespLS++;
    // addpLSTrackCandidateToMemory (SDL/TrackCandidate.h:112-124): hits in (x, z, y, w) order
    ev.tcType[idx] = 8;
    ev.tcDirectObjectIndices[idx] = i;
    ev.tcPixelSeedIndex[idx] = (int)ev.plsSeedIdx[i];
    ev.tcObjectIndices[2 * idx] = i;
    ev.tcObjectIndices[2 * idx + 1] = i;
    ev.tcHitIndices[14 * idx + 0] = ev.plsHitsIdxs[4 * i + 0];
    ev.tcHitIndices[14 * idx + 1] = ev.plsHitsIdxs[4 * i + 2];
    ev.tcHitIndices[14 * idx + 2] = ev.plsHitsIdxs[4 * i + 1];
    ev.tcHitIndices[14 * idx + 3] = ev.plsHitsIdxs[4 * i + 3];
  }
}

// SDL/LST.cc:366-420
void event_get_output(Event& ev) {
  ev.outHitIdxs.clear(); ev.outHitOffsets.assign(1, 0); ev.outLen.clear(); ev.outSeedIdx.clear(); ev.outType.clear();
  for (unsigned t = 0; t < ev.nTrackCandidates; t++) {
    const int16_t type = ev.tcType[t];
    unsigned maxN = 0;
    if (type == 7) maxN = 14;
    else if (type == 5) maxN = 10;
    else if (type == 4) maxN = 10;
    else if (type == 8) maxN = 4;
    std::vector<unsigned> hits;
    for (unsigned i = 0; i < maxN; i++) {
      const unsigned h = ev.tcHitIndices[14 * t + i];
      const unsigned hitIdx = type == 8 ? h : ev.idxs[h];
      if (type != 4 && hits.size() == 3 && hits.back() == hitIdx) continue;
      hits.push_back(hitIdx);
    }
    ev.outHitIdxs.insert(ev.outHitIdxs.end(), hits.begin(), hits.end());
    ev.outHitOffsets.push_back(ev.outHitIdxs.size());
    ev.outLen.push_back(hits.size());
    ev.outSeedIdx.push_back(ev.tcPixelSeedIndex[t]);
    ev.outType.push_back(type);
  }
}

}  // namespace orc

// The chi-squared cut tables by name, for the tests that probe them against
// the thresholds extracted from the reference: -1 for an unknown name.
extern "C" int orc_pixel_cut_table(const char* name, const int* l, float value) {
  using namespace orc;
  const std::string n(name);
  if (n == "passPT3RZChiSquaredCuts") return passPT3RZChiSquaredCuts(l[0], l[1], l[2], value);
  if (n == "passPT3RPhiChiSquaredCuts") return passPT3RPhiChiSquaredCuts(l[0], l[1], l[2], value);
  if (n == "passPT3RPhiChiSquaredInwardsCuts") return passPT3RPhiChiSquaredInwardsCuts(l[0], l[1], l[2], value);
  if (n == "passPT5RZChiSquaredCuts") return passPT5RZChiSquaredCuts(l[0], l[1], l[2], l[3], l[4], value);
  if (n == "passPT5RPhiChiSquaredCuts") return passPT5RPhiChiSquaredCuts(l[0], l[1], l[2], l[3], l[4], value);
  if (n == "passPT5RPhiChiSquaredInwardsCuts")
    return passPT5RPhiChiSquaredInwardsCuts(l[0], l[1], l[2], l[3], l[4], value);
  return -1;
}

