// ORACLE — test infrastructure only (see orc_common.h).
// Line segments (SDL/Segment.h:177-1029, SDL/Event.cc:469-510) and
// triplets (SDL/Triplet.h:153-1123, SDL/Event.cc:512-618), serial order.
#include "orc_event.h"

namespace orc {
namespace {

// SDL/Segment.h:177-200 (isTighterTiltedModules_seg) and :202-262 (moduleGapSize_seg)
float gapSizeSeg(const Geometry& g, unsigned m) {
  static const float tilted[3] = {0.26f, 0.26f, 0.26f};
  static const float flat[6] = {0.26f, 0.16f, 0.16f, 0.18f, 0.18f, 0.18f};
  static const float looseTilted[3] = {0.4f, 0.4f, 0.4f};
  static const float endcap[5][15] = {
      {0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f},
      {0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f},
      {0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f},
      {0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f},
      {0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.4f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f, 0.18f}};
  unsigned iL = g.layers[m] - 1, iR = g.rings[m] - 1;
  short subdet = g.subdets[m], side = g.sides[m], layer = g.layers[m], rod = g.rods[m];
  bool tighter = (subdet == Barrel) &&
                 (((side != Center) && (layer == 3)) ||
                  ((side == NegZ) && (((layer == 2) && (rod > 5)) || ((layer == 1) && (rod > 9)))) ||
                  ((side == PosZ) && (((layer == 2) && (rod < 8)) || ((layer == 1) && (rod < 4)))));
  if (subdet == Barrel && side == Center) return flat[iL];
  if (tighter) return tilted[iL];
  if (subdet == Endcap) return endcap[iL][iR];
  return looseTilted[iL];
}

// SDL/Segment.h:264-356
void dAlphaThreshold(const Event& ev, float* out, float xIn, float yIn, float zIn, float rtIn, float xOut, float yOut,
                     float zOut, float rtOut, uint16_t inM, uint16_t outM, unsigned inMD, unsigned outMD) {
  const Geometry& g = *ev.g;
  const float ptCut = ev.cfg.ptCut;
  float sdMuls = (g.subdets[inM] == Barrel) ? miniMulsPtScaleBarrel[g.layers[inM] - 1] * 3.f / ptCut
                                            : miniMulsPtScaleEndcap[g.layers[inM] - 1] * 3.f / ptCut;
  float segmentDr = fsqrt((yOut - yIn) * (yOut - yIn) + (xOut - xIn) * (xOut - xIn));
  const float dAlpha_Bfield = fasin(fminf_(segmentDr * k2Rinv1GeVf / ptCut, sinAlphaMax));
  bool isInnerTilted = g.subdets[inM] == Barrel && g.sides[inM] != Center;
  bool isOuterTilted = g.subdets[outM] == Barrel && g.sides[outM] != Center;
  const float drdzInner = g.drdzs[inM];
  const float drdzOuter = g.drdzs[outM];
  float innerGap = gapSizeSeg(g, inM);
  float outerGap = gapSizeSeg(g, outM);
  const float innerminiTilt =
      isInnerTilted ? (0.5f * pixelPSZpitch * drdzInner / fsqrt(1.f + drdzInner * drdzInner) / innerGap) : 0;
  const float outerminiTilt =
      isOuterTilted ? (0.5f * pixelPSZpitch * drdzOuter / fsqrt(1.f + drdzOuter * drdzOuter) / outerGap) : 0;
  float miniDelta = innerGap;
  float sdLumIn, sdLumOut;
  if (g.subdets[inM] == Barrel) sdLumIn = innerminiTilt * dAlpha_Bfield;
  else sdLumIn = ev.mdDphis[inMD] * 15.0f / ev.mdDzs[inMD];
  if (g.subdets[outM] == Barrel) sdLumOut = outerminiTilt * dAlpha_Bfield;
  else sdLumOut = ev.mdDphis[outMD] * 15.0f / ev.mdDzs[outMD];
  float resIn = 0.02f / miniDelta * (g.subdets[inM] == Barrel ? 1.0f : std::fabs(zIn) / rtIn);
  float resOut = 0.02f / miniDelta * (g.subdets[outM] == Barrel ? 1.0f : std::fabs(zOut) / rtOut);
  float res = resIn + resOut;
  if (g.subdets[inM] == Barrel && g.sides[inM] == Center)
    out[0] = dAlpha_Bfield + fsqrt(res * res + sdMuls * sdMuls);
  else
    out[0] = dAlpha_Bfield + fsqrt(res * res + sdMuls * sdMuls + sdLumIn * sdLumIn);
  if (g.subdets[outM] == Barrel && g.sides[outM] == Center)
    out[1] = dAlpha_Bfield + fsqrt(res * res + sdMuls * sdMuls);
  else
    out[1] = dAlpha_Bfield + fsqrt(res * res + sdMuls * sdMuls + sdLumOut * sdLumOut);
  out[2] = dAlpha_Bfield + fsqrt(res * res + sdMuls * sdMuls);
}

struct SegOut {
  float dPhi = 0, dPhiMin = 0, dPhiMax = 0, dPhiChange = 0, dPhiChangeMin = 0, dPhiChangeMax = 0;
};

// SDL/Segment.h:454-558
bool segBarrel(const Event& ev, uint16_t inM, uint16_t outM, unsigned inMD, unsigned outMD, SegOut& o) {
  const Geometry& g = *ev.g;
  const float ptCut = ev.cfg.ptCut;
  float sdMuls = (g.subdets[inM] == Barrel) ? miniMulsPtScaleBarrel[g.layers[inM] - 1] * 3.f / ptCut
                                            : miniMulsPtScaleEndcap[g.layers[inM] - 1] * 3.f / ptCut;
  float xIn = ev.anchorX[inMD], yIn = ev.anchorY[inMD], zIn = ev.anchorZ[inMD], rtIn = ev.anchorRt[inMD];
  float xOut = ev.anchorX[outMD], yOut = ev.anchorY[outMD], zOut = ev.anchorZ[outMD], rtOut = ev.anchorRt[outMD];
  float sdSlope = fasin(fminf_(rtOut * k2Rinv1GeVf / ptCut, sinAlphaMax));
  float sdPVoff = 0.1f / rtOut;
  float dzDrtScale = ftan(sdSlope) / sdSlope;
  const float zGeom = g.layers[inM] <= 2 ? 2.f * pixelPSZpitch : 2.f * strip2SZpitch;
  float zLo = zIn + (zIn - deltaZLum) * (rtOut / rtIn - 1.f) * (zIn > 0.f ? 1.f : dzDrtScale) - zGeom;
  float zHi = zIn + (zIn + deltaZLum) * (rtOut / rtIn - 1.f) * (zIn < 0.f ? 1.f : dzDrtScale) + zGeom;
  if (!((zOut >= zLo) && (zOut <= zHi))) return false;
  float sdCut = sdSlope + fsqrt(sdMuls * sdMuls + sdPVoff * sdPVoff);
  o.dPhi = phi_mpi_pi(ev.anchorPhi[outMD] - ev.anchorPhi[inMD]);
  if (!(std::fabs(o.dPhi) <= sdCut)) return false;
  o.dPhiChange = phi_mpi_pi(phi(xOut - xIn, yOut - yIn) - ev.anchorPhi[inMD]);
  if (!(std::fabs(o.dPhiChange) <= sdCut)) return false;
  float th[3];
  dAlphaThreshold(ev, th, xIn, yIn, zIn, rtIn, xOut, yOut, zOut, rtOut, inM, outM, inMD, outMD);
  float inA = ev.mdDphichanges[inMD], outA = ev.mdDphichanges[outMD];
  if (!(std::fabs(inA - o.dPhiChange) < th[0])) return false;
  if (!(std::fabs(outA - o.dPhiChange) < th[1])) return false;
  return std::fabs(inA - outA) < th[2];
}

// SDL/Segment.h:560-704
bool segEndcap(const Event& ev, uint16_t inM, uint16_t outM, unsigned inMD, unsigned outMD, SegOut& o) {
  const Geometry& g = *ev.g;
  const float ptCut = ev.cfg.ptCut;
  float xIn = ev.anchorX[inMD], yIn = ev.anchorY[inMD], zIn = ev.anchorZ[inMD], rtIn = ev.anchorRt[inMD];
  float xOut = ev.anchorX[outMD], yOut = ev.anchorY[outMD], zOut = ev.anchorZ[outMD], rtOut = ev.anchorRt[outMD];
  bool outerLayerEndcapTwoS = (g.subdets[outM] == Endcap) && (g.moduleType[outM] == TwoS);
  float sdSlope = fasin(fminf_(rtOut * k2Rinv1GeVf / ptCut, sinAlphaMax));
  float disks2SMinRadius = 60.f;
  float rtGeom = ((rtIn < disks2SMinRadius && rtOut < disks2SMinRadius)
                      ? (2.f * pixelPSZpitch)
                      : ((rtIn < disks2SMinRadius || rtOut < disks2SMinRadius) ? (pixelPSZpitch + strip2SZpitch)
                                                                               : (2.f * strip2SZpitch)));
  if (!(zIn * zOut >= 0)) return false;
  float dz = zOut - zIn;
  float dLum = copysignf_(deltaZLum, zIn);
  float drtDzScale = sdSlope / ftan(sdSlope);
  float rtLo = fmaxf_(rtIn * (1.f + dz / (zIn + dLum) * drtDzScale) - rtGeom, rtIn - 0.5f * rtGeom);
  float rtHi = rtIn * (zOut - dLum) / (zIn - dLum) + rtGeom;
  if (!((rtOut >= rtLo) && (rtOut <= rtHi))) return false;
  o.dPhi = phi_mpi_pi(ev.anchorPhi[outMD] - ev.anchorPhi[inMD]);
  float sdCut = sdSlope;
  if (outerLayerEndcapTwoS) {
    float hi = phi_mpi_pi(ev.anchorHighEdgePhi[outMD] - ev.anchorPhi[inMD]);
    float lo = phi_mpi_pi(ev.anchorLowEdgePhi[outMD] - ev.anchorPhi[inMD]);
    o.dPhiMax = std::fabs(hi) > std::fabs(lo) ? hi : lo;
    o.dPhiMin = std::fabs(hi) > std::fabs(lo) ? lo : hi;
  } else {
    o.dPhiMax = o.dPhi;
    o.dPhiMin = o.dPhi;
  }
  if (!(std::fabs(o.dPhi) <= sdCut)) return false;
  float dzFrac = dz / zIn;
  o.dPhiChange = o.dPhi / dzFrac * (1.f + dzFrac);
  o.dPhiChangeMin = o.dPhiMin / dzFrac * (1.f + dzFrac);
  o.dPhiChangeMax = o.dPhiMax / dzFrac * (1.f + dzFrac);
  if (!(std::fabs(o.dPhiChange) <= sdCut)) return false;
  float th[3];
  dAlphaThreshold(ev, th, xIn, yIn, zIn, rtIn, xOut, yOut, zOut, rtOut, inM, outM, inMD, outMD);
  float inA = ev.mdDphichanges[inMD], outA = ev.mdDphichanges[outMD];
  if (!(std::fabs(inA - o.dPhiChange) < th[0])) return false;
  if (!(std::fabs(outA - o.dPhiChange) < th[1])) return false;
  return std::fabs(inA - outA) < th[2];
}

}  // namespace

// SDL/Segment.h:798-910 and :1010-1029
void event_create_segments(Event& ev) {
  const Geometry& g = *ev.g;
  for (uint16_t inM = 0; inM < g.nLowerModules; inM++) {
    unsigned nInner = ev.nMDs[inM];
    if (nInner == 0) continue;
    unsigned nConn = g.nConnectedModules[inM];
    for (unsigned c = 0; c < nConn; c++) {
      uint16_t outM = g.moduleMap[(size_t)inM * g.stride + c];
      unsigned nOuter = ev.nMDs[outM];
      unsigned limit = nInner * nOuter;
      for (unsigned h = 0; h < limit; h++) {
        unsigned inMD = ev.mdRanges[inM * 2] + h / nOuter;
        unsigned outMD = ev.mdRanges[outM * 2] + h % nOuter;
        SegOut o;
        bool pass = (g.subdets[inM] == Barrel && g.subdets[outM] == Barrel) ? segBarrel(ev, inM, outM, inMD, outMD, o)
                                                                            : segEndcap(ev, inM, outM, inMD, outMD, o);
        if (!pass) continue;
        unsigned occ = ev.totOccupancySegments[inM]++;
        if ((int)occ >= ev.segmentModuleOccupancy[inM]) continue;
        unsigned idx = ev.segmentModuleIndices[inM] + ev.nSegments[inM]++;
        ev.segMdIndices[idx * 2] = inMD;
        ev.segMdIndices[idx * 2 + 1] = outMD;
        ev.segInnerLowerModule[idx] = inM;
        ev.segOuterLowerModule[idx] = outM;
        ev.segInnerAnchorHit[idx] = ev.mdAnchorHitIndices[inMD];
        ev.segOuterAnchorHit[idx] = ev.mdAnchorHitIndices[outMD];
        ev.segDPhis[idx] = o.dPhi; ev.segDPhiMins[idx] = o.dPhiMin; ev.segDPhiMaxs[idx] = o.dPhiMax;
        ev.segDPhiChanges[idx] = o.dPhiChange; ev.segDPhiChangeMins[idx] = o.dPhiChangeMin;
        ev.segDPhiChangeMaxs[idx] = o.dPhiChangeMax;
      }
    }
  }
  for (unsigned i = 0; i < g.nLowerModules; i++) {
    if (ev.nSegments[i] == 0) {
      ev.segmentRanges[i * 2] = -1; ev.segmentRanges[i * 2 + 1] = -1;
    } else {
      ev.segmentRanges[i * 2] = ev.segmentModuleIndices[i];
      ev.segmentRanges[i * 2 + 1] = ev.segmentModuleIndices[i] + ev.nSegments[i] - 1;
    }
  }
}

namespace {

// SDL/Triplet.h:241-298
bool passRZConstraint(const Event& ev, uint16_t m1, uint16_t m2, uint16_t m3, unsigned md1, unsigned md2, unsigned md3) {
  const Geometry& g = *ev.g;
  const float r1 = ev.anchorRt[md1], r2 = ev.anchorRt[md2], r3 = ev.anchorRt[md3];
  const float z1 = ev.anchorZ[md1], z2 = ev.anchorZ[md2], z3 = ev.anchorZ[md3];
  const int l1 = g.sdlLayers[m1], l2 = g.sdlLayers[m2], l3 = g.sdlLayers[m3];
  const float residual = z2 - ((z3 - z1) / (r3 - r1) * (r2 - r1) + z1);
  const float a = std::fabs(residual);
  if (l1 == 12 && l2 == 13 && l3 == 14) return false;
  if (l1 == 1 && l2 == 2 && l3 == 3) return a < 0.53f;
  if (l1 == 1 && l2 == 2 && l3 == 7) return a < 1;
  if (l1 == 13 && l2 == 14 && l3 == 15) return false;
  if (l1 == 14 && l2 == 15 && l3 == 16) return false;
  if (l1 == 1 && l2 == 7 && l3 == 8) return a < 1;
  if (l1 == 2 && l2 == 3 && l3 == 4) return a < 1.21f;
  if (l1 == 2 && l2 == 3 && l3 == 7) return a < 1.f;
  if (l1 == 2 && l2 == 7 && l3 == 8) return a < 1.f;
  if (l1 == 3 && l2 == 4 && l3 == 5) return a < 2.7f;
  if (l1 == 4 && l2 == 5 && l3 == 6) return a < 3.06f;
  if (l1 == 7 && l2 == 8 && l3 == 9) return a < 1;
  if (l1 == 8 && l2 == 9 && l3 == 10) return a < 1;
  if (l1 == 9 && l2 == 10 && l3 == 11) return a < 1;
  return a < 5;
}

// SDL/Triplet.h:300-402
bool pointingBBB(const Event& ev, uint16_t m1, uint16_t m3, unsigned md1, unsigned md2, unsigned md3, unsigned inSeg,
                 float& betaIn) {
  const Geometry& g = *ev.g;
  const float ptCut = ev.cfg.ptCut;
  bool isPSIn = g.moduleType[m1] == PS, isPSOut = g.moduleType[m3] == PS;
  float rtIn = ev.anchorRt[md1], rtMid = ev.anchorRt[md2], rtOut = ev.anchorRt[md3];
  float zIn = ev.anchorZ[md1], zMid = ev.anchorZ[md2], zOut = ev.anchorZ[md3];
  float alpha1GeVOut = fasin(fminf_(rtOut * k2Rinv1GeVf / ptCut, sinAlphaMax));
  float rtRatio = rtOut / rtIn;
  float dzDrtScale = ftan(alpha1GeVOut) / alpha1GeVOut;
  float zpitchIn = isPSIn ? pixelPSZpitch : strip2SZpitch;
  float zpitchOut = isPSOut ? pixelPSZpitch : strip2SZpitch;
  const float zHi = zIn + (zIn + deltaZLum) * (rtRatio - 1.f) * (zIn < 0.f ? 1.f : dzDrtScale) + (zpitchIn + zpitchOut);
  const float zLo = zIn + (zIn - deltaZLum) * (rtRatio - 1.f) * (zIn > 0.f ? 1.f : dzDrtScale) - (zpitchIn + zpitchOut);
  if (!((zOut >= zLo) && (zOut <= zHi))) return false;
  float drt_OutIn = rtOut - rtIn;
  float r3In = fsqrt(zIn * zIn + rtIn * rtIn);
  float drt_InSeg = rtMid - rtIn;
  float dz_InSeg = zMid - zIn;
  float dr3_InSeg = fsqrt(rtMid * rtMid + zMid * zMid) - fsqrt(rtIn * rtIn + zIn * zIn);
  float coshEta = dr3_InSeg / drt_InSeg;
  float dzErr = (zpitchIn + zpitchOut) * (zpitchIn + zpitchOut) * 2.f;
  float sdlThetaMulsF = 0.015f * fsqrt(0.1f + 0.2f * (rtOut - rtIn) / 50.f) * fsqrt(r3In / rtIn);
  float sdlMuls = sdlThetaMulsF * 3.f / ptCut * 4.f;
  dzErr += sdlMuls * sdlMuls * drt_OutIn * drt_OutIn / 3.f * coshEta * coshEta;
  dzErr = fsqrt(dzErr);
  const float dzMean = dz_InSeg / drt_InSeg * drt_OutIn;
  const float zWindow = dzErr / drt_InSeg * drt_OutIn + (zpitchIn + zpitchOut);
  const float zLoPointed = zIn + dzMean * (zIn > 0.f ? 1.f : dzDrtScale) - zWindow;
  const float zHiPointed = zIn + dzMean * (zIn < 0.f ? 1.f : dzDrtScale) + zWindow;
  bool pass = (zOut >= zLoPointed) && (zOut <= zHiPointed);
  float alpha_InLo = ev.segDPhiChanges[inSeg];
  float tlx = ev.anchorX[md3] - ev.anchorX[md1], tly = ev.anchorY[md3] - ev.anchorY[md1];
  betaIn = alpha_InLo - phi_mpi_pi(phi(tlx, tly) - ev.anchorPhi[md1]);
  float drt_tl_axis = fsqrt(tlx * tlx + tly * tly);
  float dx = ev.anchorX[md2] - ev.anchorX[md1], dy = ev.anchorY[md2] - ev.anchorY[md1];
  const float rt_InSeg = fsqrt(dx * dx + dy * dy);
  float betaInCut = fasin(fminf_((-rt_InSeg + drt_tl_axis) * k2Rinv1GeVf / ptCut, sinAlphaMax)) + (0.02f / drt_InSeg);
  return pass && (std::fabs(betaIn) < betaInCut);
}

// SDL/Triplet.h:404-526
bool pointingBBE(const Event& ev, uint16_t m1, uint16_t m3, unsigned md1, unsigned md2, unsigned md3, unsigned inSeg,
                 float& betaIn) {
  const Geometry& g = *ev.g;
  const float ptCut = ev.cfg.ptCut;
  bool isPSIn = g.moduleType[m1] == PS, isPSOut = g.moduleType[m3] == PS;
  float rtIn = ev.anchorRt[md1], rtMid = ev.anchorRt[md2], rtOut = ev.anchorRt[md3];
  float zIn = ev.anchorZ[md1], zMid = ev.anchorZ[md2], zOut = ev.anchorZ[md3];
  float alpha1GeV_OutLo = fasin(fminf_(rtOut * k2Rinv1GeVf / ptCut, sinAlphaMax));
  float dzDrtScale = ftan(alpha1GeV_OutLo) / alpha1GeV_OutLo;
  float zpitchIn = isPSIn ? pixelPSZpitch : strip2SZpitch;
  float zpitchOut = isPSOut ? pixelPSZpitch : strip2SZpitch;
  float zGeom = zpitchIn + zpitchOut;
  if (!(zIn * zOut > 0)) return false;
  float dLum = copysignf_(deltaZLum, zIn);
  bool isOutSgInnerMDPS = g.moduleType[m3] == PS;
  float rtGeom1 = isOutSgInnerMDPS ? pixelPSZpitch : strip2SZpitch;
  float zGeom1 = copysignf_(zGeom, zIn);
  float rtLo = rtIn * (1.f + (zOut - zIn - zGeom1) / (zIn + zGeom1 + dLum) / dzDrtScale) - rtGeom1;
  float zInForHi = zIn - zGeom1 - dLum;
  if (zInForHi * zIn < 0) zInForHi = copysignf_(0.1f, zIn);
  float rtHi = rtIn * (1.f + (zOut - zIn + zGeom1) / zInForHi) + rtGeom1;
  if (!((rtOut >= rtLo) && (rtOut <= rtHi))) return false;
  float rIn = fsqrt(zIn * zIn + rtIn * rtIn);
  const float drtSDIn = rtMid - rtIn;
  const float dzSDIn = zMid - zIn;
  const float dr3SDIn = fsqrt(rtMid * rtMid + zMid * zMid) - fsqrt(rtIn * rtIn + zIn * zIn);
  const float coshEta = dr3SDIn / drtSDIn;
  const float dzOutInAbs = std::fabs(zOut - zIn);
  const float multDzDr = dzOutInAbs * coshEta / (coshEta * coshEta - 1.f);
  const float zGeom1_another = pixelPSZpitch;
  const float kZ = (zOut - zIn) / dzSDIn;
  float drtErr = zGeom1_another * zGeom1_another * drtSDIn * drtSDIn / dzSDIn / dzSDIn * (1.f - 2.f * kZ + 2.f * kZ * kZ);
  // double-precision site: `0.2 *` is a double literal (SDL/Triplet.h:484)
  const float sdlThetaMulsF =
      0.015f * std::sqrt(0.1f + 0.2 * (rtOut - rtIn) / 50.f) * fsqrt(rIn / rtIn);
  const float sdlMuls = sdlThetaMulsF * 3.f / ptCut * 4.f;
  drtErr += sdlMuls * sdlMuls * multDzDr * multDzDr / 3.f * coshEta * coshEta;
  drtErr = fsqrt(drtErr);
  (void)drtErr;
  bool pass = (kZ >= 0) && (rtOut >= rtLo) && (rtOut <= rtHi);
  float rt_InLo = ev.anchorRt[md1], rt_InOut = ev.anchorRt[md2];
  float sdIn_alpha = ev.segDPhiChanges[inSeg];
  float tlx = ev.anchorX[md3] - ev.anchorX[md1], tly = ev.anchorY[md3] - ev.anchorY[md1];
  betaIn = sdIn_alpha - phi_mpi_pi(phi(tlx, tly) - ev.anchorPhi[md1]);
  float betaInRHmin = betaIn, betaInRHmax = betaIn;
  if (std::fabs(betaInRHmin) > std::fabs(betaInRHmax)) std::swap(betaInRHmin, betaInRHmax);
  float dx = ev.anchorX[md2] - ev.anchorX[md1], dy = ev.anchorY[md2] - ev.anchorY[md1];
  float sdIn_dr = fsqrt(dx * dx + dy * dy);
  float sdIn_d = rt_InOut - rt_InLo;
  float dr = fsqrt(tlx * tlx + tly * tly);
  float betaInCut = fasin(fminf_((-sdIn_dr + dr) * k2Rinv1GeVf / ptCut, sinAlphaMax)) + (0.02f / sdIn_d);
  return pass && (std::fabs(betaInRHmin) < betaInCut);
}

// SDL/Triplet.h:528-663
bool pointingEEE(const Event& ev, uint16_t m1, uint16_t m3, unsigned md1, unsigned md2, unsigned md3, unsigned inSeg,
                 float& betaIn) {
  const Geometry& g = *ev.g;
  const float ptCut = ev.cfg.ptCut;
  float rtIn = ev.anchorRt[md1], rtMid = ev.anchorRt[md2], rtOut = ev.anchorRt[md3];
  float zIn = ev.anchorZ[md1], zMid = ev.anchorZ[md2], zOut = ev.anchorZ[md3];
  float alpha1GeV_Out = fasin(fminf_(rtOut * k2Rinv1GeVf / ptCut, sinAlphaMax));
  float dzDrtScale = ftan(alpha1GeV_Out) / alpha1GeV_Out;
  if (!(zIn * zOut > 0)) return false;
  float dLum = copysignf_(deltaZLum, zIn);
  bool isOutSgOuterMDPS = g.moduleType[m3] == PS;
  bool isInSgInnerMDPS = g.moduleType[m1] == PS;
  float rtGeom = (isInSgInnerMDPS && isOutSgOuterMDPS)   ? 2.f * pixelPSZpitch
                 : (isInSgInnerMDPS || isOutSgOuterMDPS) ? pixelPSZpitch + strip2SZpitch
                                                         : 2.f * strip2SZpitch;
  float dz = zOut - zIn;
  const float rtLo = rtIn * (1.f + dz / (zIn + dLum) / dzDrtScale) - rtGeom;
  const float rtHi = rtIn * (1.f + dz / (zIn - dLum)) + rtGeom;
  if (!((rtOut >= rtLo) && (rtOut <= rtHi))) return false;
  bool isInSgOuterMDPS = g.moduleType[m3] == PS;
  float drtSDIn = rtMid - rtIn;
  float dzSDIn = zMid - zIn;
  float dr3SDIn = fsqrt(rtMid * rtMid + zMid * zMid) - fsqrt(rtIn * rtIn + zIn * zIn);
  float coshEta = dr3SDIn / drtSDIn;
  float dzOutInAbs = std::fabs(zOut - zIn);
  float multDzDr = dzOutInAbs * coshEta / (coshEta * coshEta - 1.f);
  float kZ = (zOut - zIn) / dzSDIn;
  float sdlThetaMulsF = 0.015f * fsqrt(0.1f + 0.2f * (rtOut - rtIn) / 50.f);
  float sdlMuls = sdlThetaMulsF * 3.f / ptCut * 4.f;
  float drtErr = fsqrt(pixelPSZpitch * pixelPSZpitch * 2.f / (dzSDIn * dzSDIn) * (dzOutInAbs * dzOutInAbs) +
                       sdlMuls * sdlMuls * multDzDr * multDzDr / 3.f * coshEta * coshEta);
  float drtMean = drtSDIn * dzOutInAbs / std::fabs(dzSDIn);
  float rtWindow = drtErr + rtGeom;
  float rtLo_point = rtIn + drtMean / dzDrtScale - rtWindow;
  float rtHi_point = rtIn + drtMean + rtWindow;
  bool pass = true;
  if (isInSgInnerMDPS && isInSgOuterMDPS) pass = (kZ >= 0) && (rtOut >= rtLo_point) && (rtOut <= rtHi_point);
  float rt_InLo = ev.anchorRt[md1], rt_InOut = ev.anchorRt[md2];
  float sdIn_alpha = ev.segDPhiChanges[inSeg];
  float tlx = ev.anchorX[md3] - ev.anchorX[md1], tly = ev.anchorY[md3] - ev.anchorY[md1];
  betaIn = sdIn_alpha - phi_mpi_pi(phi(tlx, tly) - ev.anchorPhi[md1]);
  float sdIn_alphaRHmin = ev.segDPhiChangeMins[inSeg];
  float sdIn_alphaRHmax = ev.segDPhiChangeMaxs[inSeg];
  float betaInRHmin = betaIn + sdIn_alphaRHmin - sdIn_alpha;
  float betaInRHmax = betaIn + sdIn_alphaRHmax - sdIn_alpha;
  if (std::fabs(betaInRHmin) > std::fabs(betaInRHmax)) std::swap(betaInRHmin, betaInRHmax);
  float dx = ev.anchorX[md2] - ev.anchorX[md1], dy = ev.anchorY[md2] - ev.anchorY[md1];
  float sdIn_dr = fsqrt(dx * dx + dy * dy);
  float sdIn_d = rt_InOut - rt_InLo;
  float dr = fsqrt(tlx * tlx + tly * tly);
  float betaInCut = fasin(fminf_((-sdIn_dr + dr) * k2Rinv1GeVf / ptCut, sinAlphaMax)) + (0.02f / sdIn_d);
  return pass && (std::fabs(betaInRHmin) < betaInCut);
}

}  // namespace

// SDL/Triplet.h:762-794
float computeRadiusFromThreeAnchorHits(float x1, float y1, float x2, float y2, float x3, float y3, float& g, float& f) {
  float radius = 0.f;
  float denomInv = 1.0f / ((y1 - y3) * (x2 - x3) - (x1 - x3) * (y2 - y3));
  float xy1sqr = x1 * x1 + y1 * y1;
  float xy2sqr = x2 * x2 + y2 * y2;
  float xy3sqr = x3 * x3 + y3 * y3;
  g = 0.5f * ((y3 - y2) * xy1sqr + (y1 - y3) * xy2sqr + (y2 - y1) * xy3sqr) * denomInv;
  f = 0.5f * ((x2 - x3) * xy1sqr + (x3 - x1) * xy2sqr + (x1 - x2) * xy3sqr) * denomInv;
  float c = ((x2 * y3 - x3 * y2) * xy1sqr + (x3 * y1 - x1 * y3) * xy2sqr + (x1 * y2 - x2 * y1) * xy3sqr) * denomInv;
  if (((y1 - y3) * (x2 - x3) - (x1 - x3) * (y2 - y3) == 0) || (g * g + f * f - c < 0)) radius = -1.f;
  else radius = fsqrt(g * g + f * f - c);
  return radius;
}

// SDL/Event.cc:512-618, SDL/Triplet.h:796-1123
void event_create_triplets(Event& ev) {
  const Geometry& g = *ev.g;
  const unsigned nL = g.nLowerModules;
  // createTripletArrayRanges: SDL/Triplet.h:1007-1102
  int total = 0;
  for (unsigned i = 0; i < nL; i++) {
    if (ev.nSegments[i] == 0) {
      ev.tripletModuleIndices[i] = total;
      ev.tripletModuleOccupancy[i] = 0;
      continue;
    }
    int en = 0, occ;
    short rings = g.rings[i], layers = g.layers[i], subdets = g.subdets[i];
    float e = std::fabs(g.eta[i]);
    int cat;
    if (layers <= 3 && subdets == 5) cat = 0;
    else if (layers >= 4 && subdets == 5) cat = 1;
    else if (layers <= 2 && subdets == 4 && rings >= 11) cat = 2;
    else if (layers >= 3 && subdets == 4 && rings >= 8) cat = 2;
    else if (layers <= 2 && subdets == 4 && rings <= 10) cat = 3;
    else if (layers >= 3 && subdets == 4 && rings <= 7) cat = 3;
    else cat = -1;
    if (e < 0.75) en = 0;
    else if (e > 0.75 && e < 1.5) en = 1;
    else if (e > 1.5 && e < 2.25) en = 2;
    else if (e > 2.25 && e < 3) en = 3;
    else en = -1;
    if (cat == 0 && en == 0) occ = 543;
    else if (cat == 0 && en == 1) occ = 235;
    else if (cat == 0 && en == 2) occ = 88;
    else if (cat == 0 && en == 3) occ = 46;
    else if (cat == 1 && en == 0) occ = 755;
    else if (cat == 1 && en == 1) occ = 347;
    else if (cat == 2 && en == 1) occ = 0;
    else if (cat == 2 && en == 2) occ = 0;
    else if (cat == 3 && en == 1) occ = 38;
    else if (cat == 3 && en == 2) occ = 46;
    else if (cat == 3 && en == 3) occ = 39;
    else occ = 0;
    ev.tripletModuleOccupancy[i] = occ;
    ev.tripletModuleIndices[i] = total;
    total += occ;
  }
  ev.nTotalTrips = total;
  const unsigned n = total;
  ev.t3Capacity = n;
  ev.nTriplets.assign(nL, 0); ev.totOccupancyTriplets.assign(nL, 0);
  ev.t3SegmentIndices.assign(2 * n, 0); ev.t3HitIndices.assign(6 * n, 0); ev.t3LowerModuleIndices.assign(3 * n, 0);
  ev.t3LogicalLayers.assign(3 * n, 0);
  ev.t3PartOfPT5.assign(n, 0); ev.t3PartOfT5.assign(n, 0); ev.t3PartOfPT3.assign(n, 0);
  ev.t3BetaIn.assign(n, 0); ev.t3CircleRadius.assign(n, 0); ev.t3CircleCenterX.assign(n, 0); ev.t3CircleCenterY.assign(n, 0);

  // nonZeroModules host loop: SDL/Event.cc:568-576 (ascending module order)
  for (uint16_t m1 = 0; m1 < nL; m1++) {
    if (g.nConnectedModules[m1] == 0 || ev.nSegments[m1] == 0) continue;
    unsigned nInnerSeg = ev.nSegments[m1];
    for (unsigned a = 0; a < nInnerSeg; a++) {
      unsigned inSeg = ev.segmentRanges[m1 * 2] + a;
      uint16_t m2 = ev.segOuterLowerModule[inSeg];
      unsigned nOuterSeg = ev.nSegments[m2];
      for (unsigned b = 0; b < nOuterSeg; b++) {
        unsigned outSeg = ev.segmentRanges[2 * m2] + b;
        uint16_t m3 = ev.segOuterLowerModule[outSeg];
        // runTripletConstraintsAndAlgo: SDL/Triplet.h:796-872
        if (ev.segMdIndices[2 * inSeg + 1] != ev.segMdIndices[2 * outSeg]) continue;
        unsigned md1 = ev.segMdIndices[2 * inSeg], md2 = ev.segMdIndices[2 * outSeg], md3 = ev.segMdIndices[2 * outSeg + 1];
        if (!passRZConstraint(ev, m1, m2, m3, md1, md2, md3)) continue;
        short s1 = g.subdets[m1], s2 = g.subdets[m2], s3 = g.subdets[m3];
        float betaIn = 0;
        bool ok;
        if (s1 == Barrel && s2 == Barrel && s3 == Barrel) ok = pointingBBB(ev, m1, m3, md1, md2, md3, inSeg, betaIn);
        else if (s1 == Barrel && s2 == Barrel && s3 == Endcap) ok = pointingBBE(ev, m1, m3, md1, md2, md3, inSeg, betaIn);
        else if (s1 == Barrel && s2 == Endcap && s3 == Endcap) ok = pointingBBE(ev, m1, m3, md1, md2, md3, inSeg, betaIn);
        else if (s1 == Endcap && s2 == Endcap && s3 == Endcap) ok = pointingEEE(ev, m1, m3, md1, md2, md3, inSeg, betaIn);
        else ok = false;
        if (!ok) continue;
        float cx, cy;
        float radius = computeRadiusFromThreeAnchorHits(ev.anchorX[md1], ev.anchorY[md1], ev.anchorX[md2], ev.anchorY[md2],
                                                        ev.anchorX[md3], ev.anchorY[md3], cx, cy);
        unsigned occ = ev.totOccupancyTriplets[m1]++;
        if ((int)occ >= ev.tripletModuleOccupancy[m1]) continue;
        unsigned t = ev.tripletModuleIndices[m1] + ev.nTriplets[m1]++;
        // addTripletToMemory: SDL/Triplet.h:153-239
        ev.t3SegmentIndices[2 * t] = inSeg; ev.t3SegmentIndices[2 * t + 1] = outSeg;
        ev.t3LowerModuleIndices[3 * t] = m1; ev.t3LowerModuleIndices[3 * t + 1] = m2; ev.t3LowerModuleIndices[3 * t + 2] = m3;
        ev.t3BetaIn[t] = betaIn; ev.t3CircleRadius[t] = radius; ev.t3CircleCenterX[t] = cx; ev.t3CircleCenterY[t] = cy;
        ev.t3LogicalLayers[3 * t] = g.layers[m1] + (g.subdets[m1] == 4) * 6;
        ev.t3LogicalLayers[3 * t + 1] = g.layers[m2] + (g.subdets[m2] == 4) * 6;
        ev.t3LogicalLayers[3 * t + 2] = g.layers[m3] + (g.subdets[m3] == 4) * 6;
        unsigned mdA = ev.segMdIndices[2 * inSeg], mdB = ev.segMdIndices[2 * inSeg + 1], mdC = ev.segMdIndices[2 * outSeg + 1];
        ev.t3HitIndices[6 * t] = ev.mdAnchorHitIndices[mdA]; ev.t3HitIndices[6 * t + 1] = ev.mdOuterHitIndices[mdA];
        ev.t3HitIndices[6 * t + 2] = ev.mdAnchorHitIndices[mdB]; ev.t3HitIndices[6 * t + 3] = ev.mdOuterHitIndices[mdB];
        ev.t3HitIndices[6 * t + 4] = ev.mdAnchorHitIndices[mdC]; ev.t3HitIndices[6 * t + 5] = ev.mdOuterHitIndices[mdC];
      }
    }
  }
  // addTripletRangesToEventExplicit: SDL/Triplet.h:1104-1123
  for (unsigned i = 0; i < nL; i++) {
    if (ev.nTriplets[i] == 0) {
      ev.tripletRanges[i * 2] = -1; ev.tripletRanges[i * 2 + 1] = -1;
    } else {
      ev.tripletRanges[i * 2] = ev.tripletModuleIndices[i];
      ev.tripletRanges[i * 2 + 1] = ev.tripletModuleIndices[i] + ev.nTriplets[i] - 1;
    }
  }
}

}  // namespace orc
