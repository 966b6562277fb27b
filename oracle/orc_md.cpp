// ORACLE — test infrastructure only (see orc_common.h).
// Hits (SDL/Hit.h:194-270, SDL/Event.cc:151-215), pixel segments
// (SDL/Segment.h:392-452,1031-1117, SDL/Event.cc:217-391) and mini-doublets
// (SDL/MiniDoublet.h:193-1104, SDL/Event.cc:393-467), serial order.
#include "orc_event.h"
#include <algorithm>
#include <climits>

namespace orc {

// SDL/Hit.h:217-270 + SDL/Hit.h:194-215
bool event_add_hits(Event& ev, const Input& in) {
  const Geometry& g = *ev.g;
  const unsigned n = in.x.size();
  ev.nHits = n;
  ev.xs = in.x; ev.ys = in.y; ev.zs = in.z; ev.detid = in.detId; ev.idxs = in.idx;
  ev.rts.assign(n, 0); ev.phis.assign(n, 0); ev.etas.assign(n, 0);
  ev.highEdgeXs.assign(n, 0); ev.highEdgeYs.assign(n, 0); ev.lowEdgeXs.assign(n, 0); ev.lowEdgeYs.assign(n, 0);
  ev.hitModuleIndices.assign(n, 0);
  ev.hitRanges.assign(2 * g.nModules, -1);
  ev.hitRangesLower.assign(g.nModules, -1); ev.hitRangesUpper.assign(g.nModules, -1);
  ev.hitRangesnLower.assign(g.nModules, -1); ev.hitRangesnUpper.assign(g.nModules, -1);
  for (unsigned i = 0; i < n; i++) {
    float x = ev.xs[i], y = ev.ys[i], z = ev.zs[i];
    int det = ev.detid[i];
    ev.rts[i] = fsqrt(x * x + y * y);
    ev.phis[i] = phi(x, y);
    ev.etas[i] = ((z > 0) - (z < 0)) * facosh(fsqrt(x * x + y * y + z * z) / ev.rts[i]);
    int found = binary_search(g.mapdetId.data(), det, g.nModules);
    if (found < 0) return false;  // the reference would index mapIdx[-1]
    uint16_t mod = g.mapIdx[found];
    ev.hitModuleIndices[i] = mod;
    if (g.subdets[mod] == Endcap && g.moduleType[mod] == TwoS) {
      int f2 = binary_search(g.geoMapDetId.data(), det, g.geoMapDetId.size());
      float p = g.geoMapPhi[f2];
      float c = fcos(p);
      ev.highEdgeXs[i] = x + 2.5f * c;
      ev.lowEdgeXs[i] = x - 2.5f * c;
      float s = fsin(p);
      ev.highEdgeYs[i] = y + 2.5f * s;
      ev.lowEdgeYs[i] = y - 2.5f * s;
    }
    if (ev.hitRanges[2 * mod] == -1) ev.hitRanges[2 * mod] = i;
    else ev.hitRanges[2 * mod] = std::min(ev.hitRanges[2 * mod], (int)i);
    ev.hitRanges[2 * mod + 1] = std::max(ev.hitRanges[2 * mod + 1], (int)i);
  }
  for (unsigned lo = 0; lo < g.nLowerModules; lo++) {
    uint16_t up = g.partnerModuleIndices[lo];
    if (ev.hitRanges[lo * 2] != -1 && ev.hitRanges[up * 2] != -1) {
      ev.hitRangesLower[lo] = ev.hitRanges[lo * 2];
      ev.hitRangesUpper[lo] = ev.hitRanges[up * 2];
      ev.hitRangesnLower[lo] = ev.hitRanges[lo * 2 + 1] - ev.hitRanges[lo * 2] + 1;
      ev.hitRangesnUpper[lo] = ev.hitRanges[up * 2 + 1] - ev.hitRanges[up * 2] + 1;
    }
  }
  return true;
}

namespace {

int occupancy_category(const Geometry& g, unsigned i, int& eta_number) {
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
  if (e < 0.75) eta_number = 0;
  else if (e > 0.75 && e < 1.5) eta_number = 1;
  else if (e > 1.5 && e < 2.25) eta_number = 2;
  else if (e > 2.25 && e < 3) eta_number = 3;
  else eta_number = -1;
  return cat;
}

// createMDArrayRangesGPU: SDL/MiniDoublet.h:994-1082
void md_array_ranges(Event& ev) {
  const Geometry& g = *ev.g;
  int total = 0;
  for (unsigned i = 0; i < g.nLowerModules; i++) {
    int en, occ;
    int cat = occupancy_category(g, i, en);
    if (cat == 0 && en == 0) occ = 49;
    else if (cat == 0 && en == 1) occ = 42;
    else if (cat == 0 && en == 2) occ = 37;
    else if (cat == 0 && en == 3) occ = 41;
    else if (cat == 1) occ = 100;
    else if (cat == 2 && en == 1) occ = 16;
    else if (cat == 2 && en == 2) occ = 19;
    else if (cat == 3 && en == 1) occ = 14;
    else if (cat == 3 && en == 2) occ = 20;
    else if (cat == 3 && en == 3) occ = 25;
    else occ = 0;
    ev.miniDoubletModuleIndices[i] = total;
    ev.miniDoubletModuleOccupancy[i] = occ;
    total += occ;
  }
  ev.miniDoubletModuleIndices[g.nLowerModules] = total;
  ev.nTotalMDs = total;
}

// createSegmentArrayRanges: SDL/Segment.h:912-1008
void segment_array_ranges(Event& ev) {
  const Geometry& g = *ev.g;
  int total = 0;
  for (unsigned i = 0; i < g.nLowerModules; i++) {
    if (g.nConnectedModules[i] == 0) {
      ev.segmentModuleIndices[i] = total;
      ev.segmentModuleOccupancy[i] = 0;
      continue;
    }
    int en, occ;
    int cat = occupancy_category(g, i, en);
    if (cat == 0 && en == 0) occ = 572;
    else if (cat == 0 && en == 1) occ = 300;
    else if (cat == 0 && en == 2) occ = 183;
    else if (cat == 0 && en == 3) occ = 62;
    else if (cat == 1 && en == 0) occ = 191;
    else if (cat == 1 && en == 1) occ = 128;
    else if (cat == 2 && en == 1) occ = 107;
    else if (cat == 2 && en == 2) occ = 102;
    else if (cat == 3 && en == 1) occ = 64;
    else if (cat == 3 && en == 2) occ = 79;
    else if (cat == 3 && en == 3) occ = 85;
    else occ = 0;
    ev.segmentModuleIndices[i] = total;
    ev.segmentModuleOccupancy[i] = occ;
    total += occ;
  }
  ev.segmentModuleIndices[g.nLowerModules] = total;
  ev.nTotalSegs = total;
}

// addMDToMemory: SDL/MiniDoublet.h:193-262
void add_md(Event& ev, unsigned lowerHit, unsigned upperHit, uint16_t mod, float dz, float dPhi, float dPhiChange,
            float sx, float sy, float sz, float nsDz, float nsDphi, float nsDphiChange, unsigned idx) {
  const Geometry& g = *ev.g;
  ev.mdModuleIndices[idx] = mod;
  unsigned a, o;
  if (g.moduleType[mod] == PS && g.moduleLayerType[mod] == Strip) { a = upperHit; o = lowerHit; }
  else { a = lowerHit; o = upperHit; }
  ev.mdAnchorHitIndices[idx] = a;
  ev.mdOuterHitIndices[idx] = o;
  ev.mdDphichanges[idx] = dPhiChange;
  ev.mdDphis[idx] = dPhi;
  ev.mdDzs[idx] = dz;
  ev.mdShiftedXs[idx] = sx; ev.mdShiftedYs[idx] = sy; ev.mdShiftedZs[idx] = sz;
  ev.mdNoShiftedDzs[idx] = nsDz; ev.mdNoShiftedDphis[idx] = nsDphi; ev.mdNoShiftedDphiChanges[idx] = nsDphiChange;
  ev.anchorX[idx] = ev.xs[a]; ev.anchorY[idx] = ev.ys[a]; ev.anchorZ[idx] = ev.zs[a];
  ev.anchorRt[idx] = ev.rts[a]; ev.anchorPhi[idx] = ev.phis[a]; ev.anchorEta[idx] = ev.etas[a];
  ev.anchorHighEdgeX[idx] = ev.highEdgeXs[a]; ev.anchorHighEdgeY[idx] = ev.highEdgeYs[a];
  ev.anchorLowEdgeX[idx] = ev.lowEdgeXs[a]; ev.anchorLowEdgeY[idx] = ev.lowEdgeYs[a];
  ev.anchorHighEdgePhi[idx] = fatan2(ev.anchorHighEdgeY[idx], ev.anchorHighEdgeX[idx]);
  ev.anchorLowEdgePhi[idx] = fatan2(ev.anchorLowEdgeY[idx], ev.anchorLowEdgeX[idx]);
  ev.outerX[idx] = ev.xs[o]; ev.outerY[idx] = ev.ys[o]; ev.outerZ[idx] = ev.zs[o];
  ev.outerRt[idx] = ev.rts[o]; ev.outerPhi[idx] = ev.phis[o]; ev.outerEta[idx] = ev.etas[o];
  ev.outerHighEdgeX[idx] = ev.highEdgeXs[o]; ev.outerHighEdgeY[idx] = ev.highEdgeYs[o];
  ev.outerLowEdgeX[idx] = ev.lowEdgeXs[o]; ev.outerLowEdgeY[idx] = ev.lowEdgeYs[o];
}

void alloc_mds(Event& ev) {
  const Geometry& g = *ev.g;
  unsigned n = ev.nTotalMDs + kNMaxPixelMDs;
  ev.mdCapacity = n;
  ev.nMDs.assign(g.nLowerModules + 1, 0); ev.totOccupancyMDs.assign(g.nLowerModules + 1, 0);
  for (auto* v : {&ev.mdAnchorHitIndices, &ev.mdOuterHitIndices}) v->assign(n, 0);
  ev.mdModuleIndices.assign(n, 0);
  for (auto* v : {&ev.mdDphichanges, &ev.mdDzs, &ev.mdDphis, &ev.mdShiftedXs, &ev.mdShiftedYs, &ev.mdShiftedZs,
                  &ev.mdNoShiftedDzs, &ev.mdNoShiftedDphis, &ev.mdNoShiftedDphiChanges, &ev.anchorX, &ev.anchorY,
                  &ev.anchorZ, &ev.anchorRt, &ev.anchorPhi, &ev.anchorEta, &ev.anchorHighEdgeX, &ev.anchorHighEdgeY,
                  &ev.anchorLowEdgeX, &ev.anchorLowEdgeY, &ev.anchorHighEdgePhi, &ev.anchorLowEdgePhi, &ev.outerX,
                  &ev.outerY, &ev.outerZ, &ev.outerRt, &ev.outerPhi, &ev.outerEta, &ev.outerHighEdgeX,
                  &ev.outerHighEdgeY, &ev.outerLowEdgeX, &ev.outerLowEdgeY})
    v->assign(n, 0.f);
}

void alloc_segments(Event& ev) {
  const Geometry& g = *ev.g;
  unsigned n = ev.nTotalSegs + kNMaxPixelSegments;
  ev.segCapacity = n;
  ev.nSegments.assign(g.nLowerModules + 1, 0); ev.totOccupancySegments.assign(g.nLowerModules + 1, 0);
  ev.segMdIndices.assign(2 * n, 0); ev.segInnerAnchorHit.assign(n, 0); ev.segOuterAnchorHit.assign(n, 0);
  ev.segInnerLowerModule.assign(n, 0); ev.segOuterLowerModule.assign(n, 0);
  for (auto* v : {&ev.segDPhis, &ev.segDPhiMins, &ev.segDPhiMaxs, &ev.segDPhiChanges, &ev.segDPhiChangeMins,
                  &ev.segDPhiChangeMaxs})
    v->assign(n, 0.f);
  const unsigned p = kNMaxPixelSegments;
  ev.plsSeedIdx.assign(p, 0); ev.plsHitsIdxs.assign(4 * p, 0); ev.plsCharge.assign(p, 0); ev.plsSuperbin.assign(p, 0);
  ev.plsPixelType.assign(p, 0); ev.plsIsQuad.assign(p, 0); ev.plsIsDup.assign(p, 0); ev.plsPartOfPT5.assign(p, 0);
  for (auto* v : {&ev.plsPtIn, &ev.plsPtErr, &ev.plsPx, &ev.plsPy, &ev.plsPz, &ev.plsEtaErr, &ev.plsEta, &ev.plsPhi,
                  &ev.plsScore, &ev.plsCircleCenterX, &ev.plsCircleCenterY, &ev.plsCircleRadius})
    v->assign(p, 0.f);
}

}  // namespace

// SDL/Event.cc:217-391 with SDL/Segment.h:1031-1117 and :392-452
void event_add_pixel_segments(Event& ev, const Input& in) {
  const Geometry& g = *ev.g;
  const unsigned nL = g.nLowerModules;
  ev.mdRanges.assign(2 * g.nModules, -1); ev.segmentRanges.assign(2 * g.nModules, -1);
  ev.tripletRanges.assign(2 * g.nModules, -1); ev.quintupletRanges.assign(2 * g.nModules, -1);
  ev.miniDoubletModuleIndices.assign(nL + 1, 0); ev.miniDoubletModuleOccupancy.assign(nL + 1, 0);
  ev.segmentModuleIndices.assign(nL + 1, 0); ev.segmentModuleOccupancy.assign(nL + 1, 0);
  ev.tripletModuleIndices.assign(nL, 0); ev.tripletModuleOccupancy.assign(nL, 0);
  ev.quintupletModuleIndices.assign(nL, -1); ev.quintupletModuleOccupancy.assign(nL, 0);
  ev.indicesOfEligibleT5Modules.assign(nL, 0);

  unsigned size = in.ptIn.size();
  if (size > kNMaxPixelSegments) size = kNMaxPixelSegments;  // SDL/Event.cc:237-244
  const uint16_t pix = g.pixelModuleIndex;
  ev.miniDoubletModuleOccupancy[nL] = kNMaxPixelMDs;
  md_array_ranges(ev);
  alloc_mds(ev);
  segment_array_ranges(ev);
  alloc_segments(ev);
  ev.nPLS = size;
  for (unsigned t = 0; t < size; t++) {
    ev.plsPtIn[t] = in.ptIn[t]; ev.plsPtErr[t] = in.ptErr[t]; ev.plsPx[t] = in.px[t]; ev.plsPy[t] = in.py[t];
    ev.plsPz[t] = in.pz[t]; ev.plsEtaErr[t] = in.etaErr[t]; ev.plsIsQuad[t] = in.isQuad[t]; ev.plsEta[t] = in.eta[t];
    ev.plsPhi[t] = in.phi[t]; ev.plsCharge[t] = in.charge[t]; ev.plsSeedIdx[t] = in.seedIdx[t];
    ev.plsSuperbin[t] = in.superbin[t]; ev.plsPixelType[t] = in.pixelType[t];
  }
  ev.nSegments[pix] = size; ev.totOccupancySegments[pix] = size;
  ev.nMDs[pix] = 2 * size; ev.totOccupancyMDs[pix] = 2 * size;

  for (unsigned tid = 0; tid < size; tid++) {
    unsigned innerMD = ev.miniDoubletModuleIndices[pix] + 2 * tid;
    unsigned outerMD = ev.miniDoubletModuleIndices[pix] + 2 * tid + 1;
    unsigned pls = ev.segmentModuleIndices[pix] + tid;
    add_md(ev, in.hitIdx0[tid], in.hitIdx1[tid], pix, 0, 0, 0, 0, 0, 0, 0, 0, 0, innerMD);
    add_md(ev, in.hitIdx2[tid], in.hitIdx3[tid], pix, 0, 0, 0, 0, 0, 0, 0, 0, 0, outerMD);
    float slope = fsinh(ev.ys[ev.mdOuterHitIndices[innerMD]]);
    float intercept = ev.zs[ev.mdAnchorHitIndices[innerMD]] - slope * ev.rts[ev.mdAnchorHitIndices[innerMD]];
    float score = (ev.rts[ev.mdAnchorHitIndices[outerMD]] * slope + intercept) - ev.zs[ev.mdAnchorHitIndices[outerMD]];
    score = score * score;
    unsigned hits1[4] = {ev.idxs[ev.mdAnchorHitIndices[innerMD]], ev.idxs[ev.mdAnchorHitIndices[outerMD]],
                         ev.idxs[ev.mdOuterHitIndices[innerMD]], ev.idxs[ev.mdOuterHitIndices[outerMD]]};
    // addPixelSegmentToMemory: SDL/Segment.h:392-452
    ev.segMdIndices[pls * 2] = innerMD; ev.segMdIndices[pls * 2 + 1] = outerMD;
    ev.segInnerLowerModule[pls] = pix; ev.segOuterLowerModule[pls] = pix;
    ev.segInnerAnchorHit[pls] = in.hitIdx0[tid]; ev.segOuterAnchorHit[pls] = in.hitIdx2[tid];
    ev.segDPhiChanges[pls] = in.dPhiChange[tid];
    ev.plsIsDup[tid] = 0;
    ev.plsScore[tid] = score;
    for (int k = 0; k < 4; k++) ev.plsHitsIdxs[4 * tid + k] = hits1[k];
    float circleRadius = ev.outerX[innerMD] / (2 * k2Rinv1GeVf);
    float circlePhi = ev.outerZ[innerMD];
    float cx[2] = {ev.anchorX[innerMD] + circleRadius * fsin(circlePhi), ev.anchorX[innerMD] - circleRadius * fsin(circlePhi)};
    float cy[2] = {ev.anchorY[innerMD] - circleRadius * fcos(circlePhi), ev.anchorY[innerMD] + circleRadius * fcos(circlePhi)};
    float best = SDL_INF;
    size_t bi = 0;
    for (size_t i = 0; i < 2; i++) {
      float dx = ev.anchorX[outerMD] - cx[i], dy = ev.anchorY[outerMD] - cy[i];
      float chi = std::fabs(fsqrt(dx * dx + dy * dy) - circleRadius);
      if (chi < best) { best = chi; bi = i; }
    }
    ev.plsCircleCenterX[tid] = cx[bi]; ev.plsCircleCenterY[tid] = cy[bi]; ev.plsCircleRadius[tid] = circleRadius;
  }
}

namespace {

// SDL/MiniDoublet.h:264-284
bool isTighterTiltedModules(const Geometry& g, uint16_t m) {
  short subdet = g.subdets[m], layer = g.layers[m], side = g.sides[m], rod = g.rods[m];
  return (subdet == Barrel && side != Center && layer == 3) ||
         (subdet == Barrel && side == NegZ && layer == 2 && rod > 5) ||
         (subdet == Barrel && side == PosZ && layer == 2 && rod < 8) ||
         (subdet == Barrel && side == NegZ && layer == 1 && rod > 9) ||
         (subdet == Barrel && side == PosZ && layer == 1 && rod < 4);
}

// SDL/MiniDoublet.h:286-337 (table equals SDL/Segment.h:233-240)
float moduleGapSize(const Geometry& g, uint16_t m) {
  static const float tilted[3] = {0.26f, 0.26f, 0.26f};
  static const float flat[6] = {0.26f, 0.16f, 0.16f, 0.18f, 0.18f, 0.18f};
  static const float looseTilted[3] = {0.4f, 0.4f, 0.4f};
  unsigned iL = g.layers[m] - 1, iR = g.rings[m] - 1;
  short subdet = g.subdets[m], side = g.sides[m];
  if (subdet == Barrel && side == Center) return flat[iL];
  if (isTighterTiltedModules(g, m)) return tilted[iL];
  if (subdet == Endcap) {
    unsigned lim = (iL <= 1) ? 10 : (iL <= 3 ? 8 : 9);
    return iR < lim ? 0.4f : 0.18f;
  }
  return looseTilted[iL];
}

// SDL/MiniDoublet.h:339-401
float dPhiThreshold(const Geometry& g, float ptCut, float rt, uint16_t m, float dPhi = 0, float dz = 0) {
  unsigned iL = g.layers[m] - 1;
  const float miniSlope = fasin(fminf_(rt * k2Rinv1GeVf / ptCut, sinAlphaMax));
  const float rLayNominal = (g.subdets[m] == Barrel) ? miniRminMeanBarrel[iL] : miniRminMeanEndcap[iL];
  const float miniPVoff = 0.1f / rLayNominal;
  const float miniMuls = (g.subdets[m] == Barrel) ? miniMulsPtScaleBarrel[iL] * 3.f / ptCut
                                                  : miniMulsPtScaleEndcap[iL] * 3.f / ptCut;
  const bool isTilted = g.subdets[m] == Barrel && g.sides[m] != Center;
  float drdz;
  if (isTilted) {
    if (g.moduleType[m] == PS && g.moduleLayerType[m] == Strip) drdz = g.drdzs[m];
    else drdz = g.drdzs[g.partnerModuleIndices[m]];
  } else {
    drdz = 0;
  }
  const float miniTilt = isTilted ? 0.5f * pixelPSZpitch * drdz / fsqrt(1.f + drdz * drdz) / moduleGapSize(g, m) : 0;
  const float miniLum = std::fabs(dPhi * deltaZLum / dz);
  if (g.subdets[m] == Barrel && g.sides[m] == Center)
    return miniSlope + fsqrt(miniMuls * miniMuls + miniPVoff * miniPVoff);
  else if (g.subdets[m] == Barrel && g.sides[m] != Center)
    return miniSlope + fsqrt(miniMuls * miniMuls + miniPVoff * miniPVoff + miniTilt * miniTilt * miniSlope * miniSlope);
  return miniSlope + fsqrt(miniMuls * miniMuls + miniPVoff * miniPVoff + miniLum * miniLum);
}

// SDL/MiniDoublet.h:403-567
void shiftStripHits(const Geometry& g, uint16_t lowerM, float* sc, float xLower, float yLower, float zLower,
                    float rtLower, float xUpper, float yUpper, float zUpper, float rtUpper) {
  float xp, yp, zp, rtp, xo, yo;
  const float slope = g.dxdys[lowerM];
  const float drdz_ = g.drdzs[lowerM];
  if (g.moduleType[lowerM] == PS) {
    if (g.subdets[lowerM] == Barrel ? g.moduleLayerType[lowerM] != Pixel : g.moduleLayerType[lowerM] == Pixel) {
      xo = xUpper; yo = yUpper; xp = xLower; yp = yLower; zp = zLower; rtp = rtLower;
    } else {
      xo = xLower; yo = yLower; xp = xUpper; yp = yUpper; zp = zUpper; rtp = rtUpper;
    }
  } else {
    xo = xUpper; yo = yUpper; xp = xLower; yp = yLower; zp = zLower; rtp = rtLower;
  }
  (void)rtp;
  bool isEndcap = g.subdets[lowerM] == Endcap;
  float angleA = std::fabs(fatan(rtp / zp));
  float angleB = isEndcap ? float(M_PI) / 2.f : fatan(drdz_);
  float moduleSeparation = moduleGapSize(g, lowerM);
  if (g.moduleType[lowerM] == PS && g.moduleLayerType[lowerM] != Pixel) moduleSeparation *= -1;
  float drprime = (moduleSeparation / fsin(angleA + angleB)) * fsin(angleA);
  float absArctanSlope = (slope != SDL_INF) ? std::fabs(fatan(slope)) : float(M_PI) / 2.f;
  float angleM;
  if (xp > 0 && yp > 0) angleM = absArctanSlope;
  else if (xp > 0 && yp < 0) angleM = float(M_PI) - absArctanSlope;
  else if (xp < 0 && yp < 0) angleM = float(M_PI) + absArctanSlope;
  else angleM = 2.f * float(M_PI) - absArctanSlope;
  float drprime_x = drprime * fsin(angleM);
  float drprime_y = drprime * fcos(angleM);
  float xa = xp + drprime_x, ya = yp + drprime_y;
  float xn, yn;
  if (slope == SDL_INF) { xn = xa; yn = yo; }
  else if (slope == 0) { xn = xo; yn = ya; }
  else {
    xn = (slope * xa + (1.f / slope) * xo - ya + yo) / (slope + (1.f / slope));
    yn = (xn - xa) * slope + ya;
  }
  float absdzprime = std::fabs(moduleSeparation / fsin(angleA + angleB) * fcos(angleA));
  float abszn;
  if (g.moduleLayerType[lowerM] == Pixel) abszn = std::fabs(zp) + absdzprime;
  else abszn = std::fabs(zp) - absdzprime;
  float zn = abszn * ((zp > 0) ? 1 : -1);
  sc[0] = xn; sc[1] = yn; sc[2] = zn;
}

// SDL/MiniDoublet.h:645-772
bool mdBarrel(const Geometry& g, float ptCut, uint16_t lowerM, float& dz, float& dPhi, float& dPhiChange,
              float& shiftedX, float& shiftedY, float& shiftedZ, float& noShiftedDz, float& noShiftedDphi,
              float& noShiftedDphiChange, float xLower, float yLower, float zLower, float rtLower, float xUpper,
              float yUpper, float zUpper, float rtUpper) {
  dz = zLower - zUpper;
  const float dzCut = g.moduleType[lowerM] == PS ? 2.f : 10.f;
  const float sign = ((dz > 0) - (dz < 0)) * ((zLower > 0) - (zLower < 0));
  const float invertedcrossercut = (std::fabs(dz) > 2) * sign;
  if (!((std::fabs(dz) < dzCut) && (invertedcrossercut <= 0))) return false;
  float miniCut = g.moduleLayerType[lowerM] == Pixel ? dPhiThreshold(g, ptCut, rtLower, lowerM)
                                                      : dPhiThreshold(g, ptCut, rtUpper, lowerM);
  float xn = 0.f, yn = 0.f;
  float shiftedRt = 0;
  if (g.sides[lowerM] != Center) {
    float sc[3];
    shiftStripHits(g, lowerM, sc, xLower, yLower, zLower, rtLower, xUpper, yUpper, zUpper, rtUpper);
    xn = sc[0]; yn = sc[1];
    if (g.moduleLayerType[lowerM] == Pixel) {
      shiftedX = xn; shiftedY = yn; shiftedZ = zUpper;
      shiftedRt = fsqrt(xn * xn + yn * yn);
      dPhi = deltaPhi(xLower, yLower, shiftedX, shiftedY);
      noShiftedDphi = deltaPhi(xLower, yLower, xUpper, yUpper);
    } else {
      shiftedX = xn; shiftedY = yn; shiftedZ = zLower;
      shiftedRt = fsqrt(xn * xn + yn * yn);
      dPhi = deltaPhi(shiftedX, shiftedY, xUpper, yUpper);
      noShiftedDphi = deltaPhi(xLower, yLower, xUpper, yUpper);
    }
  } else {
    shiftedX = 0; shiftedY = 0; shiftedZ = 0;
    dPhi = deltaPhi(xLower, yLower, xUpper, yUpper);
    noShiftedDphi = dPhi;
  }
  if (!(std::fabs(dPhi) < miniCut)) return false;
  if (g.sides[lowerM] != Center) {
    if (g.moduleLayerType[lowerM] != Pixel) {
      dPhiChange = (rtLower < shiftedRt) ? deltaPhiChange(xLower, yLower, shiftedX, shiftedY)
                                         : deltaPhiChange(shiftedX, shiftedY, xLower, yLower);
    } else {
      dPhiChange = (shiftedRt < rtUpper) ? deltaPhiChange(shiftedX, shiftedY, xUpper, yUpper)
                                         : deltaPhiChange(xUpper, yUpper, shiftedX, shiftedY);
    }
    noShiftedDphiChange = rtLower < rtUpper ? deltaPhiChange(xLower, yLower, xUpper, yUpper)
                                            : deltaPhiChange(xUpper, yUpper, xLower, yLower);
  } else {
    dPhiChange = deltaPhiChange(xLower, yLower, xUpper, yUpper);
    noShiftedDphiChange = dPhiChange;
  }
  noShiftedDz = 0;
  return std::fabs(dPhiChange) < miniCut;
}

// SDL/MiniDoublet.h:774-894
bool mdEndcap(const Geometry& g, float ptCut, uint16_t lowerM, float& drt, float& dPhi, float& dPhiChange,
              float& shiftedX, float& shiftedY, float& shiftedZ, float& noShiftedDz, float& noShiftedDphi,
              float& noShiftedDphiChange, float xLower, float yLower, float zLower, float rtLower, float xUpper,
              float yUpper, float zUpper, float rtUpper) {
  float dz = zLower - zUpper;
  if (!(std::fabs(dz) < 1.f)) return false;
  const float drtCut = g.moduleType[lowerM] == PS ? 2.f : 10.f;
  drt = rtLower - rtUpper;
  if (!(std::fabs(drt) < drtCut)) return false;
  float sc[3];
  shiftStripHits(g, lowerM, sc, xLower, yLower, zLower, rtLower, xUpper, yUpper, zUpper, rtUpper);
  float xn = sc[0], yn = sc[1], zn = sc[2];
  if (g.moduleType[lowerM] == PS) {
    if (g.moduleLayerType[lowerM] == Pixel) {
      shiftedX = xn; shiftedY = yn; shiftedZ = zUpper;
      dPhi = deltaPhi(xLower, yLower, shiftedX, shiftedY);
      noShiftedDphi = deltaPhi(xLower, yLower, xUpper, yUpper);
    } else {
      shiftedX = xn; shiftedY = yn; shiftedZ = zLower;
      dPhi = deltaPhi(shiftedX, shiftedY, xUpper, yUpper);
      noShiftedDphi = deltaPhi(xLower, yLower, xUpper, yUpper);
    }
  } else {
    shiftedX = xn; shiftedY = yn; shiftedZ = zUpper;
    dPhi = deltaPhi(xLower, yLower, xn, yn);
    noShiftedDphi = deltaPhi(xLower, yLower, xUpper, yUpper);
  }
  if (g.moduleType[lowerM] == PS) dz = g.moduleLayerType[lowerM] == Pixel ? zLower - zn : zUpper - zn;
  float miniCut = g.moduleLayerType[lowerM] == Pixel ? dPhiThreshold(g, ptCut, rtLower, lowerM, dPhi, dz)
                                                      : dPhiThreshold(g, ptCut, rtUpper, lowerM, dPhi, dz);
  if (!(std::fabs(dPhi) < miniCut)) return false;
  float dzFrac = std::fabs(dz) / std::fabs(zLower);
  dPhiChange = dPhi / dzFrac * (1.f + dzFrac);
  noShiftedDphiChange = noShiftedDphi / dzFrac * (1.f + dzFrac);
  noShiftedDz = 0;
  return std::fabs(dPhiChange) < miniCut;
}

}  // namespace

// SDL/Event.cc:393-467, SDL/MiniDoublet.h:896-992 and :1084-1104
void event_create_mini_doublets(Event& ev) {
  const Geometry& g = *ev.g;
  ev.miniDoubletModuleOccupancy[g.nLowerModules] = kNMaxPixelMDs;
  md_array_ranges(ev);
  for (uint16_t lowerM = 0; lowerM < g.nLowerModules; lowerM++) {
    int nLo = ev.hitRangesnLower[lowerM], nUp = ev.hitRangesnUpper[lowerM];
    if (ev.hitRangesLower[lowerM] == -1) continue;
    unsigned upBase = ev.hitRangesUpper[lowerM], loBase = ev.hitRangesLower[lowerM];
    int limit = nUp * nLo;
    for (int h = 0; h < limit; h++) {
      int li = h / nUp, ui = h % nUp;
      if (ui >= nUp) continue;  // SDL/MiniDoublet.h:919-923: counts wrapped negative skip every pair
      if (li >= nLo) continue;
      unsigned lh = loBase + li, uh = upBase + ui;
      float xL = ev.xs[lh], yL = ev.ys[lh], zL = ev.zs[lh], rL = ev.rts[lh];
      float xU = ev.xs[uh], yU = ev.ys[uh], zU = ev.zs[uh], rU = ev.rts[uh];
      float dz = 0, dphi = 0, dphichange = 0, sx = 0, sy = 0, sz = 0, nsdz = 0, nsdphi = 0, nsdphichange = 0;
      bool ok = (g.subdets[lowerM] == Barrel)
                    ? mdBarrel(g, ev.cfg.ptCut, lowerM, dz, dphi, dphichange, sx, sy, sz, nsdz, nsdphi, nsdphichange,
                               xL, yL, zL, rL, xU, yU, zU, rU)
                    : mdEndcap(g, ev.cfg.ptCut, lowerM, dz, dphi, dphichange, sx, sy, sz, nsdz, nsdphi, nsdphichange,
                               xL, yL, zL, rL, xU, yU, zU, rU);
      if (!ok) continue;
      unsigned occ = ev.totOccupancyMDs[lowerM]++;
      if ((int)occ >= ev.miniDoubletModuleOccupancy[lowerM]) continue;
      unsigned k = ev.nMDs[lowerM]++;
      add_md(ev, lh, uh, lowerM, dz, dphi, dphichange, sx, sy, sz, nsdz, nsdphi, nsdphichange,
             ev.miniDoubletModuleIndices[lowerM] + k);
    }
  }
  for (unsigned i = 0; i < g.nLowerModules; i++) {
    if (ev.nMDs[i] == 0 || ev.hitRanges[i * 2] == -1) {
      ev.mdRanges[i * 2] = -1; ev.mdRanges[i * 2 + 1] = -1;
    } else {
      ev.mdRanges[i * 2] = ev.miniDoubletModuleIndices[i];
      ev.mdRanges[i * 2 + 1] = ev.miniDoubletModuleIndices[i] + ev.nMDs[i] - 1;
    }
  }
}

}  // namespace orc
