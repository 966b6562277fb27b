// ORACLE — test infrastructure only (see orc_common.h).
// C entry points used by oracle/orc.py (ctypes).  Not part of the product.
#include "orc_prepare.h"
#include "orc_t5.h"
#include <cstring>
#include <string>

using namespace orc;

namespace {
enum DType { F32 = 0, U32 = 1, I32 = 2, U16 = 3, I16 = 4, U8 = 5, I8 = 6, F64 = 7 };
template <class T> constexpr int dt();
template <> constexpr int dt<float>() { return F32; }
template <> constexpr int dt<unsigned>() { return U32; }
template <> constexpr int dt<int>() { return I32; }
template <> constexpr int dt<uint16_t>() { return U16; }
template <> constexpr int dt<short>() { return I16; }
template <> constexpr int dt<uint8_t>() { return U8; }
template <> constexpr int dt<int8_t>() { return I8; }
template <> constexpr int dt<char>() { return I8; }

struct Handle {
  Event ev;
  Input in;
};

template <class T>
bool out(const std::vector<T>& v, const void** p, int* t, long* n) {
  *p = v.data(); *t = dt<T>(); *n = (long)v.size();
  return true;
}
}  // namespace

#define F(name, vec) if (!std::strcmp(key, name)) return out(vec, p, t, n)

extern "C" {

void* orc_geometry_load(const char* dir, unsigned stride, char* err, int errlen) {
  auto* g = new Geometry();
  std::string e;
  if (!load_geometry(dir, stride, *g, e)) {
    if (err && errlen > 0) { std::strncpy(err, e.c_str(), errlen - 1); err[errlen - 1] = 0; }
    delete g;
    return nullptr;
  }
  return g;
}
void orc_geometry_free(void* g) { delete static_cast<Geometry*>(g); }

int orc_geometry_field(void* gp, const char* key, const void** p, int* t, long* n) {
  const Geometry& g = *static_cast<Geometry*>(gp);
  F("detIds", g.detIds); F("moduleMap", g.moduleMap); F("mapdetId", g.mapdetId); F("mapIdx", g.mapIdx);
  F("nConnectedModules", g.nConnectedModules); F("drdzs", g.drdzs); F("dxdys", g.dxdys);
  F("partnerModuleIndices", g.partnerModuleIndices); F("layers", g.layers); F("rings", g.rings);
  F("modules", g.modules); F("rods", g.rods); F("subdets", g.subdets); F("sides", g.sides); F("eta", g.eta);
  F("r", g.r); F("isInverted", g.isInverted); F("isLower", g.isLower); F("isAnchor", g.isAnchor);
  F("moduleType", g.moduleType); F("moduleLayerType", g.moduleLayerType); F("sdlLayers", g.sdlLayers);
  F("connectedPixels", g.connectedPixels); F("connectedPixelsIndex", g.cpIndex);
  F("connectedPixelsSizes", g.cpSizes); F("connectedPixelsIndexPos", g.cpIndexPos);
  F("connectedPixelsSizesPos", g.cpSizesPos); F("connectedPixelsIndexNeg", g.cpIndexNeg);
  F("connectedPixelsSizesNeg", g.cpSizesNeg); F("geoMapDetId", g.geoMapDetId); F("geoMapPhi", g.geoMapPhi);
  return 0;
}

unsigned orc_geometry_scalar(void* gp, const char* key) {
  const Geometry& g = *static_cast<Geometry*>(gp);
  if (!std::strcmp(key, "nModules")) return g.nModules;
  if (!std::strcmp(key, "nLowerModules")) return g.nLowerModules;
  if (!std::strcmp(key, "nPixels")) return g.nPixels;
  if (!std::strcmp(key, "pixelModuleIndex")) return g.pixelModuleIndex;
  if (!std::strcmp(key, "stride")) return g.stride;
  return 0;
}

void* orc_event_new(void* g, float ptCut, int t5Mode) {
  auto* h = new Handle();
  h->ev.g = static_cast<Geometry*>(g);
  h->ev.cfg.ptCut = ptCut;
  h->ev.cfg.t5Mode = t5Mode;
  return h;
}
void orc_event_free(void* h) { delete static_cast<Handle*>(h); }

// Arguments of SDL::LST::run (SDL/LST.h:29-48), flattened.
int orc_prepare(void* hp, unsigned nSeeds, const float* px, const float* py, const float* pz, const float* dxy,
                const float* dz, const float* ptErr, const float* etaErr, const float* gx, const float* gy,
                const float* gz, const float* gpx, const float* gpy, const float* gpz, const int* q,
                const int* hitIdx, const unsigned* hitIdxOff, unsigned nHits, const unsigned* detId,
                const float* x, const float* y, const float* z) {
  auto* h = static_cast<Handle*>(hp);
  SeedInput s;
  auto cp = [](auto& v, const auto* src, size_t n) { v.assign(src, src + n); };
  cp(s.see_px, px, nSeeds); cp(s.see_py, py, nSeeds); cp(s.see_pz, pz, nSeeds); cp(s.see_dxy, dxy, nSeeds);
  cp(s.see_dz, dz, nSeeds); cp(s.see_ptErr, ptErr, nSeeds); cp(s.see_etaErr, etaErr, nSeeds);
  cp(s.see_stateTrajGlbX, gx, nSeeds); cp(s.see_stateTrajGlbY, gy, nSeeds); cp(s.see_stateTrajGlbZ, gz, nSeeds);
  cp(s.see_stateTrajGlbPx, gpx, nSeeds); cp(s.see_stateTrajGlbPy, gpy, nSeeds); cp(s.see_stateTrajGlbPz, gpz, nSeeds);
  cp(s.see_q, q, nSeeds);
  cp(s.see_hitIdx_off, hitIdxOff, nSeeds + 1);
  cp(s.see_hitIdx, hitIdx, hitIdxOff[nSeeds]);
  cp(s.ph2_detId, detId, nHits); cp(s.ph2_x, x, nHits); cp(s.ph2_y, y, nHits); cp(s.ph2_z, z, nHits);
  prepare_input(s, 0.8, h->in);
  return 0;
}

// Stage sequence of SDL::LST::run (SDL/LST.cc:34-160); runs stages [0, upto).
// 0 hits, 1 pixel segments, 2 MD, 3 LS, 4 T3, 5 T5, 6 pLS cleaning, 7 pT5,
// 8 pT3, 9 track candidates and the output lists.
int orc_run(void* hp, int upto) {
  auto* h = static_cast<Handle*>(hp);
  Event& ev = h->ev;
  if (upto > 0 && !event_add_hits(ev, h->in)) return -1;
  if (upto > 1) event_add_pixel_segments(ev, h->in);
  if (upto > 2) event_create_mini_doublets(ev);
  if (upto > 3) event_create_segments(ev);
  if (upto > 4) event_create_triplets(ev);
  if (upto > 5) event_create_quintuplets(ev);
  if (upto > 6) event_pixel_segment_cleaning(ev);
  if (upto > 7) event_create_pixel_quintuplets(ev);
  if (upto > 8) event_create_pixel_triplets(ev);
  if (upto > 9) {
    event_create_track_candidates(ev);
    event_get_output(ev);
  }
  return 0;
}

unsigned orc_event_scalar(void* hp, const char* key) {
  const Event& ev = static_cast<Handle*>(hp)->ev;
  if (!std::strcmp(key, "nHits")) return ev.nHits;
  if (!std::strcmp(key, "nPLS")) return ev.nPLS;
  if (!std::strcmp(key, "nTotalMDs")) return ev.nTotalMDs;
  if (!std::strcmp(key, "nTotalSegs")) return ev.nTotalSegs;
  if (!std::strcmp(key, "nTotalTrips")) return ev.nTotalTrips;
  if (!std::strcmp(key, "mdCapacity")) return ev.mdCapacity;
  if (!std::strcmp(key, "segCapacity")) return ev.segCapacity;
  if (!std::strcmp(key, "t3Capacity")) return ev.t3Capacity;
  if (!std::strcmp(key, "t5Capacity")) return ev.t5Capacity;
  if (!std::strcmp(key, "nEligibleT5Modules")) return ev.nEligibleT5Modules;
  if (!std::strcmp(key, "nTotalQuints")) return ev.nTotalQuints;
  if (!std::strcmp(key, "nPixelQuintuplets")) return ev.nPixelQuintuplets;
  if (!std::strcmp(key, "totOccupancyPixelQuintuplets")) return ev.totOccupancyPixelQuintuplets;
  if (!std::strcmp(key, "nPixelTriplets")) return ev.nPixelTriplets;
  if (!std::strcmp(key, "totOccupancyPixelTriplets")) return ev.totOccupancyPixelTriplets;
  if (!std::strcmp(key, "nTrackCandidates")) return ev.nTrackCandidates;
  if (!std::strcmp(key, "nTrackCandidatespT3")) return ev.nTrackCandidatespT3;
  if (!std::strcmp(key, "nTrackCandidatespT5")) return ev.nTrackCandidatespT5;
  if (!std::strcmp(key, "nTrackCandidatespLS")) return ev.nTrackCandidatespLS;
  if (!std::strcmp(key, "nTrackCandidatesT5")) return ev.nTrackCandidatesT5;
  return 0;
}

int orc_event_field(void* hp, const char* key, const void** p, int* t, long* n) {
  auto* h = static_cast<Handle*>(hp);
  const Event& ev = h->ev;
  const Input& in = h->in;
  // prepared inputs (LST::prepareInput outputs)
  F("in_x", in.x); F("in_y", in.y); F("in_z", in.z); F("in_detId", in.detId); F("in_idx", in.idx);
  F("in_hitIdx0", in.hitIdx0); F("in_hitIdx1", in.hitIdx1); F("in_hitIdx2", in.hitIdx2); F("in_hitIdx3", in.hitIdx3);
  F("in_dPhiChange", in.dPhiChange); F("in_ptIn", in.ptIn); F("in_ptErr", in.ptErr); F("in_px", in.px);
  F("in_py", in.py); F("in_pz", in.pz); F("in_eta", in.eta); F("in_etaErr", in.etaErr); F("in_phi", in.phi);
  F("in_charge", in.charge); F("in_superbin", in.superbin); F("in_seedIdx", in.seedIdx);
  F("in_pixelType", in.pixelType); F("in_isQuad", in.isQuad);
  // hits
  F("hits.xs", ev.xs); F("hits.ys", ev.ys); F("hits.zs", ev.zs); F("hits.rts", ev.rts); F("hits.phis", ev.phis);
  F("hits.etas", ev.etas); F("hits.highEdgeXs", ev.highEdgeXs); F("hits.highEdgeYs", ev.highEdgeYs);
  F("hits.lowEdgeXs", ev.lowEdgeXs); F("hits.lowEdgeYs", ev.lowEdgeYs); F("hits.detid", ev.detid);
  F("hits.idxs", ev.idxs); F("hits.moduleIndices", ev.hitModuleIndices); F("hits.hitRanges", ev.hitRanges);
  F("hits.hitRangesLower", ev.hitRangesLower); F("hits.hitRangesUpper", ev.hitRangesUpper);
  F("hits.hitRangesnLower", ev.hitRangesnLower); F("hits.hitRangesnUpper", ev.hitRangesnUpper);
  // ranges
  F("ranges.mdRanges", ev.mdRanges); F("ranges.segmentRanges", ev.segmentRanges);
  F("ranges.tripletRanges", ev.tripletRanges); F("ranges.quintupletRanges", ev.quintupletRanges);
  F("ranges.miniDoubletModuleIndices", ev.miniDoubletModuleIndices);
  F("ranges.miniDoubletModuleOccupancy", ev.miniDoubletModuleOccupancy);
  F("ranges.segmentModuleIndices", ev.segmentModuleIndices);
  F("ranges.segmentModuleOccupancy", ev.segmentModuleOccupancy);
  F("ranges.tripletModuleIndices", ev.tripletModuleIndices);
  F("ranges.tripletModuleOccupancy", ev.tripletModuleOccupancy);
  F("ranges.quintupletModuleIndices", ev.quintupletModuleIndices);
  F("ranges.quintupletModuleOccupancy", ev.quintupletModuleOccupancy);
  F("ranges.indicesOfEligibleT5Modules", ev.indicesOfEligibleT5Modules);
  // mini-doublets
  F("mds.nMDs", ev.nMDs); F("mds.totOccupancyMDs", ev.totOccupancyMDs);
  F("mds.anchorHitIndices", ev.mdAnchorHitIndices); F("mds.outerHitIndices", ev.mdOuterHitIndices);
  F("mds.moduleIndices", ev.mdModuleIndices); F("mds.dphichanges", ev.mdDphichanges); F("mds.dzs", ev.mdDzs);
  F("mds.dphis", ev.mdDphis); F("mds.shiftedXs", ev.mdShiftedXs); F("mds.shiftedYs", ev.mdShiftedYs);
  F("mds.shiftedZs", ev.mdShiftedZs); F("mds.noShiftedDzs", ev.mdNoShiftedDzs);
  F("mds.noShiftedDphis", ev.mdNoShiftedDphis); F("mds.noShiftedDphiChanges", ev.mdNoShiftedDphiChanges);
  F("mds.anchorX", ev.anchorX); F("mds.anchorY", ev.anchorY); F("mds.anchorZ", ev.anchorZ);
  F("mds.anchorRt", ev.anchorRt); F("mds.anchorPhi", ev.anchorPhi); F("mds.anchorEta", ev.anchorEta);
  F("mds.anchorHighEdgeX", ev.anchorHighEdgeX); F("mds.anchorHighEdgeY", ev.anchorHighEdgeY);
  F("mds.anchorLowEdgeX", ev.anchorLowEdgeX); F("mds.anchorLowEdgeY", ev.anchorLowEdgeY);
  F("mds.anchorHighEdgePhi", ev.anchorHighEdgePhi); F("mds.anchorLowEdgePhi", ev.anchorLowEdgePhi);
  F("mds.outerX", ev.outerX); F("mds.outerY", ev.outerY); F("mds.outerZ", ev.outerZ); F("mds.outerRt", ev.outerRt);
  F("mds.outerPhi", ev.outerPhi); F("mds.outerEta", ev.outerEta); F("mds.outerHighEdgeX", ev.outerHighEdgeX);
  F("mds.outerHighEdgeY", ev.outerHighEdgeY); F("mds.outerLowEdgeX", ev.outerLowEdgeX);
  F("mds.outerLowEdgeY", ev.outerLowEdgeY);
  // segments
  F("segs.nSegments", ev.nSegments); F("segs.totOccupancySegments", ev.totOccupancySegments);
  F("segs.mdIndices", ev.segMdIndices); F("segs.innerLowerModuleIndices", ev.segInnerLowerModule);
  F("segs.outerLowerModuleIndices", ev.segOuterLowerModule);
  F("segs.innerMiniDoubletAnchorHitIndices", ev.segInnerAnchorHit);
  F("segs.outerMiniDoubletAnchorHitIndices", ev.segOuterAnchorHit); F("segs.dPhis", ev.segDPhis);
  F("segs.dPhiMins", ev.segDPhiMins); F("segs.dPhiMaxs", ev.segDPhiMaxs); F("segs.dPhiChanges", ev.segDPhiChanges);
  F("segs.dPhiChangeMins", ev.segDPhiChangeMins); F("segs.dPhiChangeMaxs", ev.segDPhiChangeMaxs);
  F("segs.seedIdx", ev.plsSeedIdx); F("segs.pLSHitsIdxs", ev.plsHitsIdxs); F("segs.charge", ev.plsCharge);
  F("segs.superbin", ev.plsSuperbin); F("segs.pixelType", ev.plsPixelType); F("segs.isQuad", ev.plsIsQuad);
  F("segs.isDup", ev.plsIsDup); F("segs.partOfPT5", ev.plsPartOfPT5); F("segs.ptIn", ev.plsPtIn);
  F("segs.ptErr", ev.plsPtErr); F("segs.px", ev.plsPx); F("segs.py", ev.plsPy); F("segs.pz", ev.plsPz);
  F("segs.etaErr", ev.plsEtaErr); F("segs.eta", ev.plsEta); F("segs.phi", ev.plsPhi); F("segs.score", ev.plsScore);
  F("segs.circleCenterX", ev.plsCircleCenterX); F("segs.circleCenterY", ev.plsCircleCenterY);
  F("segs.circleRadius", ev.plsCircleRadius);
  // triplets
  F("t3.nTriplets", ev.nTriplets); F("t3.totOccupancyTriplets", ev.totOccupancyTriplets);
  F("t3.segmentIndices", ev.t3SegmentIndices); F("t3.hitIndices", ev.t3HitIndices);
  F("t3.lowerModuleIndices", ev.t3LowerModuleIndices); F("t3.logicalLayers", ev.t3LogicalLayers);
  F("t3.betaIn", ev.t3BetaIn); F("t3.circleRadius", ev.t3CircleRadius); F("t3.circleCenterX", ev.t3CircleCenterX);
  F("t3.circleCenterY", ev.t3CircleCenterY); F("t3.partOfPT5", ev.t3PartOfPT5); F("t3.partOfT5", ev.t3PartOfT5);
  F("t3.partOfPT3", ev.t3PartOfPT3);
  F("dnn.features", ev.dnnFeatures); F("dnn.score", ev.dnnScore); F("dnn.t3Pair", ev.dnnT3Pair);
  F("dnn.hitIdxs", ev.dnnHitIdxs);
  // quintuplets
  F("t5.nQuintuplets", ev.nQuintuplets); F("t5.totOccupancyQuintuplets", ev.totOccupancyQuintuplets);
  F("t5.tripletIndices", ev.t5TripletIndices); F("t5.lowerModuleIndices", ev.t5LowerModuleIndices);
  F("t5.hitIndices", ev.t5HitIndices); F("t5.logicalLayers", ev.t5LogicalLayers); F("t5.layer", ev.t5Layer);
  F("t5.isDup", ev.t5IsDup); F("t5.TightCutFlag", ev.t5TightCutFlag); F("t5.partOfPT5", ev.t5PartOfPT5);
  F("t5.innerRadius", ev.t5InnerRadius); F("t5.bridgeRadius", ev.t5BridgeRadius);
  F("t5.outerRadius", ev.t5OuterRadius); F("t5.pt", ev.t5Pt); F("t5.eta", ev.t5Eta); F("t5.phi", ev.t5Phi);
  F("t5.score_rphisum", ev.t5ScoreRphiSum); F("t5.regressionRadius", ev.t5RegressionRadius);
  F("t5.regressionG", ev.t5RegressionG); F("t5.regressionF", ev.t5RegressionF);
  F("t5.rzChiSquared", ev.t5RzChiSquared); F("t5.chiSquared", ev.t5ChiSquared);
  F("t5.nonAnchorChiSquared", ev.t5NonAnchorChiSquared);
  F("ranges.indicesOfEligibleT5Modules", ev.indicesOfEligibleT5Modules);
  // pixel quintuplets
  F("pt5.pixelIndices", ev.pt5PixelIndices); F("pt5.T5Indices", ev.pt5T5Indices); F("pt5.isDup", ev.pt5IsDup);
  F("pt5.score", ev.pt5Score); F("pt5.eta", ev.pt5Eta); F("pt5.phi", ev.pt5Phi);
  F("pt5.logicalLayers", ev.pt5LogicalLayers); F("pt5.hitIndices", ev.pt5HitIndices);
  F("pt5.lowerModuleIndices", ev.pt5LowerModuleIndices); F("pt5.pixelRadius", ev.pt5PixelRadius);
  F("pt5.quintupletRadius", ev.pt5QuintupletRadius); F("pt5.centerX", ev.pt5CenterX); F("pt5.centerY", ev.pt5CenterY);
  F("pt5.rzChiSquared", ev.pt5RzChiSquared); F("pt5.rPhiChiSquared", ev.pt5RPhiChiSquared);
  F("pt5.rPhiChiSquaredInwards", ev.pt5RPhiChiSquaredInwards);
  // pixel triplets
  F("pt3.pixelSegmentIndices", ev.pt3PixelSegmentIndices); F("pt3.tripletIndices", ev.pt3TripletIndices);
  F("pt3.pixelRadius", ev.pt3PixelRadius); F("pt3.tripletRadius", ev.pt3TripletRadius); F("pt3.pt", ev.pt3Pt);
  F("pt3.eta", ev.pt3Eta); F("pt3.phi", ev.pt3Phi); F("pt3.eta_pix", ev.pt3EtaPix); F("pt3.phi_pix", ev.pt3PhiPix);
  F("pt3.score", ev.pt3Score); F("pt3.isDup", ev.pt3IsDup); F("pt3.logicalLayers", ev.pt3LogicalLayers);
  F("pt3.hitIndices", ev.pt3HitIndices); F("pt3.lowerModuleIndices", ev.pt3LowerModuleIndices);
  F("pt3.centerX", ev.pt3CenterX); F("pt3.centerY", ev.pt3CenterY); F("pt3.rzChiSquared", ev.pt3RzChiSquared);
  F("pt3.rPhiChiSquared", ev.pt3RPhiChiSquared); F("pt3.rPhiChiSquaredInwards", ev.pt3RPhiChiSquaredInwards);
  // track candidates
  F("tc.trackCandidateType", ev.tcType); F("tc.directObjectIndices", ev.tcDirectObjectIndices);
  F("tc.objectIndices", ev.tcObjectIndices); F("tc.hitIndices", ev.tcHitIndices);
  F("tc.pixelSeedIndex", ev.tcPixelSeedIndex); F("tc.logicalLayers", ev.tcLogicalLayers);
  F("tc.lowerModuleIndices", ev.tcLowerModuleIndices); F("tc.centerX", ev.tcCenterX); F("tc.centerY", ev.tcCenterY);
  F("tc.radius", ev.tcRadius);
  // LST output lists
  F("out.hitIdxs", ev.outHitIdxs); F("out.hitOffsets", ev.outHitOffsets); F("out.len", ev.outLen);
  F("out.seedIdx", ev.outSeedIdx); F("out.trackCandidateType", ev.outType);
  return 0;
}

}  // extern "C"

// Scalar helpers exposed for known-answer tests of the restatement.
extern "C" {
float orc_phi_mpi_pi(float x) { return orc::phi_mpi_pi(x); }
float orc_phi(float x, float y) { return orc::phi(x, y); }
float orc_delta_phi(float x1, float y1, float x2, float y2) { return orc::deltaPhi(x1, y1, x2, y2); }
float orc_eta(float x, float y, float z) { return orc::eta(x, y, z); }
int orc_binary_search(const unsigned* data, unsigned v, unsigned n) { return orc::binary_search(data, v, n); }
float orc_circle_radius(float x1, float y1, float x2, float y2, float x3, float y3, float* g, float* f) {
  return orc::computeRadiusFromThreeAnchorHits(x1, y1, x2, y2, x3, y3, *g, *f);
}
// weighted algebraic circle fit of the T5 (SDL/Quintuplet.h:1223-1320) and the
// chi-squared of points against a circle (:1322-1367), n <= 16
float orc_regression_circle(unsigned n, const float* xs, const float* ys, const float* d1, const float* d2,
                            const float* slopes, const unsigned char* flat, float* g, float* f, float* chi2) {
  bool fl[16];
  for (unsigned i = 0; i < n && i < 16; i++) fl[i] = flat[i] != 0;
  return orc::regressionCircle(n, xs, ys, d1, d2, slopes, fl, *g, *f, *chi2);
}
// T5 DNN forward pass on 38 inputs (SDL/NeuralNetwork.h:116-162)
float orc_t5dnn_forward(const float* in) { return orc::t5dnnForward(in); }
// replace the DNN parameters (packed as in the engine's lst_t5dnn_set_weights)
void orc_t5dnn_set_weights(const float* p) { orc::t5dnnSetWeights(p); }
// the 38 DNN inputs of stored quintuplet slot t5 (SDL/NeuralNetwork.h:36-114); -1 if out of range
int orc_t5dnn_features(void* hp, unsigned t5, float* out) {
  const orc::Event& ev = static_cast<Handle*>(hp)->ev;
  if (2 * (size_t)t5 + 1 >= ev.t5TripletIndices.size()) return -1;
  const unsigned in = ev.t5TripletIndices[2 * t5], ou = ev.t5TripletIndices[2 * t5 + 1];
  const unsigned s1 = ev.t3SegmentIndices[2 * in], s2 = ev.t3SegmentIndices[2 * in + 1];
  const unsigned s3 = ev.t3SegmentIndices[2 * ou], s4 = ev.t3SegmentIndices[2 * ou + 1];
  const unsigned md[5] = {ev.segMdIndices[2 * s1], ev.segMdIndices[2 * s2], ev.segMdIndices[2 * s2 + 1],
                          ev.segMdIndices[2 * s3 + 1], ev.segMdIndices[2 * s4 + 1]};
  uint16_t mods[5];
  for (int k = 0; k < 5; k++) mods[k] = ev.t5LowerModuleIndices[5 * t5 + k];
  orc::t5dnnFeatures(ev, md, mods, in, ev.t5InnerRadius[t5], ev.t5OuterRadius[t5], ev.t5BridgeRadius[t5], out);
  return 0;
}
float orc_circle_chi2(unsigned n, const float* xs, const float* ys, const float* d1, const float* d2,
                      const float* slopes, const unsigned char* flat, float g, float f, float r) {
  bool fl[16];
  for (unsigned i = 0; i < n && i < 16; i++) fl[i] = flat[i] != 0;
  return orc::chiSquaredOfCircle(n, xs, ys, d1, d2, slopes, fl, g, f, r);
}
}
