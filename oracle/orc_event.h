// ORACLE — test infrastructure only (see orc_common.h).
// Per-event SoA state of the serial restatement.  Field names and layouts
// follow the reference structs (SDL/Hit.h:13-56, SDL/Module.h:21-85,
// SDL/MiniDoublet.h:16-190, SDL/Segment.h:17-175, SDL/Triplet.h:17-151,
// SDL/Quintuplet.h:20-141, SDL/PixelTriplet.h:20-132,1871-1960,
// SDL/TrackCandidate.h:20-109) so tests can compare them by name.
#pragma once
#include "orc_common.h"

namespace orc {

struct Config {
  float ptCut = 0.8f;    // SDL/Constants.h:102-105 (PT_CUT)
  int t5Mode = 0;        // 0: USE_RZCHI2+USE_RPHICHI2 ("plain"), 1: USE_RZCHI2+USE_T5_DNN
  bool tcPlsTriplets = false;   // TC_PLS_TRIPLETS
  bool noPlsDupClean = false;   // NOPLSDUPCLEAN
};

struct Input {  // what SDL::Event::addHitToEvent/addPixelSegmentToEvent receive
  std::vector<float> x, y, z;
  std::vector<unsigned> detId, idx;
  std::vector<unsigned> hitIdx0, hitIdx1, hitIdx2, hitIdx3;
  std::vector<float> dPhiChange, ptIn, ptErr, px, py, pz, eta, etaErr, phi;
  std::vector<int> charge, superbin;
  std::vector<unsigned> seedIdx;
  std::vector<int8_t> pixelType;
  std::vector<char> isQuad;
};

struct Event {
  const Geometry* g = nullptr;
  Config cfg;
  // hits
  unsigned nHits = 0;
  std::vector<float> xs, ys, zs, rts, phis, etas, highEdgeXs, highEdgeYs, lowEdgeXs, lowEdgeYs;
  std::vector<unsigned> detid, idxs;
  std::vector<uint16_t> hitModuleIndices;
  std::vector<int> hitRanges, hitRangesLower, hitRangesUpper;
  std::vector<int8_t> hitRangesnLower, hitRangesnUpper;
  // object ranges
  std::vector<int> mdRanges, segmentRanges, tripletRanges, quintupletRanges;
  std::vector<int> miniDoubletModuleIndices, miniDoubletModuleOccupancy;
  std::vector<int> segmentModuleIndices, segmentModuleOccupancy;
  std::vector<int> tripletModuleIndices, tripletModuleOccupancy;
  std::vector<int> quintupletModuleIndices, quintupletModuleOccupancy;
  uint16_t nEligibleT5Modules = 0;
  std::vector<uint16_t> indicesOfEligibleT5Modules;
  unsigned nTotalMDs = 0, nTotalSegs = 0, nTotalTrips = 0, nTotalQuints = 0;
  // mini-doublets
  unsigned mdCapacity = 0;
  std::vector<unsigned> nMDs, totOccupancyMDs, mdAnchorHitIndices, mdOuterHitIndices;
  std::vector<uint16_t> mdModuleIndices;
  std::vector<float> mdDphichanges, mdDzs, mdDphis, mdShiftedXs, mdShiftedYs, mdShiftedZs,
      mdNoShiftedDzs, mdNoShiftedDphis, mdNoShiftedDphiChanges;
  std::vector<float> anchorX, anchorY, anchorZ, anchorRt, anchorPhi, anchorEta, anchorHighEdgeX,
      anchorHighEdgeY, anchorLowEdgeX, anchorLowEdgeY, anchorHighEdgePhi, anchorLowEdgePhi;
  std::vector<float> outerX, outerY, outerZ, outerRt, outerPhi, outerEta, outerHighEdgeX,
      outerHighEdgeY, outerLowEdgeX, outerLowEdgeY;
  // segments (OT + pLS)
  unsigned segCapacity = 0, nPLS = 0;
  std::vector<unsigned> nSegments, totOccupancySegments, segMdIndices, segInnerAnchorHit, segOuterAnchorHit;
  std::vector<uint16_t> segInnerLowerModule, segOuterLowerModule;
  std::vector<float> segDPhis, segDPhiMins, segDPhiMaxs, segDPhiChanges, segDPhiChangeMins, segDPhiChangeMaxs;
  // pLS-only arrays, indexed by pLS number
  std::vector<unsigned> plsSeedIdx, plsHitsIdxs;  // plsHitsIdxs: 4 per pLS (x,y,z,w)
  std::vector<int> plsCharge, plsSuperbin;
  std::vector<int8_t> plsPixelType;
  std::vector<char> plsIsQuad, plsIsDup;
  std::vector<uint8_t> plsPartOfPT5;
  std::vector<float> plsPtIn, plsPtErr, plsPx, plsPy, plsPz, plsEtaErr, plsEta, plsPhi, plsScore,
      plsCircleCenterX, plsCircleCenterY, plsCircleRadius;
  // triplets
  unsigned t3Capacity = 0;
  std::vector<unsigned> nTriplets, totOccupancyTriplets, t3SegmentIndices, t3HitIndices;
  std::vector<uint16_t> t3LowerModuleIndices;
  std::vector<uint8_t> t3LogicalLayers, t3PartOfPT5, t3PartOfT5, t3PartOfPT3;
  std::vector<float> t3BetaIn, t3CircleRadius, t3CircleCenterX, t3CircleCenterY;
  // every T5 candidate at the DNN call site in DNN mode (SDL/Quintuplet.h:2921-2936), in serial
  // order: 38 inputs, score, (inner, outer) T3 pair, 10 hit indices in input numbering
  mutable std::vector<float> dnnFeatures, dnnScore;
  mutable std::vector<unsigned> dnnT3Pair, dnnHitIdxs;
  // quintuplets
  unsigned t5Capacity = 0;
  std::vector<unsigned> nQuintuplets, totOccupancyQuintuplets, t5TripletIndices, t5HitIndices;
  std::vector<uint16_t> t5LowerModuleIndices;
  std::vector<uint8_t> t5LogicalLayers, t5Layer, t5TightCutFlag, t5PartOfPT5;
  std::vector<int8_t> t5IsDup;
  std::vector<float> t5InnerRadius, t5BridgeRadius, t5OuterRadius, t5Pt, t5Eta, t5Phi, t5ScoreRphiSum,
      t5RegressionRadius, t5RegressionG, t5RegressionF, t5RzChiSquared, t5ChiSquared, t5NonAnchorChiSquared;
  // pixel quintuplets (SDL/PixelTriplet.h:1871-1913), capacity N_MAX_PIXEL_QUINTUPLETS
  unsigned nPixelQuintuplets = 0, totOccupancyPixelQuintuplets = 0;
  std::vector<unsigned> pt5PixelIndices, pt5T5Indices, pt5HitIndices;
  std::vector<uint8_t> pt5IsDup, pt5LogicalLayers;
  std::vector<uint16_t> pt5LowerModuleIndices;
  std::vector<float> pt5Score, pt5Eta, pt5Phi, pt5PixelRadius, pt5QuintupletRadius, pt5CenterX, pt5CenterY,
      pt5RzChiSquared, pt5RPhiChiSquared, pt5RPhiChiSquaredInwards;
  // pixel triplets (SDL/PixelTriplet.h:20-75), capacity N_MAX_PIXEL_TRIPLETS
  unsigned nPixelTriplets = 0, totOccupancyPixelTriplets = 0;
  std::vector<unsigned> pt3PixelSegmentIndices, pt3TripletIndices, pt3HitIndices;
  std::vector<uint8_t> pt3IsDup, pt3LogicalLayers;
  std::vector<uint16_t> pt3LowerModuleIndices;
  std::vector<float> pt3PixelRadius, pt3TripletRadius, pt3Pt, pt3Eta, pt3Phi, pt3EtaPix, pt3PhiPix, pt3Score,
      pt3CenterX, pt3CenterY, pt3RzChiSquared, pt3RPhiChiSquared, pt3RPhiChiSquaredInwards;
  // track candidates (SDL/TrackCandidate.h:20-111), capacity 31000
  unsigned nTrackCandidates = 0, nTrackCandidatespT3 = 0, nTrackCandidatespT5 = 0, nTrackCandidatespLS = 0,
           nTrackCandidatesT5 = 0;
  std::vector<int16_t> tcType;
  std::vector<unsigned> tcDirectObjectIndices, tcObjectIndices, tcHitIndices;
  std::vector<int> tcPixelSeedIndex;
  std::vector<uint8_t> tcLogicalLayers;
  std::vector<uint16_t> tcLowerModuleIndices;
  std::vector<float> tcCenterX, tcCenterY, tcRadius;
  // LST::getOutput (SDL/LST.cc:366-420): ragged hit lists, flattened
  std::vector<unsigned> outHitIdxs, outHitOffsets, outLen;
  std::vector<int> outSeedIdx;
  std::vector<int16_t> outType;
};

bool event_add_hits(Event& ev, const Input& in);
void event_add_pixel_segments(Event& ev, const Input& in);
void event_create_mini_doublets(Event& ev);
void event_create_segments(Event& ev);
void event_create_triplets(Event& ev);
void event_create_quintuplets(Event& ev);
void event_pixel_segment_cleaning(Event& ev);
void event_create_pixel_quintuplets(Event& ev);
void event_create_pixel_triplets(Event& ev);
void event_create_track_candidates(Event& ev);
void event_get_output(Event& ev);
void runDeltaBetaIterationsT5(float& betaIn, float& betaOut, float& betaAv, float& pt_beta, float sdIn_dr,
                              float sdOut_dr, float dr, float lIn);
float computeRadiusFromThreeAnchorHits(float x1, float y1, float x2, float y2, float x3, float y3, float& g, float& f);

}  // namespace orc
