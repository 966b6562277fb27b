// ORACLE — test infrastructure only (see orc_common.h).
#pragma once
#include "orc_event.h"

namespace orc {
struct T5Out {
  float innerRadius = 0, bridgeRadius = 0, outerRadius = 0, regressionG = 0, regressionF = 0, regressionRadius = 0,
        rzChiSquared = 0, chiSquared = 0, nonAnchorChiSquared = 0;
  bool tightCutFlag = false;
  unsigned md[5] = {0, 0, 0, 0, 0};
};
bool runQuintupletDefaultAlgo(const Event& ev, const uint16_t mods[5], unsigned innerT3, unsigned outerT3, T5Out& o);
void runDeltaBetaIterationsT5(float& betaIn, float& betaOut, float& betaAv, float& pt_beta, float sdIn_dr,
                              float sdOut_dr, float dr, float lIn);
void sigmasForRegression(const Geometry& G, const uint16_t* mods, float* delta1, float* delta2, float* slopes,
                         bool* isFlat, unsigned nPoints, bool anchorHits);
float regressionCircle(unsigned nPoints, const float* xs, const float* ys, const float* delta1, const float* delta2,
                       const float* slopes, const bool* isFlat, float& g, float& f, float& chiSquared);
float chiSquaredOfCircle(unsigned nPoints, const float* xs, const float* ys, const float* delta1, const float* delta2,
                         const float* slopes, const bool* isFlat, float g, float f, float radius);
float t5dnnForward(const float in[38]);
void t5dnnSetWeights(const float* packed);
void t5dnnFeatures(const Event& ev, const unsigned md[5], const uint16_t mods[5], unsigned innerT3, float innerRadius,
                   float outerRadius, float bridgeRadius, float out[38]);
}  // namespace orc
