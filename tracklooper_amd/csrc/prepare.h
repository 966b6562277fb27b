// Host-side input preparation of SDL::LST::run (SDL/LST.cc:175-365).
#pragma once
#include <cstdint>
#include <vector>
#include "../../include/lst_amd.h"

namespace lst {

struct PreparedEvent {
  // hits handed to addHitToEvent: outer-tracker hits then pLS pseudo-hits
  std::vector<float> x, y, z;
  std::vector<uint32_t> detId, idx;
  // the 18 pixel-segment vectors of addPixelSegmentToEvent
  std::vector<uint32_t> h0, h1, h2, h3, seedIdx;
  std::vector<float> dPhiChange, ptIn, ptErr, px, py, pz, eta, etaErr, phi;
  std::vector<int32_t> charge, superbin;
  std::vector<int8_t> pixelType, isQuad;
  lst_pixel_segments view() const;
};

// Seeds with pT above 0.8 - 2 ptErr (double literal of SDL/LST.cc:245) become
// pLSs; every pLS appends 3 or 4 pseudo-hits with detId 1.
void prepare_event(const lst_input& in, PreparedEvent& out);

}  // namespace lst
