// ORACLE — test infrastructure only (see orc_common.h).
#pragma once
#include "orc_event.h"

namespace orc {
// Arguments of SDL::LST::run (SDL/LST.h:29-48); see_hitIdx is flattened with offsets.
struct SeedInput {
  std::vector<float> see_px, see_py, see_pz, see_dxy, see_dz, see_ptErr, see_etaErr, see_stateTrajGlbX,
      see_stateTrajGlbY, see_stateTrajGlbZ, see_stateTrajGlbPx, see_stateTrajGlbPy, see_stateTrajGlbPz;
  std::vector<int> see_q, see_hitIdx;
  std::vector<unsigned> see_hitIdx_off;  // n_seeds + 1 offsets into see_hitIdx
  std::vector<unsigned> ph2_detId;
  std::vector<float> ph2_x, ph2_y, ph2_z;
};
// plsPtCut is the double literal 0.8 of SDL/LST.cc:245,261
void prepare_input(const SeedInput& s, double plsPtCut, Input& out);
}  // namespace orc
