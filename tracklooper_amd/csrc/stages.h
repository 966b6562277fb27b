// Host-callable launchers of the LST stage kernels (one file per stage group).
#pragma once
#include <hip/hip_runtime.h>
#include "es.h"
#include "soa.h"

namespace lst {

struct EventView {  // everything a stage kernel needs, by value
  ModulesView M;
  HitsSoA H;
  MDsSoA D;
  SegsSoA S;
  PLSSoA P;
  T3SoA T;
  T5SoA Q;
  PT5SoA P5;
  PT3SoA P3;
  TCSoA C;
  PixelWork W;
  T3Work W3;
  T5Work W5;
  RangesSoA R;
  const int32_t* mdModuleIndices;    // [nLower + 1], geometry-only
  const int32_t* mdModuleOccupancy;  // [nLower + 1]
  const int32_t* segModuleIndices;   // [nLower + 1]
  const int32_t* segModuleOccupancy; // [nLower + 1]
  const int32_t* t3OccIfAny;         // [nLower] T3 capacity of a module that has segments
  const int32_t* t5OccIfAny;         // [nLower] T5 capacity of a seed module, -1 if none
  uint32_t nT5SeedModules;           // geometry bound on the eligible T5 module count
  const uint32_t* pixIndex[3];       // pixel map: connected-module start per superbin, by pLS type
  const uint32_t* pixSizes[3];
  int tcPlsTriplets;                 // TC_PLS_TRIPLETS: triplet pLS may become track candidates
  int perModuleBuilders;             // run the per-module T3/T5 builders instead of the flattened ones
  uint32_t nHits, nPLS, pixelModule;
  float ptCut;
};

// Optional hipEvent pair recorded around the stage's main kernel (timing).
struct KernelTimer {
  hipEvent_t start = nullptr, stop = nullptr;
  void begin(hipStream_t s) const { if (start) (void)hipEventRecord(start, s); }
  void end(hipStream_t s) const { if (stop) (void)hipEventRecord(stop, s); }
};

void launch_hits(const EventView& v, hipStream_t s, const KernelTimer& t = {});
void launch_pixel_segments(const EventView& v, const uint32_t* hitIdx[4], const float* dPhiChange, hipStream_t s,
                           const KernelTimer& t = {});
void launch_mini_doublets(const EventView& v, hipStream_t s, const KernelTimer& t = {});
void launch_segments(const EventView& v, hipStream_t s, const KernelTimer& t = {});
void launch_triplets(const EventView& v, hipStream_t s, const KernelTimer& t = {});
void launch_quintuplets(const EventView& v, hipStream_t s, const KernelTimer& t = {});
void launch_pixel_segment_cleaning(const EventView& v, hipStream_t s, const KernelTimer& t = {});
void launch_pixel_quintuplets(const EventView& v, hipStream_t s, const KernelTimer& t = {});
void launch_pixel_triplets(const EventView& v, hipStream_t s, const KernelTimer& t = {});
void launch_track_candidates(const EventView& v, hipStream_t s, const KernelTimer& t = {});

}  // namespace lst
