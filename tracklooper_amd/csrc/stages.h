// Host-callable launchers of the LST stage kernels (one file per stage group).
#pragma once
#include <hip/hip_runtime.h>
#include "es.h"
#include "soa.h"

namespace lst {

// Optional record of every T5 candidate at the DNN call site (DNN mode,
// flattened builder): the 38 inputs, the score, the (inner, outer) T3 pair
// and the 10 hit indices in input numbering.  Row k is parked result k.
struct T5DnnDump {
  float* features = nullptr;  // [capacity * 38]; null: no record
  float* score = nullptr;     // [capacity]
  uint32_t* t3Pair = nullptr; // [capacity * 2]
  uint32_t* hitIdxs = nullptr;// [capacity * 10]
  uint32_t capacity = 0;
};

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
  FlatScan FS;
  MDWork W2;
  LSWork WL;
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
  uint32_t t5Cap;                    // T5 arena slots
  const uint32_t* pixIndex[3];       // pixel map: connected-module start per superbin, by pLS type
  const uint32_t* pixSizes[3];
  int tcPlsTriplets;                 // TC_PLS_TRIPLETS: triplet pLS may become track candidates
  int perModuleBuilders;             // run the per-module T3/T5 builders instead of the flattened ones
  int largePaths;                    // per-MD lists and eta binning take the paths of oversized modules/events
  int t5Dnn;                         // T5 selection by the DNN (USE_T5_DNN) instead of the r-phi chi2 cut
  T5DnnDump dnnDump;
  uint32_t nHits, nPLS, pixelModule;
  float ptCut;
};

// Optional hipEvent pairs recorded around the hot kernels (timing).  Slot k
// brackets one kernel; `recorded` (null: timing off) collects the slots used.
enum TimerId : int {
  kTmHits = 0,          // k_hits
  kTmPixelSegments,     // k_pixel_segments
  kTmMiniDoublets,      // k_md_eval
  kTmSegments,          // k_ls_eval
  kTmT3Select,          // k_t3_select
  kTmT5Final,           // k_t5_final
  kTmPlsDup,            // k_pls_pairs<false> (pLS cleaning, first pass)
  kTmPT5Build,          // k_pix_stage_a<PT5Kind> (pixel quintuplets, cheap cuts)
  kTmPT3Build,          // k_pix_stage_a<PT3Kind> (pixel triplets, cheap cuts)
  kTmT5DupTC,           // k_t5_tc_pairs
  kTmPT5Dup,            // k_pt5_pairs
  kTmPT3Dup,            // k_pt3_pairs
  kTmCrossPLS,          // k_cross_pls_pairs
  kTmPlsDup2,           // k_pls_pairs<true> (second pass)
  kTmT5DupAfter,        // k_t5_dup_after_build
  kTmT3Filter,          // k_t3_enum (triplet candidates sharing the middle mini-doublet)
  kTmT5Dnn,             // k_t5_dnn (DNN mode only)
  kNTimers
};
struct KernelTimers {
  hipEvent_t start[kNTimers] = {}, stop[kNTimers] = {};
  uint32_t* recorded = nullptr;
  void begin(hipStream_t s, int k) const {
    if (!recorded) return;
    (void)hipEventRecord(start[k], s);
    *recorded |= 1u << k;
  }
  void end(hipStream_t s, int k) const {
    if (recorded) (void)hipEventRecord(stop[k], s);
  }
};
inline const KernelTimers& no_timers() {
  static const KernelTimers off{};
  return off;
}

// Device-to-device copy of up to kMaxCopies 4-byte aligned buffers in one launch.
constexpr int kMaxCopies = 24;
struct CopyList {
  const void* src[kMaxCopies];
  void* dst[kMaxCopies];
  uint32_t bytes[kMaxCopies];
  int n = 0;
};
void launch_copy_multi(const CopyList& c, hipStream_t s);
// writes one device word (the event's hit / pLS counts) in stream order
void launch_set_scalar(uint32_t* p, uint32_t value, hipStream_t s);

// An event's prepared inputs in device memory (hits, then pixel line
// segments), described in device memory so that one captured launch sequence
// can read whichever event the descriptor names.
struct InputDesc {
  const float *x, *y, *z;
  const uint32_t *det, *idx;
  uint32_t nHits;
  const uint32_t* hitIdx[4];
  const float *dPhiChange, *ptIn, *ptErr, *px, *py, *pz, *eta, *etaErr, *phi;
  const int32_t *charge, *superbin;
  const uint32_t* seedIdx;
  const int8_t *pixelType, *isQuad;
  uint32_t nPLS;
};
// copies the described inputs into the event arena and sets the event's counts
void launch_stage_in(const EventView& v, const InputDesc* d, uint32_t* const plsHitIdx[4], float* plsDPhiChange,
                     hipStream_t s);

void launch_hits(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
void launch_pixel_segments(const EventView& v, const uint32_t* hitIdx[4], const float* dPhiChange, hipStream_t s,
                           const KernelTimers& t = no_timers());
void launch_mini_doublets(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
void launch_segments(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
void launch_triplets(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
void launch_quintuplets(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
void launch_pixel_segment_cleaning(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
void launch_pixel_quintuplets(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
void launch_pixel_triplets(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
void launch_track_candidates(const EventView& v, hipStream_t s, const KernelTimers& t = no_timers());
// all-pairs duplicate removal and cross cleaning (stages_clean.hip)
void launch_pls_dup(const EventView& v, hipStream_t s, bool second, const KernelTimers& t);
void launch_pt5_dup(const EventView& v, hipStream_t s, const KernelTimers& t);
void launch_pt3_dup(const EventView& v, hipStream_t s, const KernelTimers& t);
void launch_cross_pt3(const EventView& v, hipStream_t s);
void launch_t5_list(const EventView& v, hipStream_t s);
void launch_t5_dup_after_build(const EventView& v, hipStream_t s, const KernelTimers& t);
void launch_t5_dup_before_tc(const EventView& v, hipStream_t s, const KernelTimers& t);
void launch_cross_t5(const EventView& v, hipStream_t s);
void launch_cross_pls(const EventView& v, hipStream_t s, const KernelTimers& t);

}  // namespace lst
