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
  int noFallback;                    // a batch slot: no per-module builder runs after the flattened ones
  T5DnnDump dnnDump;
  uint32_t* plsInHitIdx[4];          // [pLS capacity] the four pseudo-hit indices of each pLS, as added
  float* plsInDPhiChange;            // [pLS capacity] its deltaPhiChange, as added
  uint32_t outCap;                   // track-candidate records W.outPacked holds
  uint32_t nHits, nPLS, pixelModule;
  float ptCut;
};

// Kernels read their event's view from a device array: view blockIdx.z.  One
// launch sequence therefore serves one event (gridDim.z = 1) or a batch of
// events, each with its own arena, in the same launches (gridDim.z = events).
#define LST_VIEWS const EventView* __restrict__ vs_
#define LST_BIND const EventView& v = vs_[blockIdx.z]

// Grid-stride kernels whose work-group count a batch may choose (autotuner).
// Every kernel listed here is correct for any gridDim.x >= 1.
enum TuneId : int {
  kTnStageIn = 0, kTnEventInit, kTnHits, kTnModuleRanges, kTnHitRangesFix, kTnPixelSegments,
  kTnMdCandFill, kTnMdEval, kTnMdScanBlocks, kTnMdPlace,
  kTnLsCandCount, kTnLsCandFill, kTnLsEval, kTnLsScanBlocks, kTnLsPlace,
  kTnMdSegLists, kTnT3CandCount, kTnT3Enum, kTnT3Select, kTnT3ScanBlocks, kTnT3Place,
  kTnMdT3Lists, kTnT5CandCount, kTnT5Enum, kTnT5FilterT4, kTnT5Final, kTnT5Dnn, kTnT5ScanBlocks, kTnT5Place,
  kTnT5ListFill, kTnT5DupAfter, kTnPlsPairs, kTnPlsConnected, kTnT3ListFill, kTnPixRecords,
  kTnPT5StageA, kTnPT5StageB, kTnPT5Write, kTnPT5FullWrite, kTnPT5BucketCount, kTnPT5BucketFill, kTnPT5Pairs,
  kTnPT3StageA, kTnPT3StageB, kTnPT3Write, kTnPT3FullWrite, kTnPT3Pairs, kTnCrossPT3,
  kTnT5TcPairs, kTnCrossT5, kTnTcT5Count, kTnTcT5Write, kTnTcPrep, kTnCrossPls, kTnTcOutput, kTnOutCopy,
  kNTune
};
extern const char* const kTuneNames[kNTune];
extern const uint32_t kBatchGrid[kNTune];  // batch launch-set defaults (0: the one-event default)
constexpr uint32_t kMaxBatchSlots = 128;  // events of one batch launch set

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

// One enqueue of (part of) the stage sequence.
struct Launch {
  const EventView& v;            // host copy of the first event's view: geometry sizes and mode flags, which
                                 //   every event of a launch shares
  const EventView* dv;           // device array of nEv views
  uint32_t nEv;                  // events of the launch (gridDim.z)
  hipStream_t s;
  const KernelTimers& t;         // per-kernel timers (one-event launches)
  const uint32_t* grid = nullptr;  // [kNTune] chosen work-group counts (0: default), null: defaults
  bool fallbacks = true;         // launch the per-module builders that take events overflowing the flattened ones
  const hipEvent_t* tuneEvents = nullptr;  // [2 kNTune] recorded before / after each tunable kernel (autotuner)
  uint32_t* tuneRecorded = nullptr;        // bit k of word k / 32: kernel k was bracketed
  uint32_t gx(int id, uint32_t dflt) const {
    if (grid && grid[id]) return grid[id];
    return nEv > 1 && kBatchGrid[id] ? kBatchGrid[id] : dflt;
  }
  void tune_begin(int id) const {
    if (!tuneEvents) return;
    (void)hipEventRecord(tuneEvents[2 * id], s);
    tuneRecorded[id / 32] |= 1u << (id % 32);
  }
  void tune_end(int id) const {
    if (tuneEvents) (void)hipEventRecord(tuneEvents[2 * id + 1], s);
  }
};
// launch kernel K over every event of L; gridDim = (GX, GY, events)
#define LST_GO(L, K, GX, GY, BLOCK) hipLaunchKernelGGL(K, dim3((GX), (GY), (L).nEv), dim3(BLOCK), 0, (L).s, (L).dv)
// the same for a grid-stride kernel whose work-group count the launch may override
#define LST_GO_T(L, ID, K, GX, BLOCK)                                                              \
  do {                                                                                             \
    const uint32_t gx_ = (L).gx(ID, (GX));                                                         \
    (L).tune_begin(ID);                                                                            \
    hipLaunchKernelGGL(K, dim3(gx_, 1, (L).nEv), dim3(BLOCK), 0, (L).s, (L).dv);                   \
    (L).tune_end(ID);                                                                              \
  } while (0)

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
  uint32_t* outHost;  // where k_out_copy puts the block's header (pinned host memory), null: no output
  uint32_t* outDev;   // where k_out_copy puts the whole packed block (device memory, read back by the host)
  uint32_t outCap;    // track-candidate records outDev holds
};
// copies the described inputs into the event arenas (descriptor blockIdx.z of
// the device array d) and sets the events' counts
void launch_stage_in(const Launch& L, const InputDesc* d);

void launch_hits(const Launch& L);
void launch_pixel_segments(const Launch& L);
void launch_mini_doublets(const Launch& L);
void launch_segments(const Launch& L);
void launch_triplets(const Launch& L);
void launch_quintuplets(const Launch& L);
void launch_pixel_segment_cleaning(const Launch& L);
void launch_pixel_quintuplets(const Launch& L);
void launch_pixel_triplets(const Launch& L);
void launch_track_candidates(const Launch& L);
// copies each event's packed output block (header and the records it holds room for) to the
// device block its descriptor names, and the header alone to its host block
void launch_out_copy(const Launch& L, const InputDesc* d);
// all-pairs duplicate removal and cross cleaning (stages_clean.hip)
void launch_pls_dup(const Launch& L, bool second);
void launch_pt5_dup(const Launch& L);
void launch_pt3_dup(const Launch& L);
void launch_cross_pt3(const Launch& L);
void launch_t5_list(const Launch& L);
void launch_t5_dup_after_build(const Launch& L);
void launch_t5_dup_before_tc(const Launch& L);
void launch_cross_t5(const Launch& L);
void launch_cross_pls(const Launch& L);

}  // namespace lst
