/* C ABI of the MI355X Line Segment Tracking engine (liblst_amd.so).
 *
 * Drop-in boundary for the reference's LST path: the entry points below
 * replace, one for one, the reference interfaces cited on each declaration
 * (paths relative to the SegmentLinking/TrackLooper tree).  Plain pointers and
 * sizes only; the engine owns its device memory, the caller owns its inputs
 * (read during the call only) and output buffers.
 *
 * Threading: an lst_es handle is immutable after creation and may be shared by
 * any number of events; an lst_event handle is single-threaded.
 */
#ifndef LST_AMD_H
#define LST_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lst_es lst_es;        /* geometry / event setup (LSTESDeviceData) */
typedef struct lst_event lst_event;  /* one event's device state (SDL::Event)    */

/* Status codes.  LST_W_TRUNCATED is a warning bit OR-ed into LST_OK when a
 * capacity truncated objects exactly as the reference would: more than 50000
 * pixel segments, or a per-module mini-doublet / segment / triplet /
 * quintuplet capacity, or the pT5 / pT3 capacity (the reference only prints a
 * banner, SDL/Event.cc:237-244,759-781, SDL/MiniDoublet.h:960-965). */
typedef enum {
  LST_OK = 0,
  LST_E_IO = 1,        /* geometry file missing or unreadable (reference: exit(2), SDL/LSTESData.cc:31-38) */
  LST_E_FORMAT = 2,    /* malformed geometry file (reference: std::runtime_error, SDL/ModuleConnectionMap.cc:14-46) */
  LST_E_INPUT = 3,     /* bad input: a detId absent from the geometry (reference: indexes mapIdx[-1],
                          SDL/Hit.h:245-246), more than 127 hits on a module (its int8 counts wrap,
                          SDL/Hit.h:31-32), a module's hits not contiguous (SDL/Hit.h:261-267) */
  LST_E_ORDER = 4,     /* stage called out of order */
  LST_E_DEVICE = 5,    /* HIP runtime error */
  LST_E_CAPACITY = 6,  /* output buffer too small */
  LST_W_TRUNCATED = 0x100
} lst_status;

/* Message of the last error on the calling thread ("" if none). */
const char* lst_last_error(void);

/* Replaces SDL::loadAndFillESHost() + SDL::loadAndFillESDevice(queue, host)
 * (SDL/LSTESData.h:79-80, SDL/LSTESData.cc:78-113).  The geometry directory
 * and pT cut are parameters (the reference hard-codes pt0.8 under $LST_BASE,
 * SDL/LSTESData.cc:46-57,98-99).  module_map_stride = 0 uses the maximum
 * fan-out of the geometry (the reference fixes 40, SDL/Constants.h:108, which
 * the pt0.6 map overflows). */
lst_status lst_es_create(const char* geom_dir, float pt_cut, uint32_t module_map_stride, lst_es** out);
void lst_es_destroy(lst_es* es);
/* Geometry scalars: "nModules", "nLowerModules", "nPixels", "pixelModuleIndex", "stride". */
uint32_t lst_es_scalar(const lst_es* es, const char* name);

/* T5 selection modes (reference build flags, bin/sdl_make_tracklooper:157-162). */
enum { LST_T5_PLAIN_CHI2 = 0 /* -DUSE_RZCHI2 -DUSE_RPHICHI2 */, LST_T5_DNN = 1 /* -DUSE_RZCHI2 -DUSE_T5_DNN */ };

/* Replaces SDL::Event<Acc>(verbose, queue, deviceESData) (SDL/Event.h:100-111).
 * hip_stream is a hipStream_t (NULL: the engine creates its own).  Device
 * arenas are allocated once here and reused by every event run on the handle. */
lst_status lst_event_create(lst_es* es, void* hip_stream, int t5_mode, lst_event** out);
void lst_event_destroy(lst_event* ev);

/* Replaces SDL::Event::addHitToEvent(x, y, z, detId, idxInNtuple)
 * (SDL/Event.h:115-119, SDL/Event.cc:151-215).  on_device != 0: the pointers
 * are device pointers (inputs already resident in HBM). */
lst_status lst_event_add_hits(lst_event* ev, const float* x, const float* y, const float* z, const uint32_t* det_id,
                              const uint32_t* idx_in_ntuple, uint32_t n_hits, int on_device);

/* The 18 vectors of SDL::Event::addPixelSegmentToEvent (SDL/Event.h:120-137,
 * SDL/Event.cc:217-391), as arrays of n_pls entries. */
typedef struct {
  const uint32_t *hit_indices0, *hit_indices1, *hit_indices2, *hit_indices3;
  const float *delta_phi_change, *pt_in, *pt_err, *px, *py, *pz, *eta, *eta_err, *phi;
  const int32_t* charge;
  const uint32_t* seed_idx;
  const int32_t* superbin;
  const int8_t* pixel_type;
  const int8_t* is_quad;
  uint32_t n_pls;
} lst_pixel_segments;
lst_status lst_event_add_pixel_segments(lst_event* ev, const lst_pixel_segments* pls, int on_device);

/* Stage calls, in this order (SDL/Event.h:138-150, SDL/LST.cc:74-148). */
lst_status lst_event_create_mini_doublets(lst_event* ev);        /* createMiniDoublets */
lst_status lst_event_create_segments(lst_event* ev);             /* createSegmentsWithModuleMap */
lst_status lst_event_create_triplets(lst_event* ev);             /* createTriplets */
/* createQuintuplets (SDL/Event.cc:903-991): eligible-module list, T5 builder,
 * after-build duplicate flags.  The T5 selection is the event's mode: the r-phi
 * chi2 cut (LST_T5_PLAIN_CHI2) or the T5 DNN at its working point
 * (LST_T5_DNN). */
lst_status lst_event_create_quintuplets(lst_event* ev);
/* pixelLineSegmentCleaning (SDL/Event.cc:993-1007): first checkHitspLS pass. */
lst_status lst_event_pixel_line_segment_cleaning(lst_event* ev);
/* createPixelQuintuplets (SDL/Event.cc:1009-1146): pT5 builder, pT5 duplicate
 * flags, pT5 track candidates. */
lst_status lst_event_create_pixel_quintuplets(lst_event* ev);
/* createPixelTriplets (SDL/Event.cc:784-901): pT3 builder and duplicate flags. */
lst_status lst_event_create_pixel_triplets(lst_event* ev);
/* createTrackCandidates (SDL/Event.cc:620-782): cross cleaning, pT3/T5/pLS
 * track candidates, second checkHitspLS pass. */
lst_status lst_event_create_track_candidates(lst_event* ev);

/* Replaces SDL::LST<Acc>::getOutput / getHitIdxs (SDL/LST.cc:366-420) after
 * createTrackCandidates (waits for the event's stream).  LST_E_INPUT if the
 * device-side hit checks found bad input (hits given as device pointers are
 * checked there; host hits are checked by lst_run and lst_event_stage_input
 * before any upload), LST_W_TRUNCATED if a capacity truncated objects.
 * *n_tc receives the number of track candidates; when
 * any array is given (each may be NULL) and capacity >= *n_tc, hit_idxs
 * receives 14 entries per candidate (the first len[i] are the candidate's
 * ntuple hit indices, the reference's out_tc_hitIdxs_), len the list lengths,
 * seed_idx the pixel seed index (-1 for T5 candidates) and type the candidate
 * type (4 T5, 5 pT3, 7 pT5, 8 pLS).  LST_E_CAPACITY if capacity is too small. */
lst_status lst_event_get_output(lst_event* ev, uint32_t* n_tc, uint32_t* hit_idxs, uint32_t* len, int32_t* seed_idx,
                                int16_t* type, uint32_t capacity);

/* Replaces SDL::Event::resetEvent (SDL/Event.cc:45-149); keeps the arenas. */
lst_status lst_event_reset(lst_event* ev);
/* Waits for the event's stream; then the status of the device-side checks
 * (as lst_event_get_output). */
lst_status lst_event_synchronize(lst_event* ev);

/* Replaces SDL::Event::getNumberOf* (SDL/Event.h:152-183): "hits", "pixelSegments",
 * "miniDoublets", "segments", "triplets", "quintuplets" (outer-tracker totals),
 * "pixelQuintuplets", "pixelTriplets", "trackCandidates", "pT5TrackCandidates",
 * "pT3TrackCandidates", "T5TrackCandidates", "pLSTrackCandidates".  Diagnostics
 * with no counterpart in the reference: "tripletCandidates" (segment pairs the
 * triplet builder enumerates), "tripletSharedMDCandidates" (those sharing the
 * middle mini-doublet), "quintupletCandidates" (triplet pairs enumerated),
 * "miniDoubletCandidates" (hit pairs enumerated), "segmentCandidates"
 * (mini-doublet pairs enumerated), "pixelQuintupletSurvivors" and
 * "pixelTripletSurvivors" (pLS-object pairs passing the cheap cuts), and
 * "arenaMiB" (device memory of the event handle, MiB). */
uint32_t lst_event_count(lst_event* ev, const char* what);

/* Replaces the host-copy getters SDL::Event::get{Hits,MiniDoublets,Segments,
 * Triplets,Ranges}() (SDL/Event.h:185-204) field by field: copies the named
 * SoA field ("hits.xs", "mds.anchorHitIndices", "segs.mdIndices",
 * "t3.circleRadius", "ranges.mdRanges", ...) into dst.  With dst == NULL only
 * *dtype (0 f32, 1 u32, 2 i32, 3 u16, 4 i16, 5 u8, 6 i8) and *count are set. */
lst_status lst_event_get_field(lst_event* ev, const char* name, void* dst, size_t dst_bytes, int* dtype,
                               size_t* count);

/* Arguments of SDL::LST<Acc>::run (SDL/LST.h:29-48); see_hitIdx is flattened
 * with n_seeds + 1 offsets. */
typedef struct {
  const float *see_px, *see_py, *see_pz, *see_dxy, *see_dz, *see_ptErr, *see_etaErr;
  const float *see_stateTrajGlbX, *see_stateTrajGlbY, *see_stateTrajGlbZ;
  const float *see_stateTrajGlbPx, *see_stateTrajGlbPy, *see_stateTrajGlbPz;
  const int32_t* see_q;
  const int32_t* see_hitIdx;
  const uint32_t* see_hitIdx_offsets;
  uint32_t n_seeds;
  const uint32_t* ph2_detId;
  const float *ph2_x, *ph2_y, *ph2_z;
  uint32_t n_hits;
} lst_input;

/* Replaces SDL::LST<Acc>::run (SDL/LST.cc:12-161) up to and including
 * `last_stage` (0 hits, 1 pixel segments, 2 mini-doublets, 3 segments,
 * 4 triplets, 5 quintuplets, 6 pLS cleaning, 7 pixel quintuplets, 8 pixel
 * triplets, 9 track candidates): prepareInput on the host (SDL/LST.cc:175-365),
 * the hit checks of LST_E_INPUT, then the stage sequence on the device. */
lst_status lst_run(lst_event* ev, const lst_input* in, int last_stage);

/* Throughput path (the reference preloads its inputs on the host,
 * bin/sdl.cc:352-384, and its timed loop copies them to the device,
 * code/core/trkCore.cc:877-879): prepare an event on the host once (with the
 * hit checks) and keep it in pinned host memory in input slot `slot` ... */
lst_status lst_event_stage_input(lst_event* ev, const lst_input* in, uint32_t slot);
/* ... then copy it to the device in one transfer and run the chain on it
 * (asynchronous on the event's stream; same results as lst_run). */
lst_status lst_event_run_staged(lst_event* ev, uint32_t slot, int last_stage);

/* T5 DNN training rows (no counterpart in the reference, whose DNN score is
 * not stored; the rows are the candidates at its DNN call site,
 * SDL/Quintuplet.h:2921-2936).  lst_event_record_t5dnn(ev, capacity) on an
 * LST_T5_DNN event makes later createQuintuplets calls record, per candidate
 * that reaches the DNN, its 38 inputs (SDL/NeuralNetwork.h:36-114), its score,
 * its (inner, outer) triplet pair and its 10 hit indices in input numbering;
 * capacity 0 stops recording.  lst_event_get_t5dnn_record copies up to
 * `capacity` rows (any array may be NULL) and sets *n to the rows recorded;
 * row order is arbitrary.  LST_W_TRUNCATED if rows were dropped;
 * LST_E_CAPACITY if the event took the per-module quintuplet builder. */
lst_status lst_event_record_t5dnn(lst_event* ev, uint32_t capacity);
lst_status lst_event_get_t5dnn_record(lst_event* ev, uint32_t* n, float* features, float* scores, uint32_t* t3_pairs,
                                      uint32_t* hit_idxs, uint32_t capacity);

/* T5 DNN parameters and training (no counterpart in the reference, which
 * compiles its trained weights in, SDL/NeuralNetworkWeights.h).  Parameters
 * are packed as bias_0[32], wgtT_0[38][32], bias_2[32], wgtT_2[32][32],
 * bias_4[1], wgtT_4[32][1] (LST_T5DNN_NPARAMS floats).
 * lst_t5dnn_set_weights / get_weights: the parameters every event's DNN mode
 * uses (process-wide, on the current device; synchronises it).
 * lst_t5dnn_forward / backward: one training step's kernels on device arrays
 * (x [n][38]; activations h1, h2 [n][32]; logit, dlogit [n]; work d1, d2
 * [n][32]; grad [LST_T5DNN_NPARAMS], overwritten), enqueued on `stream`. */
#define LST_T5DNN_NPARAMS 2337
lst_status lst_t5dnn_set_weights(const float* params);
lst_status lst_t5dnn_get_weights(float* params);
/* The score a T5 candidate must exceed in LST_T5_DNN mode (process-wide, on
 * the current device; default the reference's LSTWP2 = 0.6177366,
 * SDL/Constants.h:145, which SDL/Quintuplet.h:2937-2938 applies), for
 * retrained parameters calibrated to another working point. */
lst_status lst_t5dnn_set_working_point(float wp);
float lst_t5dnn_get_working_point(void);
lst_status lst_t5dnn_forward(const float* x, const float* params, float* h1, float* h2, float* logit, uint32_t n,
                             void* stream);
lst_status lst_t5dnn_backward(const float* x, const float* params, const float* h1, const float* h2,
                              const float* dlogit, float* d2, float* d1, float* grad, uint32_t n, void* stream);
/* One training step on the matrix cores: forward, binary cross-entropy of the
 * logits against labels y [n] (0 or 1), and backward, every GEMM in BF16 with
 * FP32 accumulation (the 32 -> 1 output layer and the element-wise steps in
 * FP32).  grad [LST_T5DNN_NPARAMS] receives the gradient of the mean loss,
 * *loss (device pointer, may be NULL) the mean loss.  Enqueued on `stream`;
 * uses a workspace the library allocates per device (kept for the process);
 * deterministic (partials of 64-sample chunks summed in a fixed order). */
lst_status lst_t5dnn_train_step_bf16(const float* x, const float* y, const float* params, float* grad, float* loss,
                                     uint32_t n, void* stream);
/* lst_t5dnn_train_step_bf16 with every GEMM operand split into two BF16 terms
 * (hi = BF16(v), lo = BF16(v - hi)) and three MFMAs per product (lo*hi +
 * hi*lo + hi*hi, Henry, Tang and Heinecke, ARITH 2019): GEMMs of about FP32
 * accuracy on the matrix cores, for inputs such as the raw T5 features
 * (z up to 270 cm) whose BF16 rounding is visible in the loss.  Same arguments
 * and workspace as lst_t5dnn_train_step_bf16. */
lst_status lst_t5dnn_train_step_bf16x3(const float* x, const float* y, const float* params, float* grad, float* loss,
                                       uint32_t n, void* stream);
/* lst_t5dnn_train_step_bf16 without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight training loops, harnesses that forbid hipMalloc).  The
 * chunk partials live in a fixed per-device array of the library
 * (LST_T5DNN_NOALLOC_MAX_ROWS / 64 chunks, about 4.8 MB, held until the
 * process ends); n above LST_T5DNN_NOALLOC_MAX_ROWS returns LST_E_CAPACITY.
 * Same kernels and bit-identical results as lst_t5dnn_train_step_bf16.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * streams, other threads) race. */
#define LST_T5DNN_NOALLOC_MAX_ROWS 32768u
lst_status lst_t5dnn_train_step_bf16_noalloc(const float* x, const float* y, const float* params, float* grad,
                                             float* loss, uint32_t n, void* stream);
/* Adam (Kingma and Ba 2015) on the packed parameters in place: m, v the
 * moment estimates (zero before step 1), step = 1, 2, ... */
lst_status lst_t5dnn_adam(float* params, const float* grad, float* m, float* v, uint32_t step, float lr, float beta1,
                          float beta2, float eps, void* stream);

/* ---- Batched engine.  The reference runs one SDL::Event per queue with one
 * OpenMP thread per queue (bin/sdl.cc:389-419) and LST::run one event at a
 * time (SDL/LST.cc:12-161).  A batch holds `n_slots` event arenas on one GPU
 * and runs up to n_slots events in one launch set: each kernel of the stage
 * sequence is launched once for all of them, and the sequence is one captured
 * HIP graph.  Inputs are prepared on the host and copied to the device inside
 * the run, overlapped with the previous launch set; outputs come back to
 * pinned host memory.  An event whose candidates overflow the flattened
 * builders' work arrays is run again by a per-event handle (per-module
 * builders).  Per event the results equal lst_run's. */
typedef struct lst_batch lst_batch;

/* Output lists of one event, as LST::hits()/len()/seedIdx()/trackCandidateType()
 * (SDL/LST.cc:366-420).  The caller sets `capacity` and the arrays (any may be
 * NULL); n_tc is set to the event's track-candidate count.  When n_tc exceeds
 * capacity the first `capacity` entries are written and the call returns
 * LST_OK | LST_W_TRUNCATED. */
typedef struct {
  uint32_t capacity;  /* in: entries the arrays hold */
  uint32_t n_tc;      /* out */
  uint32_t* hit_idxs; /* [capacity][14], input hit indices, entries past len are 0 */
  uint32_t* len;      /* [capacity] */
  int32_t* seed_idx;  /* [capacity] */
  int16_t* type;      /* [capacity] 4 T5, 5 pT3, 7 pT5, 8 pLS */
  int32_t status;     /* out: the event's lst_status (LST_E_INPUT, LST_W_TRUNCATED, ...) */
} lst_output;

/* n_slots: events per launch set, 1..128; each slot holds one event arena. */
lst_status lst_batch_create(lst_es* es, uint32_t n_slots, int t5_mode, lst_batch** out);
void lst_batch_destroy(lst_batch* b);
uint32_t lst_batch_contexts(const lst_batch* b);
/* Launch sets kept in flight (1, the default, or 2; no reference counterpart:
 * the reference runs one event per stream and OpenMP thread,
 * bin/sdl.cc:389-419).  With 2, each set has its own n_slots arenas, view
 * array and compute stream, and a launch set runs beside the previous one,
 * filling the GPU while the other's small kernels and tails drain; the device
 * memory of the arenas doubles.  Synchronises the batch; the results are the
 * same either way. */
lst_status lst_batch_set_sets_in_flight(lst_batch* b, uint32_t sets);
/* n_ev events from host arrays (prepareInput on the host per event, then
 * launch sets of n_slots events); out: NULL or n_ev output descriptors.
 * Returns the first event error (each event's own status is in out[i].status),
 * with LST_W_TRUNCATED if any event was truncated. */
lst_status lst_run_batch(lst_batch* b, const lst_input* events, uint32_t n_ev, lst_output* out);
/* Throughput path: prepare n_ev events once and keep them in pinned host
 * memory (staged events 0..n_ev-1) ... */
lst_status lst_batch_stage(lst_batch* b, const lst_input* events, uint32_t n_ev);
/* ... then run n_ev events from staged event `first` on (inputs copied to the
 * device inside the run, no host preparation); event i is staged event
 * (first + i) modulo the number staged. */
lst_status lst_batch_run_staged(lst_batch* b, uint32_t first, uint32_t n_ev, lst_output* out);
/* Autotuner (no reference counterpart): on one launch set of staged events
 * from `first` on, times every grid-stride kernel of the sequence (the
 * lst_tune_count() kernels named by lst_tune_name) at each of a list of
 * work-group counts, `repeats` times, and keeps the fastest count per kernel
 * for the batch's later runs; *default_ms / *tuned_ms (may be NULL) get the
 * summed kernel times at the default and the chosen counts. */
lst_status lst_batch_autotune(lst_batch* b, uint32_t first, uint32_t repeats, float* default_ms, float* tuned_ms);
/* Device time in ms of each tunable kernel (n entries, in lst_tune_name
 * order; 0 for a kernel the sequence does not run) over one launch set of
 * staged events from `first` on, launched without the graph (synchronises). */
lst_status lst_batch_kernel_times(lst_batch* b, uint32_t first, float* ms, uint32_t n);
/* The chosen work-group counts (0: the kernel's default), n entries. */
lst_status lst_batch_get_tuning(const lst_batch* b, uint32_t* grid, uint32_t n);
lst_status lst_batch_set_tuning(lst_batch* b, const uint32_t* grid, uint32_t n);
uint32_t lst_tune_count(void);
const char* lst_tune_name(uint32_t k);

/* Per-kernel device timing (hipEvents on the event's stream).  After a run
 * with timing on, lst_event_kernel_times writes the durations in ms of the
 * last launch of: 0 hits, 1 pixel segments, 2 mini-doublets, 3 segments,
 * 4 triplet selection, 5 quintuplet selection, 6 pLS cleaning (first pass),
 * 7 pixel-quintuplet candidates, 8 pixel-triplet candidates, 9 T5 duplicate
 * removal before the track candidates, 10 pT5 duplicate removal, 11 pT3
 * duplicate removal, 12 pLS cross cleaning, 13 pLS cleaning (second pass),
 * 14 T5 duplicate removal after build, 15 triplet shared-MD filter, 16 T5
 * DNN (LST_T5_DNN mode only); 0 for a kernel the run did not reach
 * (n <= 17 entries; synchronises the stream). */
lst_status lst_event_set_timing(lst_event* ev, int on);
lst_status lst_event_kernel_times(lst_event* ev, float* ms, int n);
/* The hipStream_t the event runs on. */
void* lst_event_stream(lst_event* ev);

#ifdef __cplusplus
}
#endif
#endif /* LST_AMD_H */
