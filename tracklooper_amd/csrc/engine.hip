// Event orchestration and the C ABI of the LST engine.
//
// Device memory: one arena per event handle, carved once into the SoA of every
// object kind at reference capacities and reused by every event run on the
// handle (the reference allocates and frees ~0.4 GB per event,
// SDL/Event.cc:45-149).  No stage makes a host round trip: capacities that the
// reference reads back (SDL/Event.cc:252-320,520-541) are computed on the
// device or fixed by the geometry.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lst_amd.h"
#include "es.h"
#include "prepare.h"
#include "stages.h"

namespace {

thread_local std::string g_err;

lst_status fail(lst_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

#define HIP_OK(call)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return fail(LST_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

enum DType { F32 = 0, U32 = 1, I32 = 2, U16 = 3, I16 = 4, U8 = 5, I8 = 6 };
template <class T> constexpr int dtype_of();
template <> constexpr int dtype_of<float>() { return F32; }
template <> constexpr int dtype_of<uint32_t>() { return U32; }
template <> constexpr int dtype_of<int32_t>() { return I32; }
template <> constexpr int dtype_of<uint16_t>() { return U16; }
template <> constexpr int dtype_of<int16_t>() { return I16; }
template <> constexpr int dtype_of<uint8_t>() { return U8; }
template <> constexpr int dtype_of<int8_t>() { return I8; }

struct FieldRef {
  void* ptr;
  int dtype;
  size_t elem;
  size_t count;
};

// Input copies of one stage: from device memory, the 4-byte aligned buffers go
// in one multi-copy launch (others and host sources take hipMemcpyAsync).
struct CopyBatch {
  bool onDevice;
  hipStream_t stream;
  lst::CopyList list;
  hipError_t err = hipSuccess;
  CopyBatch(bool d, hipStream_t s) : onDevice(d), stream(s) {}
  void add(void* dst, const void* src, size_t bytes) {
    if (!bytes || err != hipSuccess) return;
    const bool aligned = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 3u) == 0;
    if (onDevice && aligned && bytes < (1ull << 32) && list.n < lst::kMaxCopies) {
      list.src[list.n] = src;
      list.dst[list.n] = dst;
      list.bytes[list.n] = (uint32_t)bytes;
      list.n++;
    } else {
      err = hipMemcpyAsync(dst, src, bytes, onDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream);
    }
  }
  hipError_t flush() {
    if (err != hipSuccess) return err;
    lst::launch_copy_multi(list, stream);
    list.n = 0;
    return hipGetLastError();
  }
};

// Bump allocator over one device allocation; 256-byte aligned sub-buffers.
struct Carve {
  size_t off = 0;
  std::vector<std::pair<void**, size_t>> todo;
  template <class T>
  void take(T*& p, size_t n) {
    todo.push_back({reinterpret_cast<void**>(&p), off});
    off += (n * sizeof(T) + 255) & ~size_t(255);
  }
  void bind(char* base) {
    for (auto& t : todo) *t.first = base + t.second;
  }
};

}  // namespace

struct lst_es {
  lst::HostGeometry host;
  lst::DeviceGeometry dev;
  float ptCut = 0.8f;
};

struct lst_event {
  lst_es* es = nullptr;
  hipStream_t stream = nullptr;
  bool ownStream = false;
  int t5Mode = 0;
  int stage = -1;  // last completed stage
  lst::EventView v{};
  // capacities
  uint32_t hitCap = 0, mdCap = 0, segCap = 0, plsCap = lst::kNMaxPixelSegments, t3Cap = 0, t5Cap = 0;
  static constexpr uint32_t pt5Cap = lst::kNMaxPixelQuintuplets, pt3Cap = lst::kNMaxPixelTriplets,
                            tcCap = lst::kNMaxPixelTrackCandidates + lst::kNMaxNonPixelTrackCandidates;
  char* arena = nullptr;
  size_t arenaBytes = 0;
  uint32_t* plsHitIdx[4] = {nullptr, nullptr, nullptr, nullptr};
  float* plsDPhiChange = nullptr;
  std::vector<std::pair<std::string, FieldRef>> fields;
  lst::PreparedEvent prepared;
  // timing of the hot kernels (slots of lst::TimerId)
  bool timing = false;
  lst::KernelTimers timers;
  uint32_t recorded = 0;  // timer slots recorded since timing was switched on
  const lst::KernelTimers& tm() {
    timers.recorded = timing ? &recorded : nullptr;
    return timers;
  }
  // captured launch sequence of the whole chain on described device inputs
  // (batched engine): recaptured whenever the arena is reallocated
  void* dnnRecordMem = nullptr;        // T5 DNN training rows (lst_event_record_t5dnn)
  lst::InputDesc* descDev = nullptr;   // device copy of the current event's descriptor
  lst::InputDesc* descHost = nullptr;  // pinned staging of that descriptor
  hipGraphExec_t graph = nullptr;
  uint32_t arenaGen = 0, graphGen = 0;
  // device-resident prepared inputs (throughput path)
  struct Staged {
    char* mem = nullptr;
    size_t bytes = 0;
    uint32_t nHits = 0, nPLS = 0;
    bool truncated = false;  // the event had more pixel segments than kNMaxPixelSegments
    float *x = nullptr, *y = nullptr, *z = nullptr;
    uint32_t *det = nullptr, *idx = nullptr;
    uint32_t *h[4] = {nullptr, nullptr, nullptr, nullptr}, *seed = nullptr;
    float *dphi = nullptr, *pt = nullptr, *pterr = nullptr, *px = nullptr, *py = nullptr, *pz = nullptr,
          *eta = nullptr, *etaerr = nullptr, *phi = nullptr;
    int32_t *charge = nullptr, *superbin = nullptr;
    int8_t *ptype = nullptr, *quad = nullptr;
  };
  std::vector<Staged> staged;

  lst_status allocate(uint32_t nHitsNeeded);
  lst_status add_hits(const float* x, const float* y, const float* z, const uint32_t* det, const uint32_t* idx,
                      uint32_t n, bool onDevice);
  lst_status add_pixel_segments(const lst_pixel_segments& p, bool onDevice);
  const FieldRef* field(const char* name);
};

lst_status lst_event::allocate(uint32_t nHitsNeeded) {
  if (arena && nHitsNeeded <= hitCap) return LST_OK;
  if (arena) {
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipFree(arena));
    arena = nullptr;
  }
  const lst::DeviceGeometry& g = es->dev;
  const uint32_t nMod = g.view.nModules, nL = g.view.nLowerModules;
  hitCap = nHitsNeeded < (1u << 18) ? (1u << 18) : nHitsNeeded + nHitsNeeded / 4;
  mdCap = g.nTotalMDCap + lst::kNMaxPixelMDs;
  segCap = g.nTotalSegCap + lst::kNMaxPixelSegments;
  t3Cap = (uint32_t)g.t3CapBound;
  t5Cap = (uint32_t)g.t5CapBound;
  Carve c;
  lst::EventView& e = v;
#define CARVE(TY, NM, MU) c.take(e.H.NM, (size_t)hitCap * MU);
  LST_HIT_FIELDS(CARVE)
#undef CARVE
#define CARVE(TY, NM, MU) c.take(e.D.NM, (size_t)mdCap * MU);
  LST_MD_FIELDS(CARVE)
#undef CARVE
#define CARVE(TY, NM, MU) c.take(e.S.NM, (size_t)segCap * MU);
  LST_SEG_FIELDS(CARVE)
#undef CARVE
#define CARVE(TY, NM, MU) c.take(e.P.NM, (size_t)plsCap * MU);
  LST_PLS_FIELDS(CARVE)
#undef CARVE
#define CARVE(TY, NM, MU) c.take(e.T.NM, (size_t)(t3Cap ? t3Cap : 1) * MU);
  LST_T3_FIELDS(CARVE)
#undef CARVE
#define CARVE(TY, NM, MU) c.take(e.Q.NM, (size_t)(t5Cap ? t5Cap : 1) * MU);
  LST_T5_FIELDS(CARVE)
#undef CARVE
#define CARVE(TY, NM, MU) c.take(e.P5.NM, (size_t)pt5Cap * MU);
  LST_PT5_FIELDS(CARVE)
#undef CARVE
#define CARVE(TY, NM, MU) c.take(e.P3.NM, (size_t)pt3Cap * MU);
  LST_PT3_FIELDS(CARVE)
#undef CARVE
#define CARVE(TY, NM, MU) c.take(e.C.NM, (size_t)tcCap * MU);
  LST_TC_FIELDS(CARVE)
#undef CARVE
  c.take(e.W.connBase, plsCap); c.take(e.W.connSize, plsCap);
  c.take(e.W.pt5Count, plsCap + 1); c.take(e.W.pt3Count, plsCap + 1);
  c.take(e.W.t5TcCount, nL + 1);
  c.take(e.W.outHits, (size_t)tcCap * 14); c.take(e.W.outLen, tcCap);
  c.take(e.W.outPacked, lst::kOutHeaderWords + (size_t)tcCap * lst::kOutRecWords);
  c.take(e.W.t5ListOff, nL + 1); c.take(e.W.t5List, t5Cap ? t5Cap : 1); c.take(e.W.t5ListMod, t5Cap ? t5Cap : 1);
  c.take(e.W.tcEta, tcCap); c.take(e.W.tcPhi, tcCap); c.take(e.W.tcKind, tcCap);
  c.take(e.W.plsBinned, plsCap); c.take(e.W.plsBinOff, lst::kEtaBinOffsets);
  c.take(e.W.plsBinRec, plsCap); c.take(e.W.plsBinHits, plsCap);
  c.take(e.W.t5Binned, t5Cap ? t5Cap : 1); c.take(e.W.t5BinOff, lst::kEtaBinOffsets);
  c.take(e.W.t5BinRec, t5Cap ? t5Cap : 1); c.take(e.W.t5ListRec, t5Cap ? t5Cap : 1);
  c.take(e.W.pixBinned, pt5Cap + pt3Cap); c.take(e.W.pixBinOff, lst::kEtaBinOffsets);
  c.take(e.W.tcBinned, tcCap); c.take(e.W.tcBinOff, lst::kEtaBinOffsets); c.take(e.W.tcHitSet, lst::kTCHitSet);
  c.take(e.W.pt5Mask, lst::kPixMaskWords); c.take(e.W.pt5MaskBase, plsCap);
  c.take(e.W.pt3Mask, lst::kPixMaskWords); c.take(e.W.pt3MaskBase, plsCap);
  c.take(e.W.pt5Shard, 2 * lst::kPixShards); c.take(e.W.pt3Shard, 2 * lst::kPixShards);
  c.take(e.W.pixWords, plsCap); c.take(e.W.pixFull, plsCap);
  c.take(e.W.pixSurv, (size_t)lst::kSurvWords * lst::kPixSurvMax);
  c.take(e.W.pixParkIdx, (size_t)64 * lst::kPixMaskWords);
  c.take(e.W.t3ListOff, nL); c.take(e.W.t3List, t3Cap ? t3Cap : 1);
  c.take(e.W.t5Rec, (size_t)lst::kT5RecWords * (t5Cap ? t5Cap : 1));
  c.take(e.W.t3Rec, (size_t)lst::kT3RecWords * (t3Cap ? t3Cap : 1));
  c.take(e.W.pt5BucketOff, lst::kPT5HashBuckets + 1); c.take(e.W.pt5BucketEnt, 14 * (size_t)pt5Cap);
  c.take(e.FS.pre, lst::kScanCandMax); c.take(e.FS.sums, lst::kScanCandMax / 4096 + 1);
  c.take(e.W2.candOff, nL + 1); c.take(e.W2.resIdx, lst::kMDCandMax); c.take(e.W2.hitSlope, hitCap);
  c.take(e.W2.park, (size_t)lst::kMDParkWords * lst::kMDCandMax); c.take(e.W2.cand, lst::kMDCandMax);
  c.take(e.WL.connPre, (size_t)nL * (g.view.stride + 1)); c.take(e.WL.candOff, nL + 1);
  c.take(e.WL.mdSlope, mdCap);
  c.take(e.WL.resIdx, lst::kLSCandMax); c.take(e.WL.park, (size_t)lst::kLSParkWords * lst::kLSCandMax);
  c.take(e.WL.cand, lst::kLSCandMax);
  c.take(e.W3.candPre, segCap); c.take(e.W3.modOff, nL + 1);
  c.take(e.W3.resIdx, lst::kT3CandMax); c.take(e.W3.list, 2 * (size_t)lst::kT3CandMax);
  c.take(e.W3.results, (size_t)lst::kT3ResultWords * lst::kT3CandMax);
  c.take(e.W3.survPre, segCap); c.take(e.W3.survOff, nL + 1); c.take(e.W3.segByMD, segCap); c.take(e.W3.segTmp, segCap); c.take(e.W3.mdSegOff, mdCap); c.take(e.W3.mdSegEnd, mdCap);
  c.take(e.W5.candPre, t3Cap ? t3Cap : 1); c.take(e.W5.modOff, nL + 2);
  c.take(e.W5.resIdx, lst::kT5CandMax);
  c.take(e.W5.listA, 3 * (size_t)lst::kT5CandMax); c.take(e.W5.listB, 3 * (size_t)lst::kT5CandMax);
  c.take(e.W5.results, (size_t)lst::kT5ResultWords * lst::kT5ResultMax);
  c.take(e.W5.survPre, t3Cap ? t3Cap : 1); c.take(e.W5.survOff, nL + 2); c.take(e.W5.t3ByMD, t3Cap ? t3Cap : 1); c.take(e.W5.t3Tmp, t3Cap ? t3Cap : 1);
  c.take(e.W5.mdT3Off, mdCap); c.take(e.W5.mdT3End, mdCap);
  c.take(e.R.hitRanges, 2 * nMod); c.take(e.R.hitRangesLower, nMod); c.take(e.R.hitRangesUpper, nMod);
  c.take(e.R.hitRangesnLower, nMod); c.take(e.R.hitRangesnUpper, nMod);
  c.take(e.R.mdRanges, 2 * nMod); c.take(e.R.segmentRanges, 2 * nMod); c.take(e.R.tripletRanges, 2 * nMod);
  c.take(e.R.quintupletRanges, 2 * nMod);
  c.take(e.R.tripletModuleIndices, nL); c.take(e.R.tripletModuleOccupancy, nL);
  c.take(e.R.nMDs, nL + 1); c.take(e.R.totOccupancyMDs, nL + 1);
  c.take(e.R.nSegments, nL + 1); c.take(e.R.totOccupancySegments, nL + 1);
  c.take(e.R.nTriplets, nL); c.take(e.R.totOccupancyTriplets, nL);
  c.take(e.R.quintupletModuleIndices, nL); c.take(e.R.quintupletModuleOccupancy, nL);
  c.take(e.R.nQuintuplets, nL); c.take(e.R.totOccupancyQuintuplets, nL);
  c.take(e.R.indicesOfEligibleT5Modules, nL);
  c.take(e.R.scalars, lst::kNScalars);
  for (int k = 0; k < 4; k++) c.take(plsHitIdx[k], plsCap);
  c.take(plsDPhiChange, plsCap);
  arenaBytes = c.off;
  HIP_OK(hipMalloc(&arena, arenaBytes));
  c.bind(arena);
  arenaGen++;
  e.M = g.view;
  e.mdModuleIndices = g.mdIndices;
  e.mdModuleOccupancy = g.mdOccupancy;
  e.segModuleIndices = g.segIndices;
  e.segModuleOccupancy = g.segOccupancy;
  e.t3OccIfAny = g.t3OccIfAny;
  e.t5OccIfAny = g.t5OccIfAny;
  e.nT5SeedModules = g.nT5SeedModules;
  e.t5Cap = t5Cap ? t5Cap : 1;
  for (int k = 0; k < 3; k++) {
    e.pixIndex[k] = g.pixIndex[k];
    e.pixSizes[k] = g.pixSizes[k];
  }
  e.tcPlsTriplets = 0;
  // LST_PER_MODULE_BUILDERS=1 selects the per-module T3/T5 builders (same
  // results; they are the path taken when an event overflows the flattened one)
  {
    const char* pm = std::getenv("LST_PER_MODULE_BUILDERS");
    e.perModuleBuilders = (pm && pm[0] == '1') ? 1 : 0;
    // LST_LARGE_PATHS=1: the per-mini-doublet lists take the path of modules
    // too large for the LDS counting sort, and the eta binnings the path of
    // counts too large to keep in registers (same results)
    const char* lc = std::getenv("LST_LARGE_PATHS");
    e.largePaths = (lc && lc[0] == '1') ? 1 : 0;
  }
  e.t5Dnn = t5Mode == LST_T5_DNN ? 1 : 0;
  e.pixelModule = g.pixelModuleIndex;
  e.ptCut = es->ptCut;

  // by-name field table (host getters)
  fields.clear();
  auto add = [&](const std::string& name, auto* p, size_t n) {
    using T = std::remove_pointer_t<decltype(p)>;
    fields.push_back({name, FieldRef{(void*)p, dtype_of<T>(), sizeof(T), n}});
  };
#define REG(PFX, OBJ, CAP) [&] { (void)0; }
#define FIELD(TY, NM, MU) add(prefix + #NM, cur->NM, capacity * MU);
  {
    std::string prefix = "hits.";
    size_t capacity = hitCap;
    auto* cur = &e.H;
    LST_HIT_FIELDS(FIELD)
  }
  {
    std::string prefix = "mds.";
    size_t capacity = mdCap;
    auto* cur = &e.D;
    LST_MD_FIELDS(FIELD)
  }
  {
    std::string prefix = "segs.";
    size_t capacity = segCap;
    auto* cur = &e.S;
    LST_SEG_FIELDS(FIELD)
  }
  {
    std::string prefix = "segs.";
    size_t capacity = plsCap;
    auto* cur = &e.P;
    LST_PLS_FIELDS(FIELD)
  }
  {
    std::string prefix = "t3.";
    size_t capacity = t3Cap;
    auto* cur = &e.T;
    LST_T3_FIELDS(FIELD)
  }
  {
    std::string prefix = "t5.";
    size_t capacity = t5Cap;
    auto* cur = &e.Q;
    LST_T5_FIELDS(FIELD)
  }
  {
    std::string prefix = "pt5.";
    size_t capacity = pt5Cap;
    auto* cur = &e.P5;
    LST_PT5_FIELDS(FIELD)
  }
  {
    std::string prefix = "pt3.";
    size_t capacity = pt3Cap;
    auto* cur = &e.P3;
    LST_PT3_FIELDS(FIELD)
  }
  {
    std::string prefix = "tc.";
    size_t capacity = tcCap;
    auto* cur = &e.C;
    LST_TC_FIELDS(FIELD)
  }
#undef FIELD
#undef REG
  add("hits.hitRanges", e.R.hitRanges, 2 * nMod);
  add("hits.hitRangesLower", e.R.hitRangesLower, nMod);
  add("hits.hitRangesUpper", e.R.hitRangesUpper, nMod);
  add("hits.hitRangesnLower", e.R.hitRangesnLower, nMod);
  add("hits.hitRangesnUpper", e.R.hitRangesnUpper, nMod);
  add("ranges.mdRanges", e.R.mdRanges, 2 * nMod);
  add("ranges.segmentRanges", e.R.segmentRanges, 2 * nMod);
  add("ranges.tripletRanges", e.R.tripletRanges, 2 * nMod);
  add("ranges.quintupletRanges", e.R.quintupletRanges, 2 * nMod);
  add("ranges.miniDoubletModuleIndices", const_cast<int32_t*>(g.mdIndices), nL + 1);
  add("ranges.miniDoubletModuleOccupancy", const_cast<int32_t*>(g.mdOccupancy), nL + 1);
  add("ranges.segmentModuleIndices", const_cast<int32_t*>(g.segIndices), nL + 1);
  add("ranges.segmentModuleOccupancy", const_cast<int32_t*>(g.segOccupancy), nL + 1);
  add("ranges.tripletModuleIndices", e.R.tripletModuleIndices, nL);
  add("ranges.tripletModuleOccupancy", e.R.tripletModuleOccupancy, nL);
  add("mds.nMDs", e.R.nMDs, nL + 1);
  add("mds.totOccupancyMDs", e.R.totOccupancyMDs, nL + 1);
  add("segs.nSegments", e.R.nSegments, nL + 1);
  add("segs.totOccupancySegments", e.R.totOccupancySegments, nL + 1);
  add("t3.nTriplets", e.R.nTriplets, nL);
  add("t3.totOccupancyTriplets", e.R.totOccupancyTriplets, nL);
  add("ranges.quintupletModuleIndices", e.R.quintupletModuleIndices, nL);
  add("ranges.quintupletModuleOccupancy", e.R.quintupletModuleOccupancy, nL);
  add("ranges.indicesOfEligibleT5Modules", e.R.indicesOfEligibleT5Modules, nL);
  add("t5.nQuintuplets", e.R.nQuintuplets, nL);
  add("t5.totOccupancyQuintuplets", e.R.totOccupancyQuintuplets, nL);
  add("scalars", e.R.scalars, lst::kNScalars);
  add("out.hitIdxs", e.W.outHits, (size_t)tcCap * 14);
  add("out.len", e.W.outLen, tcCap);
  return LST_OK;
}

const FieldRef* lst_event::field(const char* name) {
  for (auto& f : fields)
    if (f.first == name) return &f.second;
  return nullptr;
}

lst_status lst_event::add_hits(const float* x, const float* y, const float* z, const uint32_t* det,
                               const uint32_t* idx, uint32_t n, bool onDevice) {
  lst_status s = allocate(n);
  if (s != LST_OK) return s;
  if (n) {
    CopyBatch cb(onDevice, stream);
    cb.add(v.H.xs, x, n * sizeof(float));
    cb.add(v.H.ys, y, n * sizeof(float));
    cb.add(v.H.zs, z, n * sizeof(float));
    cb.add(v.H.detid, det, n * sizeof(uint32_t));
    cb.add(v.H.idxs, idx, n * sizeof(uint32_t));
    HIP_OK(cb.flush());
  }
  v.nHits = n;
  v.nPLS = 0;
  lst::launch_set_scalar(v.R.scalars + lst::kEvNHits, n, stream);
  lst::launch_set_scalar(v.R.scalars + lst::kEvNPLS, 0, stream);
  lst::launch_hits(v, stream, tm());
  HIP_OK(hipGetLastError());
  stage = 0;
  return LST_OK;
}

lst_status lst_event::add_pixel_segments(const lst_pixel_segments& p, bool onDevice) {
  if (stage != 0) return fail(LST_E_ORDER, "addPixelSegmentToEvent must follow addHitToEvent");
  uint32_t n = p.n_pls;
  lst_status warn = LST_OK;
  if (n > lst::kNMaxPixelSegments) {  // SDL/Event.cc:237-244
    n = lst::kNMaxPixelSegments;
    warn = LST_W_TRUNCATED;
  }
  if (!onDevice) {  // pseudo-hit indices must address hits of this event
    const uint32_t* hi[4] = {p.hit_indices0, p.hit_indices1, p.hit_indices2, p.hit_indices3};
    for (int q = 0; q < 4; q++)
      for (uint32_t t = 0; t < n; t++)
        if (hi[q][t] >= v.nHits) return fail(LST_E_INPUT, "pixel segment hit index out of range");
  }
  if (n) {
    CopyBatch cb(onDevice, stream);
    cb.add(plsHitIdx[0], p.hit_indices0, n * 4);
    cb.add(plsHitIdx[1], p.hit_indices1, n * 4);
    cb.add(plsHitIdx[2], p.hit_indices2, n * 4);
    cb.add(plsHitIdx[3], p.hit_indices3, n * 4);
    cb.add(plsDPhiChange, p.delta_phi_change, n * 4);
    cb.add(v.P.ptIn, p.pt_in, n * 4);
    cb.add(v.P.ptErr, p.pt_err, n * 4);
    cb.add(v.P.px, p.px, n * 4);
    cb.add(v.P.py, p.py, n * 4);
    cb.add(v.P.pz, p.pz, n * 4);
    cb.add(v.P.eta, p.eta, n * 4);
    cb.add(v.P.etaErr, p.eta_err, n * 4);
    cb.add(v.P.phi, p.phi, n * 4);
    cb.add(v.P.charge, p.charge, n * 4);
    cb.add(v.P.seedIdx, p.seed_idx, n * 4);
    cb.add(v.P.superbin, p.superbin, n * 4);
    cb.add(v.P.pixelType, p.pixel_type, n);
    cb.add(v.P.isQuad, p.is_quad, n);
    HIP_OK(cb.flush());
  }
  v.nPLS = n;
  lst::launch_set_scalar(v.R.scalars + lst::kEvNPLS, n, stream);
  const uint32_t* hi[4] = {plsHitIdx[0], plsHitIdx[1], plsHitIdx[2], plsHitIdx[3]};
  lst::launch_pixel_segments(v, hi, plsDPhiChange, stream, tm());
  HIP_OK(hipGetLastError());
  stage = 1;
  return warn;
}

// prepareInput on the host, then the prepared arrays into one device allocation
static lst_status stage_input(const lst_input& in, lst_event::Staged& st) {
  lst::PreparedEvent p;
  lst::prepare_event(in, p);
  if (st.mem) {
    HIP_OK(hipFree(st.mem));
    st = lst_event::Staged();
  }
  const uint32_t nh = p.x.size();
  const uint32_t np = std::min<uint32_t>(p.ptIn.size(), lst::kNMaxPixelSegments);
  st.truncated = p.ptIn.size() > lst::kNMaxPixelSegments;  // SDL/Event.cc:237-244
  Carve c;
  c.take(st.x, nh); c.take(st.y, nh); c.take(st.z, nh); c.take(st.det, nh); c.take(st.idx, nh);
  for (int k = 0; k < 4; k++) c.take(st.h[k], np);
  c.take(st.seed, np); c.take(st.dphi, np); c.take(st.pt, np); c.take(st.pterr, np); c.take(st.px, np);
  c.take(st.py, np); c.take(st.pz, np); c.take(st.eta, np); c.take(st.etaerr, np); c.take(st.phi, np);
  c.take(st.charge, np); c.take(st.superbin, np); c.take(st.ptype, np); c.take(st.quad, np);
  HIP_OK(hipMalloc(&st.mem, c.off ? c.off : 256));
  st.bytes = c.off;
  c.bind(st.mem);
  st.nHits = nh;
  st.nPLS = np;
  auto up = [&](void* d, const void* h, size_t b) { return b ? hipMemcpy(d, h, b, hipMemcpyHostToDevice) : hipSuccess; };
  HIP_OK(up(st.x, p.x.data(), nh * 4)); HIP_OK(up(st.y, p.y.data(), nh * 4)); HIP_OK(up(st.z, p.z.data(), nh * 4));
  HIP_OK(up(st.det, p.detId.data(), nh * 4)); HIP_OK(up(st.idx, p.idx.data(), nh * 4));
  HIP_OK(up(st.h[0], p.h0.data(), np * 4)); HIP_OK(up(st.h[1], p.h1.data(), np * 4));
  HIP_OK(up(st.h[2], p.h2.data(), np * 4)); HIP_OK(up(st.h[3], p.h3.data(), np * 4));
  HIP_OK(up(st.seed, p.seedIdx.data(), np * 4)); HIP_OK(up(st.dphi, p.dPhiChange.data(), np * 4));
  HIP_OK(up(st.pt, p.ptIn.data(), np * 4)); HIP_OK(up(st.pterr, p.ptErr.data(), np * 4));
  HIP_OK(up(st.px, p.px.data(), np * 4)); HIP_OK(up(st.py, p.py.data(), np * 4)); HIP_OK(up(st.pz, p.pz.data(), np * 4));
  HIP_OK(up(st.eta, p.eta.data(), np * 4)); HIP_OK(up(st.etaerr, p.etaErr.data(), np * 4));
  HIP_OK(up(st.phi, p.phi.data(), np * 4)); HIP_OK(up(st.charge, p.charge.data(), np * 4));
  HIP_OK(up(st.superbin, p.superbin.data(), np * 4)); HIP_OK(up(st.ptype, p.pixelType.data(), np));
  HIP_OK(up(st.quad, p.isQuad.data(), np));
  return LST_OK;
}

// the chain on inputs already in device memory, enqueued on the event's stream
static lst_status run_staged(lst_event* ev, const lst_event::Staged& st, int last_stage) {
  lst_status s = ev->add_hits(st.x, st.y, st.z, st.det, st.idx, st.nHits, true);
  if (s != LST_OK || last_stage < 1) return s;
  lst_pixel_segments p;
  p.hit_indices0 = st.h[0]; p.hit_indices1 = st.h[1]; p.hit_indices2 = st.h[2]; p.hit_indices3 = st.h[3];
  p.delta_phi_change = st.dphi; p.pt_in = st.pt; p.pt_err = st.pterr; p.px = st.px; p.py = st.py; p.pz = st.pz;
  p.eta = st.eta; p.eta_err = st.etaerr; p.phi = st.phi; p.charge = st.charge; p.seed_idx = st.seed;
  p.superbin = st.superbin; p.pixel_type = st.ptype; p.is_quad = st.quad; p.n_pls = st.nPLS;
  const lst_status warn = st.truncated ? (lst_status)LST_W_TRUNCATED : LST_OK;
  if ((s = ev->add_pixel_segments(p, true)) != LST_OK || last_stage < 2) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_mini_doublets(ev)) != LST_OK || last_stage < 3) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_segments(ev)) != LST_OK || last_stage < 4) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_triplets(ev)) != LST_OK || last_stage < 5) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_quintuplets(ev)) != LST_OK || last_stage < 6) return s == LST_OK ? warn : s;
  if ((s = lst_event_pixel_line_segment_cleaning(ev)) != LST_OK || last_stage < 7) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_pixel_quintuplets(ev)) != LST_OK || last_stage < 8) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_pixel_triplets(ev)) != LST_OK || last_stage < 9) return s == LST_OK ? warn : s;
  s = lst_event_create_track_candidates(ev);
  return s == LST_OK ? warn : s;
}

// the whole chain on the inputs ev->descDev describes: kernel launches only,
// no host reads, so that it can be captured once and replayed per event
static void enqueue_described_chain(lst_event* ev) {
  const lst::EventView& v = ev->v;
  const hipStream_t s = ev->stream;
  lst::launch_stage_in(v, ev->descDev, ev->plsHitIdx, ev->plsDPhiChange, s);
  lst::launch_hits(v, s);
  const uint32_t* hi[4] = {ev->plsHitIdx[0], ev->plsHitIdx[1], ev->plsHitIdx[2], ev->plsHitIdx[3]};
  lst::launch_pixel_segments(v, hi, ev->plsDPhiChange, s);
  lst::launch_mini_doublets(v, s);
  lst::launch_segments(v, s);
  lst::launch_triplets(v, s);
  lst::launch_quintuplets(v, s);
  lst::launch_pixel_segment_cleaning(v, s);
  lst::launch_pixel_quintuplets(v, s);
  lst::launch_pixel_triplets(v, s);
  lst::launch_track_candidates(v, s);
}

// run a staged event through the context's captured chain (captured on first use)
static lst_status run_staged_graph(lst_event* ev, const lst_event::Staged& st) {
  lst_status s = ev->allocate(st.nHits);
  if (s != LST_OK) return s;
  if (!ev->descDev) {
    HIP_OK(hipMalloc(&ev->descDev, sizeof(lst::InputDesc)));
    HIP_OK(hipHostMalloc((void**)&ev->descHost, sizeof(lst::InputDesc), hipHostMallocDefault));
  }
  if (!ev->graph || ev->graphGen != ev->arenaGen) {
    if (ev->graph) HIP_OK(hipGraphExecDestroy(ev->graph));
    ev->graph = nullptr;
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(ev->stream, hipStreamCaptureModeThreadLocal));
    enqueue_described_chain(ev);
    HIP_OK(hipStreamEndCapture(ev->stream, &g));
    const hipError_t e = hipGraphInstantiate(&ev->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(LST_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    ev->graphGen = ev->arenaGen;
  }
  lst::InputDesc& d = *ev->descHost;  // the stream is idle here (the batch waited for this context)
  d.x = st.x; d.y = st.y; d.z = st.z; d.det = st.det; d.idx = st.idx; d.nHits = st.nHits;
  for (int k = 0; k < 4; k++) d.hitIdx[k] = st.h[k];
  d.dPhiChange = st.dphi; d.ptIn = st.pt; d.ptErr = st.pterr; d.px = st.px; d.py = st.py; d.pz = st.pz;
  d.eta = st.eta; d.etaErr = st.etaerr; d.phi = st.phi; d.charge = st.charge; d.superbin = st.superbin;
  d.seedIdx = st.seed; d.pixelType = st.ptype; d.isQuad = st.quad; d.nPLS = st.nPLS;
  HIP_OK(hipMemcpyAsync(ev->descDev, ev->descHost, sizeof(lst::InputDesc), hipMemcpyHostToDevice, ev->stream));
  HIP_OK(hipGraphLaunch(ev->graph, ev->stream));
  ev->v.nHits = st.nHits;
  ev->v.nPLS = st.nPLS;
  ev->stage = 9;
  return st.truncated ? (lst_status)LST_W_TRUNCATED : LST_OK;
}

extern "C" {

const char* lst_last_error(void) { return g_err.c_str(); }

lst_status lst_es_create(const char* geom_dir, float pt_cut, uint32_t stride, lst_es** out) {
  *out = nullptr;
  auto es = std::make_unique<lst_es>();
  std::string err;
  if (!lst::load_host_geometry(geom_dir ? geom_dir : "", stride, es->host, err))
    return fail(err.rfind("cannot read", 0) == 0 ? LST_E_IO : LST_E_FORMAT, err);
  if (es->host.stride > 159) return fail(LST_E_FORMAT, "module fan-out above 159 is not supported");
  for (int k = 0; k < 3; k++)
    for (uint32_t sz : es->host.pixSizes[k])
      if (sz > 64) return fail(LST_E_FORMAT, "a pixel superbin connects to more than 64 modules");
  for (uint32_t m : es->host.connectedPixels)
    if (m >= es->host.nLowerModules) return fail(LST_E_FORMAT, "pixel map names a module that is not a lower module");
  if (!lst::upload_geometry(es->host, es->dev, err)) return fail(LST_E_DEVICE, err);
  es->ptCut = pt_cut;
  *out = es.release();
  return LST_OK;
}

void lst_es_destroy(lst_es* es) {
  if (!es) return;
  lst::free_geometry(es->dev);
  delete es;
}

uint32_t lst_es_scalar(const lst_es* es, const char* name) {
  if (!es || !name) return 0;
  const auto& h = es->host;
  if (!std::strcmp(name, "nModules")) return h.nModules;
  if (!std::strcmp(name, "nLowerModules")) return h.nLowerModules;
  if (!std::strcmp(name, "nPixels")) return h.nPixels;
  if (!std::strcmp(name, "pixelModuleIndex")) return h.pixelModuleIndex;
  if (!std::strcmp(name, "stride")) return h.stride;
  return 0;
}

lst_status lst_event_create(lst_es* es, void* hip_stream, int t5_mode, lst_event** out) {
  *out = nullptr;
  if (!es) return fail(LST_E_INPUT, "null event setup");
  if (t5_mode != LST_T5_PLAIN_CHI2 && t5_mode != LST_T5_DNN) return fail(LST_E_INPUT, "unknown T5 selection mode");
  auto ev = std::make_unique<lst_event>();
  ev->es = es;
  ev->t5Mode = t5_mode;
  if (hip_stream) {
    ev->stream = static_cast<hipStream_t>(hip_stream);
  } else {
    HIP_OK(hipStreamCreateWithFlags(&ev->stream, hipStreamNonBlocking));
    ev->ownStream = true;
  }
  lst_status s = ev->allocate(0);
  if (s != LST_OK) return s;
  *out = ev.release();
  return LST_OK;
}

void lst_event_destroy(lst_event* ev) {
  if (!ev) return;
  if (ev->stream) hipStreamSynchronize(ev->stream);
  if (ev->arena) hipFree(ev->arena);
  for (auto& st : ev->staged)
    if (st.mem) hipFree(st.mem);
  for (int k = 0; k < lst::kNTimers; k++) {
    if (ev->timers.start[k]) hipEventDestroy(ev->timers.start[k]);
    if (ev->timers.stop[k]) hipEventDestroy(ev->timers.stop[k]);
  }
  if (ev->dnnRecordMem) (void)hipFree(ev->dnnRecordMem);
  if (ev->graph) (void)hipGraphExecDestroy(ev->graph);
  if (ev->descDev) (void)hipFree(ev->descDev);
  if (ev->descHost) (void)hipHostFree(ev->descHost);
  if (ev->ownStream) hipStreamDestroy(ev->stream);
  delete ev;
}

lst_status lst_event_add_hits(lst_event* ev, const float* x, const float* y, const float* z, const uint32_t* det_id,
                              const uint32_t* idx, uint32_t n, int on_device) {
  if (!ev) return fail(LST_E_INPUT, "null event");
  return ev->add_hits(x, y, z, det_id, idx, n, on_device != 0);
}

lst_status lst_event_add_pixel_segments(lst_event* ev, const lst_pixel_segments* p, int on_device) {
  if (!ev || !p) return fail(LST_E_INPUT, "null argument");
  return ev->add_pixel_segments(*p, on_device != 0);
}

lst_status lst_event_create_mini_doublets(lst_event* ev) {
  if (!ev || ev->stage != 1) return fail(LST_E_ORDER, "createMiniDoublets needs hits and pixel segments");
  lst::launch_mini_doublets(ev->v, ev->stream, ev->tm());
  HIP_OK(hipGetLastError());
  ev->stage = 2;
  return LST_OK;
}

lst_status lst_event_create_segments(lst_event* ev) {
  if (!ev || ev->stage != 2) return fail(LST_E_ORDER, "createSegmentsWithModuleMap needs mini-doublets");
  lst::launch_segments(ev->v, ev->stream, ev->tm());
  HIP_OK(hipGetLastError());
  ev->stage = 3;
  return LST_OK;
}

lst_status lst_event_create_triplets(lst_event* ev) {
  if (!ev || ev->stage != 3) return fail(LST_E_ORDER, "createTriplets needs segments");
  lst::launch_triplets(ev->v, ev->stream, ev->tm());
  HIP_OK(hipGetLastError());
  ev->stage = 4;
  return LST_OK;
}

lst_status lst_event_create_quintuplets(lst_event* ev) {
  if (!ev || ev->stage != 4) return fail(LST_E_ORDER, "createQuintuplets needs triplets");
  lst::launch_quintuplets(ev->v, ev->stream, ev->tm());
  HIP_OK(hipGetLastError());
  ev->stage = 5;
  return LST_OK;
}

lst_status lst_event_pixel_line_segment_cleaning(lst_event* ev) {
  if (!ev || ev->stage != 5) return fail(LST_E_ORDER, "pixelLineSegmentCleaning needs quintuplets");
  lst::launch_pixel_segment_cleaning(ev->v, ev->stream, ev->tm());
  HIP_OK(hipGetLastError());
  ev->stage = 6;
  return LST_OK;
}

lst_status lst_event_create_pixel_quintuplets(lst_event* ev) {
  if (!ev || ev->stage != 6) return fail(LST_E_ORDER, "createPixelQuintuplets needs pixelLineSegmentCleaning");
  lst::launch_pixel_quintuplets(ev->v, ev->stream, ev->tm());
  HIP_OK(hipGetLastError());
  ev->stage = 7;
  return LST_OK;
}

lst_status lst_event_create_pixel_triplets(lst_event* ev) {
  if (!ev || ev->stage != 7) return fail(LST_E_ORDER, "createPixelTriplets needs pixel quintuplets");
  lst::launch_pixel_triplets(ev->v, ev->stream, ev->tm());
  HIP_OK(hipGetLastError());
  ev->stage = 8;
  return LST_OK;
}

lst_status lst_event_create_track_candidates(lst_event* ev) {
  if (!ev || ev->stage != 8) return fail(LST_E_ORDER, "createTrackCandidates needs pixel triplets");
  lst::launch_track_candidates(ev->v, ev->stream, ev->tm());
  HIP_OK(hipGetLastError());
  ev->stage = 9;
  return LST_OK;
}

lst_status lst_event_get_output(lst_event* ev, uint32_t* n_tc, uint32_t* hit_idxs, uint32_t* len, int32_t* seed_idx,
                                int16_t* type, uint32_t capacity) {
  if (!ev || !n_tc) return fail(LST_E_INPUT, "null argument");
  if (ev->stage != 9) return fail(LST_E_ORDER, "getOutput needs track candidates");
  uint32_t scal[lst::kNScalars];
  HIP_OK(hipMemcpyAsync(scal, ev->v.R.scalars, sizeof(scal), hipMemcpyDeviceToHost, ev->stream));
  HIP_OK(hipStreamSynchronize(ev->stream));
  const uint32_t n = scal[lst::kNTrackCandidates];
  *n_tc = n;
  if (!hit_idxs && !len && !seed_idx && !type) return LST_OK;
  if (capacity < n) return fail(LST_E_CAPACITY, "output arrays hold fewer track candidates than were built");
  if (n) {
    if (hit_idxs) HIP_OK(hipMemcpyAsync(hit_idxs, ev->v.W.outHits, (size_t)n * 14 * 4, hipMemcpyDeviceToHost, ev->stream));
    if (len) HIP_OK(hipMemcpyAsync(len, ev->v.W.outLen, (size_t)n * 4, hipMemcpyDeviceToHost, ev->stream));
    if (seed_idx) HIP_OK(hipMemcpyAsync(seed_idx, ev->v.C.pixelSeedIndex, (size_t)n * 4, hipMemcpyDeviceToHost, ev->stream));
    if (type) HIP_OK(hipMemcpyAsync(type, ev->v.C.trackCandidateType, (size_t)n * 2, hipMemcpyDeviceToHost, ev->stream));
    HIP_OK(hipStreamSynchronize(ev->stream));
  }
  return LST_OK;
}

lst_status lst_event_reset(lst_event* ev) {
  if (!ev) return fail(LST_E_INPUT, "null event");
  ev->stage = -1;
  ev->v.nHits = 0;
  ev->v.nPLS = 0;
  return LST_OK;
}

lst_status lst_event_synchronize(lst_event* ev) {
  if (!ev) return fail(LST_E_INPUT, "null event");
  HIP_OK(hipStreamSynchronize(ev->stream));
  return LST_OK;
}

uint32_t lst_event_count(lst_event* ev, const char* what) {
  if (!ev || !what) return 0;
  if (!std::strcmp(what, "hits")) return ev->v.nHits;
  if (!std::strcmp(what, "pixelSegments")) return ev->v.nPLS;
  if (!std::strcmp(what, "arenaMiB")) return (uint32_t)(ev->arenaBytes >> 20);
  const uint32_t nL = ev->v.M.nLowerModules;
  const char* key = nullptr;
  uint32_t n = nL;
  if (!std::strcmp(what, "miniDoublets")) key = "mds.nMDs";
  else if (!std::strcmp(what, "segments")) key = "segs.nSegments";
  else if (!std::strcmp(what, "triplets")) key = "t3.nTriplets";
  else if (!std::strcmp(what, "quintuplets")) key = "t5.nQuintuplets";
  else {
    static const struct { const char* name; int idx; } sc[] = {
        {"pixelQuintuplets", lst::kNPixelQuintuplets}, {"pixelTriplets", lst::kNPixelTriplets},
        {"trackCandidates", lst::kNTrackCandidates}, {"pT3TrackCandidates", lst::kNTrackCandidatespT3},
        {"pT5TrackCandidates", lst::kNTrackCandidatespT5}, {"pLSTrackCandidates", lst::kNTrackCandidatespLS},
        {"T5TrackCandidates", lst::kNTrackCandidatesT5}, {"tripletCandidates", lst::kT3CandTotal},
        {"tripletSharedMDCandidates", lst::kT3List}, {"quintupletCandidates", lst::kT5CandTotal},
        {"miniDoubletCandidates", lst::kMDCandTotal}, {"segmentCandidates", lst::kLSCandTotal},
        {"pixelQuintupletSurvivors", lst::kPT5Survivors}, {"pixelTripletSurvivors", lst::kPT3Survivors}};
    for (const auto& e : sc)
      if (!std::strcmp(what, e.name)) {
        uint32_t val = 0;
        if (hipMemcpyAsync(&val, ev->v.R.scalars + e.idx, 4, hipMemcpyDeviceToHost, ev->stream) != hipSuccess) return 0;
        if (hipStreamSynchronize(ev->stream) != hipSuccess) return 0;
        return val;
      }
  }
  if (!key) return 0;
  const FieldRef* f = ev->field(key);
  std::vector<uint32_t> h(n);
  if (hipMemcpyAsync(h.data(), f->ptr, n * 4, hipMemcpyDeviceToHost, ev->stream) != hipSuccess) return 0;
  if (hipStreamSynchronize(ev->stream) != hipSuccess) return 0;
  uint64_t tot = 0;
  for (auto x : h) tot += x;
  return (uint32_t)tot;
}

lst_status lst_event_get_field(lst_event* ev, const char* name, void* dst, size_t dst_bytes, int* dtype,
                               size_t* count) {
  if (!ev || !name) return fail(LST_E_INPUT, "null argument");
  const FieldRef* f = ev->field(name);
  if (!f) return fail(LST_E_INPUT, std::string("unknown field ") + name);
  if (dtype) *dtype = f->dtype;
  if (count) *count = f->count;
  if (!dst) return LST_OK;
  const size_t bytes = f->count * f->elem;
  if (dst_bytes < bytes) return fail(LST_E_CAPACITY, "destination too small");
  HIP_OK(hipMemcpyAsync(dst, f->ptr, bytes, hipMemcpyDeviceToHost, ev->stream));
  HIP_OK(hipStreamSynchronize(ev->stream));
  return LST_OK;
}

lst_status lst_run(lst_event* ev, const lst_input* in, int last_stage) {
  if (!ev || !in) return fail(LST_E_INPUT, "null argument");
  lst::prepare_event(*in, ev->prepared);
  const auto& p = ev->prepared;
  lst_status s = ev->add_hits(p.x.data(), p.y.data(), p.z.data(), p.detId.data(), p.idx.data(),
                              (uint32_t)p.x.size(), false);
  if (s != LST_OK || last_stage < 1) return s;
  const lst_pixel_segments pv = p.view();
  lst_status warn = ev->add_pixel_segments(pv, false);
  if ((warn & 0xff) != LST_OK || last_stage < 2) return warn;
  if ((s = lst_event_create_mini_doublets(ev)) != LST_OK || last_stage < 3) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_segments(ev)) != LST_OK || last_stage < 4) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_triplets(ev)) != LST_OK || last_stage < 5) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_quintuplets(ev)) != LST_OK || last_stage < 6) return s == LST_OK ? warn : s;
  if ((s = lst_event_pixel_line_segment_cleaning(ev)) != LST_OK || last_stage < 7) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_pixel_quintuplets(ev)) != LST_OK || last_stage < 8) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_pixel_triplets(ev)) != LST_OK || last_stage < 9) return s == LST_OK ? warn : s;
  if ((s = lst_event_create_track_candidates(ev)) != LST_OK) return s;
  return warn;
}

lst_status lst_event_stage_input(lst_event* ev, const lst_input* in, uint32_t slot) {
  if (!ev || !in) return fail(LST_E_INPUT, "null argument");
  if (ev->staged.size() <= slot) ev->staged.resize(slot + 1);
  lst_event::Staged& st = ev->staged[slot];
  if (st.mem) HIP_OK(hipStreamSynchronize(ev->stream));
  lst_status s = stage_input(*in, st);
  if (s != LST_OK) return s;
  return ev->allocate(st.nHits);
}

lst_status lst_event_run_staged(lst_event* ev, uint32_t slot, int last_stage) {
  if (!ev || slot >= ev->staged.size() || !ev->staged[slot].mem) return fail(LST_E_INPUT, "input slot not staged");
  return run_staged(ev, ev->staged[slot], last_stage);
}

lst_status lst_event_set_timing(lst_event* ev, int on) {
  if (!ev) return fail(LST_E_INPUT, "null event");
  if (on && !ev->timers.start[0])
    for (int k = 0; k < lst::kNTimers; k++) {
      HIP_OK(hipEventCreate(&ev->timers.start[k]));
      HIP_OK(hipEventCreate(&ev->timers.stop[k]));
    }
  ev->timing = on != 0;
  ev->recorded = 0;
  return LST_OK;
}

lst_status lst_event_kernel_times(lst_event* ev, float* ms, int n) {
  if (!ev || !ms) return fail(LST_E_INPUT, "null argument");
  if (!ev->timers.start[0]) return fail(LST_E_ORDER, "timing was never enabled");
  HIP_OK(hipStreamSynchronize(ev->stream));
  for (int k = 0; k < n && k < lst::kNTimers; k++) {
    ms[k] = 0.f;
    // only timers recorded by a stage run hold a valid interval; querying the
    // others would leave an error behind for the next hipGetLastError
    if ((ev->recorded >> k) & 1u) HIP_OK(hipEventElapsedTime(&ms[k], ev->timers.start[k], ev->timers.stop[k]));
  }
  return LST_OK;
}

void* lst_event_stream(lst_event* ev) { return ev ? (void*)ev->stream : nullptr; }

lst_status lst_event_record_t5dnn(lst_event* ev, uint32_t capacity) {
  if (!ev) return fail(LST_E_INPUT, "null event");
  if (ev->t5Mode != LST_T5_DNN) return fail(LST_E_INPUT, "the T5 DNN record needs an event in LST_T5_DNN mode");
  HIP_OK(hipStreamSynchronize(ev->stream));
  if (ev->dnnRecordMem) HIP_OK(hipFree(ev->dnnRecordMem));
  ev->dnnRecordMem = nullptr;
  ev->v.dnnDump = lst::T5DnnDump{};
  if (capacity == 0) return LST_OK;
  if (capacity > lst::kT5ResultMax) capacity = lst::kT5ResultMax;
  Carve c;
  lst::T5DnnDump& d = ev->v.dnnDump;
  c.take(d.features, 38 * (size_t)capacity); c.take(d.score, capacity);
  c.take(d.t3Pair, 2 * (size_t)capacity); c.take(d.hitIdxs, 10 * (size_t)capacity);
  HIP_OK(hipMalloc(&ev->dnnRecordMem, c.off));
  c.bind(static_cast<char*>(ev->dnnRecordMem));
  d.capacity = capacity;
  return LST_OK;
}

lst_status lst_event_get_t5dnn_record(lst_event* ev, uint32_t* n, float* features, float* scores, uint32_t* t3_pairs,
                                      uint32_t* hit_idxs, uint32_t capacity) {
  if (!ev || !n) return fail(LST_E_INPUT, "null argument");
  const lst::T5DnnDump& d = ev->v.dnnDump;
  if (!d.features) return fail(LST_E_ORDER, "no T5 DNN record was requested");
  if (ev->stage < 5) return fail(LST_E_ORDER, "the T5 DNN record needs createQuintuplets");
  uint32_t scal[lst::kNScalars];
  HIP_OK(hipMemcpyAsync(scal, ev->v.R.scalars, sizeof(scal), hipMemcpyDeviceToHost, ev->stream));
  HIP_OK(hipStreamSynchronize(ev->stream));
  if (ev->v.perModuleBuilders || scal[lst::kT5CandTotal] > lst::kT5CandMax || scal[lst::kT5Results] > lst::kT5ResultMax)
    return fail(LST_E_CAPACITY, "the event took the per-module quintuplet builder, which records no DNN rows");
  const uint32_t rows = scal[lst::kT5Results];
  const uint32_t kept = std::min(rows, d.capacity);
  *n = kept;
  const uint32_t m = std::min(kept, capacity);
  if (m) {
    if (features) HIP_OK(hipMemcpyAsync(features, d.features, (size_t)m * 38 * 4, hipMemcpyDeviceToHost, ev->stream));
    if (scores) HIP_OK(hipMemcpyAsync(scores, d.score, (size_t)m * 4, hipMemcpyDeviceToHost, ev->stream));
    if (t3_pairs) HIP_OK(hipMemcpyAsync(t3_pairs, d.t3Pair, (size_t)m * 8, hipMemcpyDeviceToHost, ev->stream));
    if (hit_idxs) HIP_OK(hipMemcpyAsync(hit_idxs, d.hitIdxs, (size_t)m * 40, hipMemcpyDeviceToHost, ev->stream));
    HIP_OK(hipStreamSynchronize(ev->stream));
  }
  return rows > d.capacity || kept > capacity ? (lst_status)LST_W_TRUNCATED : LST_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// Several event contexts (arena + stream each) run a batch of events round
// robin; the host waits on a context only when it hands it the next event, so
// the other contexts keep the GPU busy meanwhile.  Outputs travel through a
// pinned buffer per context.
struct lst_batch {
  lst_es* es = nullptr;
  std::vector<lst_event*> ctx;
  struct Out {
    char* pinned = nullptr;  // n_tc, then hits [cap*14], len [cap], seed [cap], type [cap]
    uint32_t cap = 0;
    int64_t pending = -1;    // batch position whose output is in flight
    lst_output* dst = nullptr;
  };
  std::vector<Out> outs;
  std::vector<lst_event::Staged> staged;
  bool graphs = true;  // staged events run through each context's captured chain
};

namespace {
constexpr uint32_t kTCCap = lst::kNMaxPixelTrackCandidates + lst::kNMaxNonPixelTrackCandidates;

size_t out_bytes(uint32_t cap) { return 4 * (lst::kOutHeaderWords + (size_t)cap * lst::kOutRecWords); }

// enqueue the output copies of context c's event into its pinned buffer
lst_status enqueue_output(lst_batch* b, uint32_t c, lst_output* dst) {
  lst_batch::Out& o = b->outs[c];
  lst_event* ev = b->ctx[c];
  const uint32_t cap = std::min(dst->capacity, kTCCap);
  if (!o.pinned || o.cap < cap) {
    if (o.pinned) HIP_OK(hipHostFree(o.pinned));
    o.pinned = nullptr;
    HIP_OK(hipHostMalloc((void**)&o.pinned, out_bytes(cap), hipHostMallocDefault));
    o.cap = cap;
  }
  // the packed block k_tc_output wrote: count, then one record per track candidate
  HIP_OK(hipMemcpyAsync(o.pinned, ev->v.W.outPacked, out_bytes(cap), hipMemcpyDeviceToHost, ev->stream));
  o.dst = dst;
  return LST_OK;
}

// wait for context c and hand its finished event's output to the caller
lst_status finish_context(lst_batch* b, uint32_t c, lst_status& warn) {
  lst_batch::Out& o = b->outs[c];
  lst_event* ev = b->ctx[c];
  HIP_OK(hipStreamSynchronize(ev->stream));
  if (o.pending >= 0 && o.dst) {
    lst_output* d = o.dst;
    const uint32_t cap = std::min(d->capacity, kTCCap);
    uint32_t n;
    std::memcpy(&n, o.pinned, 4);
    d->n_tc = n;
    const uint32_t m = std::min(n, cap);
    if (n > d->capacity) warn = LST_W_TRUNCATED;
    const uint32_t* rec = reinterpret_cast<const uint32_t*>(o.pinned) + lst::kOutHeaderWords;
    for (uint32_t t = 0; t < m; t++, rec += lst::kOutRecWords) {
      if (d->hit_idxs) std::memcpy(d->hit_idxs + 14 * (size_t)t, rec, 56);
      if (d->len) d->len[t] = rec[14];
      if (d->seed_idx) d->seed_idx[t] = (int32_t)rec[15];
      if (d->type) d->type[t] = (int16_t)(int32_t)rec[16];
    }
  }
  o.pending = -1;
  o.dst = nullptr;
  return LST_OK;
}

// run events [0, n) through the contexts; `launch(ctx, i)` enqueues event i
template <class F>
lst_status run_round_robin(lst_batch* b, uint32_t n, lst_output* out, F launch) {
  const uint32_t K = (uint32_t)b->ctx.size();
  lst_status warn = LST_OK, s;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t c = i % K;
    if (b->outs[c].pending >= 0 && (s = finish_context(b, c, warn)) != LST_OK) return s;
    s = launch(b->ctx[c], i);
    if ((s & 0xff) != LST_OK) return s;
    if (s != LST_OK) warn = s;
    if (out && (s = enqueue_output(b, c, &out[i])) != LST_OK) return s;
    b->outs[c].pending = i;
  }
  for (uint32_t c = 0; c < K; c++)
    if ((s = finish_context(b, c, warn)) != LST_OK) return s;
  return warn;
}
}  // namespace

extern "C" {

lst_status lst_batch_create(lst_es* es, uint32_t n_contexts, int t5_mode, lst_batch** out) {
  if (!out) return fail(LST_E_INPUT, "null argument");
  *out = nullptr;
  if (!es) return fail(LST_E_INPUT, "null event setup");
  if (n_contexts == 0 || n_contexts > 64) return fail(LST_E_INPUT, "a batch needs 1..64 event contexts");
  auto b = std::make_unique<lst_batch>();
  b->es = es;
  for (uint32_t c = 0; c < n_contexts; c++) {
    lst_event* ev = nullptr;
    lst_status s = lst_event_create(es, nullptr, t5_mode, &ev);
    if (s != LST_OK) {
      lst_batch_destroy(b.release());
      return s;
    }
    b->ctx.push_back(ev);
  }
  b->outs.resize(n_contexts);
  const char* g = std::getenv("LST_BATCH_GRAPHS");  // "0": launch every kernel instead
  b->graphs = !(g && g[0] == '0');
  *out = b.release();
  return LST_OK;
}

void lst_batch_destroy(lst_batch* b) {
  if (!b) return;
  for (lst_event* ev : b->ctx) lst_event_destroy(ev);
  for (auto& o : b->outs)
    if (o.pinned) (void)hipHostFree(o.pinned);
  for (auto& st : b->staged)
    if (st.mem) (void)hipFree(st.mem);
  delete b;
}

lst_status lst_run_batch(lst_batch* b, const lst_input* events, uint32_t n_ev, lst_output* out) {
  if (!b || (n_ev && !events)) return fail(LST_E_INPUT, "null argument");
  return run_round_robin(b, n_ev, out, [&](lst_event* ev, uint32_t i) { return lst_run(ev, &events[i], 9); });
}

lst_status lst_batch_stage(lst_batch* b, const lst_input* events, uint32_t n_ev) {
  if (!b || (n_ev && !events)) return fail(LST_E_INPUT, "null argument");
  for (lst_event* ev : b->ctx) HIP_OK(hipStreamSynchronize(ev->stream));
  if (b->staged.size() < n_ev) b->staged.resize(n_ev);
  uint32_t maxHits = 0;
  for (uint32_t i = 0; i < n_ev; i++) {
    lst_status s = stage_input(events[i], b->staged[i]);
    if (s != LST_OK) return s;
    maxHits = std::max(maxHits, b->staged[i].nHits);
  }
  for (lst_event* ev : b->ctx) {  // size every arena for the largest staged event up front
    lst_status s = ev->allocate(maxHits);
    if (s != LST_OK) return s;
  }
  return LST_OK;
}

lst_status lst_batch_run_staged(lst_batch* b, uint32_t first, uint32_t n_ev, lst_output* out) {
  if (!b) return fail(LST_E_INPUT, "null batch");
  const size_t ns = b->staged.size();
  if (ns == 0 || first >= ns) return fail(LST_E_INPUT, "batch slots not staged");
  for (const auto& st : b->staged)
    if (!st.mem) return fail(LST_E_INPUT, "batch slot not staged");
  return run_round_robin(b, n_ev, out, [&](lst_event* ev, uint32_t i) {
    const lst_event::Staged& st = b->staged[(first + i) % ns];
    return b->graphs ? run_staged_graph(ev, st) : run_staged(ev, st, 9);
  });
}

uint32_t lst_batch_contexts(const lst_batch* b) { return b ? (uint32_t)b->ctx.size() : 0; }

}  // extern "C"
