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
  size_t stride = 0;  // bytes between records when the field is in records (0: a plain array)
  size_t mult = 1;    // elements of the field per record (4 bytes apart)
};

// dst[i] = element i of a record field: word i % mult of the field in record i / mult (records `stride`
// bytes apart), i < count; elem is 1, 2, 4 or 8 bytes
__global__ void k_gather_field(const char* __restrict__ src, uint32_t stride, uint32_t elem, uint32_t mult,
                               size_t count, char* __restrict__ dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const char* p = src + (i / mult) * stride + (i % mult) * 4;
    if (elem == 8) reinterpret_cast<uint2*>(dst)[i] = *reinterpret_cast<const uint2*>(p);
    else if (elem == 4) reinterpret_cast<uint32_t*>(dst)[i] = *reinterpret_cast<const uint32_t*>(p);
    else if (elem == 2) reinterpret_cast<uint16_t*>(dst)[i] = *reinterpret_cast<const uint16_t*>(p);
    else reinterpret_cast<uint8_t*>(dst)[i] = *reinterpret_cast<const uint8_t*>(p);
  }
}

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

// An event prepared on the host (prepareInput, SDL/LST.cc:175-365) and kept
// in one pinned block, so that its inputs cross to the device in one copy per
// event: x, y, z, detId, idx per hit, then the pLS arrays.
struct PreparedInput {
  char* host = nullptr;  // pinned
  size_t bytes = 0;
  uint32_t nHits = 0, nPLS = 0;
  bool truncated = false;  // the event had more pixel segments than kNMaxPixelSegments (SDL/Event.cc:237-244)
  enum { X, Y, Z, DET, IDX, H0, H1, H2, H3, SEED, DPHI, PT, PTERR, PX, PY, PZ, ETA, ETAERR, PHI, CHARGE, SUPERBIN,
         PTYPE, QUAD, NF };
  size_t off[NF] = {};
  void release() {
    if (host) (void)hipHostFree(host);
    host = nullptr;
    bytes = 0;
  }
  // the descriptor of this event's arrays at device address `base` (or at the host block itself)
  lst::InputDesc desc(const char* base) const {
    lst::InputDesc d{};
    auto at = [&](int f) { return base + off[f]; };
    d.x = (const float*)at(X); d.y = (const float*)at(Y); d.z = (const float*)at(Z);
    d.det = (const uint32_t*)at(DET); d.idx = (const uint32_t*)at(IDX); d.nHits = nHits;
    for (int k = 0; k < 4; k++) d.hitIdx[k] = (const uint32_t*)at(H0 + k);
    d.dPhiChange = (const float*)at(DPHI); d.ptIn = (const float*)at(PT); d.ptErr = (const float*)at(PTERR);
    d.px = (const float*)at(PX); d.py = (const float*)at(PY); d.pz = (const float*)at(PZ);
    d.eta = (const float*)at(ETA); d.etaErr = (const float*)at(ETAERR); d.phi = (const float*)at(PHI);
    d.charge = (const int32_t*)at(CHARGE); d.superbin = (const int32_t*)at(SUPERBIN);
    d.seedIdx = (const uint32_t*)at(SEED); d.pixelType = (const int8_t*)at(PTYPE); d.isQuad = (const int8_t*)at(QUAD);
    d.nPLS = nPLS;
    d.outHost = nullptr;
    d.outDev = nullptr;
    d.outCap = 0;
    return d;
  }
};

// prepareInput on the host, then the prepared arrays into one pinned block
static lst_status prepare_input(const lst_input& in, PreparedInput& p) {
  lst::PreparedEvent e;
  lst::prepare_event(in, e);
  p.release();
  const uint32_t nh = (uint32_t)e.x.size();
  const uint32_t np = std::min<uint32_t>((uint32_t)e.ptIn.size(), lst::kNMaxPixelSegments);
  p.truncated = e.ptIn.size() > lst::kNMaxPixelSegments;
  p.nHits = nh;
  p.nPLS = np;
  const void* src[PreparedInput::NF] = {e.x.data(), e.y.data(), e.z.data(), e.detId.data(), e.idx.data(),
                                        e.h0.data(), e.h1.data(), e.h2.data(), e.h3.data(), e.seedIdx.data(),
                                        e.dPhiChange.data(), e.ptIn.data(), e.ptErr.data(), e.px.data(), e.py.data(),
                                        e.pz.data(), e.eta.data(), e.etaErr.data(), e.phi.data(), e.charge.data(),
                                        e.superbin.data(), e.pixelType.data(), e.isQuad.data()};
  size_t len[PreparedInput::NF];
  size_t off = 0;
  for (int f = 0; f < PreparedInput::NF; f++) {
    len[f] = f < PreparedInput::H0 ? (size_t)nh * 4 : f < PreparedInput::PTYPE ? (size_t)np * 4 : (size_t)np;
    p.off[f] = off;
    off += (len[f] + 255) & ~size_t(255);
  }
  p.bytes = off ? off : 256;
  HIP_OK(hipHostMalloc((void**)&p.host, p.bytes, hipHostMallocDefault));
  for (int f = 0; f < PreparedInput::NF; f++)
    if (len[f]) std::memcpy(p.host + p.off[f], src[f], len[f]);
  return LST_OK;
}

struct lst_event {
  lst_es* es = nullptr;
  hipStream_t stream = nullptr;
  bool ownStream = false;
  int t5Mode = 0;
  int stage = -1;  // last completed stage
  lst::EventView v{};
  lst::EventView* vDev = nullptr;  // device copy of v (kernels read their view from device memory)
  // capacities
  uint32_t hitCap = 0, mdCap = 0, segCap = 0, plsCap = lst::kNMaxPixelSegments, t3Cap = 0, t5Cap = 0;
  static constexpr uint32_t pt5Cap = lst::kNMaxPixelQuintuplets, pt3Cap = lst::kNMaxPixelTriplets,
                            tcCap = lst::kNMaxPixelTrackCandidates + lst::kNMaxNonPixelTrackCandidates;
  char* arena = nullptr;
  size_t arenaBytes = 0;
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
  // the launch context of this handle's one event
  lst::Launch launch() { return lst::Launch{v, vDev, 1, stream, tm()}; }
  void* dnnRecordMem = nullptr;        // T5 DNN training rows (lst_event_record_t5dnn)
  char* gatherMem = nullptr;           // a record field gathered contiguous for lst_event_get_field
  size_t gatherBytes = 0;
  // prepared inputs kept in pinned host memory (throughput path): copied to the
  // device, described by descDev, at each run
  std::vector<PreparedInput> staged;
  char* inDev = nullptr;               // device copy of the running event's prepared block
  size_t inDevBytes = 0;
  lst::InputDesc* descDev = nullptr;   // device descriptor of that block
  lst::InputDesc* descHost = nullptr;  // pinned staging of the descriptor

  lst_status allocate(uint32_t nHitsNeeded);
  lst_status upload_view();
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
  char* mdHot = nullptr;  // the mini-doublet records, hot and cold parts (fields set after binding)
  char* mdCold = nullptr;
  c.take(mdHot, (size_t)mdCap * lst::kMDHotBytes);
  c.take(mdCold, (size_t)mdCap * lst::kMDColdBytes);
  char* segRec = nullptr;  // the segment records (fields set after binding)
  c.take(segRec, (size_t)segCap * lst::kSegRecBytes);
#define CARVE(TY, NM, MU) c.take(e.P.NM, (size_t)plsCap * MU);
  LST_PLS_FIELDS(CARVE)
#undef CARVE
  char* t3Rec = nullptr;  // the triplet records (fields set after binding)
  c.take(t3Rec, (size_t)(t3Cap ? t3Cap : 1) * lst::kT3ObjBytes);
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
  c.take(e.W.pixBinRec, pt5Cap + pt3Cap); c.take(e.W.pixBinOff, lst::kEtaBinOffsets);
  c.take(e.W.tcBinRec, tcCap); c.take(e.W.tcBinOff, lst::kEtaBinOffsets); c.take(e.W.tcHitSet, lst::kTCHitSet);
  c.take(e.W.pt5Mask, lst::kPixMaskWords); c.take(e.W.pt5MaskBase, plsCap);
  c.take(e.W.pt3Mask, lst::kPixMaskWords); c.take(e.W.pt3MaskBase, plsCap);
  c.take(e.W.pt5Shard, 2 * lst::kPixShards); c.take(e.W.pt3Shard, 2 * lst::kPixShards);
  c.take(e.W.pixWords, plsCap); c.take(e.W.pixFull, plsCap);
  c.take(e.W.pixSurv, (size_t)lst::kSurvWords * lst::kPixSurvMax);
  c.take(e.W.t3ListOff, nL); c.take(e.W.t3List, t3Cap ? t3Cap : 1);
  c.take(e.W.t5Rec, (size_t)lst::kT5RecWords * (t5Cap ? t5Cap : 1));
  c.take(e.W.t3Rec, (size_t)lst::kT3RecWords * (t3Cap ? t3Cap : 1));
  c.take(e.W.pt5BucketOff, lst::kPT5HashBuckets + 1); c.take(e.W.pt5BucketEnt, 14 * (size_t)pt5Cap);
  c.take(e.FS.mask, lst::kScanCandMax / 64); c.take(e.FS.pre, lst::kScanCandMax / 64);
  c.take(e.FS.sums, lst::kScanCandMax / 4096 + 1);
  c.take(e.FS.pass, lst::kScanCandMax);
  c.take(e.W2.candOff, nL + 1); c.take(e.W2.hitSlope, hitCap); c.take(e.W2.shiftPoint, hitCap);
  c.take(e.W2.park, (size_t)lst::kMDParkWords * lst::kMDCandMax); c.take(e.W2.cand, lst::kMDCandMax);
  c.take(e.WL.connPre, (size_t)nL * (g.view.stride + 1)); c.take(e.WL.candOff, nL + 1);
  c.take(e.WL.park, (size_t)lst::kLSParkWords * lst::kLSCandMax);
  c.take(e.WL.cand, lst::kLSCandMax);
  c.take(e.W3.modOff, nL + 1);
  c.take(e.W3.list, 2 * (size_t)lst::kT3CandMax);
  c.take(e.W3.results, (size_t)lst::kT3ResultWords * lst::kT3CandMax);
  c.take(e.W3.survOff, nL + 1); c.take(e.W3.segByMD, segCap); c.take(e.W3.segTmp, segCap); c.take(e.W3.mdSegOff, mdCap); c.take(e.W3.mdSegEnd, mdCap);
  c.take(e.W5.modOff, nL + 2);
  c.take(e.W5.resIdx, lst::kT5CandMax);
  c.take(e.W5.listA, 3 * (size_t)lst::kT5CandMax); c.take(e.W5.listB, 3 * (size_t)lst::kT5CandMax);
  c.take(e.W5.results, (size_t)lst::kT5ResultWords * lst::kT5ResultMax);
  c.take(e.W5.survOff, nL + 2); c.take(e.W5.t3ByMD, t3Cap ? t3Cap : 1); c.take(e.W5.t3Tmp, t3Cap ? t3Cap : 1);
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
  for (int k = 0; k < 4; k++) c.take(e.plsInHitIdx[k], plsCap);
  c.take(e.plsInDPhiChange, plsCap);
  arenaBytes = c.off;
  HIP_OK(hipMalloc(&arena, arenaBytes));
  c.bind(arena);
  {
    uint32_t k = 0;
#define SLOT(TY, NM, MU) \
  e.D.NM.p = (k < lst::kMDHotFields ? mdHot : mdCold) + 4 * lst::md_slot(k); \
  k++;
    LST_MD_FIELDS(SLOT)
#undef SLOT
    if (k != lst::kMDFields) return fail(LST_E_DEVICE, "mini-doublet record layout");
    e.WL.mdSlope.p = mdHot + 4 * lst::kMDSlopeWord;
    uint32_t w = 0;
#define SLOT(TY, NM, MU) e.S.NM.p = segRec + 4 * w; w += MU;
    LST_SEG_FIELDS(SLOT)
#undef SLOT
    if (w != lst::kSegRecWords) return fail(LST_E_DEVICE, "segment record layout");
    w = 0;
#define SLOT(TY, NM, MU) e.T.NM.p = t3Rec + 4 * w; w += MU;
    LST_T3_FIELDS(SLOT)
#undef SLOT
    if (w > lst::kT3ObjWords) return fail(LST_E_DEVICE, "triplet record layout");
  }

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
  e.outCap = tcCap;

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
  auto addRec = [&](const std::string& name, auto f, size_t n, size_t stride, size_t mult) {
    using T = std::remove_reference_t<decltype(f[0])>;
    fields.push_back({name, FieldRef{(void*)f.p, dtype_of<T>(), sizeof(T), n, stride, mult}});
  };
  {
#define MDFIELD(TY, NM, MU) addRec("mds." #NM, e.D.NM, mdCap, lst::md_rec_bytes(lst::kMDw_##NM), 1);
    LST_MD_FIELDS(MDFIELD)
#undef MDFIELD
  }
  {
#define SEGFIELD(TY, NM, MU) addRec("segs." #NM, e.S.NM, (size_t)segCap * MU, lst::kSegRecBytes, MU);
    LST_SEG_FIELDS(SEGFIELD)
#undef SEGFIELD
  }
  {
    std::string prefix = "segs.";
    size_t capacity = plsCap;
    auto* cur = &e.P;
    LST_PLS_FIELDS(FIELD)
  }
  {
#define T3FIELD(TY, NM, MU) addRec("t3." #NM, e.T.NM, (size_t)t3Cap * MU, lst::kT3ObjBytes, MU);
    LST_T3_FIELDS(T3FIELD)
#undef T3FIELD
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
  return upload_view();
}

// the device copy of the view, which every kernel of this handle reads
lst_status lst_event::upload_view() {
  if (!vDev) HIP_OK(hipMalloc(&vDev, sizeof(lst::EventView)));
  HIP_OK(hipStreamSynchronize(stream));
  HIP_OK(hipMemcpy(vDev, &v, sizeof(lst::EventView), hipMemcpyHostToDevice));
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
  lst::launch_hits(launch());
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
    cb.add(v.plsInHitIdx[0], p.hit_indices0, n * 4);
    cb.add(v.plsInHitIdx[1], p.hit_indices1, n * 4);
    cb.add(v.plsInHitIdx[2], p.hit_indices2, n * 4);
    cb.add(v.plsInHitIdx[3], p.hit_indices3, n * 4);
    cb.add(v.plsInDPhiChange, p.delta_phi_change, n * 4);
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
  lst::launch_pixel_segments(launch());
  HIP_OK(hipGetLastError());
  stage = 1;
  return warn;
}

// Host-side checks of the hits the reference's hit kernels assume (SDL/Hit.h:194-270): every detId is a
// module of the geometry, and each module's hits are contiguous and at most 127 (its int8 counts).
static lst_status check_hits(const lst::HostGeometry& g, const uint32_t* det, uint32_t n) {
  uint32_t prev = ~0u, run = 0;
  std::vector<uint8_t> seen;
  for (uint32_t i = 0; i < n; i++) {
    if (det[i] == prev) {
      if (++run > 127 && det[i] != 1u)
        return fail(LST_E_INPUT, "more than 127 hits on module " + std::to_string(det[i]));
      continue;
    }
    const int m = g.module_of(det[i]);
    if (m < 0) return fail(LST_E_INPUT, "hit " + std::to_string(i) + " has detId " + std::to_string(det[i]) +
                                            ", which is no module of the geometry");
    if (seen.empty()) seen.assign(g.nModules, 0);
    if (seen[m]) return fail(LST_E_INPUT, "the hits of module " + std::to_string(det[i]) + " are not contiguous");
    seen[m] = 1;
    prev = det[i];
    run = 1;
  }
  return LST_OK;
}

// the stage sequence of one event, launched stage by stage as the SDL::Event methods would
static lst_status run_stages(lst_event* ev, int last_stage, lst_status warn) {
  lst_status s = LST_OK;
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

// a prepared event through the handle: its block copied to the device in one
// transfer, spread into the arena by one kernel, then the stages
static lst_status run_prepared(lst_event* ev, const PreparedInput& p, int last_stage) {
  lst_status s = ev->allocate(p.nHits);
  if (s != LST_OK) return s;
  if (!ev->descDev) {
    HIP_OK(hipMalloc(&ev->descDev, sizeof(lst::InputDesc)));
    HIP_OK(hipHostMalloc((void**)&ev->descHost, sizeof(lst::InputDesc), hipHostMallocDefault));
  }
  if (ev->inDevBytes < p.bytes) {
    HIP_OK(hipStreamSynchronize(ev->stream));
    if (ev->inDev) HIP_OK(hipFree(ev->inDev));
    ev->inDev = nullptr;
    HIP_OK(hipMalloc(&ev->inDev, p.bytes));
    ev->inDevBytes = p.bytes;
  }
  HIP_OK(hipStreamSynchronize(ev->stream));  // the pinned descriptor of the previous run has been read
  *ev->descHost = p.desc(ev->inDev);
  HIP_OK(hipMemcpyAsync(ev->inDev, p.host, p.bytes, hipMemcpyHostToDevice, ev->stream));
  HIP_OK(hipMemcpyAsync(ev->descDev, ev->descHost, sizeof(lst::InputDesc), hipMemcpyHostToDevice, ev->stream));
  const lst::Launch L = ev->launch();
  lst::launch_stage_in(L, ev->descDev);
  lst::launch_hits(L);
  ev->v.nHits = p.nHits;
  ev->v.nPLS = 0;
  ev->stage = 0;
  HIP_OK(hipGetLastError());
  if (last_stage < 1) return LST_OK;
  lst::launch_pixel_segments(L);
  ev->v.nPLS = p.nPLS;
  ev->stage = 1;
  HIP_OK(hipGetLastError());
  const lst_status warn = p.truncated ? (lst_status)LST_W_TRUNCATED : LST_OK;
  if (last_stage < 2) return warn;
  return run_stages(ev, last_stage, warn);
}

// the device-side flags the hit kernels and the last kernel leave in the
// event's scalars, as a status: bad input is an error, truncation a warning
static lst_status device_status(lst_event* ev) {
  uint32_t scal[lst::kNScalars];
  HIP_OK(hipMemcpyAsync(scal, ev->v.R.scalars, sizeof(scal), hipMemcpyDeviceToHost, ev->stream));
  HIP_OK(hipStreamSynchronize(ev->stream));
  const uint32_t bad = scal[lst::kBadHits];
  if (bad & lst::kBadDetId) return fail(LST_E_INPUT, "a hit's detId is no module of the geometry");
  if (bad & lst::kBadModuleHits) return fail(LST_E_INPUT, "more than 127 hits on one module");
  if (bad & lst::kBadNonContiguous) return fail(LST_E_INPUT, "the hits of a module are not contiguous");
  return scal[lst::kTruncated] ? (lst_status)LST_W_TRUNCATED : LST_OK;
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
  if (!lst::fill_module_terms(es->host, pt_cut, err)) return fail(LST_E_FORMAT, err);
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
  if (ev->stream) (void)hipStreamSynchronize(ev->stream);
  if (ev->arena) (void)hipFree(ev->arena);
  if (ev->vDev) (void)hipFree(ev->vDev);
  if (ev->inDev) (void)hipFree(ev->inDev);
  for (auto& st : ev->staged) st.release();
  for (int k = 0; k < lst::kNTimers; k++) {
    if (ev->timers.start[k]) (void)hipEventDestroy(ev->timers.start[k]);
    if (ev->timers.stop[k]) (void)hipEventDestroy(ev->timers.stop[k]);
  }
  if (ev->dnnRecordMem) (void)hipFree(ev->dnnRecordMem);
  if (ev->gatherMem) (void)hipFree(ev->gatherMem);
  if (ev->descDev) (void)hipFree(ev->descDev);
  if (ev->descHost) (void)hipHostFree(ev->descHost);
  if (ev->ownStream) (void)hipStreamDestroy(ev->stream);
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
  lst::launch_mini_doublets(ev->launch());
  HIP_OK(hipGetLastError());
  ev->stage = 2;
  return LST_OK;
}

lst_status lst_event_create_segments(lst_event* ev) {
  if (!ev || ev->stage != 2) return fail(LST_E_ORDER, "createSegmentsWithModuleMap needs mini-doublets");
  lst::launch_segments(ev->launch());
  HIP_OK(hipGetLastError());
  ev->stage = 3;
  return LST_OK;
}

lst_status lst_event_create_triplets(lst_event* ev) {
  if (!ev || ev->stage != 3) return fail(LST_E_ORDER, "createTriplets needs segments");
  lst::launch_triplets(ev->launch());
  HIP_OK(hipGetLastError());
  ev->stage = 4;
  return LST_OK;
}

lst_status lst_event_create_quintuplets(lst_event* ev) {
  if (!ev || ev->stage != 4) return fail(LST_E_ORDER, "createQuintuplets needs triplets");
  lst::launch_quintuplets(ev->launch());
  HIP_OK(hipGetLastError());
  ev->stage = 5;
  return LST_OK;
}

lst_status lst_event_pixel_line_segment_cleaning(lst_event* ev) {
  if (!ev || ev->stage != 5) return fail(LST_E_ORDER, "pixelLineSegmentCleaning needs quintuplets");
  lst::launch_pixel_segment_cleaning(ev->launch());
  HIP_OK(hipGetLastError());
  ev->stage = 6;
  return LST_OK;
}

lst_status lst_event_create_pixel_quintuplets(lst_event* ev) {
  if (!ev || ev->stage != 6) return fail(LST_E_ORDER, "createPixelQuintuplets needs pixelLineSegmentCleaning");
  lst::launch_pixel_quintuplets(ev->launch());
  HIP_OK(hipGetLastError());
  ev->stage = 7;
  return LST_OK;
}

lst_status lst_event_create_pixel_triplets(lst_event* ev) {
  if (!ev || ev->stage != 7) return fail(LST_E_ORDER, "createPixelTriplets needs pixel quintuplets");
  lst::launch_pixel_triplets(ev->launch());
  HIP_OK(hipGetLastError());
  ev->stage = 8;
  return LST_OK;
}

lst_status lst_event_create_track_candidates(lst_event* ev) {
  if (!ev || ev->stage != 8) return fail(LST_E_ORDER, "createTrackCandidates needs pixel triplets");
  lst::launch_track_candidates(ev->launch());
  HIP_OK(hipGetLastError());
  ev->stage = 9;
  return LST_OK;
}

lst_status lst_event_get_output(lst_event* ev, uint32_t* n_tc, uint32_t* hit_idxs, uint32_t* len, int32_t* seed_idx,
                                int16_t* type, uint32_t capacity) {
  if (!ev || !n_tc) return fail(LST_E_INPUT, "null argument");
  if (ev->stage != 9) return fail(LST_E_ORDER, "getOutput needs track candidates");
  const lst_status st = device_status(ev);
  if ((st & 0xff) != LST_OK) return st;
  uint32_t scal[lst::kNScalars];
  HIP_OK(hipMemcpyAsync(scal, ev->v.R.scalars, sizeof(scal), hipMemcpyDeviceToHost, ev->stream));
  HIP_OK(hipStreamSynchronize(ev->stream));
  const uint32_t n = scal[lst::kNTrackCandidates];
  *n_tc = n;
  if (!hit_idxs && !len && !seed_idx && !type) return st;
  if (capacity < n) return fail(LST_E_CAPACITY, "output arrays hold fewer track candidates than were built");
  if (n) {
    if (hit_idxs) HIP_OK(hipMemcpyAsync(hit_idxs, ev->v.W.outHits, (size_t)n * 14 * 4, hipMemcpyDeviceToHost, ev->stream));
    if (len) HIP_OK(hipMemcpyAsync(len, ev->v.W.outLen, (size_t)n * 4, hipMemcpyDeviceToHost, ev->stream));
    if (seed_idx) HIP_OK(hipMemcpyAsync(seed_idx, ev->v.C.pixelSeedIndex, (size_t)n * 4, hipMemcpyDeviceToHost, ev->stream));
    if (type) HIP_OK(hipMemcpyAsync(type, ev->v.C.trackCandidateType, (size_t)n * 2, hipMemcpyDeviceToHost, ev->stream));
    HIP_OK(hipStreamSynchronize(ev->stream));
  }
  return st;
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
  if (ev->stage < 0) return LST_OK;
  return device_status(ev);
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
  const void* src = f->ptr;
  if (f->stride) {  // a record field: gathered contiguous on the device first (a debug getter, not the hot path)
    if (ev->gatherBytes < bytes) {
      HIP_OK(hipStreamSynchronize(ev->stream));
      if (ev->gatherMem) HIP_OK(hipFree(ev->gatherMem));
      ev->gatherMem = nullptr;
      ev->gatherBytes = 0;
      HIP_OK(hipMalloc(&ev->gatherMem, bytes));
      ev->gatherBytes = bytes;
    }
    hipLaunchKernelGGL(k_gather_field, dim3(1024), dim3(256), 0, ev->stream, static_cast<const char*>(f->ptr),
                       (uint32_t)f->stride, (uint32_t)f->elem, (uint32_t)f->mult, f->count, ev->gatherMem);
    HIP_OK(hipGetLastError());
    src = ev->gatherMem;
  }
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ev->stream));
  HIP_OK(hipStreamSynchronize(ev->stream));
  return LST_OK;
}

lst_status lst_run(lst_event* ev, const lst_input* in, int last_stage) {
  if (!ev || !in) return fail(LST_E_INPUT, "null argument");
  if (in->n_hits && (!in->ph2_detId || !in->ph2_x || !in->ph2_y || !in->ph2_z))
    return fail(LST_E_INPUT, "null hit arrays");
  lst::prepare_event(*in, ev->prepared);
  const auto& p = ev->prepared;
  lst_status s = check_hits(ev->es->host, p.detId.data(), (uint32_t)p.x.size());
  if (s != LST_OK) return s;
  s = ev->add_hits(p.x.data(), p.y.data(), p.z.data(), p.detId.data(), p.idx.data(), (uint32_t)p.x.size(), false);
  if (s != LST_OK || last_stage < 1) return s;
  const lst_pixel_segments pv = p.view();
  lst_status warn = ev->add_pixel_segments(pv, false);
  if ((warn & 0xff) != LST_OK || last_stage < 2) return warn;
  return run_stages(ev, last_stage, warn);
}

lst_status lst_event_stage_input(lst_event* ev, const lst_input* in, uint32_t slot) {
  if (!ev || !in) return fail(LST_E_INPUT, "null argument");
  if (ev->staged.size() <= slot) ev->staged.resize(slot + 1);
  PreparedInput& st = ev->staged[slot];
  HIP_OK(hipStreamSynchronize(ev->stream));  // a run may still be copying this slot's block
  lst_status s = prepare_input(*in, st);
  if (s != LST_OK) return s;
  if ((s = check_hits(ev->es->host, (const uint32_t*)(st.host + st.off[PreparedInput::DET]), st.nHits)) != LST_OK)
    return s;
  return ev->allocate(st.nHits);
}

lst_status lst_event_run_staged(lst_event* ev, uint32_t slot, int last_stage) {
  if (!ev || slot >= ev->staged.size() || !ev->staged[slot].host) return fail(LST_E_INPUT, "input slot not staged");
  return run_prepared(ev, ev->staged[slot], last_stage);
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
  if (capacity == 0) return ev->upload_view();
  if (capacity > lst::kT5ResultMax) capacity = lst::kT5ResultMax;
  Carve c;
  lst::T5DnnDump& d = ev->v.dnnDump;
  c.take(d.features, 38 * (size_t)capacity); c.take(d.score, capacity);
  c.take(d.t3Pair, 2 * (size_t)capacity); c.take(d.hitIdxs, 10 * (size_t)capacity);
  HIP_OK(hipMalloc(&ev->dnnRecordMem, c.off));
  c.bind(static_cast<char*>(ev->dnnRecordMem));
  d.capacity = capacity;
  return ev->upload_view();
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
// A batch runs up to `slots` events in one launch set: every kernel of the
// stage sequence is launched once for all of them (gridDim.z = slots), each
// event in its own arena, and the sequence is one captured graph.  Inputs are
// prepared on the host once and kept pinned; each launch set copies its
// events' blocks to the device on a transfer stream while the previous set
// computes; its last kernel puts each event's packed output block in a device
// block and the block's header in pinned host memory, and the host copies the
// records back on the transfer stream while the next set computes.  Launch set k
// uses frame k % 3 (input blocks, descriptors, output blocks) and, with two
// sets in flight, arena set k % 2 on its own compute stream; the host consumes
// set k - (sets in flight) after launching set k.  An event whose candidates overflow a
// flattened builder's work arrays (a flag in its output header) is run again
// by a per-event handle, which has the per-module builders.
struct lst_batch {
  lst_es* es = nullptr;
  int t5Mode = 0;
  uint32_t slots = 0;
  uint32_t sets = 1;                    // launch sets in flight (1 or 2)
  std::vector<lst_event*> ctx;          // one arena per slot, per set in flight (set s: ctx[s * slots ...])
  lst_event* rerun = nullptr;           // per-event handle for events that overflow the flattened builders
  // per set in flight its compute stream and device view array (with one set, entry 1 is entry 0)
  hipStream_t stream[2] = {nullptr, nullptr};
  // input copies and output record copies (one stream: a process has few hardware queues, and a stream
  // that shares one with a compute stream waits for that stream's graphs)
  hipStream_t xferStream = nullptr;
  lst::EventView* viewsDev[2] = {nullptr, nullptr};  // [slots] each
  static constexpr int kFrames = 3;
  // per frame: the launch set's device input blocks (slots x slotBytes), its descriptors (pinned), its output
  // blocks (pinned headers and records, device blocks), and the events of its input copy and of its end
  char* inDev[kFrames] = {};
  size_t slotBytes = 0;
  lst::InputDesc* descHost[kFrames] = {};
  uint32_t* outHost[kFrames] = {};      // slots x out_words()
  uint32_t* outDev[kFrames] = {};       // slots x out_words()
  uint32_t outCapAlloc = 0;             // records per event outHost and outDev hold
  hipEvent_t copyDone[kFrames] = {}, computeDone[kFrames] = {}, outDone[kFrames] = {};
  // per arena set: the descriptors its graph reads (copied in on its stream before each launch) and the graph
  lst::InputDesc* descDev[2] = {nullptr, nullptr};  // [slots]
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  bool graphs = true;
  std::vector<PreparedInput> staged;
  uint32_t grid[lst::kNTune] = {};      // chosen work-group counts (0: the defaults)
  size_t out_words(uint32_t cap) const {
    return ((lst::kOutHeaderWords + (size_t)cap * lst::kOutRecWords) + 63) & ~size_t(63);
  }
  // the arena set of launch set k
  uint32_t set_of(uint32_t k) const { return sets == 2 ? k % 2 : 0u; }
  lst_status sync_compute() const {
    for (uint32_t s = 0; s < sets; s++)
      if (hipStreamSynchronize(stream[s]) != hipSuccess) return LST_E_DEVICE;
    return LST_OK;
  }
};

namespace {
constexpr uint32_t kTCCap = lst::kNMaxPixelTrackCandidates + lst::kNMaxNonPixelTrackCandidates;

void drop_graphs(lst_batch* b) {
  for (auto& g : b->graph)
    if (g) (void)hipGraphExecDestroy(g);
  b->graph[0] = b->graph[1] = nullptr;
}

// every slot's arena (of every set in flight) sized for nHits, and the device view arrays
lst_status batch_arenas(lst_batch* b, uint32_t nHits) {
  bool changed = false;
  for (uint32_t s = 0; s < b->sets; s++) changed |= !b->viewsDev[s];
  for (lst_event* ev : b->ctx) {
    const char* before = ev->arena;
    lst_status s = ev->allocate(nHits);
    if (s != LST_OK) return s;
    changed |= ev->arena != before;
  }
  if (!changed) return LST_OK;
  if (b->sync_compute() != LST_OK) return fail(LST_E_DEVICE, "stream synchronize failed");
  for (uint32_t s = 0; s < b->sets; s++) {
    if (!b->viewsDev[s]) HIP_OK(hipMalloc(&b->viewsDev[s], sizeof(lst::EventView) * b->slots));
    for (uint32_t i = 0; i < b->slots; i++)
      HIP_OK(hipMemcpy(b->viewsDev[s] + i, &b->ctx[s * b->slots + i]->v, sizeof(lst::EventView),
                       hipMemcpyHostToDevice));
  }
  if (b->sets == 1) b->viewsDev[1] = b->viewsDev[0];
  drop_graphs(b);
  return LST_OK;
}

// input and output buffers for events of up to `bytes` prepared bytes and `cap` output records
lst_status batch_io(lst_batch* b, size_t bytes, uint32_t cap) {
  if (b->sync_compute() != LST_OK) return fail(LST_E_DEVICE, "stream synchronize failed");
  HIP_OK(hipStreamSynchronize(b->xferStream));
  bytes = (bytes + 255) & ~size_t(255);
  if (bytes > b->slotBytes) {
    for (auto& p : b->inDev)
      if (p) HIP_OK(hipFree(p));
    for (auto& p : b->inDev) HIP_OK(hipMalloc(&p, bytes * b->slots));
    b->slotBytes = bytes;
  }
  if (!b->descHost[0])
    for (int k = 0; k < lst_batch::kFrames; k++) {
      HIP_OK(hipHostMalloc((void**)&b->descHost[k], sizeof(lst::InputDesc) * b->slots, hipHostMallocDefault));
      HIP_OK(hipEventCreateWithFlags(&b->copyDone[k], hipEventDisableTiming));
      HIP_OK(hipEventCreateWithFlags(&b->computeDone[k], hipEventDisableTiming));
      HIP_OK(hipEventCreateWithFlags(&b->outDone[k], hipEventDisableTiming));
    }
  for (int k = 0; k < 2; k++)
    if (!b->descDev[k]) HIP_OK(hipMalloc(&b->descDev[k], sizeof(lst::InputDesc) * b->slots));
  if (!b->outHost[0] || cap > b->outCapAlloc) {
    for (auto& p : b->outHost)
      if (p) HIP_OK(hipHostFree(p));
    for (auto& p : b->outDev)
      if (p) HIP_OK(hipFree(p));
    for (auto& p : b->outHost)
      HIP_OK(hipHostMalloc((void**)&p, 4 * b->out_words(cap) * b->slots, hipHostMallocDefault));
    for (auto& p : b->outDev) HIP_OK(hipMalloc(&p, 4 * b->out_words(cap) * b->slots));
    b->outCapAlloc = cap;
  }
  return LST_OK;
}

// the stage sequence of one launch set on arena set `set`, reading that set's descriptors
void enqueue_set(lst_batch* b, uint32_t set, const hipEvent_t* tuneEvents = nullptr, uint32_t* tuneRecorded = nullptr) {
  lst::Launch L{b->ctx[set * b->slots]->v, b->viewsDev[set], b->slots, b->stream[set], lst::no_timers()};
  L.grid = b->grid;
  L.fallbacks = false;  // overflowing events are run again by the per-event handle
  L.tuneEvents = tuneEvents;
  L.tuneRecorded = tuneRecorded;
  lst::launch_stage_in(L, b->descDev[set]);
  lst::launch_hits(L);
  lst::launch_pixel_segments(L);
  lst::launch_mini_doublets(L);
  lst::launch_segments(L);
  lst::launch_triplets(L);
  lst::launch_quintuplets(L);
  lst::launch_pixel_segment_cleaning(L);
  lst::launch_pixel_quintuplets(L);
  lst::launch_pixel_triplets(L);
  lst::launch_track_candidates(L);
  lst::launch_out_copy(L, b->descDev[set]);
}

lst_status capture(lst_batch* b) {
  for (uint32_t set = 0; set < b->sets; set++) {
    if (b->graph[set]) continue;
    hipGraph_t g = nullptr;
    hipStream_t st = b->stream[set];
    HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    enqueue_set(b, set);
    HIP_OK(hipStreamEndCapture(st, &g));
    const hipError_t e = hipGraphInstantiate(&b->graph[set], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(LST_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return LST_OK;
}

// The events of each launch set of `slots` events, ordered by hit count,
// largest first: every kernel hands out its work-groups event by event (x,
// then z), so the work-groups of the largest events start first and the
// launch set does not end waiting on a large event that was dispatched last
// (the summed kernel time of a 64-event uneven set fell 4 % this way).
// Returns positions into ids.
std::vector<uint32_t> largest_first(const lst_batch* b, const uint32_t* ids, size_t n, uint32_t slots) {
  std::vector<uint32_t> pos(n);
  for (size_t i = 0; i < n; i++) pos[i] = (uint32_t)i;
  for (size_t k = 0; k < n; k += slots) {
    const auto first = pos.begin() + k, last = pos.begin() + std::min(n, k + slots);
    std::stable_sort(first, last,
                     [&](uint32_t a, uint32_t c) { return b->staged[ids[a]].nHits > b->staged[ids[c]].nHits; });
  }
  return pos;
}

// the inputs of events ids[0..n) into frame f (transfer stream); the host may
// rewrite descHost[f] and the copies may overwrite inDev[f] because the set
// that last used frame f has been consumed
lst_status issue_inputs(lst_batch* b, int f, const uint32_t* ids, uint32_t n, uint32_t outCap) {
  const size_t ow = b->out_words(outCap);
  for (uint32_t i = 0; i < b->slots; i++) {
    lst::InputDesc& d = b->descHost[f][i];
    if (i < n) {
      const PreparedInput& p = b->staged[ids[i]];
      char* dst = b->inDev[f] + (size_t)i * b->slotBytes;
      d = p.desc(dst);
      HIP_OK(hipMemcpyAsync(dst, p.host, p.bytes, hipMemcpyHostToDevice, b->xferStream));
    } else {
      d = lst::InputDesc{};  // an empty event
    }
    d.outHost = b->outHost[f] + ow * i;
    d.outDev = b->outDev[f] + ow * i;
    d.outCap = outCap;
  }
  HIP_OK(hipEventRecord(b->copyDone[f], b->xferStream));
  return LST_OK;
}

// frame f's descriptors into arena set `set`'s descriptor array, on that set's stream once frame f's
// inputs are on the device (stream order: after the set's previous launch set, whose graph read them)
lst_status stage_descriptors(lst_batch* b, uint32_t set, int f) {
  hipStream_t st = b->stream[set];
  HIP_OK(hipStreamWaitEvent(st, b->copyDone[f], 0));
  HIP_OK(hipMemcpyAsync(b->descDev[set], b->descHost[f], sizeof(lst::InputDesc) * b->slots,
                        hipMemcpyHostToDevice, st));
  return LST_OK;
}

// launch set k (frame k % 3) on its arena set's stream; with two sets in flight it runs beside set k - 1
lst_status launch_set(lst_batch* b, uint32_t k) {
  const uint32_t set = b->set_of(k);
  const int f = (int)(k % lst_batch::kFrames);
  hipStream_t st = b->stream[set];
  lst_status s = stage_descriptors(b, set, f);
  if (s != LST_OK) return s;
  if (b->graphs) {
    if ((s = capture(b)) != LST_OK) return s;
    HIP_OK(hipGraphLaunch(b->graph[set], st));
  } else {
    enqueue_set(b, set);
    HIP_OK(hipGetLastError());
  }
  HIP_OK(hipEventRecord(b->computeDone[f], st));
  return LST_OK;
}

// one event's output lists from a packed block (or the per-event handle) into
// the caller's descriptor; returns the event's status
lst_status copy_output(const uint32_t* blk, lst_output* d) {
  const uint32_t n = blk[lst::kOutNTC];
  d->n_tc = n;
  const uint32_t m = std::min(n, d->capacity);
  const uint32_t* rec = blk + lst::kOutHeaderWords;
  for (uint32_t t = 0; t < m; t++, rec += lst::kOutRecWords) {
    if (d->hit_idxs) std::memcpy(d->hit_idxs + 14 * (size_t)t, rec, 56);
    if (d->len) d->len[t] = rec[14];
    if (d->seed_idx) d->seed_idx[t] = (int32_t)rec[15];
    if (d->type) d->type[t] = (int16_t)(int32_t)rec[16];
  }
  return n > d->capacity || blk[lst::kOutTruncated] ? (lst_status)LST_W_TRUNCATED : LST_OK;
}

lst_status rerun_event(lst_batch* b, const PreparedInput& p, lst_output* d) {
  if (!b->rerun) {
    lst_status s = lst_event_create(b->es, nullptr, b->t5Mode, &b->rerun);
    if (s != LST_OK) return s;
  }
  lst_event* ev = b->rerun;
  lst_status s = run_prepared(ev, p, 9);
  if ((s & 0xff) != LST_OK) return s;
  if (!d) return lst_event_synchronize(ev);
  uint32_t n = 0;
  lst_status st = lst_event_get_output(ev, &n, nullptr, nullptr, nullptr, nullptr, 0);
  if ((st & 0xff) != LST_OK) return st;
  std::vector<uint32_t> hits((size_t)n * 14), len(n);
  std::vector<int32_t> seed(n);
  std::vector<int16_t> type(n);
  if (n) {
    st = lst_event_get_output(ev, &n, hits.data(), len.data(), seed.data(), type.data(), n);
    if ((st & 0xff) != LST_OK) return st;
  }
  d->n_tc = n;
  const uint32_t m = std::min(n, d->capacity);
  for (uint32_t t = 0; t < m; t++) {
    if (d->hit_idxs) std::memcpy(d->hit_idxs + 14 * (size_t)t, hits.data() + 14 * (size_t)t, 56);
    if (d->len) d->len[t] = len[t];
    if (d->seed_idx) d->seed_idx[t] = seed[t];
    if (d->type) d->type[t] = type[t];
  }
  return n > d->capacity ? (lst_status)(st | LST_W_TRUNCATED) : st;
}

// waits for the set of frame f and queues the copies of the records the callers take, device block to
// pinned block behind the header that is already there (transfer stream, then the frame's outDone event)
lst_status collect(lst_batch* b, int f, uint32_t n, lst_output* const* outs, uint32_t outCap) {
  HIP_OK(hipEventSynchronize(b->computeDone[f]));
  const size_t ow = b->out_words(outCap);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t* blk = b->outHost[f] + ow * i;
    if (!outs[i] || blk[lst::kOutBadHits] || blk[lst::kOutOverflow]) continue;
    const size_t rec = (size_t)std::min(blk[lst::kOutNTC], outCap) * lst::kOutRecWords;
    if (!rec) continue;
    HIP_OK(hipMemcpyAsync(blk + lst::kOutHeaderWords, b->outDev[f] + ow * i + lst::kOutHeaderWords, 4 * rec,
                          hipMemcpyDeviceToHost, b->xferStream));
  }
  HIP_OK(hipEventRecord(b->outDone[f], b->xferStream));
  return LST_OK;
}

// once frame f's record copies are done, hands the set's outputs over (slot i's to outs[i], which may be
// null); the later launch sets compute meanwhile, and the next one to use frame f is issued only after this
lst_status unpack(lst_batch* b, int f, const uint32_t* ids, uint32_t n, lst_output* const* outs, uint32_t outCap,
                  lst_status& agg) {
  HIP_OK(hipEventSynchronize(b->outDone[f]));
  const size_t ow = b->out_words(outCap);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t* blk = b->outHost[f] + ow * i;
    lst_output* d = outs[i];
    const PreparedInput& p = b->staged[ids[i]];
    lst_status s;
    if (blk[lst::kOutBadHits]) {
      const uint32_t bad = blk[lst::kOutBadHits];
      s = fail(LST_E_INPUT, "event " + std::to_string(ids[i]) +
                                (bad & lst::kBadDetId ? ": a hit's detId is no module of the geometry"
                                 : bad & lst::kBadModuleHits ? ": more than 127 hits on one module"
                                                              : ": the hits of a module are not contiguous"));
      if (d) d->n_tc = 0;
    } else if (blk[lst::kOutOverflow]) {
      s = rerun_event(b, p, d);
    } else {
      s = d ? copy_output(blk, d) : (blk[lst::kOutTruncated] ? (lst_status)LST_W_TRUNCATED : LST_OK);
    }
    if (p.truncated) s = (lst_status)(s | LST_W_TRUNCATED);
    if (d) d->status = (int32_t)s;
    if ((s & 0xff) != LST_OK && (agg & 0xff) == LST_OK) agg = (lst_status)((s & 0xff) | (agg & LST_W_TRUNCATED));
    else agg = (lst_status)(agg | (s & LST_W_TRUNCATED));
  }
  return LST_OK;
}

// events ids[0..n) through launch sets of b->slots events, two sets in flight
lst_status run_ids(lst_batch* b, const std::vector<uint32_t>& ids, lst_output* out) {
  uint32_t outCap = 0;
  size_t maxBytes = 0;
  uint32_t maxHits = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    if (out) outCap = std::max(outCap, std::min(out[i].capacity, kTCCap));
    maxBytes = std::max(maxBytes, b->staged[ids[i]].bytes);
    maxHits = std::max(maxHits, b->staged[ids[i]].nHits);
  }
  lst_status s = batch_arenas(b, maxHits);
  if (s != LST_OK) return s;
  if ((s = batch_io(b, maxBytes, outCap)) != LST_OK) return s;
  const uint32_t S = b->slots, nSets = (uint32_t)((ids.size() + S - 1) / S);
  // slot order inside each launch set: largest event first; the outputs go back to the callers' order
  const std::vector<uint32_t> pos = largest_first(b, ids.data(), ids.size(), S);
  std::vector<uint32_t> sid(ids.size());
  std::vector<lst_output*> sout(ids.size());
  for (size_t i = 0; i < ids.size(); i++) {
    sid[i] = ids[pos[i]];
    sout[i] = out ? &out[pos[i]] : nullptr;
  }
  lst_status agg = LST_OK;
  auto set_n = [&](uint32_t k) { return std::min<uint32_t>(S, (uint32_t)ids.size() - k * S); };
  auto collect_set = [&](uint32_t j) {
    return collect(b, (int)(j % lst_batch::kFrames), set_n(j), sout.data() + (size_t)j * S, outCap);
  };
  auto unpack_set = [&](uint32_t j) {
    return unpack(b, (int)(j % lst_batch::kFrames), sid.data() + (size_t)j * S, set_n(j), sout.data() + (size_t)j * S,
                  outCap, agg);
  };
  // launch set k (its inputs queued first, so that they cross while the earlier sets compute), then hand
  // over set k - d (d = sets in flight), whose outputs the host unpacks while the later sets compute; frame
  // k % 3 was last used by set k - 3, handed over before this iteration
  const uint32_t d = b->sets;
  for (uint32_t k = 0; k < nSets; k++) {
    if ((s = issue_inputs(b, (int)(k % lst_batch::kFrames), sid.data() + (size_t)k * S, set_n(k), outCap)) != LST_OK)
      return s;
    if ((s = launch_set(b, k)) != LST_OK) return s;
    if (k >= d && ((s = collect_set(k - d)) != LST_OK || (s = unpack_set(k - d)) != LST_OK)) return s;
  }
  for (uint32_t j = nSets > d ? nSets - d : 0; j < nSets; j++) {
    if ((s = collect_set(j)) != LST_OK) return s;
    if ((s = unpack_set(j)) != LST_OK) return s;
  }
  return agg;
}

lst_status stage_events(lst_batch* b, const lst_input* events, uint32_t n_ev) {
  if (b->sync_compute() != LST_OK) return fail(LST_E_DEVICE, "stream synchronize failed");
  HIP_OK(hipStreamSynchronize(b->xferStream));
  for (auto& p : b->staged) p.release();
  b->staged.assign(n_ev, PreparedInput());
  for (uint32_t i = 0; i < n_ev; i++) {
    lst_status s = prepare_input(events[i], b->staged[i]);
    if (s != LST_OK) return s;
  }
  return LST_OK;
}
}  // namespace

extern "C" {

const char* lst_tune_name(uint32_t k) { return k < lst::kNTune ? lst::kTuneNames[k] : nullptr; }
uint32_t lst_tune_count(void) { return lst::kNTune; }

lst_status lst_batch_create(lst_es* es, uint32_t n_slots, int t5_mode, lst_batch** out) {
  if (!out) return fail(LST_E_INPUT, "null argument");
  *out = nullptr;
  if (!es) return fail(LST_E_INPUT, "null event setup");
  if (n_slots == 0 || n_slots > lst::kMaxBatchSlots) return fail(LST_E_INPUT, "a batch holds 1..128 events per launch set");
  if (t5_mode != LST_T5_PLAIN_CHI2 && t5_mode != LST_T5_DNN) return fail(LST_E_INPUT, "unknown T5 selection mode");
  auto b = std::make_unique<lst_batch>();
  b->es = es;
  b->t5Mode = t5_mode;
  b->slots = n_slots;
  HIP_OK(hipStreamCreateWithFlags(&b->stream[0], hipStreamNonBlocking));
  b->stream[1] = b->stream[0];
  HIP_OK(hipStreamCreateWithFlags(&b->xferStream, hipStreamNonBlocking));
  for (uint32_t c = 0; c < n_slots; c++) {
    lst_event* ev = nullptr;
    lst_status s = lst_event_create(es, b->stream[0], t5_mode, &ev);
    if (s != LST_OK) {
      lst_batch_destroy(b.release());
      return s;
    }
    ev->v.noFallback = 1;  // uploaded with the batch's view array
    b->ctx.push_back(ev);
  }
  const char* g = std::getenv("LST_BATCH_GRAPHS");  // "0": launch every kernel instead of the captured graph
  b->graphs = !(g && g[0] == '0');
  *out = b.release();
  return LST_OK;
}

void lst_batch_destroy(lst_batch* b) {
  if (!b) return;
  (void)b->sync_compute();
  if (b->xferStream) (void)hipStreamSynchronize(b->xferStream);
  drop_graphs(b);
  for (lst_event* ev : b->ctx) lst_event_destroy(ev);
  if (b->rerun) lst_event_destroy(b->rerun);
  for (auto& p : b->staged) p.release();
  for (int k = 0; k < lst_batch::kFrames; k++) {
    if (b->inDev[k]) (void)hipFree(b->inDev[k]);
    if (b->descHost[k]) (void)hipHostFree(b->descHost[k]);
    if (b->outHost[k]) (void)hipHostFree(b->outHost[k]);
    if (b->outDev[k]) (void)hipFree(b->outDev[k]);
    if (b->copyDone[k]) (void)hipEventDestroy(b->copyDone[k]);
    if (b->computeDone[k]) (void)hipEventDestroy(b->computeDone[k]);
    if (b->outDone[k]) (void)hipEventDestroy(b->outDone[k]);
  }
  for (int k = 0; k < 2; k++)
    if (b->descDev[k]) (void)hipFree(b->descDev[k]);
  if (b->viewsDev[0]) (void)hipFree(b->viewsDev[0]);
  if (b->viewsDev[1] && b->viewsDev[1] != b->viewsDev[0]) (void)hipFree(b->viewsDev[1]);
  if (b->stream[1] && b->stream[1] != b->stream[0]) (void)hipStreamDestroy(b->stream[1]);
  if (b->stream[0]) (void)hipStreamDestroy(b->stream[0]);
  if (b->xferStream) (void)hipStreamDestroy(b->xferStream);
  delete b;
}

uint32_t lst_batch_contexts(const lst_batch* b) { return b ? b->slots : 0; }

lst_status lst_batch_set_sets_in_flight(lst_batch* b, uint32_t sets) {
  if (!b) return fail(LST_E_INPUT, "null batch");
  if (sets != 1 && sets != 2) return fail(LST_E_INPUT, "a batch keeps 1 or 2 launch sets in flight");
  if (sets == b->sets) return LST_OK;
  if (b->sync_compute() != LST_OK) return fail(LST_E_DEVICE, "stream synchronize failed");
  HIP_OK(hipStreamSynchronize(b->xferStream));
  drop_graphs(b);
  if (sets == 2) {  // a second set of contexts on a stream of its own; arenas come with the next run
    hipStream_t st = nullptr;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    std::vector<lst_event*> more;
    for (uint32_t c = 0; c < b->slots; c++) {
      lst_event* ev = nullptr;
      const lst_status s = lst_event_create(b->es, st, b->t5Mode, &ev);
      if (s != LST_OK) {
        for (lst_event* e : more) lst_event_destroy(e);
        (void)hipStreamDestroy(st);
        return s;
      }
      ev->v.noFallback = 1;
      more.push_back(ev);
    }
    b->ctx.insert(b->ctx.end(), more.begin(), more.end());
    b->stream[1] = st;
    b->viewsDev[1] = nullptr;
  } else {
    for (size_t c = b->slots; c < b->ctx.size(); c++) lst_event_destroy(b->ctx[c]);
    b->ctx.resize(b->slots);
    if (b->viewsDev[1] && b->viewsDev[1] != b->viewsDev[0]) (void)hipFree(b->viewsDev[1]);
    b->viewsDev[1] = b->viewsDev[0];
    (void)hipStreamDestroy(b->stream[1]);
    b->stream[1] = b->stream[0];
  }
  b->sets = sets;
  return LST_OK;
}

lst_status lst_run_batch(lst_batch* b, const lst_input* events, uint32_t n_ev, lst_output* out) {
  if (!b || (n_ev && !events)) return fail(LST_E_INPUT, "null argument");
  lst_status s = stage_events(b, events, n_ev);
  if (s != LST_OK) return s;
  std::vector<uint32_t> ids(n_ev);
  for (uint32_t i = 0; i < n_ev; i++) ids[i] = i;
  return run_ids(b, ids, out);
}

lst_status lst_batch_stage(lst_batch* b, const lst_input* events, uint32_t n_ev) {
  if (!b || (n_ev && !events)) return fail(LST_E_INPUT, "null argument");
  lst_status s = stage_events(b, events, n_ev);
  if (s != LST_OK) return s;
  uint32_t maxHits = 0;
  size_t maxBytes = 0;
  for (const auto& p : b->staged) {
    maxHits = std::max(maxHits, p.nHits);
    maxBytes = std::max(maxBytes, p.bytes);
  }
  if ((s = batch_arenas(b, maxHits)) != LST_OK) return s;
  return batch_io(b, maxBytes, b->outCapAlloc);
}

lst_status lst_batch_run_staged(lst_batch* b, uint32_t first, uint32_t n_ev, lst_output* out) {
  if (!b) return fail(LST_E_INPUT, "null batch");
  const size_t ns = b->staged.size();
  if (ns == 0 || first >= ns) return fail(LST_E_INPUT, "batch slots not staged");
  std::vector<uint32_t> ids(n_ev);
  for (uint32_t i = 0; i < n_ev; i++) ids[i] = (uint32_t)((first + i) % ns);
  return run_ids(b, ids, out);
}

lst_status lst_batch_autotune(lst_batch* b, uint32_t first, uint32_t repeats, float* default_ms, float* tuned_ms) {
  if (!b) return fail(LST_E_INPUT, "null batch");
  const size_t ns = b->staged.size();
  if (ns == 0 || first >= ns) return fail(LST_E_INPUT, "batch slots not staged");
  if (repeats == 0) repeats = 1;
  std::vector<uint32_t> ids(std::min<size_t>(b->slots, ns));
  for (size_t i = 0; i < ids.size(); i++) ids[i] = (uint32_t)((first + i) % ns);
  uint32_t maxHits = 0;
  size_t maxBytes = 0;
  for (uint32_t id : ids) {
    maxHits = std::max(maxHits, b->staged[id].nHits);
    maxBytes = std::max(maxBytes, b->staged[id].bytes);
  }
  lst_status s = batch_arenas(b, maxHits);
  if (s != LST_OK) return s;
  if ((s = batch_io(b, maxBytes, b->outCapAlloc)) != LST_OK) return s;
  std::vector<hipEvent_t> ev(2 * lst::kNTune);
  for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  // every tunable kernel at once takes candidate count c (0: its default); the
  // kernels run one after the other on one stream, so each one's time depends
  // on its own count only
  std::vector<uint32_t> sid;  // the slot order of a run: largest event first
  for (uint32_t p : largest_first(b, ids.data(), ids.size(), b->slots)) sid.push_back(ids[p]);
  static const uint32_t cand[] = {0, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096};
  constexpr int nc = sizeof(cand) / sizeof(cand[0]);
  std::vector<float> best((size_t)lst::kNTune * nc, 1e30f);
  uint32_t saved[lst::kNTune];
  std::memcpy(saved, b->grid, sizeof(saved));
  uint32_t recordedAll[2] = {0, 0};
  lst_status st = LST_OK;
  for (int c = 0; c < nc && st == LST_OK; c++) {
    for (int k = 0; k < lst::kNTune; k++)  // the pair kernels' tuned count is of y shares, at most 256
      b->grid[k] = (k == lst::kTnPT3Pairs || k == lst::kTnCrossPT3) ? std::min(cand[c], 256u) : cand[c];
    for (uint32_t r = 0; r < repeats && st == LST_OK; r++) {
      if ((st = issue_inputs(b, 0, sid.data(), (uint32_t)sid.size(), 0)) != LST_OK) break;
      if ((st = stage_descriptors(b, 0, 0)) != LST_OK) break;
      uint32_t recorded[2] = {0, 0};
      enqueue_set(b, 0, ev.data(), recorded);
      if (hipStreamSynchronize(b->stream[0]) != hipSuccess || hipGetLastError() != hipSuccess) {
        st = fail(LST_E_DEVICE, "autotune run failed");
        break;
      }
      for (int k = 0; k < lst::kNTune; k++) {
        if (!((recorded[k / 32] >> (k % 32)) & 1u)) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) != hipSuccess) continue;
        float& bk = best[(size_t)k * nc + c];
        bk = std::min(bk, ms);
      }
      recordedAll[0] |= recorded[0];
      recordedAll[1] |= recorded[1];
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (st != LST_OK) {
    std::memcpy(b->grid, saved, sizeof(saved));
    return st;
  }
  float tDefault = 0.f, tTuned = 0.f;
  for (int k = 0; k < lst::kNTune; k++) {
    b->grid[k] = 0;
    if (!((recordedAll[k / 32] >> (k % 32)) & 1u)) continue;
    int bc = 0;
    for (int c = 1; c < nc; c++)
      if (best[(size_t)k * nc + c] < 0.97f * best[(size_t)k * nc + bc]) bc = c;  // keep the default within 3 %
    b->grid[k] = cand[bc];
    tDefault += best[(size_t)k * nc];
    tTuned += best[(size_t)k * nc + bc];
  }
  if (default_ms) *default_ms = tDefault;
  if (tuned_ms) *tuned_ms = tTuned;
  drop_graphs(b);
  return LST_OK;
}

lst_status lst_batch_kernel_times(lst_batch* b, uint32_t first, float* ms, uint32_t n) {
  if (!b || !ms) return fail(LST_E_INPUT, "null argument");
  const size_t ns = b->staged.size();
  if (ns == 0 || first >= ns) return fail(LST_E_INPUT, "batch slots not staged");
  std::vector<uint32_t> ids(std::min<size_t>(b->slots, ns));
  for (size_t i = 0; i < ids.size(); i++) ids[i] = (uint32_t)((first + i) % ns);
  uint32_t maxHits = 0;
  size_t maxBytes = 0;
  for (uint32_t id : ids) {
    maxHits = std::max(maxHits, b->staged[id].nHits);
    maxBytes = std::max(maxBytes, b->staged[id].bytes);
  }
  lst_status s = batch_arenas(b, maxHits);
  if (s != LST_OK) return s;
  if ((s = batch_io(b, maxBytes, b->outCapAlloc)) != LST_OK) return s;
  std::vector<hipEvent_t> ev(2 * lst::kNTune);
  for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  uint32_t recorded[2] = {0, 0};
  std::vector<uint32_t> sid;  // the slot order of a run: largest event first
  for (uint32_t p : largest_first(b, ids.data(), ids.size(), b->slots)) sid.push_back(ids[p]);
  s = issue_inputs(b, 0, sid.data(), (uint32_t)sid.size(), 0);
  if (s == LST_OK) s = stage_descriptors(b, 0, 0);
  if (s == LST_OK) {
    enqueue_set(b, 0, ev.data(), recorded);
    if (hipStreamSynchronize(b->stream[0]) != hipSuccess || hipGetLastError() != hipSuccess)
      s = fail(LST_E_DEVICE, "timed launch set failed");
  }
  for (uint32_t k = 0; k < n; k++) {
    ms[k] = 0.f;
    if (s == LST_OK && k < (uint32_t)lst::kNTune && ((recorded[k / 32] >> (k % 32)) & 1u))
      (void)hipEventElapsedTime(&ms[k], ev[2 * k], ev[2 * k + 1]);
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return s;
}

lst_status lst_batch_get_tuning(const lst_batch* b, uint32_t* grid, uint32_t n) {
  if (!b || !grid) return fail(LST_E_INPUT, "null argument");
  for (uint32_t k = 0; k < n && k < (uint32_t)lst::kNTune; k++) grid[k] = b->grid[k];
  return LST_OK;
}

lst_status lst_batch_set_tuning(lst_batch* b, const uint32_t* grid, uint32_t n) {
  if (!b || !grid) return fail(LST_E_INPUT, "null argument");
  if (b->sync_compute() != LST_OK) return fail(LST_E_DEVICE, "stream synchronize failed");
  for (uint32_t k = 0; k < (uint32_t)lst::kNTune; k++) b->grid[k] = k < n ? grid[k] : 0u;
  drop_graphs(b);
  return LST_OK;
}

}  // extern "C"

namespace lst {
// Work-group counts per event of the grid-stride kernels in batch launch sets
// (0: the kernel's one-event default), chosen by lst_batch_autotune on 64
// uneven PU50-250 events on an MI355X; the autotuner refines them per batch.
const uint32_t kBatchGrid[kNTune] = {
    128, 16, 128, 32, 32, 0, 32, 128, 16, 128, 32, 0,
    128, 64, 128, 128, 32, 128, 128, 16, 128, 128, 32, 128,
    64, 128, 0, 16, 128, 32, 256, 256, 16, 32, 0, 256,
    64, 512, 16, 0, 2048, 128, 256, 16, 512, 16, 32, 32,
    256, 256, 32, 0, 16, 256, 16, 256,
};
const char* const kTuneNames[kNTune] = {
    "k_stage_in", "k_event_init", "k_hits", "k_module_ranges", "k_hit_ranges_fix", "k_pixel_segments",
    "k_md_cand_fill", "k_md_eval", "k_md_scan_blocks", "k_md_place",
    "k_ls_cand_count", "k_ls_cand_fill", "k_ls_eval", "k_ls_scan_blocks", "k_ls_place",
    "k_md_seg_lists", "k_t3_cand_count", "k_t3_enum", "k_t3_select", "k_t3_scan_blocks", "k_t3_place",
    "k_md_t3_lists", "k_t5_cand_count", "k_t5_enum", "k_t5_filter_t4", "k_t5_final", "k_t5_dnn",
    "k_t5_scan_blocks", "k_t5_place", "k_t5_list_fill", "k_t5_dup_after_build", "k_pls_pairs",
    "k_pls_connected", "k_t3_list_fill", "k_pix_records",
    "k_pix_stage_a<PT5Kind>", "k_pix_stage_b<PT5Kind>", "k_pix_write<PT5Kind>", "k_pix_full_write<PT5Kind>",
    "k_pt5_bucket_count", "k_pt5_bucket_fill", "k_pt5_pairs",
    "k_pix_stage_a<PT3Kind>", "k_pix_stage_b<PT3Kind>", "k_pix_write<PT3Kind>", "k_pix_full_write<PT3Kind>",
    "k_pt3_pairs (y shares)", "k_cross_pt3_pairs (y shares)",
    "k_t5_tc_pairs", "k_cross_t5_pairs", "k_tc_t5_count", "k_tc_t5_write", "k_tc_prep", "k_cross_pls_pairs",
    "k_tc_output", "k_out_copy"};
}  // namespace lst
