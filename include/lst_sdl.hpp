// C++ facade of the MI355X LST engine with the reference's host interface:
// SDL::loadAndFillESHost / loadAndFillESDevice (SDL/LSTESData.h:79-80),
// SDL::Event (SDL/Event.h:26-205) and SDL::LST::run with its output getters
// (SDL/LST.h:15-54), over the C ABI of lst_amd.h.  Header only; link with
// -llst_amd.  The Alpaka queue argument becomes a hipStream_t passed as
// void* (nullptr: the engine's own stream).  Errors throw SDL::LSTError
// instead of printing or calling exit (SDL/LSTESData.cc:31-38).
#ifndef LST_SDL_HPP
#define LST_SDL_HPP

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lst_amd.h"

namespace SDL {

using QueueAcc = void*;  // hipStream_t

class LSTError : public std::runtime_error {
 public:
  LSTError(lst_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  lst_status status;
};

namespace detail {
inline lst_status check(lst_status s, const char* call) {
  if (s & 0xff) throw LSTError(s, std::string(call) + ": " + lst_last_error());
  return s;
}
inline const char* default_geometry() {
  const char* base = std::getenv("LST_GEOMETRY_DIR");
  return base ? base : "tracklooper_amd/data/OT800_IT615_pt0.8";
}
}  // namespace detail

// What loadAndFillESHost reads: here only where the geometry lies and the pT
// cut (the engine reads the files once, when the device data are made).
struct LSTESHostData {
  std::string geomDir;
  float ptCut = 0.8f;
};

// Geometry resident on the device, shared read-only by every event
// (LSTESDeviceData, SDL/LSTESData.h:43-77).
class LSTESDeviceData {
 public:
  explicit LSTESDeviceData(const LSTESHostData& h) {
    detail::check(lst_es_create(h.geomDir.c_str(), h.ptCut, 0, &es_), "lst_es_create");
    nModules = lst_es_scalar(es_, "nModules");
    nLowerModules = lst_es_scalar(es_, "nLowerModules");
    nPixels = lst_es_scalar(es_, "nPixels");
  }
  ~LSTESDeviceData() { lst_es_destroy(es_); }
  LSTESDeviceData(const LSTESDeviceData&) = delete;
  LSTESDeviceData& operator=(const LSTESDeviceData&) = delete;
  lst_es* handle() const { return es_; }
  uint16_t nModules = 0, nLowerModules = 0;
  unsigned int nPixels = 0;

 private:
  lst_es* es_ = nullptr;
};

// SDL/LSTESData.cc:40-76; the directory is a parameter (default: $LST_GEOMETRY_DIR or the in-tree pt0.8 data)
inline std::unique_ptr<LSTESHostData> loadAndFillESHost(const std::string& geomDir = detail::default_geometry(),
                                                        float ptCut = 0.8f) {
  return std::unique_ptr<LSTESHostData>(new LSTESHostData{geomDir, ptCut});
}
// SDL/LSTESData.cc:78-113
inline std::unique_ptr<LSTESDeviceData> loadAndFillESDevice(QueueAcc, const LSTESHostData* host) {
  return std::unique_ptr<LSTESDeviceData>(new LSTESDeviceData(*host));
}

// One event's device state and its stage methods, in the reference's order
// (SDL/Event.h:100-204).  t5Mode: LST_T5_PLAIN_CHI2 or LST_T5_DNN.
class Event {
 public:
  Event(bool verbose, QueueAcc queue, const LSTESDeviceData* deviceESData, int t5Mode = LST_T5_PLAIN_CHI2)
      : verbose_(verbose) {
    detail::check(lst_event_create(deviceESData->handle(), queue, t5Mode, &ev_), "lst_event_create");
  }
  ~Event() { lst_event_destroy(ev_); }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;

  void resetEvent() { detail::check(lst_event_reset(ev_), "resetEvent"); }

  void addHitToEvent(std::vector<float> x, std::vector<float> y, std::vector<float> z,
                     std::vector<unsigned int> detId, std::vector<unsigned int> idxInNtuple) {
    detail::check(lst_event_add_hits(ev_, x.data(), y.data(), z.data(), detId.data(), idxInNtuple.data(),
                                     (uint32_t)x.size(), 0),
                  "addHitToEvent");
  }
  void addPixelSegmentToEvent(std::vector<unsigned int> hitIndices0, std::vector<unsigned int> hitIndices1,
                              std::vector<unsigned int> hitIndices2, std::vector<unsigned int> hitIndices3,
                              std::vector<float> dPhiChange, std::vector<float> ptIn, std::vector<float> ptErr,
                              std::vector<float> px, std::vector<float> py, std::vector<float> pz,
                              std::vector<float> eta, std::vector<float> etaErr, std::vector<float> phi,
                              std::vector<int> charge, std::vector<unsigned int> seedIdx, std::vector<int> superbin,
                              std::vector<int8_t> pixelType, std::vector<char> isQuad) {
    lst_pixel_segments p{hitIndices0.data(), hitIndices1.data(), hitIndices2.data(), hitIndices3.data(),
                         dPhiChange.data(), ptIn.data(), ptErr.data(), px.data(), py.data(), pz.data(), eta.data(),
                         etaErr.data(), phi.data(), charge.data(), seedIdx.data(), superbin.data(), pixelType.data(),
                         reinterpret_cast<const int8_t*>(isQuad.data()), (uint32_t)ptIn.size()};
    detail::check(lst_event_add_pixel_segments(ev_, &p, 0), "addPixelSegmentToEvent");
  }

  void createMiniDoublets() { detail::check(lst_event_create_mini_doublets(ev_), "createMiniDoublets"); }
  void createSegmentsWithModuleMap() { detail::check(lst_event_create_segments(ev_), "createSegmentsWithModuleMap"); }
  void createTriplets() { detail::check(lst_event_create_triplets(ev_), "createTriplets"); }
  void createQuintuplets() { detail::check(lst_event_create_quintuplets(ev_), "createQuintuplets"); }
  void pixelLineSegmentCleaning() {
    detail::check(lst_event_pixel_line_segment_cleaning(ev_), "pixelLineSegmentCleaning");
  }
  void createPixelQuintuplets() { detail::check(lst_event_create_pixel_quintuplets(ev_), "createPixelQuintuplets"); }
  void createPixelTriplets() { detail::check(lst_event_create_pixel_triplets(ev_), "createPixelTriplets"); }
  void createTrackCandidates() { detail::check(lst_event_create_track_candidates(ev_), "createTrackCandidates"); }
  // waits for the event's stream; LSTError on bad input found by the device-side checks
  void wait() { detail::check(lst_event_synchronize(ev_), "wait"); }

  // the hits added, pixel pseudo-hits included (the reference's per-layer hit counters behind this getter
  // are reset but never filled, SDL/Event.cc:20-60,1264-1272)
  unsigned int getNumberOfHits() { return lst_event_count(ev_, "hits"); }
  unsigned int getNumberOfMiniDoublets() { return lst_event_count(ev_, "miniDoublets"); }
  unsigned int getNumberOfSegments() { return lst_event_count(ev_, "segments"); }
  unsigned int getNumberOfTriplets() { return lst_event_count(ev_, "triplets"); }
  unsigned int getNumberOfQuintuplets() { return lst_event_count(ev_, "quintuplets"); }
  int getNumberOfPixelTriplets() { return (int)lst_event_count(ev_, "pixelTriplets"); }
  int getNumberOfPixelQuintuplets() { return (int)lst_event_count(ev_, "pixelQuintuplets"); }
  int getNumberOfTrackCandidates() { return (int)lst_event_count(ev_, "trackCandidates"); }
  int getNumberOfPT5TrackCandidates() { return (int)lst_event_count(ev_, "pT5TrackCandidates"); }
  int getNumberOfPT3TrackCandidates() { return (int)lst_event_count(ev_, "pT3TrackCandidates"); }
  int getNumberOfT5TrackCandidates() { return (int)lst_event_count(ev_, "T5TrackCandidates"); }
  int getNumberOfPLSTrackCandidates() { return (int)lst_event_count(ev_, "pLSTrackCandidates"); }
  int getNumberOfPixelTrackCandidates() {
    return getNumberOfTrackCandidates() - getNumberOfT5TrackCandidates();  // SDL/Event.cc:1464-1476
  }

  // A host copy of one SoA field ("mds.anchorHitIndices", "t5.innerRadius", ...), the field-by-field
  // form of get{Hits,MiniDoublets,...}() (SDL/Event.h:185-204); T must match the field's type.
  template <class T>
  std::vector<T> getField(const std::string& name) {
    int dtype = 0;
    size_t count = 0;
    detail::check(lst_event_get_field(ev_, name.c_str(), nullptr, 0, &dtype, &count), "getField");
    std::vector<T> out(count);
    detail::check(lst_event_get_field(ev_, name.c_str(), out.data(), out.size() * sizeof(T), &dtype, &count),
                  "getField");
    return out;
  }

  lst_event* handle() const { return ev_; }
  bool verbose() const { return verbose_; }

 private:
  lst_event* ev_ = nullptr;
  bool verbose_ = false;
};

// SDL::LST<Acc> (SDL/LST.h:15-80): run() takes the seed and hit vectors by
// value as the reference does, and keeps the four output lists.
class LST {
 public:
  LST() = default;

  void run(QueueAcc& queue, bool verbose, const LSTESDeviceData* deviceESData, const std::vector<float> see_px,
           const std::vector<float> see_py, const std::vector<float> see_pz, const std::vector<float> see_dxy,
           const std::vector<float> see_dz, const std::vector<float> see_ptErr, const std::vector<float> see_etaErr,
           const std::vector<float> see_stateTrajGlbX, const std::vector<float> see_stateTrajGlbY,
           const std::vector<float> see_stateTrajGlbZ, const std::vector<float> see_stateTrajGlbPx,
           const std::vector<float> see_stateTrajGlbPy, const std::vector<float> see_stateTrajGlbPz,
           const std::vector<int> see_q, const std::vector<std::vector<int>> see_hitIdx,
           const std::vector<unsigned int> ph2_detId, const std::vector<float> ph2_x,
           const std::vector<float> ph2_y, const std::vector<float> ph2_z) {
    if (!event_ || es_ != deviceESData || queue_ != queue) {  // one event handle per (geometry, queue)
      event_.reset(new Event(verbose, queue, deviceESData));
      es_ = deviceESData;
      queue_ = queue;
    }
    std::vector<int32_t> flat;
    std::vector<uint32_t> off{0};
    for (const auto& s : see_hitIdx) {
      flat.insert(flat.end(), s.begin(), s.end());
      off.push_back((uint32_t)flat.size());
    }
    const lst_input in{see_px.data(), see_py.data(), see_pz.data(), see_dxy.data(), see_dz.data(),
                       see_ptErr.data(), see_etaErr.data(), see_stateTrajGlbX.data(), see_stateTrajGlbY.data(),
                       see_stateTrajGlbZ.data(), see_stateTrajGlbPx.data(), see_stateTrajGlbPy.data(),
                       see_stateTrajGlbPz.data(), see_q.data(), flat.data(), off.data(), (uint32_t)see_px.size(),
                       ph2_detId.data(), ph2_x.data(), ph2_y.data(), ph2_z.data(), (uint32_t)ph2_x.size()};
    lst_event* ev = event_->handle();
    detail::check(lst_run(ev, &in, 9), "LST::run");
    uint32_t n = 0;
    detail::check(lst_event_get_output(ev, &n, nullptr, nullptr, nullptr, nullptr, 0), "getOutput");
    std::vector<uint32_t> h(14 * (size_t)n), len(n);
    std::vector<int32_t> seed(n);
    std::vector<int16_t> type(n);
    if (n) detail::check(lst_event_get_output(ev, &n, h.data(), len.data(), seed.data(), type.data(), n), "getOutput");
    out_tc_hitIdxs_.assign(n, {});
    for (uint32_t i = 0; i < n; i++) out_tc_hitIdxs_[i].assign(h.begin() + 14 * (size_t)i, h.begin() + 14 * (size_t)i + len[i]);
    out_tc_len_.assign(len.begin(), len.end());
    out_tc_seedIdx_.assign(seed.begin(), seed.end());
    out_tc_trackCandidateType_.assign(type.begin(), type.end());
  }
  std::vector<std::vector<unsigned int>> hits() { return out_tc_hitIdxs_; }
  std::vector<unsigned int> len() { return out_tc_len_; }
  std::vector<int> seedIdx() { return out_tc_seedIdx_; }
  std::vector<short> trackCandidateType() { return out_tc_trackCandidateType_; }

 private:
  std::unique_ptr<Event> event_;
  const LSTESDeviceData* es_ = nullptr;
  QueueAcc queue_ = nullptr;
  std::vector<std::vector<unsigned int>> out_tc_hitIdxs_;
  std::vector<unsigned int> out_tc_len_;
  std::vector<int> out_tc_seedIdx_;
  std::vector<short> out_tc_trackCandidateType_;
};

}  // namespace SDL

#endif  // LST_SDL_HPP
