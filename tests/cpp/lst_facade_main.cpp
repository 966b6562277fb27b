// Drives the C++ facade (include/lst_sdl.hpp) the way the reference's callers
// drive SDL::LST and SDL::Event: reads one event's LST::run arguments and its
// prepared inputs from binary files, runs SDL::LST::run and the SDL::Event
// stage methods, and writes the output lists and object counts.
//
//   lst_facade_main <geometry dir> <run args file> <prepared inputs file> <output file>
//
// Both input files are little-endian blocks written by tests/test_facade.py.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "lst_sdl.hpp"

namespace {
struct Reader {
  std::ifstream f;
  explicit Reader(const char* path) : f(path, std::ios::binary) {
    if (!f) throw std::runtime_error(std::string("cannot read ") + path);
  }
  uint32_t u32() {
    uint32_t v = 0;
    f.read(reinterpret_cast<char*>(&v), 4);
    return v;
  }
  template <class T>
  std::vector<T> vec(size_t n) {
    std::vector<T> v(n);
    if (n) f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
    if (!f) throw std::runtime_error("short input file");
    return v;
  }
};
}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    std::cerr << "usage: " << argv[0] << " <geometry dir> <run args> <prepared inputs> <output>\n";
    return 2;
  }
  try {
    auto host = SDL::loadAndFillESHost(argv[1], 0.8f);
    SDL::QueueAcc queue = nullptr;
    auto es = SDL::loadAndFillESDevice(queue, host.get());

    // SDL::LST::run on the raw seeds and hits
    Reader r(argv[2]);
    const uint32_t nSeeds = r.u32(), nHits = r.u32();
    std::vector<std::vector<float>> sf;
    for (int k = 0; k < 13; k++) sf.push_back(r.vec<float>(nSeeds));
    const auto q = r.vec<int>(nSeeds);
    const auto off = r.vec<uint32_t>(nSeeds + 1);
    const auto flat = r.vec<int>(off.back());
    std::vector<std::vector<int>> hitIdx(nSeeds);
    for (uint32_t s = 0; s < nSeeds; s++) hitIdx[s].assign(flat.begin() + off[s], flat.begin() + off[s + 1]);
    const auto det = r.vec<unsigned int>(nHits);
    const auto x = r.vec<float>(nHits), y = r.vec<float>(nHits), z = r.vec<float>(nHits);
    SDL::LST lst;
    lst.run(queue, false, es.get(), sf[0], sf[1], sf[2], sf[3], sf[4], sf[5], sf[6], sf[7], sf[8], sf[9], sf[10],
            sf[11], sf[12], q, hitIdx, det, x, y, z);

    // SDL::Event stage by stage on the prepared inputs
    Reader p(argv[3]);
    const uint32_t nH = p.u32(), nP = p.u32();
    auto hx = p.vec<float>(nH), hy = p.vec<float>(nH), hz = p.vec<float>(nH);
    auto hdet = p.vec<unsigned int>(nH), hidx = p.vec<unsigned int>(nH);
    std::vector<std::vector<unsigned int>> h4;
    for (int k = 0; k < 4; k++) h4.push_back(p.vec<unsigned int>(nP));
    std::vector<std::vector<float>> pf;
    for (int k = 0; k < 9; k++) pf.push_back(p.vec<float>(nP));  // dPhiChange ptIn ptErr px py pz eta etaErr phi
    auto charge = p.vec<int>(nP);
    auto seedIdx = p.vec<unsigned int>(nP);
    auto superbin = p.vec<int>(nP);
    auto pixelType = p.vec<int8_t>(nP);
    auto isQuad = p.vec<char>(nP);
    SDL::Event ev(false, queue, es.get());
    ev.addHitToEvent(hx, hy, hz, hdet, hidx);
    ev.addPixelSegmentToEvent(h4[0], h4[1], h4[2], h4[3], pf[0], pf[1], pf[2], pf[3], pf[4], pf[5], pf[6], pf[7],
                              pf[8], charge, seedIdx, superbin, pixelType, isQuad);
    bool orderError = false;
    try {
      ev.createTriplets();  // out of order
    } catch (const SDL::LSTError& e) {
      orderError = (e.status & 0xff) == LST_E_ORDER;
    }
    ev.createMiniDoublets();
    ev.createSegmentsWithModuleMap();
    ev.createTriplets();
    ev.createQuintuplets();
    ev.pixelLineSegmentCleaning();
    ev.createPixelQuintuplets();
    ev.createPixelTriplets();
    ev.createTrackCandidates();
    ev.wait();

    std::ofstream o(argv[4]);
    o << "order_error " << orderError << "\n";
    o << "counts " << ev.getNumberOfHits() << " " << ev.getNumberOfMiniDoublets() << " " << ev.getNumberOfSegments()
      << " " << ev.getNumberOfTriplets() << " " << ev.getNumberOfQuintuplets() << " "
      << ev.getNumberOfPixelQuintuplets() << " " << ev.getNumberOfPixelTriplets() << " "
      << ev.getNumberOfTrackCandidates() << " " << ev.getNumberOfPixelTrackCandidates() << "\n";
    const auto mdHits = ev.getField<uint32_t>("mds.anchorHitIndices");
    o << "md_anchor_field " << mdHits.size() << "\n";
    const auto hits = lst.hits();
    const auto len = lst.len();
    const auto seed = lst.seedIdx();
    const auto type = lst.trackCandidateType();
    o << "n_tc " << hits.size() << "\n";
    for (size_t i = 0; i < hits.size(); i++) {
      o << "tc " << type[i] << " " << seed[i] << " " << len[i];
      for (unsigned h : hits[i]) o << " " << h;
      o << "\n";
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
