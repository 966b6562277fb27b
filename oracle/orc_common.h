// ORACLE — test infrastructure only. Never linked into the product library.
//
// Serial CPU restatement of the Line Segment Tracking chain of
// SegmentLinking/TrackLooper (SDL/*.h, SDL/Event.cc, SDL/LST.cc), used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// parity checker for the HIP engine.  Every function cites the reference
// file:line it restates.  Loops run in the order of the reference's CPU
// backend (createWorkDiv forces 1 block x 1 thread, SDL/Constants.h:88-92),
// so object slot numbering equals the reference CPU numbering.
//
// Deliberate, documented deviations from the reference build:
//  (1) IEEE arithmetic, -ffp-contract=off, no -Ofast (reference: SDL/Makefile:47).
//  (2) Transcendentals on float arguments are evaluated in double precision and
//      rounded once to float (correctly rounded except with probability ~1e-9
//      per call); the reference calls the float libm through alpaka::math.
//  (3) Geometry directory, pT cut and the module-map row stride are
//      parameters (reference hard-codes pt0.8 and stride 40,
//      SDL/LSTESData.cc:46-57, SDL/Constants.h:108).
//  (4) Uninitialised reference buffers (2S strip-edge coordinates of hits
//      that are not on endcap 2S modules, SDL/Hit.h:250-259) are zero here.
#pragma once
#include <cstdint>
#include <cmath>
#include <vector>
#include <string>
#include <map>

namespace orc {

// ---- capacities and physics constants: SDL/Constants.h:102-139 -------------
constexpr unsigned kMaxBlocks = 80;
constexpr unsigned kNMaxPixelSegments = 50000;                 // :112
constexpr unsigned kNMaxPixelMDs = 2 * kNMaxPixelSegments;     // :114
constexpr unsigned kNMaxPixelTriplets = 5000;                  // :116
constexpr unsigned kNMaxPixelQuintuplets = 15000;              // :117
constexpr unsigned kNMaxPixelTrackCandidates = 30000;          // :119
constexpr unsigned kNMaxNonPixelTrackCandidates = 1000;        // :120
constexpr unsigned kSizeSuperbins = 45000;                     // :122

static const float miniMulsPtScaleBarrel[6] = {0.0052, 0.0038, 0.0034, 0.0034, 0.0032, 0.0034};
static const float miniMulsPtScaleEndcap[5] = {0.006, 0.006, 0.006, 0.006, 0.006};
static const float miniRminMeanBarrel[6] = {25.007152356, 37.2186993757, 52.3104270826,
                                            68.6658656666, 85.9770373007, 108.301772384};
static const float miniRminMeanEndcap[5] = {130.992832231, 154.813883559, 185.352604327,
                                            221.635123002, 265.022076742};
static const float k2Rinv1GeVf = (2.99792458e-3 * 3.8) / 2;
static const float kR1GeVf = 1. / (2.99792458e-3 * 3.8);
static const float sinAlphaMax = 0.95;
static const float deltaZLum = 15.0;
static const float pixelPSZpitch = 0.15;
static const float strip2SZpitch = 5.0;
static const float pt_betaMax = 7.0;
static const float magnetic_field = 3.8112;
static const float SDL_INF = 123456789;
static const float kPi = float(M_PI);

enum SubDet { InnerPixel = 0, Barrel = 5, Endcap = 4 };      // SDL/Module.h:14
enum Side { NegZ = 1, PosZ = 2, Center = 3 };                // :16
enum ModuleType { PS = 0, TwoS = 1, PixelModule = 2 };        // :18
enum ModuleLayerType { Pixel = 0, Strip = 1, InnerPixelLayer = 2 };  // :20

// ---- float math policy (deviation 2) ----------------------------------------
inline float fsin(float x) { return (float)std::sin((double)x); }
inline float fcos(float x) { return (float)std::cos((double)x); }
inline float ftan(float x) { return (float)std::tan((double)x); }
inline float fasin(float x) { return (float)std::asin((double)x); }
inline float facos(float x) { return (float)std::acos((double)x); }
inline float fatan(float x) { return (float)std::atan((double)x); }
inline float fatan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline float facosh(float x) { return (float)std::acosh((double)x); }
inline float fsinh(float x) { return (float)std::sinh((double)x); }
inline float fexp(float x) { return (float)std::exp((double)x); }
inline float flog(float x) { return (float)std::log((double)x); }
inline float fsqrt(float x) { return std::sqrt(x); }  // IEEE, correctly rounded
inline float fabsf_(float x) { return std::fabs(x); }
inline float fminf_(float a, float b) { return std::fmin(a, b); }
inline float fmaxf_(float a, float b) { return std::fmax(a, b); }

// ---- SDL/Hit.h:111-192 -------------------------------------------------------
inline float temp_log10(float v) {  // Hit.h:111-114
  constexpr float ln10 = 2.302585093f;
  return flog(v) / ln10;
}
inline float eta(float x, float y, float z) {  // Hit.h:117-122
  float r3 = fsqrt(x * x + y * y + z * z);
  float rt = fsqrt(x * x + y * y);
  return ((z > 0) - (z < 0)) * facosh(r3 / rt);
}
inline float phi_mpi_pi(float x) {  // Hit.h:125-132
  if (std::fabs(x) <= kPi) return x;
  constexpr float o2pi = 1.f / (2.f * float(M_PI));
  float n = std::round(x * o2pi);
  return x - n * float(2.f * float(M_PI));
}
inline float phi(float x, float y) { return phi_mpi_pi(kPi + fatan2(-y, -x)); }  // Hit.h:135-137
inline float deltaPhi(float x1, float y1, float x2, float y2) {  // Hit.h:140-144
  float p1 = phi(x1, y1);
  float p2 = phi(x2, y2);
  return phi_mpi_pi(p2 - p1);
}
inline float deltaPhiChange(float x1, float y1, float x2, float y2) {  // Hit.h:147-149
  return deltaPhi(x1, y1, x2 - x1, y2 - y1);
}
inline float copysignf_(float a, float b) {  // Hit.h:152-156
  int sa = (a < 0) ? -1 : 1;
  int sb = (b < 0) ? -1 : 1;
  return sa * sb * a;
}
inline float calculate_dPhi(float p1, float p2) {  // Hit.h:158-170
  float d = p1 - p2;
  if (d > kPi) d -= 2 * kPi;
  else if (d < -kPi) d += 2 * kPi;
  return d;
}
inline int binary_search(const unsigned* data, unsigned v, unsigned n) {  // Hit.h:172-192
  unsigned lo = 0, hi = n - 1;
  while (lo <= hi) {
    unsigned mid = (lo + hi) / 2;
    unsigned t = data[mid];
    if (t == v) return mid;
    else if (t > v) hi = mid - 1;
    else lo = mid + 1;
  }
  return -1;
}

// ---- geometry (event setup) ---------------------------------------------------
struct Geometry {
  unsigned nModules = 0, nLowerModules = 0, nPixels = 0, stride = 40;
  std::vector<unsigned> detIds, mapdetId, connectedPixels;
  std::vector<uint16_t> moduleMap, mapIdx, nConnectedModules, partnerModuleIndices;
  std::vector<float> drdzs, dxdys, eta, r;
  std::vector<short> layers, rings, modules, rods, subdets, sides;
  std::vector<uint8_t> isInverted, isLower, isAnchor;
  std::vector<int> moduleType, moduleLayerType, sdlLayers;
  // pixel map (SDL/PixelMap.h:16-36)
  uint16_t pixelModuleIndex = 0;
  std::vector<unsigned> cpIndex, cpSizes, cpIndexPos, cpSizesPos, cpIndexNeg, cpSizesNeg;
  // endcap 2S orientation map (SDL/EndcapGeometry.cc:43-70)
  std::vector<unsigned> geoMapDetId;
  std::vector<float> geoMapPhi;
};

bool load_geometry(const std::string& dir, unsigned stride, Geometry& g, std::string& err);

}  // namespace orc
