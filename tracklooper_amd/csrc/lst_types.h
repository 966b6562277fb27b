// Shared definitions of the MI355X LST engine: capacities, physics
// constants, module enums and the device-side SoA views.  The per-module
// object layout (objects of lower module m live in
// [moduleIndices[m], moduleIndices[m] + n[m])) is the reference layout
// (SDL/Module.h:21-85) so that object indices are interchangeable with the
// reference's.
#pragma once
#include <cstdint>
#include "lst_math.h"

namespace lst {

// SDL/Constants.h:107-122
constexpr unsigned kNMaxPixelSegments = 50000;
constexpr unsigned kNMaxPixelMDs = 2 * kNMaxPixelSegments;
constexpr unsigned kNMaxPixelTriplets = 5000;
constexpr unsigned kNMaxPixelQuintuplets = 15000;
constexpr unsigned kNMaxPixelTrackCandidates = 30000;
constexpr unsigned kNMaxNonPixelTrackCandidates = 1000;
constexpr unsigned kSizeSuperbins = 45000;

// SDL/Constants.h:124-139 (same float roundings of the same literals)
#define LST_CONST static constexpr
// per-layer tables live in __constant__ memory in the kernel translation unit
LST_CONST float k2Rinv1GeVf = (2.99792458e-3 * 3.8) / 2;
LST_CONST float kR1GeVf = 1. / (2.99792458e-3 * 3.8);
LST_CONST float kSinAlphaMax = 0.95;
LST_CONST float kDeltaZLum = 15.0;
LST_CONST float kPixelPSZpitch = 0.15;
LST_CONST float kStrip2SZpitch = 5.0;
LST_CONST float kPtBetaMax = 7.0;
LST_CONST float kMagneticField = 3.8112;
LST_CONST float kSdlInf = 123456789;

enum : int16_t { kInnerPixel = 0, kBarrel = 5, kEndcap = 4 };  // SubDet
enum : int16_t { kNegZ = 1, kPosZ = 2, kCenter = 3 };          // Side
enum : uint8_t { kPS = 0, kTwoS = 1, kPixelModule = 2 };       // ModuleType
enum : uint8_t { kPixelLayer = 0, kStripLayer = 1, kInnerPixelLayer = 2 };  // ModuleLayerType

// Device view of the geometry (read-only, resident in HBM for the job).
struct ModulesView {
  const uint32_t* detIds;
  const uint16_t* moduleMap;  // [nModules * stride]
  const uint32_t* mapdetId;   // sorted detIds
  const uint16_t* mapIdx;     // module index of mapdetId[i]
  const uint16_t* nConnectedModules;
  const float* drdzs;
  const float* dxdys;
  const uint16_t* partnerModuleIndices;
  const int16_t* layers;
  const int16_t* rings;
  const int16_t* modules;
  const int16_t* rods;
  const int16_t* subdets;
  const int16_t* sides;
  const float* eta;
  const float* r;
  const uint8_t* isInverted;
  const uint8_t* isLower;
  const uint8_t* isAnchor;
  const uint8_t* moduleType;
  const uint8_t* moduleLayerType;
  const int32_t* sdlLayers;
  const float* shiftAngleB;  // [nModules] strip-hit shift terms of the module (es.cpp)
  const float* shiftSinM;    // [4 nModules] per quadrant of the pixel-side hit
  const float* shiftCosM;
  const float4* modCtx;    // [nModules] packed attributes, drdz, dxdy, partner drdz (module_terms.h)
  const float4* modTerms;  // [2 nModules] module-only terms of the segment and mini-doublet cuts
  const float2* sigRot;    // [4 nModules] cos, sin of the r-phi regression's point rotation per quadrant
  const uint32_t* connectedPixels;
  const uint32_t* geoMapDetId;  // endcap 2S orientation map, sorted
  const float* geoMapPhi;
  // open-addressing table of the detIds (linear probing, never more than a
  // quarter full): slot (detId, module index | (endcap-map index + 1) << 16)
  const uint2* detHash;
  uint32_t nModules, nLowerModules, stride, nEndCapMap, detHashBits;
};
// a load issued as a global (not flat) memory load: p addresses device memory
// (vector types, 8 or 16 bytes, are loaded as one native vector)
template <class T>
__device__ __forceinline__ T ld_global(const T* p) {
  if constexpr (sizeof(T) == 16 || sizeof(T) == 8) {
    typedef unsigned V __attribute__((ext_vector_type(sizeof(T) / 4)));
    const V x = *(const __attribute__((address_space(1))) V*)p;
    T r;
    __builtin_memcpy(&r, &x, sizeof(T));
    return r;
  } else {
    return *(const __attribute__((address_space(1))) T*)p;
  }
}

constexpr uint32_t kNoDetId = 0xFFFFFFFFu;  // an empty slot of the detId table
__host__ __device__ inline uint32_t det_hash_slot(uint32_t det, uint32_t bits) {
  return (det * 0x9E3779B1u) >> (32 - bits);
}

}  // namespace lst
