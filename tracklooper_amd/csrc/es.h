// Event-setup (geometry) data of the LST engine: host tables built once per
// job from the geometry directory, then uploaded to HBM and shared read-only
// by every event (the reference shares LSTESDeviceData the same way,
// SDL/LSTESData.h:43-77).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "lst_types.h"

namespace lst {

struct HostGeometry {
  uint32_t nModules = 0, nLowerModules = 0, nPixels = 0, stride = 0;
  uint16_t pixelModuleIndex = 0;
  std::vector<uint32_t> detIds, mapdetId, connectedPixels, geoMapDetId;
  // open-addressing table detId -> (module index | (endcap-map index + 1) << 16),
  // 2 words per slot, 2^detHashBits slots, empty slots hold kNoDetId
  std::vector<uint32_t> detHash;
  uint32_t detHashBits = 0;
  // module index of a detId, -1 if the geometry has no such module
  int module_of(uint32_t det) const {
    const uint32_t mask = (1u << detHashBits) - 1;
    uint32_t s = det_hash_slot(det, detHashBits);
    for (uint32_t n = 0; n <= mask; n++, s = (s + 1) & mask) {
      if (detHash[2 * s] == det) return (int)(detHash[2 * s + 1] & 0xFFFFu);
      if (detHash[2 * s] == kNoDetId) break;
    }
    return -1;
  }
  std::vector<uint16_t> moduleMap, mapIdx, nConnectedModules, partnerModuleIndices;
  std::vector<float> drdzs, dxdys, eta, r, geoMapPhi;
  std::vector<int16_t> layers, rings, modules, rods, subdets, sides;
  std::vector<uint8_t> isInverted, isLower, isAnchor, moduleType, moduleLayerType;
  std::vector<int32_t> sdlLayers;
  // per module, the module-only terms of the strip-hit shift (SDL/MiniDoublet.h:403-567): angle B (pi/2 in
  // the endcap, else atan(drdz)) and, per quadrant of the pixel-side hit, sin and cos of the slope angle
  std::vector<float> shiftAngleB, shiftSinM, shiftCosM;
  std::vector<float> modCtx, modTerms;  // 4 and 8 floats per module (fill_module_terms)
  std::vector<float> sigRot;  // 8 floats per module: cos, sin of the point rotation per quadrant
  // pLS superbin -> connected-module lists (SDL/PixelMap.h:16-36); the pos/neg
  // index arrays already include the offsets the reference adds per event
  // (SDL/Event.cc:813-816), so a pLS only needs (superbin, pixelType).
  std::vector<uint32_t> pixIndex[3], pixSizes[3];
  // per-lower-module capacity tables that depend on geometry only; the MD and
  // segment tables carry the pixel module as entry nLowerModules
  std::vector<int32_t> mdOccupancy, segOccupancy;      // [nLower + 1]
  std::vector<int32_t> mdIndices, segIndices;          // exclusive scans, [nLower + 1]
  std::vector<int32_t> t3OccIfAny;                     // [nLower]
  // T5 capacity of a module that holds triplets; -1 where the module can never
  // seed a T5 (barrel layers 3-6, endcap disks 2-5)
  std::vector<int32_t> t5OccIfAny;                     // [nLower]
};

// Loads the reference file formats (SDL/ModuleMethods.h:191-226,
// SDL/ModuleConnectionMap.cc:9-50, SDL/EndcapGeometry.cc:19-41,
// SDL/TiltedGeometry.cc:5-31).  stride = 0 picks the maximum fan-out.
bool load_host_geometry(const std::string& dir, uint32_t stride, HostGeometry& g, std::string& err);

struct DeviceGeometry {
  ModulesView view{};
  uint16_t pixelModuleIndex = 0;
  uint32_t nPixels = 0;
  void* arena = nullptr;
  size_t bytes = 0;
  // device copies of the pLS map and capacity tables
  const uint32_t* pixIndex[3] = {nullptr, nullptr, nullptr};
  const uint32_t* pixSizes[3] = {nullptr, nullptr, nullptr};
  const int32_t* mdOccupancy = nullptr;
  const int32_t* segOccupancy = nullptr;
  const int32_t* mdIndices = nullptr;
  const int32_t* segIndices = nullptr;
  const int32_t* t3OccIfAny = nullptr;
  const int32_t* t5OccIfAny = nullptr;
  int32_t nTotalMDCap = 0, nTotalSegCap = 0;  // outer-tracker totals (pixel excluded)
  int64_t t3CapBound = 0;                      // T3 capacity if every module has segments
  int64_t t5CapBound = 0;                      // T5 capacity if every T5 seed module has triplets
  uint32_t nT5SeedModules = 0;                 // modules with t5OccIfAny >= 0
};

// the per-module context records and cut terms (module_terms.h) for a pT cut
bool fill_module_terms(HostGeometry& g, float ptCut, std::string& err);
bool upload_geometry(const HostGeometry& h, DeviceGeometry& d, std::string& err);
void free_geometry(DeviceGeometry& d);

}  // namespace lst
