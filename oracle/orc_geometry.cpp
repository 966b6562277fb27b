// ORACLE — test infrastructure only (see orc_common.h).
// Geometry ("event setup") loader restating SDL/ModuleMethods.h:168-416,
// SDL/ModuleConnectionMap.cc:9-50, SDL/EndcapGeometry.cc:3-85,
// SDL/TiltedGeometry.cc:5-49 and SDL/LSTESData.cc:40-113, with std::map
// semantics (iteration in ascending detId) kept on purpose.
#include "orc_common.h"
#include <fstream>
#include <array>

namespace orc {
namespace {

using ConnMap = std::map<unsigned, std::vector<unsigned>>;

// {u32 key, u32 n, n x u32} records: SDL/ModuleConnectionMap.cc:9-50
bool read_conn_map(const std::string& fn, ConnMap& m, std::string& err) {
  std::ifstream f(fn, std::ios::binary);
  if (!f) { err = "cannot open " + fn; return false; }
  while (true) {
    uint32_t hdr[2];
    if (!f.read(reinterpret_cast<char*>(hdr), 8)) break;
    std::vector<unsigned> v(hdr[1]);
    if (hdr[1] && !f.read(reinterpret_cast<char*>(v.data()), 4 * (size_t)hdr[1])) { err = "truncated " + fn; return false; }
    m[hdr[0]] = std::move(v);
  }
  return true;
}

template <class Rec>
bool read_records(const std::string& fn, std::vector<Rec>& out, std::string& err) {
  std::ifstream f(fn, std::ios::binary);
  if (!f) { err = "cannot open " + fn; return false; }
  Rec r;
  while (f.read(reinterpret_cast<char*>(&r), sizeof(Rec))) out.push_back(r);
  return true;
}

struct Centroid { uint32_t detId; float x, y, z; int32_t type; };
struct EndcapRec { uint32_t detId; float dxdy, phi; };
struct TiltRec { uint32_t detId; float drdz, dxdy; };

// SDL/Module.h:196-227
bool parseIsInverted(short subdet, short side, short module, short layer) {
  if (subdet == Endcap) {
    if (side == NegZ) return module % 2 == 1;
    if (side == PosZ) return module % 2 == 0;
    return false;
  }
  if (subdet == Barrel) {
    if (side == Center) {
      if (layer <= 3) return module % 2 == 1;
      if (layer >= 4) return module % 2 == 0;
      return false;
    }
    if (side == NegZ || side == PosZ) {
      if (layer <= 2) return module % 2 == 1;
      if (layer == 3) return module % 2 == 0;
      return false;
    }
  }
  return false;
}
bool parseIsLower(bool inv, unsigned detId) { return inv ? !(detId & 1) : (detId & 1); }
unsigned parsePartner(unsigned detId, bool lower, bool inv) {
  return lower ? (inv ? detId - 1 : detId + 1) : (inv ? detId + 1 : detId - 1);
}

}  // namespace

bool load_geometry(const std::string& dir, unsigned stride, Geometry& g, std::string& err) {
  g = Geometry();
  g.stride = stride;
  // SDL/LSTESData.cc:40-75
  std::vector<EndcapRec> endcap;
  std::vector<TiltRec> tilt;
  ConnMap conn;
  if (!read_records(dir + "/endcap_orientation.bin", endcap, err)) return false;
  if (!read_records(dir + "/tilted_barrel_orientation.bin", tilt, err)) return false;
  if (!read_conn_map(dir + "/module_connection_tracing_merged.bin", conn, err)) return false;
  std::map<unsigned, float> ec_dxdy, ec_phi, tb_drdz, tb_dxdy;
  for (auto& e : endcap) { ec_dxdy[e.detId] = e.dxdy; ec_phi[e.detId] = e.phi; }
  for (auto& t : tilt) { tb_drdz[t.detId] = t.drdz; tb_dxdy[t.detId] = t.dxdy; }
  const char* connects[4] = {"_layer1_subdet5", "_layer2_subdet5", "_layer1_subdet4", "_layer2_subdet4"};
  std::array<std::array<ConnMap, 4>, 3> pls;
  for (int i = 0; i < 4; i++) {
    if (!read_conn_map(dir + "/pixelmap/pLS_map" + connects[i] + ".bin", pls[0][i], err)) return false;
    if (!read_conn_map(dir + "/pixelmap/pLS_map_pos" + connects[i] + ".bin", pls[1][i], err)) return false;
    if (!read_conn_map(dir + "/pixelmap/pLS_map_neg" + connects[i] + ".bin", pls[2][i], err)) return false;
  }

  // loadCentroidsFromFile: SDL/ModuleMethods.h:191-226
  std::vector<Centroid> cents;
  if (!read_records(dir + "/sensor_centroids.bin", cents, err)) return false;
  std::map<unsigned, uint16_t> detIdToIndex;
  std::map<unsigned, float> mx, my, mz;
  std::map<unsigned, unsigned> mt;
  uint16_t counter = 0;
  for (auto& c : cents) {
    detIdToIndex[c.detId] = counter;
    mx[c.detId] = c.x; my[c.detId] = c.y; mz[c.detId] = c.z; mt[c.detId] = c.type;
    counter++;
  }
  detIdToIndex[1] = counter;
  counter++;
  const unsigned nModules = counter;
  g.nModules = nModules;
  g.nLowerModules = (nModules - 1) / 2;  // ModuleMethods.h:285
  auto resize = [&](auto& v) { v.assign(nModules, {}); };
  resize(g.detIds); resize(g.layers); resize(g.rings); resize(g.rods); resize(g.modules);
  resize(g.subdets); resize(g.sides); resize(g.eta); resize(g.r); resize(g.isInverted);
  resize(g.isLower); resize(g.isAnchor); resize(g.moduleType); resize(g.moduleLayerType);
  resize(g.dxdys); resize(g.drdzs); resize(g.partnerModuleIndices); resize(g.sdlLayers);

  // index layout lower | pixel | upper: ModuleMethods.h:284-365
  uint16_t lowerCounter = 0, upperCounter = g.nLowerModules + 1;
  for (auto it = detIdToIndex.begin(); it != detIdToIndex.end(); ++it) {
    unsigned detId = it->first;
    float m_x = mx[detId], m_y = my[detId], m_z = mz[detId];
    unsigned m_t = mt[detId];
    unsigned short layer, ring, rod, module, subdet, side;
    float eta_, r_;
    bool inv, low;
    if (detId == 1) {
      layer = ring = rod = module = subdet = side = 0;
      inv = low = false;
      eta_ = 0; r_ = 0;
    } else {
      // setDerivedQuantities: ModuleMethods.h:168-189
      subdet = (detId & (7 << 25)) >> 25;
      side = (subdet == Endcap) ? (detId & (3 << 23)) >> 23 : (detId & (3 << 18)) >> 18;
      layer = (subdet == Endcap) ? (detId & (7 << 18)) >> 18 : (detId & (7 << 20)) >> 20;
      ring = (subdet == Endcap) ? (detId & (15 << 12)) >> 12 : 0;
      module = (detId & (127 << 2)) >> 2;
      rod = (subdet == Endcap) ? 0 : (detId & (127 << 10)) >> 10;
      r_ = std::sqrt(m_x * m_x + m_y * m_y + m_z * m_z);
      eta_ = ((m_z > 0) - (m_z < 0)) * std::acosh(r_ / std::sqrt(m_x * m_x + m_y * m_y));
      inv = parseIsInverted(subdet, side, module, layer);
      low = parseIsLower(inv, detId);
    }
    uint16_t index;
    if (low) index = lowerCounter++;
    else if (detId != 1) index = upperCounter++;
    else index = g.nLowerModules;
    it->second = index;
    g.detIds[index] = detId; g.layers[index] = layer; g.rings[index] = ring; g.rods[index] = rod;
    g.modules[index] = module; g.subdets[index] = subdet; g.sides[index] = side;
    g.eta[index] = eta_; g.r[index] = r_; g.isInverted[index] = inv; g.isLower[index] = low;
    if (detId == 1) {
      g.moduleType[index] = PixelModule;
      g.moduleLayerType[index] = InnerPixelLayer;
      g.dxdys[index] = 0; g.drdzs[index] = 0; g.isAnchor[index] = false;
    } else {
      g.moduleType[index] = (m_t == 25 ? TwoS : PS);
      g.moduleLayerType[index] = (m_t == 23 ? Pixel : Strip);
      if (g.moduleType[index] == PS && g.moduleLayerType[index] == Pixel) g.isAnchor[index] = true;
      else if (g.moduleType[index] == TwoS && g.isLower[index]) g.isAnchor[index] = true;
      else g.isAnchor[index] = false;
      auto get = [](std::map<unsigned, float>& mm, unsigned k) { auto f = mm.find(k); return f == mm.end() ? 0.f : f->second; };
      g.dxdys[index] = (subdet == Endcap) ? get(ec_dxdy, detId) : get(tb_dxdy, detId);
      g.drdzs[index] = (subdet == Barrel) ? get(tb_drdz, detId) : 0;
    }
    g.sdlLayers[index] = layer + 6 * (subdet == Endcap) + 5 * (subdet == Endcap && g.moduleType[index] == TwoS);
  }
  // partners and drdz/dxdy back-fill: ModuleMethods.h:367-382 (ascending detId order)
  for (auto it = detIdToIndex.begin(); it != detIdToIndex.end(); ++it) {
    unsigned detId = it->first;
    uint16_t index = it->second;
    if (detId == 1) continue;
    auto pit = detIdToIndex.find(parsePartner(detId, g.isLower[index], g.isInverted[index]));
    if (pit == detIdToIndex.end()) { err = "partner module missing for detId " + std::to_string(detId); return false; }
    g.partnerModuleIndices[index] = pit->second;
    if (g.drdzs[index] == 0) g.drdzs[index] = g.drdzs[pit->second];
    if (g.dxdys[index] == 0) g.dxdys[index] = g.dxdys[pit->second];
  }

  // fillPixelMap: ModuleMethods.h:33-112
  g.pixelModuleIndex = detIdToIndex[1];
  std::vector<unsigned> cAll, cPos, cNeg;
  g.cpIndex.assign(kSizeSuperbins, 0); g.cpSizes.assign(kSizeSuperbins, 0);
  g.cpIndexPos.assign(kSizeSuperbins, 0); g.cpSizesPos.assign(kSizeSuperbins, 0);
  g.cpIndexNeg.assign(kSizeSuperbins, 0); g.cpSizesNeg.assign(kSizeSuperbins, 0);
  unsigned tot = 0, totPos = 0, totNeg = 0;
  auto lookup = [](const ConnMap& m, unsigned k) -> const std::vector<unsigned>* {
    static const std::vector<unsigned> empty;
    auto f = m.find(k);
    return f == m.end() ? &empty : &f->second;
  };
  for (unsigned sb = 0; sb < kSizeSuperbins; sb++) {
    unsigned s = 0;
    for (auto& m : pls[0]) { auto* v = lookup(m, sb + kSizeSuperbins); cAll.insert(cAll.end(), v->begin(), v->end()); s += v->size(); }
    g.cpIndex[sb] = tot; g.cpSizes[sb] = s; tot += s;
    s = 0;
    for (auto& m : pls[1]) { auto* v = lookup(m, sb); cPos.insert(cPos.end(), v->begin(), v->end()); s += v->size(); }
    g.cpIndexPos[sb] = totPos; g.cpSizesPos[sb] = s; totPos += s;
    s = 0;
    for (auto& m : pls[2]) { auto* v = lookup(m, sb); cNeg.insert(cNeg.end(), v->begin(), v->end()); s += v->size(); }
    g.cpIndexNeg[sb] = totNeg; g.cpSizesNeg[sb] = s; totNeg += s;
  }
  g.nPixels = tot + totPos + totNeg;
  g.connectedPixels.resize(g.nPixels);
  auto idxOf = [&](unsigned d) -> unsigned {
    auto f = detIdToIndex.find(d);
    return f == detIdToIndex.end() ? 0u : f->second;  // operator[] would insert index 0
  };
  for (unsigned i = 0; i < tot; i++) g.connectedPixels[i] = idxOf(cAll[i]);
  for (unsigned i = 0; i < totPos; i++) g.connectedPixels[tot + i] = idxOf(cPos[i]);
  for (unsigned i = 0; i < totNeg; i++) g.connectedPixels[tot + totPos + i] = idxOf(cNeg[i]);

  // fillConnectedModuleArrayExplicit: ModuleMethods.h:114-140
  g.moduleMap.assign((size_t)nModules * stride, 0);
  g.nConnectedModules.assign(nModules, 0);
  for (auto& kv : detIdToIndex) {
    auto f = conn.find(kv.first);
    if (f == conn.end()) continue;
    if (f->second.size() > stride) { err = "module fan-out exceeds module-map stride"; return false; }
    g.nConnectedModules[kv.second] = f->second.size();
    for (size_t i = 0; i < f->second.size(); i++) g.moduleMap[(size_t)kv.second * stride + i] = idxOf(f->second[i]);
  }
  // fillMapArraysExplicit: ModuleMethods.h:142-166
  for (auto& kv : detIdToIndex) { g.mapIdx.push_back(kv.second); g.mapdetId.push_back(kv.first); }
  // EndcapGeometry::fillGeoMapArraysExplicit: SDL/EndcapGeometry.cc:43-70
  for (auto& kv : ec_phi) { g.geoMapDetId.push_back(kv.first); g.geoMapPhi.push_back(kv.second); }
  return true;
}

}  // namespace orc
