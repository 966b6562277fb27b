// Geometry loading and upload for the LST engine.
#include "es.h"
#include "lst_math.h"
#include "module_terms.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <unordered_map>

namespace lst {
namespace {

std::vector<char> slurp(const std::string& fn, bool& ok) {
  std::ifstream f(fn, std::ios::binary | std::ios::ate);
  ok = bool(f);
  if (!ok) return {};
  std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<char> buf(n);
  ok = bool(f.read(buf.data(), n));
  return buf;
}

// Connection-map records {u32 key, u32 n, n x u32}; the last record for a key wins.
struct ConnTable {
  std::unordered_map<uint32_t, std::pair<size_t, uint32_t>> at;  // key -> (offset, count)
  std::vector<uint32_t> data;
  const uint32_t* find(uint32_t key, uint32_t& n) const {
    auto it = at.find(key);
    if (it == at.end()) { n = 0; return nullptr; }
    n = it->second.second;
    return data.data() + it->second.first;
  }
};

bool read_conn(const std::string& fn, ConnTable& t, std::string& err) {
  bool ok;
  auto buf = slurp(fn, ok);
  if (!ok) { err = "cannot read " + fn; return false; }
  const size_t nw = buf.size() / 4;
  t.data.resize(nw);
  std::memcpy(t.data.data(), buf.data(), nw * 4);
  size_t i = 0;
  while (i + 2 <= nw) {
    uint32_t key = t.data[i], n = t.data[i + 1];
    if (i + 2 + n > nw) { err = "truncated connection record in " + fn; return false; }
    t.at[key] = {i + 2, n};
    i += 2 + n;
  }
  return true;
}

template <class R>
bool read_fixed(const std::string& fn, std::vector<R>& out, std::string& err) {
  bool ok;
  auto buf = slurp(fn, ok);
  if (!ok) { err = "cannot read " + fn; return false; }
  out.resize(buf.size() / sizeof(R));
  std::memcpy(out.data(), buf.data(), out.size() * sizeof(R));
  return true;
}

struct CentroidRec { uint32_t det; float x, y, z; int32_t type; };
struct OrientRec { uint32_t det; float a, b; };

// Occupancy category x |eta| bin used by every per-module capacity table
// (SDL/MiniDoublet.h:1021-1049, SDL/Segment.h:936-960, SDL/Triplet.h:1030-1054).
void category(int16_t layer, int16_t ring, int16_t subdet, float absEta, int& cat, int& eb) {
  if (layer <= 3 && subdet == 5) cat = 0;
  else if (layer >= 4 && subdet == 5) cat = 1;
  else if (layer <= 2 && subdet == 4 && ring >= 11) cat = 2;
  else if (layer >= 3 && subdet == 4 && ring >= 8) cat = 2;
  else if (layer <= 2 && subdet == 4 && ring <= 10) cat = 3;
  else if (layer >= 3 && subdet == 4 && ring <= 7) cat = 3;
  else cat = -1;
  if (absEta < 0.75) eb = 0;
  else if (absEta > 0.75 && absEta < 1.5) eb = 1;
  else if (absEta > 1.5 && absEta < 2.25) eb = 2;
  else if (absEta > 2.25 && absEta < 3) eb = 3;
  else eb = -1;
}

}  // namespace

bool load_host_geometry(const std::string& dir, uint32_t stride, HostGeometry& g, std::string& err) {
  g = HostGeometry();
  std::vector<CentroidRec> cent;
  std::vector<OrientRec> endcap, tilted;
  ConnTable conn;
  if (!read_fixed(dir + "/sensor_centroids.bin", cent, err)) return false;
  if (!read_fixed(dir + "/endcap_orientation.bin", endcap, err)) return false;
  if (!read_fixed(dir + "/tilted_barrel_orientation.bin", tilted, err)) return false;
  if (!read_conn(dir + "/module_connection_tracing_merged.bin", conn, err)) return false;

  // the last record of a detId wins, as with repeated std::map assignment
  std::unordered_map<uint32_t, CentroidRec> byDet;
  for (auto& c : cent) byDet[c.det] = c;
  std::unordered_map<uint32_t, float> ecDxdy, ecPhi, tbDrdz, tbDxdy;
  for (auto& e : endcap) { ecDxdy[e.det] = e.a; ecPhi[e.det] = e.b; }
  for (auto& t : tilted) { tbDrdz[t.det] = t.a; tbDxdy[t.det] = t.b; }

  // All detIds in ascending order with the pixel pseudo-module (detId 1) added
  // (SDL/ModuleMethods.h:223): this order defines the module index layout.
  std::vector<uint32_t> dets;
  dets.reserve(byDet.size() + 1);
  for (auto& kv : byDet) dets.push_back(kv.first);
  if (!byDet.count(1)) dets.push_back(1);
  std::sort(dets.begin(), dets.end());
  const uint32_t nMod = dets.size();
  if (nMod > 65535) { err = "more modules than a uint16 module index can address"; return false; }
  const uint32_t nLower = (nMod - 1) / 2;
  g.nModules = nMod;
  g.nLowerModules = nLower;
  auto sz = [&](auto& v) { v.assign(nMod, {}); };
  sz(g.detIds); sz(g.layers); sz(g.rings); sz(g.rods); sz(g.modules); sz(g.subdets); sz(g.sides); sz(g.eta);
  sz(g.r); sz(g.isInverted); sz(g.isLower); sz(g.isAnchor); sz(g.moduleType); sz(g.moduleLayerType);
  sz(g.dxdys); sz(g.drdzs); sz(g.partnerModuleIndices); sz(g.sdlLayers);

  std::unordered_map<uint32_t, uint16_t> index;
  index.reserve(nMod * 2);
  uint32_t nextLower = 0, nextUpper = nLower + 1;
  for (uint32_t det : dets) {
    int16_t layer = 0, ring = 0, rod = 0, mod = 0, subdet = 0, side = 0;
    float eta = 0, r = 0;
    bool inv = false, lower = false;
    const CentroidRec* c = nullptr;
    if (det != 1) {
      c = &byDet[det];
      // detId bit fields: SDL/ModuleMethods.h:180-185
      subdet = (det >> 25) & 7;
      side = subdet == kEndcap ? (det >> 23) & 3 : (det >> 18) & 3;
      layer = subdet == kEndcap ? (det >> 18) & 7 : (det >> 20) & 7;
      ring = subdet == kEndcap ? (det >> 12) & 15 : 0;
      mod = (det >> 2) & 127;
      rod = subdet == kEndcap ? 0 : (det >> 10) & 127;
      r = std::sqrt(c->x * c->x + c->y * c->y + c->z * c->z);
      eta = ((c->z > 0) - (c->z < 0)) * std::acosh(r / std::sqrt(c->x * c->x + c->y * c->y));
      // module parity rules: SDL/Module.h:196-227
      if (subdet == kEndcap) inv = side == kNegZ ? (mod % 2 == 1) : (side == kPosZ ? (mod % 2 == 0) : false);
      else if (subdet == kBarrel) {
        if (side == kCenter) inv = layer <= 3 ? (mod % 2 == 1) : (mod % 2 == 0);
        else if (side == kNegZ || side == kPosZ) inv = layer <= 2 ? (mod % 2 == 1) : (layer == 3 ? (mod % 2 == 0) : false);
      }
      lower = inv ? !(det & 1) : (det & 1);
    }
    uint32_t i = lower ? nextLower++ : (det == 1 ? nLower : nextUpper++);
    index[det] = i;
    g.detIds[i] = det; g.layers[i] = layer; g.rings[i] = ring; g.rods[i] = rod; g.modules[i] = mod;
    g.subdets[i] = subdet; g.sides[i] = side; g.eta[i] = eta; g.r[i] = r; g.isInverted[i] = inv; g.isLower[i] = lower;
    if (det == 1) {
      g.moduleType[i] = kPixelModule; g.moduleLayerType[i] = kInnerPixelLayer;
    } else {
      g.moduleType[i] = c->type == 25 ? kTwoS : kPS;
      g.moduleLayerType[i] = c->type == 23 ? kPixelLayer : kStripLayer;
      g.isAnchor[i] = (g.moduleType[i] == kPS && g.moduleLayerType[i] == kPixelLayer) ||
                      (g.moduleType[i] == kTwoS && lower);
      auto get = [](const std::unordered_map<uint32_t, float>& m, uint32_t k) {
        auto f = m.find(k);
        return f == m.end() ? 0.f : f->second;
      };
      g.dxdys[i] = subdet == kEndcap ? get(ecDxdy, det) : get(tbDxdy, det);
      g.drdzs[i] = subdet == kBarrel ? get(tbDrdz, det) : 0.f;
    }
    g.sdlLayers[i] = layer + 6 * (subdet == kEndcap) + 5 * (subdet == kEndcap && g.moduleType[i] == kTwoS);
  }
  if (nextLower != nLower) { err = "lower/upper module counts do not pair up"; return false; }
  g.pixelModuleIndex = nLower;
  // partner index and drdz/dxdy back-fill, in ascending detId order (SDL/ModuleMethods.h:367-382)
  for (uint32_t det : dets) {
    if (det == 1) continue;
    uint32_t i = index[det];
    uint32_t pd = g.isLower[i] ? (g.isInverted[i] ? det - 1 : det + 1) : (g.isInverted[i] ? det + 1 : det - 1);
    auto f = index.find(pd);
    if (f == index.end()) { err = "no partner module for detId " + std::to_string(det); return false; }
    g.partnerModuleIndices[i] = f->second;
    if (g.drdzs[i] == 0) g.drdzs[i] = g.drdzs[f->second];
    if (g.dxdys[i] == 0) g.dxdys[i] = g.dxdys[f->second];
  }
  // the module-only terms of the strip-hit shift, with the float expressions of the per-hit step
  // the rotation of a point's r-phi uncertainty in the circle regressions (SDL/Quintuplet.h:1136-1221,
  // SDL/PixelTriplet.h:419-487): by an angle of the module's slope and the quadrant of the point,
  // q0 x>0 y>0, q1 x<0 y>0, q2 x<0 y<0, q3 x>0 y<0 (a point on an axis is not rotated)
  g.sigRot.assign(8 * (size_t)nMod, 0.f);
  for (uint32_t i = 0; i < nMod; i++) {
    const float slope = g.dxdys[i];
    const float a = slope != kSdlInf ? m_abs(m_atan(slope)) : 0.5f * kPi;
    const float ang[4] = {0.5f * kPi - a, a + 0.5f * kPi, -(a + 0.5f * kPi), -(0.5f * kPi - a)};
    for (int q = 0; q < 4; q++) {
      g.sigRot[8 * (size_t)i + 2 * q] = m_cos(ang[q]);
      g.sigRot[8 * (size_t)i + 2 * q + 1] = m_sin(ang[q]);
    }
  }
  g.shiftAngleB.assign(nMod, 0.f);
  g.shiftSinM.assign(4 * (size_t)nMod, 0.f);
  g.shiftCosM.assign(4 * (size_t)nMod, 0.f);
  for (uint32_t i = 0; i < nMod; i++) {
    g.shiftAngleB[i] = g.subdets[i] == kEndcap ? kPi / 2.f : m_atan(g.drdzs[i]);
    const float slope = g.dxdys[i];
    const float a = slope != kSdlInf ? m_abs(m_atan(slope)) : kPi / 2.f;
    const float angleM[4] = {a, kPi - a, kPi + a, 2.f * kPi - a};  // x>0 y>0, x>0 y<0, x<0 y<0, otherwise
    for (int q = 0; q < 4; q++) {
      g.shiftSinM[4 * (size_t)i + q] = m_sin(angleM[q]);
      g.shiftCosM[4 * (size_t)i + q] = m_cos(angleM[q]);
    }
  }
  auto idxOf = [&](uint32_t d, bool& ok) -> uint16_t {
    auto f = index.find(d);
    if (f == index.end()) { ok = false; return 0; }
    return f->second;
  };
  // module map with row stride = max fan-out unless given (SDL/ModuleMethods.h:114-140)
  uint32_t maxFan = 0;
  for (uint32_t det : dets) { uint32_t n; conn.find(det, n); maxFan = std::max(maxFan, n); }
  if (stride == 0) stride = maxFan;
  if (maxFan > stride) { err = "module fan-out " + std::to_string(maxFan) + " exceeds the module-map stride"; return false; }
  g.stride = stride;
  g.moduleMap.assign((size_t)nMod * stride, 0);
  g.nConnectedModules.assign(nMod, 0);
  for (uint32_t det : dets) {
    uint32_t n;
    const uint32_t* c = conn.find(det, n);
    uint32_t i = index[det];
    g.nConnectedModules[i] = n;
    for (uint32_t k = 0; k < n; k++) {
      bool ok = true;
      g.moduleMap[(size_t)i * stride + k] = idxOf(c[k], ok);
      if (!ok) { err = "connection to unknown module " + std::to_string(c[k]); return false; }
    }
  }
  // sorted detId -> index lookup used by the hit kernel (SDL/ModuleMethods.h:142-166)
  g.mapdetId = dets;
  g.mapIdx.resize(nMod);
  for (uint32_t k = 0; k < nMod; k++) g.mapIdx[k] = index[dets[k]];
  // endcap 2S orientation map, sorted by detId (SDL/EndcapGeometry.cc:43-70)
  std::vector<std::pair<uint32_t, float>> ep(ecPhi.begin(), ecPhi.end());
  std::sort(ep.begin(), ep.end());
  for (auto& e : ep) { g.geoMapDetId.push_back(e.first); g.geoMapPhi.push_back(e.second); }
  // the hit kernel's detId lookup: one table probe instead of two binary searches
  g.detHashBits = 10;
  while ((1u << g.detHashBits) < 4 * nMod) g.detHashBits++;
  if (g.geoMapDetId.size() >= 0xFFFF) { err = "endcap orientation map too large"; return false; }
  g.detHash.assign(2u << g.detHashBits, kNoDetId);
  for (uint32_t k = 0; k < nMod; k++) {
    const uint32_t det = dets[k];
    if (det == kNoDetId) { err = "detId 0xFFFFFFFF is reserved"; return false; }
    const auto it = std::lower_bound(g.geoMapDetId.begin(), g.geoMapDetId.end(), det);
    const uint32_t ec = (it != g.geoMapDetId.end() && *it == det) ? (uint32_t)(it - g.geoMapDetId.begin()) + 1 : 0;
    uint32_t s = det_hash_slot(det, g.detHashBits);
    while (g.detHash[2 * s] != kNoDetId) s = (s + 1) & ((1u << g.detHashBits) - 1);
    g.detHash[2 * s] = det;
    g.detHash[2 * s + 1] = (uint32_t)g.mapIdx[k] | ec << 16;
  }

  // pLS superbin maps (SDL/LSTESData.cc:57-74, SDL/ModuleMethods.h:33-112).
  // kind 0 uses keys superbin + 45000 of the unsigned maps, kinds 1/2 the
  // pos/neg maps keyed by superbin; the four layer files are concatenated.
  static const char* layers4[4] = {"_layer1_subdet5", "_layer2_subdet5", "_layer1_subdet4", "_layer2_subdet4"};
  static const char* kinds[3] = {"", "_pos", "_neg"};
  std::vector<uint32_t> connected[3];
  for (int k = 0; k < 3; k++) {
    ConnTable t[4];
    for (int l = 0; l < 4; l++)
      if (!read_conn(dir + "/pixelmap/pLS_map" + kinds[k] + layers4[l] + ".bin", t[l], err)) return false;
    g.pixIndex[k].assign(kSizeSuperbins, 0);
    g.pixSizes[k].assign(kSizeSuperbins, 0);
    for (uint32_t sb = 0; sb < kSizeSuperbins; sb++) {
      uint32_t key = k == 0 ? sb + kSizeSuperbins : sb;
      g.pixIndex[k][sb] = connected[k].size();
      for (int l = 0; l < 4; l++) {
        uint32_t n;
        const uint32_t* c = t[l].find(key, n);
        for (uint32_t q = 0; q < n; q++) {
          bool ok = true;
          connected[k].push_back(idxOf(c[q], ok));
          if (!ok) { err = "pixel map refers to unknown module"; return false; }
        }
      }
      g.pixSizes[k][sb] = connected[k].size() - g.pixIndex[k][sb];
    }
  }
  const uint32_t off1 = connected[0].size(), off2 = off1 + connected[1].size();
  for (uint32_t sb = 0; sb < kSizeSuperbins; sb++) { g.pixIndex[1][sb] += off1; g.pixIndex[2][sb] += off2; }
  g.connectedPixels = connected[0];
  g.connectedPixels.insert(g.connectedPixels.end(), connected[1].begin(), connected[1].end());
  g.connectedPixels.insert(g.connectedPixels.end(), connected[2].begin(), connected[2].end());
  g.nPixels = g.connectedPixels.size();

  // geometry-only capacity tables (SDL/MiniDoublet.h:1051-1072, SDL/Segment.h:962-992)
  g.mdOccupancy.assign(nLower + 1, 0);
  g.segOccupancy.assign(nLower + 1, 0);
  g.t3OccIfAny.assign(nLower, 0);
  g.t5OccIfAny.assign(nLower, -1);
  for (uint32_t i = 0; i < nLower; i++) {
    int cat, eb;
    category(g.layers[i], g.rings[i], g.subdets[i], std::fabs(g.eta[i]), cat, eb);
    int md = 0;
    if (cat == 0) md = eb == 0 ? 49 : eb == 1 ? 42 : eb == 2 ? 37 : eb == 3 ? 41 : 0;
    else if (cat == 1) md = 100;
    else if (cat == 2) md = eb == 1 ? 16 : eb == 2 ? 19 : 0;
    else if (cat == 3) md = eb == 1 ? 14 : eb == 2 ? 20 : eb == 3 ? 25 : 0;
    g.mdOccupancy[i] = md;
    int ls = 0;
    if (g.nConnectedModules[i] != 0) {
      if (cat == 0) ls = eb == 0 ? 572 : eb == 1 ? 300 : eb == 2 ? 183 : eb == 3 ? 62 : 0;
      else if (cat == 1) ls = eb == 0 ? 191 : eb == 1 ? 128 : 0;
      else if (cat == 2) ls = eb == 1 ? 107 : eb == 2 ? 102 : 0;
      else if (cat == 3) ls = eb == 1 ? 64 : eb == 2 ? 79 : eb == 3 ? 85 : 0;
    }
    g.segOccupancy[i] = ls;
    // T3 table (SDL/Triplet.h:1056-1084), applied to modules that have segments
    int t3 = 0;
    if (cat == 0) t3 = eb == 0 ? 543 : eb == 1 ? 235 : eb == 2 ? 88 : eb == 3 ? 46 : 0;
    else if (cat == 1) t3 = eb == 0 ? 755 : eb == 1 ? 347 : 0;
    else if (cat == 3) t3 = eb == 1 ? 38 : eb == 2 ? 46 : eb == 3 ? 39 : 0;
    g.t3OccIfAny[i] = t3;
    // T5 seed modules and their table (SDL/Quintuplet.h:3129-3196): barrel
    // layers 1-2 and endcap disk 1 qualify once they hold a triplet
    const bool seed = (g.subdets[i] == 5 && g.layers[i] < 3) || (g.subdets[i] == 4 && g.layers[i] <= 1);
    if (seed) {
      int t5 = 0;
      if (cat == 0) t5 = eb == 0 ? 336 : eb == 1 ? 414 : eb == 2 ? 231 : eb == 3 ? 146 : 0;
      else if (cat == 3) t5 = eb == 2 ? 191 : eb == 3 ? 106 : 0;
      g.t5OccIfAny[i] = t5;
    }
  }
  // the pixel module's MD capacity is N_MAX_PIXEL_MD_PER_MODULES (SDL/Event.cc:398-407)
  g.mdOccupancy[nLower] = kNMaxPixelMDs;
  g.mdIndices.assign(nLower + 1, 0);
  g.segIndices.assign(nLower + 1, 0);
  int32_t runMD = 0, runSeg = 0;
  for (uint32_t i = 0; i <= nLower; i++) {
    g.mdIndices[i] = runMD;
    g.segIndices[i] = runSeg;
    if (i < nLower) { runMD += g.mdOccupancy[i]; runSeg += g.segOccupancy[i]; }
  }
  return true;
}

namespace {
struct Carver {
  std::vector<std::pair<size_t, std::pair<const void*, size_t>>> items;
  size_t total = 0;
  template <class T>
  size_t add(const std::vector<T>& v) {
    size_t off = total;
    size_t bytes = v.size() * sizeof(T);
    items.push_back({off, {v.data(), bytes}});
    total += (bytes + 255) & ~size_t(255);
    return off;
  }
};
}  // namespace

bool fill_module_terms(HostGeometry& g, float ptCut, std::string& err) {
  const uint32_t n = g.nModules;
  g.modCtx.assign(4 * (size_t)n, 0.f);
  g.modTerms.assign(8 * (size_t)n, 0.f);
  for (uint32_t i = 0; i < n; i++) {
    ModCtx c;
    c.subdet = g.subdets[i]; c.side = g.sides[i]; c.layer = g.layers[i]; c.ring = g.rings[i]; c.rod = g.rods[i];
    c.type = g.moduleType[i]; c.layerType = g.moduleLayerType[i];
    c.drdz = g.drdzs[i]; c.dxdy = g.dxdys[i]; c.partnerDrdz = g.drdzs[g.partnerModuleIndices[i]];
    if (!mod_packable(c, g.sdlLayers[i])) {
      err = "module " + std::to_string(i) + ": attributes outside the packed ranges";
      return false;
    }
    const uint32_t w = pack_mod(c, g.sdlLayers[i]);
    float* ctx = &g.modCtx[4 * (size_t)i];
    std::memcpy(ctx, &w, 4);
    ctx[1] = c.drdz; ctx[2] = c.dxdy; ctx[3] = c.partnerDrdz;
    if ((c.subdet == kBarrel && c.layer >= 1 && c.layer <= 6) || (c.subdet == kEndcap && c.layer >= 1 && c.layer <= 5)) {
      const ModTerms t = module_terms(c, ptCut);
      std::memcpy(&g.modTerms[8 * (size_t)i], &t, sizeof(t));
    }
  }
  return true;
}

bool upload_geometry(const HostGeometry& h, DeviceGeometry& d, std::string& err) {
  Carver c;
  size_t o_det = c.add(h.detIds), o_map = c.add(h.moduleMap), o_mdet = c.add(h.mapdetId), o_midx = c.add(h.mapIdx),
         o_ncon = c.add(h.nConnectedModules), o_drdz = c.add(h.drdzs), o_dxdy = c.add(h.dxdys),
         o_part = c.add(h.partnerModuleIndices), o_lay = c.add(h.layers), o_ring = c.add(h.rings),
         o_mod = c.add(h.modules), o_rod = c.add(h.rods), o_sub = c.add(h.subdets), o_side = c.add(h.sides),
         o_eta = c.add(h.eta), o_r = c.add(h.r), o_inv = c.add(h.isInverted), o_low = c.add(h.isLower),
         o_anc = c.add(h.isAnchor), o_mt = c.add(h.moduleType), o_mlt = c.add(h.moduleLayerType),
         o_sdl = c.add(h.sdlLayers), o_cp = c.add(h.connectedPixels), o_gd = c.add(h.geoMapDetId),
         o_gp = c.add(h.geoMapPhi), o_pi0 = c.add(h.pixIndex[0]), o_pi1 = c.add(h.pixIndex[1]),
         o_pi2 = c.add(h.pixIndex[2]), o_ps0 = c.add(h.pixSizes[0]), o_ps1 = c.add(h.pixSizes[1]),
         o_ps2 = c.add(h.pixSizes[2]), o_mdo = c.add(h.mdOccupancy), o_sgo = c.add(h.segOccupancy),
         o_mdi = c.add(h.mdIndices), o_sgi = c.add(h.segIndices), o_t3o = c.add(h.t3OccIfAny),
         o_t5o = c.add(h.t5OccIfAny), o_dh = c.add(h.detHash), o_sab = c.add(h.shiftAngleB),
         o_ssm = c.add(h.shiftSinM), o_scm = c.add(h.shiftCosM), o_mcx = c.add(h.modCtx),
         o_mtm = c.add(h.modTerms), o_sgr = c.add(h.sigRot);
  char* base = nullptr;
  if (hipMalloc(&base, c.total) != hipSuccess) { err = "hipMalloc failed for geometry"; return false; }
  for (auto& it : c.items)
    if (it.second.second &&
        hipMemcpy(base + it.first, it.second.first, it.second.second, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(base);
      err = "geometry upload failed";
      return false;
    }
  d.arena = base;
  d.bytes = c.total;
  auto P = [&](size_t o) { return (const void*)(base + o); };
  ModulesView& v = d.view;
  v.detIds = (const uint32_t*)P(o_det); v.moduleMap = (const uint16_t*)P(o_map);
  v.mapdetId = (const uint32_t*)P(o_mdet); v.mapIdx = (const uint16_t*)P(o_midx);
  v.nConnectedModules = (const uint16_t*)P(o_ncon); v.drdzs = (const float*)P(o_drdz); v.dxdys = (const float*)P(o_dxdy);
  v.partnerModuleIndices = (const uint16_t*)P(o_part); v.layers = (const int16_t*)P(o_lay);
  v.rings = (const int16_t*)P(o_ring); v.modules = (const int16_t*)P(o_mod); v.rods = (const int16_t*)P(o_rod);
  v.subdets = (const int16_t*)P(o_sub); v.sides = (const int16_t*)P(o_side); v.eta = (const float*)P(o_eta);
  v.r = (const float*)P(o_r); v.isInverted = (const uint8_t*)P(o_inv); v.isLower = (const uint8_t*)P(o_low);
  v.isAnchor = (const uint8_t*)P(o_anc); v.moduleType = (const uint8_t*)P(o_mt);
  v.moduleLayerType = (const uint8_t*)P(o_mlt); v.sdlLayers = (const int32_t*)P(o_sdl);
  v.connectedPixels = (const uint32_t*)P(o_cp); v.geoMapDetId = (const uint32_t*)P(o_gd);
  v.geoMapPhi = (const float*)P(o_gp);
  v.shiftAngleB = (const float*)P(o_sab); v.shiftSinM = (const float*)P(o_ssm); v.shiftCosM = (const float*)P(o_scm);
  v.modCtx = (const float4*)P(o_mcx); v.modTerms = (const float4*)P(o_mtm); v.sigRot = (const float2*)P(o_sgr);
  v.nModules = h.nModules; v.nLowerModules = h.nLowerModules; v.stride = h.stride; v.nEndCapMap = h.geoMapDetId.size();
  v.detHash = (const uint2*)P(o_dh); v.detHashBits = h.detHashBits;
  d.pixIndex[0] = (const uint32_t*)P(o_pi0); d.pixIndex[1] = (const uint32_t*)P(o_pi1);
  d.pixIndex[2] = (const uint32_t*)P(o_pi2); d.pixSizes[0] = (const uint32_t*)P(o_ps0);
  d.pixSizes[1] = (const uint32_t*)P(o_ps1); d.pixSizes[2] = (const uint32_t*)P(o_ps2);
  d.mdOccupancy = (const int32_t*)P(o_mdo); d.segOccupancy = (const int32_t*)P(o_sgo);
  d.mdIndices = (const int32_t*)P(o_mdi); d.segIndices = (const int32_t*)P(o_sgi);
  d.t3OccIfAny = (const int32_t*)P(o_t3o);
  d.t5OccIfAny = (const int32_t*)P(o_t5o);
  d.pixelModuleIndex = h.pixelModuleIndex;
  d.nPixels = h.nPixels;
  d.nTotalMDCap = h.mdIndices[h.nLowerModules];
  d.nTotalSegCap = h.segIndices[h.nLowerModules];
  d.t3CapBound = 0;
  for (auto v2 : h.t3OccIfAny) d.t3CapBound += v2;
  d.t5CapBound = 0;
  d.nT5SeedModules = 0;
  for (auto v2 : h.t5OccIfAny)
    if (v2 >= 0) {
      d.t5CapBound += v2;
      d.nT5SeedModules++;
    }
  return true;
}

void free_geometry(DeviceGeometry& d) {
  if (d.arena) (void)hipFree(d.arena);
  d = DeviceGeometry();
}

}  // namespace lst
