// Module attributes and the module-only terms of the mini-doublet and segment
// cuts.  The same functions run on the host, where the geometry loader
// evaluates them once per module into two device tables, and on the device.
//
// modCtx[m] (16 bytes): the integer attributes (and the SDL layer number) packed in one word, then drdz,
// dxdy and the partner module's drdz.  A selection loads a module's context
// with one 16-byte load instead of ten scattered ones.
//
// modTerms[2m], modTerms[2m + 1] (32 bytes): terms of the cuts that depend on
// the module and the pT cut only, each the float expression the cut itself
// evaluates first (C evaluation order), so the cut reads the identical float:
// first float4 (segment cuts), second float4 (mini-doublet cut):
//   muls      muls_scale(c, ptCut)
//   segTilt   0.5 pitch drdz / sqrt(1 + drdz^2) / gap, 0 unless a tilted
//             barrel module (dAlpha luminous-region term, SDL/Segment.h:286-310)
//   segRes    0.02 / gap (dAlpha resolution term, SDL/Segment.h:312-322)
//   mdBase    muls^2 + pvOff^2 with pvOff = 0.1 / rNom (SDL/MiniDoublet.h:352-401)
//   mdTilt2   tilt^2 of the mini-doublet cut (SDL/MiniDoublet.h:370-385)
//   mdRoot    sqrt(mdBase), the whole root of a flat barrel module
#pragma once
#include "lst_types.h"

namespace lst {

struct ModCtx {
  int16_t subdet, side, layer, ring, rod;
  uint8_t type, layerType;
  float drdz, dxdy, partnerDrdz;
};

// packed integer attributes: subdet [0,3) side [3,5) layer [5,9) ring [9,14) rod [14,22) type [22,24)
// layerType [24,26) sdlLayer [26,31); mod_packable says whether a module's values fit
LST_HD bool mod_packable(const ModCtx& c, int sdlLayer) {
  return c.subdet >= 0 && c.subdet < 8 && c.side >= 0 && c.side < 4 && c.layer >= 0 && c.layer < 16 &&
         c.ring >= 0 && c.ring < 32 && c.rod >= 0 && c.rod < 256 && c.type < 4 && c.layerType < 4 && sdlLayer >= 0 &&
         sdlLayer < 32;
}
LST_HD uint32_t pack_mod(const ModCtx& c, int sdlLayer) {
  return (uint32_t)c.subdet | (uint32_t)c.side << 3 | (uint32_t)c.layer << 5 | (uint32_t)c.ring << 9 |
         (uint32_t)c.rod << 14 | (uint32_t)c.type << 22 | (uint32_t)c.layerType << 24 | (uint32_t)sdlLayer << 26;
}
LST_HD void unpack_mod(uint32_t w, ModCtx& c) {
  c.subdet = (int16_t)(w & 7u);
  c.side = (int16_t)((w >> 3) & 3u);
  c.layer = (int16_t)((w >> 5) & 15u);
  c.ring = (int16_t)((w >> 9) & 31u);
  c.rod = (int16_t)((w >> 14) & 255u);
  c.type = (uint8_t)((w >> 22) & 3u);
  c.layerType = (uint8_t)((w >> 24) & 3u);
}
LST_HD int unpack_sdl_layer(uint32_t w) { return (int)((w >> 26) & 31u); }

// Separation between the two sensors of a module (SDL/MiniDoublet.h:264-337,
// same values as SDL/Segment.h:202-262).
LST_HD bool tighter_tilted(const ModCtx& c) {
  return c.subdet == kBarrel && ((c.side != kCenter && c.layer == 3) ||
                                 (c.side == kNegZ && ((c.layer == 2 && c.rod > 5) || (c.layer == 1 && c.rod > 9))) ||
                                 (c.side == kPosZ && ((c.layer == 2 && c.rod < 8) || (c.layer == 1 && c.rod < 4))));
}
LST_HD float gap_size(const ModCtx& c) {
  if (c.subdet == kBarrel && c.side == kCenter) {
    const float flat[6] = {0.26f, 0.16f, 0.16f, 0.18f, 0.18f, 0.18f};
    return flat[c.layer - 1];
  }
  if (tighter_tilted(c)) return 0.26f;
  if (c.subdet == kEndcap) {
    const int ringLimit = c.layer <= 2 ? 10 : (c.layer <= 4 ? 8 : 9);
    return (c.ring - 1) < ringLimit ? 0.4f : 0.18f;
  }
  return 0.4f;
}

// multiple-scattering scale (SDL/MiniDoublet.h:352-356, SDL/Segment.h:264-270)
LST_HD float muls_scale(const ModCtx& c, float ptCut) {
  const float barrel[6] = {0.0052, 0.0038, 0.0034, 0.0034, 0.0032, 0.0034};  // double literals, as the reference's
  const float endcap[5] = {0.006, 0.006, 0.006, 0.006, 0.006};
  return c.subdet == kBarrel ? barrel[c.layer - 1] * 3.f / ptCut : endcap[c.layer - 1] * 3.f / ptCut;
}

struct ModTerms {
  float muls, segTilt, segRes, pad0, mdBase, mdTilt2, mdRoot, pad1;
};

// the terms of a module of a barrel layer 1-6 or an endcap layer 1-5
LST_HD ModTerms module_terms(const ModCtx& c, float ptCut) {
  const float rminBarrel[6] = {25.007152356, 37.2186993757, 52.3104270826, 68.6658656666, 85.9770373007, 108.301772384};
  const float rminEndcap[5] = {130.992832231, 154.813883559, 185.352604327, 221.635123002, 265.022076742};
  ModTerms t{};
  const float gap = gap_size(c);
  t.muls = muls_scale(c, ptCut);
  const bool tilted = c.subdet == kBarrel && c.side != kCenter;
  t.segTilt = tilted ? (0.5f * kPixelPSZpitch * c.drdz / m_sqrt(1.f + c.drdz * c.drdz) / gap) : 0.f;
  t.segRes = 0.02f / gap;
  const float rNom = c.subdet == kBarrel ? rminBarrel[c.layer - 1] : rminEndcap[c.layer - 1];
  const float pvOff = 0.1f / rNom;
  t.mdBase = t.muls * t.muls + pvOff * pvOff;
  const float drdz = (c.type == kPS && c.layerType == kStripLayer) ? c.drdz : c.partnerDrdz;
  const float tilt = 0.5f * kPixelPSZpitch * drdz / m_sqrt(1.f + drdz * drdz) / gap;
  t.mdTilt2 = tilt * tilt;
  t.mdRoot = m_sqrt(t.mdBase);
  return t;
}

}  // namespace lst
