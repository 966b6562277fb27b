// Field lists of the per-event device SoA.  One X-macro list per object kind
// drives the struct members, the arena carving and the by-name getters, so the
// three can never disagree.  Names follow the reference structs
// (SDL/Hit.h:13-56, SDL/MiniDoublet.h:16-53, SDL/Segment.h:17-50,
// SDL/Triplet.h:17-31, SDL/Quintuplet.h:20-48) so tests compare fields by name against the oracle.
#pragma once
#include <cstdint>
#include "lst_types.h"

// X(type, name, per-object multiplicity)
#define LST_HIT_FIELDS(X)                                                                                  \
  X(float, xs, 1) X(float, ys, 1) X(float, zs, 1) X(uint16_t, moduleIndices, 1) X(uint32_t, idxs, 1)      \
  X(uint32_t, detid, 1) X(float, rts, 1) X(float, phis, 1) X(float, etas, 1) X(float, highEdgeXs, 1)      \
  X(float, highEdgeYs, 1) X(float, lowEdgeXs, 1) X(float, lowEdgeYs, 1)

// Mini-doublet fields in record order (4-byte words).  The first four
// 16-byte chunks hold what the selections read: chunk 0 the segment cut's
// anchor hit, module, dphichange and dz; chunk 1 the anchor point (x, y, z,
// rt) every later stage reads; chunk 2 dphi, phi and the strip-edge phis;
// chunk 3 the strip-edge points of the T4/T5 tests.
#define LST_MD_FIELDS(X)                                                                                   \
  X(uint32_t, anchorHitIndices, 1) X(uint16_t, moduleIndices, 1) X(float, dphichanges, 1) X(float, dzs, 1) \
  X(float, anchorX, 1) X(float, anchorY, 1) X(float, anchorZ, 1) X(float, anchorRt, 1)                     \
  X(float, dphis, 1) X(float, anchorPhi, 1) X(float, anchorHighEdgePhi, 1) X(float, anchorLowEdgePhi, 1)   \
  X(float, anchorHighEdgeX, 1) X(float, anchorHighEdgeY, 1) X(float, anchorLowEdgeX, 1)                    \
  X(float, anchorLowEdgeY, 1) X(uint32_t, outerHitIndices, 1) X(float, anchorEta, 1) X(float, shiftedXs, 1) \
  X(float, shiftedYs, 1) X(float, shiftedZs, 1) X(float, noShiftedDzs, 1) X(float, noShiftedDphis, 1)      \
  X(float, noShiftedDphiChanges, 1) X(float, outerX, 1) X(float, outerY, 1) X(float, outerZ, 1)            \
  X(float, outerRt, 1) X(float, outerPhi, 1) X(float, outerEta, 1) X(float, outerHighEdgeX, 1)             \
  X(float, outerHighEdgeY, 1) X(float, outerLowEdgeX, 1) X(float, outerLowEdgeY, 1)

#define LST_SEG_FIELDS(X)                                                                                  \
  X(uint32_t, mdIndices, 2) X(uint16_t, innerLowerModuleIndices, 1) X(uint16_t, outerLowerModuleIndices, 1) \
  X(uint32_t, innerMiniDoubletAnchorHitIndices, 1) X(uint32_t, outerMiniDoubletAnchorHitIndices, 1)         \
  X(float, dPhis, 1) X(float, dPhiMins, 1) X(float, dPhiMaxs, 1) X(float, dPhiChanges, 1)                  \
  X(float, dPhiChangeMins, 1) X(float, dPhiChangeMaxs, 1)

// pLS-only arrays, indexed by pLS number (SDL/Segment.h:33-55)
#define LST_PLS_FIELDS(X)                                                                                  \
  X(uint32_t, seedIdx, 1) X(uint32_t, pLSHitsIdxs, 4) X(int32_t, charge, 1) X(int32_t, superbin, 1)        \
  X(int8_t, pixelType, 1) X(int8_t, isQuad, 1) X(int8_t, isDup, 1) X(uint8_t, partOfPT5, 1)                \
  X(float, ptIn, 1) X(float, ptErr, 1) X(float, px, 1) X(float, py, 1) X(float, pz, 1) X(float, etaErr, 1) \
  X(float, eta, 1) X(float, phi, 1) X(float, score, 1) X(float, circleCenterX, 1)                          \
  X(float, circleCenterY, 1) X(float, circleRadius, 1)

#define LST_T3_FIELDS(X)                                                                                   \
  X(uint32_t, segmentIndices, 2) X(uint16_t, lowerModuleIndices, 3) X(uint8_t, logicalLayers, 3)           \
  X(uint32_t, hitIndices, 6) X(float, betaIn, 1) X(float, circleRadius, 1) X(float, circleCenterX, 1)      \
  X(float, circleCenterY, 1) X(uint8_t, partOfPT5, 1) X(uint8_t, partOfT5, 1) X(uint8_t, partOfPT3, 1)

#define LST_T5_FIELDS(X)                                                                                   \
  X(uint32_t, tripletIndices, 2) X(uint16_t, lowerModuleIndices, 5) X(uint8_t, logicalLayers, 5)           \
  X(uint32_t, hitIndices, 10) X(uint8_t, layer, 1) X(int8_t, isDup, 1) X(uint8_t, TightCutFlag, 1)        \
  X(uint8_t, partOfPT5, 1) X(float, innerRadius, 1) X(float, bridgeRadius, 1) X(float, outerRadius, 1)     \
  X(float, pt, 1) X(float, eta, 1) X(float, phi, 1) X(float, score_rphisum, 1) X(float, regressionRadius, 1) \
  X(float, regressionG, 1) X(float, regressionF, 1) X(float, rzChiSquared, 1) X(float, chiSquared, 1)      \
  X(float, nonAnchorChiSquared, 1)

// pixel quintuplets (SDL/PixelTriplet.h:1871-1913)
#define LST_PT5_FIELDS(X)                                                                                  \
  X(uint32_t, pixelIndices, 1) X(uint32_t, T5Indices, 1) X(uint8_t, isDup, 1) X(float, score, 1)           \
  X(float, eta, 1) X(float, phi, 1) X(uint8_t, logicalLayers, 7) X(uint32_t, hitIndices, 14)              \
  X(uint16_t, lowerModuleIndices, 7) X(float, pixelRadius, 1) X(float, quintupletRadius, 1)                \
  X(float, centerX, 1) X(float, centerY, 1) X(float, rzChiSquared, 1) X(float, rPhiChiSquared, 1)          \
  X(float, rPhiChiSquaredInwards, 1)

// pixel triplets (SDL/PixelTriplet.h:20-75; the never-written pixelRadiusError
// and partOfPT5 buffers are not carried)
#define LST_PT3_FIELDS(X)                                                                                  \
  X(uint32_t, pixelSegmentIndices, 1) X(uint32_t, tripletIndices, 1) X(float, pixelRadius, 1)              \
  X(float, tripletRadius, 1) X(float, pt, 1) X(float, eta, 1) X(float, phi, 1) X(float, eta_pix, 1)        \
  X(float, phi_pix, 1) X(float, score, 1) X(uint8_t, isDup, 1) X(uint8_t, logicalLayers, 5)                \
  X(uint32_t, hitIndices, 10) X(uint16_t, lowerModuleIndices, 5) X(float, centerX, 1) X(float, centerY, 1) \
  X(float, rzChiSquared, 1) X(float, rPhiChiSquared, 1) X(float, rPhiChiSquaredInwards, 1)

// track candidates (SDL/TrackCandidate.h:20-111)
#define LST_TC_FIELDS(X)                                                                                   \
  X(int16_t, trackCandidateType, 1) X(uint32_t, directObjectIndices, 1) X(uint32_t, objectIndices, 2)      \
  X(uint8_t, logicalLayers, 7) X(uint32_t, hitIndices, 14) X(int32_t, pixelSeedIndex, 1)                  \
  X(uint16_t, lowerModuleIndices, 7) X(float, centerX, 1) X(float, centerY, 1) X(float, radius, 1)

#define LST_DECL_FIELD(T, N, M) T* N;

namespace lst {
struct HitsSoA { LST_HIT_FIELDS(LST_DECL_FIELD) };
// Mini-doublets are stored as records in two arrays: the hot record (80 bytes,
// five 16-byte chunks) holds what the later selections read, the cold record
// (64 bytes) the rest.  Hot words: fields 0-15 (words 0-15), the slope pair
// (words 16-17), fields 16-17 (words 18-19); cold words: fields 18-33.  A
// gather of a mini-doublet then reads whole chunks of one short record
// instead of single words of a long one; D.<field>[i] addresses field
// <field> of record i in its array.
constexpr uint32_t kMDFields = 34;
constexpr uint32_t kMDHotFields = 18;                 // fields 0 .. 17 are in the hot record
constexpr uint32_t kMDSlopeWord = 16;                 // float2 slope pair, hot words 16, 17
constexpr uint32_t kMDHotBytes = 80, kMDColdBytes = 64;
static_assert(4 * (kMDHotFields + 2) == kMDHotBytes && 4 * (kMDFields - kMDHotFields) == kMDColdBytes,
              "every hot and cold word holds a field (or the slope pair)");
// word of field k in its record, and its record's size
constexpr uint32_t md_slot(uint32_t k) { return k < 16 ? k : k < kMDHotFields ? k + 2 : k - kMDHotFields; }
constexpr uint32_t md_rec_bytes(uint32_t k) { return k < kMDHotFields ? kMDHotBytes : kMDColdBytes; }
template <class T, uint32_t REC>
struct MDField {
  char* p;  // this field's slot in record 0
  __host__ __device__ __forceinline__ T& operator[](size_t i) const {
    return *reinterpret_cast<T*>(p + i * REC);
  }
};
// field number of each field (order of LST_MD_FIELDS)
#define LST_MD_WORD(T, N, M) kMDw_##N,
enum MDWord : uint32_t { LST_MD_FIELDS(LST_MD_WORD) kMDWordCount };
static_assert(kMDWordCount == kMDFields, "one word per mini-doublet field");
#define LST_DECL_MD_FIELD(T, N, M) MDField<T, md_rec_bytes(kMDw_##N)> N;
struct MDsSoA { LST_MD_FIELDS(LST_DECL_MD_FIELD) };
static_assert(kMDw_anchorHitIndices == 0 && kMDw_moduleIndices == 1 && kMDw_dphichanges == 2 && kMDw_dzs == 3 &&
                  kMDw_anchorX == 4 && kMDw_anchorY == 5 && kMDw_anchorZ == 6 && kMDw_anchorRt == 7 &&
                  kMDw_dphis == 8 && kMDw_anchorPhi == 9 && kMDw_anchorHighEdgePhi == 10 &&
                  kMDw_anchorLowEdgePhi == 11 && kMDw_anchorHighEdgeX == 12 && kMDw_anchorHighEdgeY == 13 &&
                  kMDw_anchorLowEdgeX == 14 && kMDw_anchorLowEdgeY == 15 && kMDw_outerHitIndices == 16 &&
                  kMDw_anchorEta == 17 && kMDw_shiftedXs == kMDHotFields,
              "the chunks md_chunk readers expect, and the first cold field");
// the slope pair of mini-doublet record i (hot words 16, 17), one global load
__device__ __forceinline__ float2 md_slope(const MDsSoA& D, uint32_t i) {
  return ld_global(reinterpret_cast<const float2*>(D.anchorHitIndices.p + (size_t)i * kMDHotBytes + 4 * kMDSlopeWord));
}
// outer hit x, y of mini-doublet record i (adjacent cold words), one global load
static_assert(md_slot(kMDw_outerY) == md_slot(kMDw_outerX) + 1 && md_slot(kMDw_outerX) % 2 == 0 &&
                  kMDw_outerX >= kMDHotFields,
              "outer x and y are an aligned pair of cold words");
__device__ __forceinline__ float2 md_outer_xy(const MDsSoA& D, uint32_t i) {
  return ld_global(reinterpret_cast<const float2*>(&D.outerX[i]));
}
// 16-byte chunk q (hot words 4q .. 4q + 3) of mini-doublet record i, one global load: 0 {anchor hit, module,
// dphichange, dz}, 1 {anchor x, y, z, rt}, 2 {dphi, anchor phi, edge phis}, 3 {edge x, y}, 4 {slope pair,
// outer hit, anchor eta}
__device__ __forceinline__ uint4 md_chunk(const MDsSoA& D, uint32_t i, uint32_t q) {
  return ld_global(reinterpret_cast<const uint4*>(D.anchorHitIndices.p + (size_t)i * kMDHotBytes) + q);
}
// Segments are records too: 12 four-byte words per segment (the two MD
// indices, then one word per other field, in LST_SEG_FIELDS order; the u16
// module indices in the low half of their word), 48 bytes.  The triplet, T5
// and pixel selections read two to six fields of a segment they reach
// through an index: one cache line instead of one per field.  Element i of a
// field with M entries per segment (mdIndices: M = 2) is word i % M of the
// field in record i / M.
template <class T, uint32_t M, uint32_t REC>
struct RecField {
  char* p;  // the field's first word in record 0
  __host__ __device__ __forceinline__ T& operator[](size_t i) const {
    return *reinterpret_cast<T*>(p + (i / M) * REC + (i % M) * 4);
  }
};
#define LST_SEG_MULT(T, N, M) M,
constexpr uint32_t kSegMult[] = {LST_SEG_FIELDS(LST_SEG_MULT)};
#define LST_SEG_INDEX(T, N, M) kSegF_##N,
enum SegFieldIndex : uint32_t { LST_SEG_FIELDS(LST_SEG_INDEX) kSegFieldCount };
// first word of field number f in a segment record
constexpr uint32_t seg_word(uint32_t f) {
  uint32_t w = 0;
  for (uint32_t i = 0; i < f; i++) w += kSegMult[i];
  return w;
}
constexpr uint32_t kSegRecWords = seg_word(kSegFieldCount);
constexpr uint32_t kSegRecBytes = 4 * kSegRecWords;
static_assert(kSegRecBytes % 16 == 0, "segment records stay 16-byte aligned");
#define LST_DECL_SEG_FIELD(T, N, M) RecField<T, M, kSegRecBytes> N;
struct SegsSoA { LST_SEG_FIELDS(LST_DECL_SEG_FIELD) };
static_assert(seg_word(kSegF_innerLowerModuleIndices) == 2 && seg_word(kSegF_outerLowerModuleIndices) == 3 &&
                  seg_word(kSegF_dPhiMins) == 7 && seg_word(kSegF_dPhiMaxs) == 8 && seg_word(kSegF_dPhiChanges) == 9,
              "the chunks seg_chunk readers expect");
// 16-byte chunk q of segment record s: 0 {inner MD, outer MD, inner module, outer module}, 1 {anchor hits,
// dPhi, dPhiMin}, 2 {dPhiMax, dPhiChange, dPhiChangeMin, dPhiChangeMax}; one global load
__device__ __forceinline__ uint4 seg_chunk(const SegsSoA& S, uint32_t s, uint32_t q) {
  return ld_global(reinterpret_cast<const uint4*>(S.mdIndices.p + (size_t)s * kSegRecBytes) + q);
}
struct PLSSoA { LST_PLS_FIELDS(LST_DECL_FIELD) };
// Triplets as records as well: one 4-byte word per element (u16 and u8
// fields in the low bytes of their words), 21 words padded to 24 (96 bytes).
#define LST_T3_MULT(T, N, M) M,
constexpr uint32_t kT3Mult[] = {LST_T3_FIELDS(LST_T3_MULT)};
#define LST_T3_INDEX(T, N, M) kT3F_##N,
enum T3FieldIndex : uint32_t { LST_T3_FIELDS(LST_T3_INDEX) kT3FieldCount };
constexpr uint32_t t3_word(uint32_t f) {
  uint32_t w = 0;
  for (uint32_t i = 0; i < f; i++) w += kT3Mult[i];
  return w;
}
constexpr uint32_t kT3ObjWords = (t3_word(kT3FieldCount) + 3) & ~3u;
constexpr uint32_t kT3ObjBytes = 4 * kT3ObjWords;
#define LST_DECL_T3_FIELD(T, N, M) RecField<T, M, kT3ObjBytes> N;
struct T3SoA { LST_T3_FIELDS(LST_DECL_T3_FIELD) };
static_assert(t3_word(kT3F_lowerModuleIndices) == 2 && t3_word(kT3F_circleRadius) == 15 &&
                  t3_word(kT3F_circleCenterX) == 16 && t3_word(kT3F_circleCenterY) == 17,
              "the chunks t3_chunk readers expect");
// 16-byte chunk q of triplet record t: 0 {inner segment, outer segment, module 0, module 1}, 1 {module 2,
// logical layers}, 2 {hits 0-3}, 3 {hits 4-5, betaIn, circle radius}, 4 {circle centre x, y, flags}; one
// global load
__device__ __forceinline__ uint4 t3_chunk(const T3SoA& T, uint32_t t, uint32_t q) {
  return ld_global(reinterpret_cast<const uint4*>(T.segmentIndices.p + (size_t)t * kT3ObjBytes) + q);
}
struct T5SoA { LST_T5_FIELDS(LST_DECL_FIELD) };
struct PT5SoA { LST_PT5_FIELDS(LST_DECL_FIELD) };
struct PT3SoA { LST_PT3_FIELDS(LST_DECL_FIELD) };
struct TCSoA { LST_TC_FIELDS(LST_DECL_FIELD) };

// Pass-flag prefix shared by the flattened builders (one builder at a time).
constexpr uint32_t kScanCandMax = 1u << 23;  // >= every builder's candidate maximum
struct FlatScan {
  uint64_t* mask;  // [kScanCandMax / 64] pass flags, candidate L is bit L % 64 of word L / 64
  uint32_t* pre;   // [kScanCandMax / 64] per flag word: passing candidates before it in its scan block
  uint32_t* sums;  // [kScanCandMax / 4096 + 1] exclusive prefix of the block totals, then the total
  uint32_t* pass;  // [kScanCandMax] per scan block, its passing candidates in order (block b from b * 4096)
};

// Work arrays of the flattened mini-doublet builder: every (lower hit, upper
// hit) pair of every module gets a global index L (module, then lower hit,
// then upper hit: the reference's serial order); passing pairs park their
// MDOut, and a wave per module places them in L order.
constexpr uint32_t kMDCandMax = 1u << 21;
constexpr int kMDParkWords = 9;
struct MDWork {
  uint32_t* candOff;  // [nLower + 1] first candidate of each module, then the total
  uint32_t* park;     // [kMDParkWords * kMDCandMax] MDOut of passing candidate L at L
  uint4* cand;        // [kMDCandMax] candidate L: (lower hit, upper hit, module, 0)
  float* hitSlope;    // [hit capacity] per hit: asin(min(rt * k2Rinv1GeVf / ptCut, kSinAlphaMax))
  float4* shiftPoint; // [hit capacity] per pixel-side hit of a module pair whose strip hits are shifted:
                      // the shift point (xa, ya) and the shifted z (md_shift_point)
};

// Work arrays of the flattened line-segment builder: candidates of module m
// are its (connected module k, inner MD, outer MD) triples in that order;
// connPre holds, per module, the first candidate of each connected module.
constexpr uint32_t kLSCandMax = 1u << 22;
constexpr int kLSParkWords = 6;
struct LSWork {
  uint32_t* connPre;  // [nLower * (stride + 1)] per module: candidate prefix over its connected modules
  uint32_t* candOff;  // [nLower + 1]
  uint32_t* park;     // [kLSParkWords * kLSCandMax] SegOut of passing candidate L at L
  uint2* cand;        // [kLSCandMax] candidate L: (inner MD, outer MD); the modules are in the MD records
  MDField<float2, kMDHotBytes> mdSlope;  // in the hot MD record: asin(min(rt * k2Rinv1GeVf / ptCut, kSinAlphaMax)), its tan
};

// Work arrays of the flattened T3 builder (same scheme as the T5 one below).
constexpr uint32_t kT3CandMax = 1u << 22;
constexpr int kT3ResultWords = 6;
struct T3Work {
  uint32_t* modOff;   // [nLower + 1] per lower module: candidate count, then offsets
  uint32_t* list;     // [2 * kT3CandMax] (module, inner/outer segment offsets) of the candidates passing the
                      //   shared-MD check, in serial order
  uint32_t* results;  // [kT3ResultWords * kT3CandMax] selection results of list entry i at i
  uint32_t* survOff;  // [nLower + 1] per lower module: shared-MD candidate count, then offsets
  uint32_t* segByMD;  // [segment capacity] each module's segments grouped by inner mini-doublet, in slot order
  uint32_t* segTmp;   // [segment capacity] the same grouping before the ordering
  uint32_t* mdSegOff; // [MD capacity] per mini-doublet: first position of its segments in segByMD
  uint32_t* mdSegEnd; // [MD capacity] per mini-doublet: end of its segments in segByMD
};

// Work arrays of the flattened T5 builder: every (inner T3, outer T3)
// candidate of the event gets a global index L (eligible module, then inner
// T3, then outer T3: the reference's serial order).
constexpr uint32_t kT5CandMax = 1u << 21;    // candidates per event handled by the flattened path
constexpr uint32_t kT5ResultMax = 1u << 19;  // passing candidates per event parked by the flattened path
struct T5Work {
  uint32_t* modOff;   // [nLower + 2] per eligible-list position: candidate count, then offsets
  uint32_t* resIdx;   // [kT5CandMax] per list-A entry: its result slot when it passes, ~0u otherwise
  uint32_t* listA;    // [3 * kT5CandMax] (eligible-list position, inner T3, outer T3) of the candidates passing
                      //   the shared-MD check, in serial order
  uint32_t* listB;    // [3 * kT5CandMax] (list-A position, inner T3, outer T3) passing both T4-style tests, any order
  uint32_t* results;  // [kT5ResultWords * kT5ResultMax] selection results of passing candidates
  uint32_t* survOff;  // [nLower + 2] per eligible-list position: shared-MD candidate count, then offsets
  uint32_t* t3ByMD;   // [T3 capacity] each module's T3s grouped by first mini-doublet, in slot order
  uint32_t* t3Tmp;    // [T3 capacity] the same grouping before the ordering
  uint32_t* mdT3Off;  // [MD capacity] per mini-doublet: first position of the T3s starting at it in t3ByMD
  uint32_t* mdT3End;  // [MD capacity] per mini-doublet: end of those T3s in t3ByMD
};
constexpr int kT5ResultWords = 12;

// Per-pLS and per-module work arrays of the pixel stages.
struct PixelWork {
  uint32_t* connBase;     // [kNMaxPixelSegments] first connected-module slot in the pixel map
  uint32_t* connSize;     // [kNMaxPixelSegments]
  uint32_t* pt5Count;     // [kNMaxPixelSegments + 1] passing pT5 candidates per pLS, then offsets
  uint32_t* pt3Count;     // [kNMaxPixelSegments + 1]
  uint32_t* t5TcCount;    // [nLower + 1] T5 track candidates per eligible module, then offsets
  uint32_t* outHits;      // [TC capacity * 14] output hit lists
  uint32_t* outLen;       // [TC capacity]
  uint32_t* outPacked;    // [kOutHeaderWords + TC capacity * kOutRecWords] the count, then per track
                          // candidate its 14 hit slots, length, seed index and type: one copy to the host
  uint32_t* t5ListOff;    // [nLower + 1] first list position of each eligible T5 module, then the total
  uint32_t* t5List;       // [T5 capacity] T5 slots of the eligible modules
  uint32_t* t5ListMod;    // [T5 capacity] eligible-list position of the module of each listed T5
  float4* t5ListRec;      // [T5 capacity] (eta, phi, score, slot bits) of each listed T5
  float* tcEta;           // [TC capacity] direction a pLS is compared with in crossCleanpLS
  float* tcPhi;           // [TC capacity]
  uint8_t* tcKind;        // [TC capacity] 0: not compared, 1: T5, 2: pT3 or pT5
  uint32_t* plsBinned;    // [pLS capacity] pLS indices listed eta bin by eta bin
  uint32_t* plsBinOff;    // [kEtaBinOffsets] first list position of each eta bin, then the end
  float4* plsBinRec;      // [pLS capacity] (eta, phi, score, index | isQuad << 31) of each binned pLS
  uint4* plsBinHits;      // [pLS capacity] its four hits
  uint32_t* t5Binned;     // [T5 capacity] listed T5 slots, eta bin by eta bin
  uint32_t* t5BinOff;     // [kEtaBinOffsets]
  float4* t5BinRec;       // [T5 capacity] the same records, eta bin by eta bin
  float4* tcBinRec;       // [TC capacity] per track candidate, eta bin by eta bin: (eta, phi, the dR2 bound
                          // of crossCleanpLS, 0), the bound -1 for one that is not compared
  uint32_t* tcBinOff;     // [kEtaBinOffsets]
  int32_t* tcHitSet;      // [kTCHitSet] open-addressing set of the pLS hits of the pT3/pT5 candidates
  float2* pixBinRec;      // [pT5 + pT3 capacity] (eta, phi) of pT5 j and pT3 (n5 + j), eta bin by eta bin
  uint32_t* pixBinOff;    // [kEtaBinOffsets]
  unsigned long long* pt5Mask;  // [kPixMaskWords] pass mask per 64-candidate chunk of each pLS (pT5)
  uint32_t* pt5MaskBase;        // [pLS capacity] first mask word of the pLS, ~0: none
  unsigned long long* pt3Mask;  // [kPixMaskWords] (pT3)
  uint32_t* pt3MaskBase;        // [pLS capacity]
  uint32_t* pt5Shard;           // [2 * kPixShards] per shard: mask words taken, survivors taken (pT5)
  uint32_t* pt3Shard;           // [2 * kPixShards] (pT3)
  uint32_t* pixWords;           // [pLS capacity] mask words of the pLS (the builder running)
  uint8_t* pixFull;             // [pLS capacity] 1: the pLS's survivors did not all fit, evaluated in full
  uint32_t* pixSurv;            // [kSurvWords * kPixSurvMax] survivors of the cheap cuts: pLS, mask
                                //   position, object, partial result
  uint32_t* t3ListOff;          // [nLower] first list position of each module's T3s (pixel-triplet modules)
  uint32_t* t3List;             // [T3 capacity] T3 slots of those modules
  float* t5Rec;                 // [kT5RecWords * T5 capacity] pLS-independent quantities per T5
  float* t3Rec;                 // [kT3RecWords * T3 capacity] the same per T3
  uint32_t* pt5BucketOff;       // [kPT5HashBuckets + 1] entry counts of the hit buckets, then first positions
  uint32_t* pt5BucketEnt;       // [14 * pT5 capacity] (pT5, position) entries 14 * pT5 + position, by bucket
};
constexpr uint32_t kPT5HashBuckets = 1u << 15;
constexpr uint32_t kTCHitSet = 1u << 18;  // > 4 * pixel TC capacity (120000): the probes stay short
constexpr uint32_t kPixMaskWords = 1u << 16;  // 4 Mi candidates per builder and event; beyond: evaluated in full
constexpr uint32_t kPixSurvMax = 1u << 18;    // survivors of the cheap cuts per builder; beyond: evaluated in full
constexpr uint32_t kSurvWords = 10;           // pLS, mask position, object, up to 7 result words
// pLS i takes its mask words and survivor records from shard i % kPixShards,
// each shard owning an equal slice of both arrays, so that no counter is
// shared by more than a few dozen waves
constexpr uint32_t kPixShards = 64;
constexpr uint32_t kShardWords = kPixMaskWords / kPixShards;
constexpr uint32_t kShardSurv = kPixSurvMax / kPixShards;
constexpr int kT5RecWords = 32;               // per-T5 record of the pixel-quintuplet selection
constexpr int kT3RecWords = 16;               // per-T3 record of the pixel-triplet selection
constexpr uint32_t kOutHeaderWords = 64, kOutRecWords = 17;  // packed output: 256-byte header, 68-byte records
constexpr int kEtaBins = 64;                    // finite eta bins of width 1/8 over [-4, 4)
constexpr int kNanEtaBin = kEtaBins;            // the bin of NaN eta
constexpr int kEtaBinOffsets = kEtaBins + 2;    // starts of bins 0..kNanEtaBin, then the end

// Per-module counters and ranges (SDL/Module.h:21-85, SDL/Hit.h:28-32).
struct RangesSoA {
  int32_t* hitRanges;        // [2 nModules]
  int32_t* hitRangesLower;   // [nModules]
  int32_t* hitRangesUpper;
  int8_t* hitRangesnLower;
  int8_t* hitRangesnUpper;
  int32_t* mdRanges;         // [2 nModules]
  int32_t* segmentRanges;
  int32_t* tripletRanges;
  int32_t* quintupletRanges;
  int32_t* tripletModuleIndices;    // [nLower]
  int32_t* tripletModuleOccupancy;  // [nLower]
  uint32_t* nMDs;            // [nLower + 1]
  uint32_t* totOccupancyMDs;
  uint32_t* nSegments;       // [nLower + 1]
  uint32_t* totOccupancySegments;
  uint32_t* nTriplets;       // [nLower]
  uint32_t* totOccupancyTriplets;
  int32_t* quintupletModuleIndices;    // [nLower], -1 where no T5 can start
  int32_t* quintupletModuleOccupancy;  // [nLower]
  uint32_t* nQuintuplets;              // [nLower]
  uint32_t* totOccupancyQuintuplets;   // [nLower]
  uint16_t* indicesOfEligibleT5Modules;  // [nLower]
  uint32_t* scalars;         // [16] device-side totals (see Scalar)
};
enum Scalar : int {
  kNTotalTrips = 0,
  kBadHits = 1,
  kNEligibleT5 = 2,
  kNTotalQuints = 3,
  kNPixelQuintuplets = 4,
  kTotOccPixelQuintuplets = 5,
  kNPixelTriplets = 6,
  kTotOccPixelTriplets = 7,
  kNTrackCandidates = 8,
  kNTrackCandidatespT3 = 9,
  kNTrackCandidatespT5 = 10,
  kNTrackCandidatespLS = 11,
  kNTrackCandidatesT5 = 12,
  kT5CandTotal = 13,
  kT5ListA = 14,
  kT5ListB = 15,
  kT5Results = 16,
  kT3CandTotal = 17,
  kT3List = 18,
  kT5Listed = 20,
  kPT5Survivors = 23,
  kPT3Survivors = 24,
  kMDCandTotal = 25,
  kLSCandTotal = 27,
  kEvNHits = 29,  // hits of the event (set before the hit kernels run)
  kEvNPLS = 30,   // pixel line segments of the event
  kPixFull = 31,  // bit 0: some pLS evaluated in full by the pT5 builder, bit 1: by the pT3 builder
  kT3Listed = 32,  // T3 slots listed for the pixel-triplet records
  kTruncated = 33,  // bits of TruncBit: objects dropped at a capacity, as the reference drops them
  kNScalars = 40
};
// kBadHits bits: input the reference cannot represent (it indexes out of bounds or wraps a count there)
enum BadHitBit : uint32_t {
  kBadDetId = 1,         // a hit's detId is no module of the geometry (SDL/Hit.h:245-246 would index mapIdx[-1])
  kBadModuleHits = 2,    // more than 127 hits on one module (the int8 counts of SDL/Hit.h:31-32 wrap)
  kBadNonContiguous = 4  // a module's hits are not contiguous in the input (the ranges of SDL/Hit.h:261-267 assume so)
};
// kTruncated bits
enum TruncBit : uint32_t {
  kTruncMD = 1, kTruncLS = 2, kTruncT3 = 4, kTruncT5 = 8, kTruncPT5 = 16, kTruncPT3 = 32
};
// kOutOverflow bits: the flattened builders' work arrays could not hold the event's candidates, so the stage
// needs the per-module builder (run by the per-event path)
enum OverflowBit : uint32_t { kOvfMD = 1, kOvfLS = 2, kOvfT3 = 4, kOvfT5 = 8, kOvfPerModule = 16 };
// words of the packed output header
enum OutHeader : int { kOutNTC = 0, kOutOverflow = 1, kOutTruncated = 2, kOutBadHits = 3 };
}  // namespace lst
namespace lst {
static_assert(kScanCandMax >= kMDCandMax && kScanCandMax >= kLSCandMax && kScanCandMax >= kT3CandMax &&
                  kScanCandMax >= kT5CandMax,
              "the shared pass-flag scan must cover every flattened builder");
}  // namespace lst
