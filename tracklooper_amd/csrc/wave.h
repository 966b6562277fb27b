// Wave64 / workgroup primitives used for deterministic, order-preserving
// compaction: every builder keeps the reference's serial slot order by ranking
// passing candidates with ballot + mbcnt instead of taking slots with atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lst {

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// number of set bits of `mask` in lanes below the calling lane
__device__ __forceinline__ unsigned lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Stable wave compaction of one 64-candidate chunk: returns this lane's rank
// among the passing lanes and the chunk's pass count.
__device__ __forceinline__ unsigned wave_rank(bool pass, unsigned& count) {
  const uint64_t m = __ballot(pass);
  count = __popcll(m);
  return lanes_below(m);
}

// Inclusive prefix sum across the 64 lanes of a wave (every lane active):
// a Hillis-Steele scan inside each row of 16 lanes with DPP row shifts by 1,
// 2, 4 and 8 (a lane whose source is outside its row adds 0), then the row
// totals carried over with the DPP row broadcasts of lane 15 (into rows 1
// and 3) and lane 31 (into rows 2 and 3).  Register moves only, no LDS.
__device__ __forceinline__ unsigned wave_inclusive_scan(unsigned v) {
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// Row-of-16 primitives: a wave's four rows of 16 lanes work as independent
// groups (each row's lanes all active).  Inclusive prefix sum inside each row
// (the first four steps of the wave scan), the value of the row's last lane,
// and the row's sum in every lane of the row.
constexpr unsigned kRow = 16;
__device__ __forceinline__ unsigned row_inclusive_scan(unsigned v) {
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  return v;
}
__device__ __forceinline__ unsigned row_last(unsigned v) {
  return (unsigned)__shfl((int)v, (int)((__lane_id() & ~(kRow - 1)) | (kRow - 1)), 64);
}
__device__ __forceinline__ unsigned row_sum(unsigned v) {
  for (unsigned off = kRow / 2; off > 0; off >>= 1) v += (unsigned)__shfl_xor((int)v, (int)off, 64);
  return v;
}
// lane `k` of the calling lane's row (k < 16)
__device__ __forceinline__ unsigned row_shfl(unsigned v, unsigned k) {
  return (unsigned)__shfl((int)v, (int)((__lane_id() & ~(kRow - 1)) | k), 64);
}

// value of lane 63 (every lane active)
__device__ __forceinline__ unsigned wave_last(unsigned v) { return (unsigned)__builtin_amdgcn_readlane((int)v, 63); }

// bitwise OR over the 64 lanes of the wave (every lane must take part)
__device__ __forceinline__ unsigned wave_or(unsigned x) {
  for (int off = 32; off > 0; off >>= 1) x |= (unsigned)__shfl_xor((int)x, off, 64);
  return x;
}

// Exclusive prefix sum of a workgroup of BLOCK threads (BLOCK multiple of 64);
// `lds` must hold BLOCK/64 + 1 words.  Returns the exclusive prefix and the
// group total through `total`.
template <int BLOCK>
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* lds, unsigned& total) {
  const unsigned lane = lane_id(), wave = threadIdx.x >> 6;
  const unsigned incl = wave_inclusive_scan(v);
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned run = 0;
    for (int w = 0; w < BLOCK / 64; w++) {
      unsigned t = lds[w];
      lds[w] = run;
      run += t;
    }
    lds[BLOCK / 64] = run;
  }
  __syncthreads();
  const unsigned r = lds[wave] + incl - v;
  total = lds[BLOCK / 64];
  __syncthreads();
  return r;
}

// kN unsigned values scanned side by side
template <int kN>
struct UVec {
  unsigned v[kN];
};

// kN exclusive scans of n values each by one 1024-lane work-group.  Wave w
// owns the 64 * kPer consecutive values from 64 * kPer * w and loads them row
// by row (coalesced, every load in flight before any use), scans the rows with
// a running carry, and wave 0 scans the 16 wave totals to join the waves (n
// beyond 1024 * kPer takes further rounds); the kN scans run interleaved.
// get(i) gives the values of i (a UVec<kN>); put(i, prefixes, values)
// receives their exclusive prefixes.  Returns the totals.  Every lane must
// call it; lds needs 17 * kN entries.
template <int kPer, int kN, class Get, class Put>
__device__ __forceinline__ UVec<kN> scan_wide_n(uint32_t n, Get get, Put put, unsigned* lds) {
  const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  UVec<kN> carry;
#pragma unroll
  for (int c = 0; c < kN; c++) carry.v[c] = 0;
  for (uint32_t b = 0; b < n; b += 1024u * kPer) {
    const uint32_t base = b + 64u * kPer * w + lane;
    unsigned x[kN][kPer], pre[kN][kPer], run[kN];
#pragma unroll
    for (int k = 0; k < kPer; k++) {  // get() of a clamped index, so that no branch splits the loads
      const uint32_t i = base + 64u * k;
      const UVec<kN> g = get(i < n ? i : n - 1);
#pragma unroll
      for (int c = 0; c < kN; c++) x[c][k] = i < n ? g.v[c] : 0u;
    }
#pragma unroll
    for (int c = 0; c < kN; c++) run[c] = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
#pragma unroll
      for (int c = 0; c < kN; c++) {
        const unsigned incl = wave_inclusive_scan(x[c][k]);
        pre[c][k] = run[c] + incl - x[c][k];
        run[c] += wave_last(incl);
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < kN; c++) lds[17 * c + w] = run[c];
    }
    __syncthreads();
    if (threadIdx.x < 64) {
#pragma unroll
      for (int c = 0; c < kN; c++) {
        const unsigned t = lane < 16 ? lds[17 * c + lane] : 0u;
        const unsigned inc = wave_inclusive_scan(t);
        if (lane < 16) lds[17 * c + lane] = inc - t;
        if (lane == 15) lds[17 * c + 16] = inc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = base + 64u * k;
      if (i < n) {
        UVec<kN> p, val;
#pragma unroll
        for (int c = 0; c < kN; c++) {
          p.v[c] = carry.v[c] + lds[17 * c + w] + pre[c][k];
          val.v[c] = x[c][k];
        }
        put(i, p, val);
      }
    }
#pragma unroll
    for (int c = 0; c < kN; c++) carry.v[c] += lds[17 * c + 16];
    __syncthreads();
  }
  return carry;
}

// One exclusive scan (scan_wide_n with kN = 1): get(i) gives value i,
// put(i, prefix, value) receives its exclusive prefix; returns the total.
// lds needs 17 entries.
template <int kPer, class Get, class Put>
__device__ __forceinline__ unsigned scan_wide(uint32_t n, Get get, Put put, unsigned* lds) {
  return scan_wide_n<kPer, 1>(
             n, [&](uint32_t i) { return UVec<1>{{get(i)}}; },
             [&](uint32_t i, const UVec<1>& p, const UVec<1>& x) { put(i, p.v[0], x.v[0]); }, lds)
      .v[0];
}
constexpr int kScanPer = 16;  // values per lane of scan_wide over module-sized arrays

// orders this wave's LDS writes before its later LDS reads of other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per mini-doublet of one module, the module's objects (segments or T3s: the
// slots base .. base + n - 1) that start at it, as the range off[X] .. end[X]
// of byMD: lists in mini-doublet order, slot order inside a list.  One wave
// per module, with no work-group barrier (other waves of the work-group work
// on other modules); first(k) is the global index of the mini-doublet object
// base + k starts at, one of the module's nMD from mdBase.  Up to kListMDs
// mini-doublets and kListCap objects: a counting sort in the wave's LDS
// (counts, offsets, scatter, then each object ranked by slot inside its short
// list).  A larger module counts positions directly (objects with a lower
// mini-doublet plus those of the same one with a lower slot) over its keys
// staged kListCap at a time, with partial counts kept in tmp, off and end
// (byCount sends every module there).
constexpr uint32_t kListCap = 512, kListMDs = 128;
struct ListLds {
  uint32_t x[kListCap + 4];  // mini-doublet of each object, relative to mdBase
  uint32_t tmp[kListCap];    // objects grouped by mini-doublet
  uint32_t cnt[kListMDs];    // counts, then scatter cursors (list ends afterwards)
  uint32_t off[kListMDs];    // list starts, relative to base
};
template <class First>
__device__ void md_lists(uint32_t nMD, uint32_t mdBase, uint32_t n, uint32_t base, First first, uint32_t* off,
                         uint32_t* end, uint32_t* byMD, uint32_t* tmp, ListLds& L, bool byCount) {
  const unsigned lane = threadIdx.x & 63u;
  if (nMD <= kListMDs && n <= kListCap && !byCount) {
    wave_lds_sync();
    for (uint32_t k = lane; k < n; k += 64) L.x[k] = first(k) - mdBase;
    for (uint32_t k = lane; k < nMD; k += 64) L.cnt[k] = 0;
    wave_lds_sync();
    for (uint32_t k = lane; k < n; k += 64) atomicAdd(&L.cnt[L.x[k]], 1u);
    wave_lds_sync();
    {
      const uint32_t b0 = 2 * lane, b1 = b0 + 1;
      const uint32_t c0 = b0 < nMD ? L.cnt[b0] : 0u, c1 = b1 < nMD ? L.cnt[b1] : 0u;
      const uint32_t incl = wave_inclusive_scan(c0 + c1), ex = incl - c0 - c1;
      wave_lds_sync();
      if (b0 < nMD) {
        L.off[b0] = L.cnt[b0] = ex;
        off[mdBase + b0] = base + ex;
        end[mdBase + b0] = base + ex + c0;
      }
      if (b1 < nMD) {
        L.off[b1] = L.cnt[b1] = ex + c0;
        off[mdBase + b1] = base + ex + c0;
        end[mdBase + b1] = base + incl;
      }
    }
    wave_lds_sync();
    for (uint32_t k = lane; k < n; k += 64) L.tmp[atomicAdd(&L.cnt[L.x[k]], 1u)] = k;
    wave_lds_sync();
    for (uint32_t k = lane; k < n; k += 64) {
      const uint32_t X = L.x[k], b = L.off[X], e = L.cnt[X];
      uint32_t r = b;
      for (uint32_t q = b; q < e; q++) r += L.tmp[q] < k ? 1u : 0u;
      byMD[base + r] = base + k;
    }
    return;
  }
  const uint4* s4 = reinterpret_cast<const uint4*>(L.x);
  for (uint32_t c0 = 0; c0 == 0 || c0 < n; c0 += kListCap) {
    const uint32_t cn = n - c0 < kListCap ? n - c0 : kListCap;
    wave_lds_sync();
    for (uint32_t k = lane; k < cn + 4; k += 64) L.x[k] = k < cn ? first(c0 + k) - mdBase : ~0u;  // ~0: none
    wave_lds_sync();
    for (uint32_t k = lane; k < n; k += 64) {
      const uint32_t X = first(k) - mdBase;
      uint32_t r = 0;
      for (uint32_t q = 0; q < cn; q += 4) {
        const uint4 y = s4[q / 4];
        const uint32_t g = c0 + q;
        r += (y.x < X) + (y.y < X) + (y.z < X) + (y.w < X);
        r += ((y.x == X) & (g < k)) + ((y.y == X) & (g + 1 < k)) + ((y.z == X) & (g + 2 < k)) +
             ((y.w == X) & (g + 3 < k));
      }
      tmp[base + k] = (c0 ? tmp[base + k] : 0u) + r;
    }
    for (uint32_t k = lane; k < nMD; k += 64) {
      uint32_t lo = 0, eq = 0;
      for (uint32_t q = 0; q < cn; q += 4) {
        const uint4 y = s4[q / 4];
        lo += (y.x < k) + (y.y < k) + (y.z < k) + (y.w < k);
        eq += (y.x == k) + (y.y == k) + (y.z == k) + (y.w == k);
      }
      off[mdBase + k] = (c0 ? off[mdBase + k] : 0u) + lo;
      end[mdBase + k] = (c0 ? end[mdBase + k] : 0u) + eq;
    }
  }
  for (uint32_t k = lane; k < n; k += 64) byMD[base + tmp[base + k]] = base + k;
  for (uint32_t k = lane; k < nMD; k += 64) {
    const uint32_t lo = off[mdBase + k];
    off[mdBase + k] = base + lo;
    end[mdBase + k] = base + lo + end[mdBase + k];
  }
}

// The counting sort of md_lists by one 16-lane row, for a module of at most
// kRowListMDs mini-doublets and kRowListCap objects, in the row's quarter of
// the wave's list LDS (the row's lanes all active).
constexpr uint32_t kRowListCap = 128, kRowListMDs = 2 * kRow;
struct RowLds {
  uint32_t x[kRowListCap];    // mini-doublet of each object, relative to mdBase
  uint32_t tmp[kRowListCap];  // objects grouped by mini-doublet
  uint32_t cnt[kRowListMDs];  // counts, then scatter cursors (list ends afterwards)
  uint32_t off[kRowListMDs];  // list starts, relative to base
};
union WaveLists {  // one wave's list LDS: the whole of it for one module, or a quarter per row
  ListLds full;
  RowLds rows[4];
};
template <class First>
__device__ void md_lists_row(uint32_t nMD, uint32_t mdBase, uint32_t n, uint32_t base, First first, uint32_t* off,
                             uint32_t* end, uint32_t* byMD, RowLds& L) {
  const unsigned gl = __lane_id() % kRow;
  wave_lds_sync();
  for (uint32_t k = gl; k < n; k += kRow) L.x[k] = first(k) - mdBase;
  for (uint32_t k = gl; k < nMD; k += kRow) L.cnt[k] = 0;
  wave_lds_sync();
  for (uint32_t k = gl; k < n; k += kRow) atomicAdd(&L.cnt[L.x[k]], 1u);
  wave_lds_sync();
  {
    const uint32_t b0 = 2 * gl, b1 = b0 + 1;
    const uint32_t c0 = b0 < nMD ? L.cnt[b0] : 0u, c1 = b1 < nMD ? L.cnt[b1] : 0u;
    const uint32_t incl = row_inclusive_scan(c0 + c1), ex = incl - c0 - c1;
    wave_lds_sync();
    if (b0 < nMD) {
      L.off[b0] = L.cnt[b0] = ex;
      off[mdBase + b0] = base + ex;
      end[mdBase + b0] = base + ex + c0;
    }
    if (b1 < nMD) {
      L.off[b1] = L.cnt[b1] = ex + c0;
      off[mdBase + b1] = base + ex + c0;
      end[mdBase + b1] = base + incl;
    }
  }
  wave_lds_sync();
  for (uint32_t k = gl; k < n; k += kRow) L.tmp[atomicAdd(&L.cnt[L.x[k]], 1u)] = k;
  wave_lds_sync();
  for (uint32_t k = gl; k < n; k += kRow) {
    const uint32_t X = L.x[k], b = L.off[X], e = L.cnt[X];
    uint32_t r = b;
    for (uint32_t q = b; q < e; q++) r += L.tmp[q] < k ? 1u : 0u;
    byMD[base + r] = base + k;
  }
}

// md_lists for every lower module, four modules per wave: when all four fit a
// row (kRowListMDs, kRowListCap, byCount off; most modules hold a few
// objects), each 16-lane row sorts one; otherwise the whole wave takes them
// one after the other.  info(m, nMD, mdBase, n, base) describes module m;
// firstOf(object) is the mini-doublet the object starts at.
template <class Info, class FirstOf>
__device__ void md_lists_all(uint32_t nL, Info info, FirstOf firstOf, uint32_t* off, uint32_t* end, uint32_t* byMD,
                             uint32_t* tmp, WaveLists& W, bool byCount) {
  const unsigned g = __lane_id() / kRow;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  const uint32_t nWaves = gridDim.x * blockDim.x / 64;
  for (uint32_t m0 = 4 * wave; m0 < nL; m0 += 4 * nWaves) {
    const uint32_t m = m0 + g;
    uint32_t nMD = 0, mdBase = 0, n = 0, base = 0;
    if (m < nL) info(m, nMD, mdBase, n, base);
    const bool fits = nMD <= kRowListMDs && n <= kRowListCap;
    if (!byCount && __ballot(!fits) == 0ull) {  // the same branch in every lane
      if (nMD != 0)
        md_lists_row(nMD, mdBase, n, base, [&](uint32_t k) { return firstOf(base + k); }, off, end, byMD, W.rows[g]);
    } else {
      for (uint32_t r = 0; r < 4 && m0 + r < nL; r++) {
        uint32_t nMDr = 0, mdBaser = 0, nr = 0, baser = 0;
        info(m0 + r, nMDr, mdBaser, nr, baser);
        if (nMDr == 0) continue;
        md_lists(nMDr, mdBaser, nr, baser, [&](uint32_t k) { return firstOf(baser + k); }, off, end, byMD, tmp, W.full,
                 byCount);
      }
    }
  }
}

}  // namespace lst
