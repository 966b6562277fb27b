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

// inclusive prefix sum across the 64 lanes of a wave
__device__ __forceinline__ unsigned wave_inclusive_scan(unsigned v) {
  const unsigned lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    unsigned t = __shfl_up(v, d, 64);
    if (lane >= (unsigned)d) v += t;
  }
  return v;
}

__device__ __forceinline__ unsigned wave_sum(unsigned v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
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

// Exclusive scan of n values by one 1024-lane work-group.  Wave w owns the
// 64 * kPer consecutive values from 64 * kPer * w and loads them row by row
// (coalesced, every load in flight before any use), scans the rows with a
// running carry, and one work-group scan of the 16 wave totals joins the
// waves (n beyond 1024 * kPer takes further rounds).  get(i) gives value i;
// put(i, prefix, value) receives its exclusive prefix.  Returns the total.
// Every lane must call it; lds needs 17 entries.
template <int kPer, class Get, class Put>
__device__ __forceinline__ unsigned scan_wide(uint32_t n, Get get, Put put, unsigned* lds) {
  const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  unsigned carry = 0;
  for (uint32_t b = 0; b < n; b += 1024u * kPer) {
    const uint32_t base = b + 64u * kPer * w + lane;
    unsigned x[kPer], pre[kPer], run = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) x[k] = base + 64u * k < n ? get(base + 64u * k) : 0u;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const unsigned incl = wave_inclusive_scan(x[k]);
      pre[k] = run + incl - x[k];
      run += __shfl(incl, 63, 64);
    }
    if (lane == 0) lds[w] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned acc = 0;
      for (int k = 0; k < 16; k++) {
        const unsigned t = lds[k];
        lds[k] = acc;
        acc += t;
      }
      lds[16] = acc;
    }
    __syncthreads();
    const unsigned off = carry + lds[w];
#pragma unroll
    for (int k = 0; k < kPer; k++)
      if (base + 64u * k < n) put(base + 64u * k, off + pre[k], x[k]);
    carry += lds[16];
    __syncthreads();
  }
  return carry;
}
constexpr int kScanPer = 16;  // values per lane of scan_wide over module-sized arrays

}  // namespace lst
