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

}  // namespace lst
