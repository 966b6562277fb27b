// Device-wide exclusive scan of pass flags over a flattened candidate list,
// used by the flattened builders to give passing candidates their slots.
//
// A builder numbers its candidates L = 0 .. total-1 in the reference's serial
// order (module by module).  After the selection, flag(L) says whether
// candidate L passed.  Its slot in module m is the number of passing
// candidates in [first(m), L), i.e. rank(L) - rank(first(m)) with
//   rank(L) = pre[L] + sums[L / kScanBlockItems],
// where pre is the exclusive prefix inside a block of kScanBlockItems
// candidates and sums the exclusive prefix of the block totals.  Every
// passing candidate can then be placed independently, so no module is walked
// serially.  rank(total) is the grand total (kept in sums[nBlocks]).
#pragma once
#include <hip/hip_runtime.h>

#include "soa.h"
#include "wave.h"

namespace lst {

constexpr uint32_t kScanBlockItems = 4096;  // 1024 lanes x 4 candidates
constexpr unsigned kScanGrid = 512;         // work-groups of the block pass (they stride over the blocks)

// pass 1: per block of kScanBlockItems candidates, the in-block exclusive
// prefix of the flags, and the block total
template <class Flag>
__device__ __forceinline__ void scan_flag_blocks(uint32_t total, Flag flag, uint32_t* pre, uint32_t* sums,
                                                 uint32_t* pass) {
  __shared__ unsigned lds[17];
  const uint32_t nBlocks = (total + kScanBlockItems - 1) / kScanBlockItems;
  for (uint32_t b = blockIdx.x; b < nBlocks; b += gridDim.x) {
    const uint32_t L0 = b * kScanBlockItems + 4 * threadIdx.x;
    unsigned f[4], c = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      f[k] = (L0 + k < total && flag(L0 + k)) ? 1u : 0u;
      c += f[k];
    }
    unsigned tot;
    unsigned run = block_exclusive_scan<1024>(c, lds, tot);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (L0 + k < total) pre[L0 + k] = run;
      if (f[k]) pass[b * kScanBlockItems + run] = L0 + k;  // its place among the block's passing candidates
      run += f[k];
    }
    if (threadIdx.x == 0) sums[b] = tot;
  }
}

// pass 2 (one 1024-lane work-group): exclusive scan of the block totals;
// sums[nBlocks] gets the grand total
__device__ __forceinline__ void scan_flag_sums(uint32_t total, uint32_t* sums) {
  __shared__ unsigned lds[17];
  const uint32_t nBlocks = (total + kScanBlockItems - 1) / kScanBlockItems;
  unsigned carry = 0;
  for (uint32_t b0 = 0; b0 < nBlocks; b0 += 1024) {
    const uint32_t b = b0 + threadIdx.x;
    const unsigned c = b < nBlocks ? sums[b] : 0u;
    unsigned tot;
    const unsigned ex = block_exclusive_scan<1024>(c, lds, tot);
    if (b < nBlocks) sums[b] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) sums[nBlocks] = carry;
}

// fn(L, rank(L)) for every passing candidate, then each(j) for j < nEach:
// the lanes take the work items grid-stride, first the passing ranks
// r = 0, 1, ... (each lane finds its rank's scan block, the last b with
// sums[b] <= r, never an empty block, by binary search over the block totals),
// then the nEach other items.  One index space, so a lane does one kind of
// item per step and the two chains of dependent loads run side by side in
// different lanes instead of one after the other in each.
template <class Fn, class Each>
__device__ __forceinline__ void for_each_passing(const FlatScan& s, uint32_t total, uint32_t nEach, Fn fn,
                                                 Each each) {
  const uint32_t nBlocks = (total + kScanBlockItems - 1) / kScanBlockItems;
  const uint32_t nPass = s.sums[nBlocks];
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nPass + nEach; r += gridDim.x * blockDim.x) {
    if (r >= nPass) {
      each(r - nPass);
      continue;
    }
    uint32_t lo = 0, hi = nBlocks;  // sums[lo] <= r < sums[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s.sums[mid] <= r) lo = mid;
      else hi = mid;
    }
    fn(s.pass[lo * kScanBlockItems + (r - s.sums[lo])], r);
  }
}

// number of passing candidates before L (L <= total)
__device__ __forceinline__ uint32_t flag_rank(const FlatScan& s, uint32_t L, uint32_t total) {
  if (L >= total) return s.sums[(total + kScanBlockItems - 1) / kScanBlockItems];
  return s.pre[L] + s.sums[L / kScanBlockItems];
}

}  // namespace lst
