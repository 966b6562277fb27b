// Device-wide exclusive scan of pass flags over a flattened candidate list,
// used by the flattened builders to give passing candidates their slots.
//
// A builder numbers its candidates L = 0 .. total-1 in the reference's serial
// order (module by module).  The pass flags are kept as 64-bit words, one per
// 64 consecutive candidates: a selection kernel whose waves each evaluate an
// aligned run of 64 candidates stores each word as one ballot.  Candidate L's
// slot in module m is the number of passing candidates in [first(m), L), i.e.
// rank(L) - rank(first(m)) with
//   rank(L) = sums[L / 4096] + pre[L / 64] + popcount(mask[L / 64] below bit L % 64),
// where pre is the exclusive prefix of the word counts inside a scan block of
// 4096 candidates (64 words) and sums the exclusive prefix of the block totals.
// Every passing candidate can then be placed independently, so no module is
// walked serially.  rank(total) is the grand total (kept in sums[nBlocks]).
#pragma once
#include <hip/hip_runtime.h>

#include "soa.h"
#include "wave.h"

namespace lst {

constexpr uint32_t kScanBlockItems = 4096;  // 64 flag words
constexpr uint32_t kScanBlockWords = kScanBlockItems / 64;
constexpr unsigned kScanGrid = 512;         // work-groups of the block pass (they stride over the blocks)

// the flag word of the calling wave's candidates L0 .. L0 + 63 (L0 a multiple of 64; lane l evaluated
// candidate L0 + l, lanes past the total pass nothing; every lane of the wave takes part), stored when
// L0 < total
__device__ __forceinline__ void store_flag_word(const FlatScan& s, uint32_t L0, uint32_t total, bool pass) {
  const uint64_t w = __ballot(pass);
  if (lane_id() == 0 && L0 < total) s.mask[L0 / 64] = w;
}

// the word pass over scan block b by one wave, lane j holding word b * 64 + j (0 past the last of nWords):
// the words' in-block prefixes, the block total, and the block's passing candidates in order (the words
// with passing candidates one at a time, each lane placing the candidate of its bit)
__device__ __forceinline__ void scan_block_words(const FlatScan& s, uint32_t b, uint32_t nWords, uint64_t m) {
  const unsigned lane = lane_id();
  const uint32_t w = b * kScanBlockWords + lane;
  const unsigned c = (unsigned)__popcll(m);
  const unsigned incl = wave_inclusive_scan(c), ex = incl - c;
  if (w < nWords) s.pre[w] = ex;
  if (lane == 63) s.sums[b] = incl;
  uint32_t* out = s.pass + (size_t)b * kScanBlockItems;
  uint64_t nz = __ballot(m != 0ull);
  while (nz) {
    const int j = __ffsll((unsigned long long)nz) - 1;
    nz &= nz - 1;
    const uint64_t mj = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(m >> 32), j) << 32) |
                        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)m, j);
    const unsigned exj = (unsigned)__builtin_amdgcn_readlane((int)ex, j);
    if ((mj >> lane) & 1ull) out[exj + lanes_below(mj)] = (b * kScanBlockWords + (uint32_t)j) * 64 + lane;
  }
}

// pass 1 over flag words a selection kernel stored: one wave per scan block
__device__ __forceinline__ void scan_mask_blocks(const FlatScan& s, uint32_t total) {
  const uint32_t nWords = (total + 63) / 64, nBlocks = (total + kScanBlockItems - 1) / kScanBlockItems;
  const uint32_t nWaves = gridDim.x * blockDim.x / 64;
  for (uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) / 64; b < nBlocks; b += nWaves) {
    const uint32_t w = b * kScanBlockWords + lane_id();
    scan_block_words(s, b, nWords, w < nWords ? s.mask[w] : 0ull);
  }
}

// pass 1 over flags given per candidate (flag(L)): a 1024-lane work-group per
// scan block forms the block's 64 words (16 waves x 4 ballots, also stored as
// the flag words), then its first wave runs the word pass
template <class Flag>
__device__ __forceinline__ void scan_flag_blocks(const FlatScan& s, uint32_t total, Flag flag) {
  __shared__ uint64_t words[kScanBlockWords];
  const uint32_t nWords = (total + 63) / 64, nBlocks = (total + kScanBlockItems - 1) / kScanBlockItems;
  const unsigned lane = lane_id(), wave = threadIdx.x / 64;
  for (uint32_t b = blockIdx.x; b < nBlocks; b += gridDim.x) {
#pragma unroll
    for (unsigned k = 0; k < 4; k++) {
      const uint32_t word = b * kScanBlockWords + k * 16 + wave, L = word * 64 + lane;
      const bool f = L < total && flag(L);
      const uint64_t m = __ballot(f);
      if (lane == 0) {
        words[k * 16 + wave] = m;
        if (word < nWords) s.mask[word] = m;
      }
    }
    __syncthreads();
    if (wave == 0) scan_block_words(s, b, nWords, words[lane]);
    __syncthreads();
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
  const uint32_t w = L / 64;
  const uint64_t below = s.mask[w] & ((1ull << (L % 64)) - 1ull);
  return s.sums[L / kScanBlockItems] + s.pre[w] + (uint32_t)__popcll(below);
}

}  // namespace lst
