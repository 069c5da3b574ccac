// pk_count.h -- exact accounting for two 16-bit counters in one 32-bit word.
//
// The packed LDS tile of the bucketed histogram keeps bin 2w in the low half
// and bin 2w + 1 in the high half of LDS word w, so a 2^16-bin bucket fits in
// 128 KiB.  The word only ever receives atomic adds of (inc << 16 * half)
// with 1 <= inc < 2^16, so its value is the sum of every add so far, mod 2^32,
// whatever order the lanes' adds reached it in.  What one add did to the two
// halves is therefore fixed by the value that add returned, and the lane that
// issued it can account for it alone, at any later time:
//   - a low-half add whose old low value plus inc reaches 2^16 wrapped the low
//     half: the low bin lost 2^16, which it is owed;
//   - that wrap carried 1 into the high half, which the high bin now holds in
//     excess (it owes 1), and when the high half held 0xffff the carry wrapped
//     it too (it is owed 2^16 as well, and the carry left the word);
//   - a high-half add whose old high value plus inc reaches 2^16 carried out of
//     the word: the high bin is owed 2^16.
// Each half's final value plus what its bin is owed is the bin's count, mod
// 2^32.  A segment adds fewer than 2^32 updates to a tile, so that is exact.
// No half is ever corrected in place, so no add depends on another lane's
// correction having landed first.
//
// Compiled by hipcc into the apply kernels, and by a host C++ compiler (with
// __host__ and __device__ defined empty) into the CPU test that replays
// adversarial add orders through it.
#pragma once
#include <stdint.h>

namespace bale {

// What one add owes the bins of its word, mod 2^32 (add these to the bins'
// tallies; 0 means nothing owed).
struct PkOwed {
  uint32_t own;   // owed to the bin the add was for
  uint32_t high;  // owed to the high bin of the word (low-half adds only)
};

// Whether an add of inc to `half` wrapped that half: the only case in which
// pk_owed owes anything (a low half's carry needs its own wrap).
__host__ __device__ inline bool pk_wrapped(uint32_t old, uint32_t half, uint32_t inc) {
  return ((old >> (16u * half)) & 0xffffu) + inc >= 0x10000u;
}

// old: the word's value the atomic add returned; half: 0 for the low bin (even
// bin number), 1 for the high bin; inc: what was added to that half (< 2^16).
__host__ __device__ inline PkOwed pk_owed(uint32_t old, uint32_t half, uint32_t inc) {
  const uint32_t lo = old & 0xffffu, hi = old >> 16;
  if (half) return PkOwed{hi + inc >= 0x10000u ? 0x10000u : 0u, 0u};
  if (lo + inc < 0x10000u) return PkOwed{0u, 0u};
  // low wrapped; its carry is 1 in excess on the high bin, or, when it wrapped
  // the high half as well, 1 short of a whole 2^16
  return PkOwed{0x10000u, hi == 0xffffu ? 0xffffu : 0xffffffffu};
}

// A bin's count from its half of the final word and its tally.
__host__ __device__ inline uint32_t pk_read(uint32_t word, uint32_t half, uint32_t tally) {
  return ((word >> (16u * half)) & 0xffffu) + tally;
}

}  // namespace bale
