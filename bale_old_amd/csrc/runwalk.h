// runwalk.h -- one workgroup's walk over a region's entries when entries are
// grouped by region inside chunks (one u16 run start per chunk and region,
// transposed to one row per region: offT[r][c] is where region r's run starts
// in chunk c's entries, offT[r + 1][c] where it ends).  Used by the random
// permutation's slot regions and the gather's LDS-served table regions.
#pragma once
#include "bale_internal.h"

namespace bale {

// Walk of region r's entries by the workgroup's NT / 64 waves, each on its
// own: wave w takes the groups of 64 chunks g = w, w + NT / 64, ...; lane l
// holds chunk 64 g + l's run (its start and the entries of the group's earlier
// runs), and the group's entries f = 0, 1, ... go to lanes in turn, Q per lane
// at a time (consecutive entries on consecutive lanes, so a load instruction
// touches about one line per run).  An entry's run is the last lane whose
// prefix is <= f: its block of 8 lanes from the eight block prefixes held in
// scalar registers, then three cross-lane steps inside the block.  The next
// group's run starts are loaded while the current one is served.  No barrier:
// the waves meet only where the caller syncs.  Chunk c's entries start at
// c << CB.  setup(chunk) runs on each lane for its chunk (chunk >= C past the
// end); fn(pos[Q], val[Q]) gets Q entry positions (-1: none) and setup's
// value for each one's chunk.
template <int NT, int CB, int Q, typename Setup, typename Fn>
__device__ inline void region_walk(const uint16_t* __restrict__ offT, int64_t C, int r, Setup setup, Fn fn) {
  constexpr int kWaves = NT / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t ng = (C + 63) / 64;
  const uint16_t* rowa = offT + (int64_t)r * C;
  const uint16_t* rowb = offT + (int64_t)(r + 1) * C;
  uint32_t na = 0, nb = 0;
  {
    const int64_t c = (int64_t)wv * 64 + lane;
    if (c < C) {
      na = rowa[c];
      nb = rowb[c];
    }
  }
  for (int64_t g = wv; g < ng; g += kWaves) {
    const int64_t c = g * 64 + lane;
    const uint32_t a = na, len = nb - na;
    const int64_t cn = c + (int64_t)kWaves * 64;
    na = nb = 0;
    if (cn < C) {
      na = rowa[cn];
      nb = rowb[cn];
    }
    const uint32_t val = setup(c);
    uint32_t incl = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o);
      if (lane >= o) incl += t;
    }
    const uint32_t pre = incl - len;  // lanes past the group's end: pre = total
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    uint32_t bp[8];
#pragma unroll
    for (int k = 0; k < 8; k++) bp[k] = (uint32_t)__builtin_amdgcn_readlane((int)pre, 8 * k);
    // entry f of run l sits at chunk l's first entry + a_l + (f - pre_l)
    const int32_t delta = (int32_t)(a - pre);
    const int64_t gbase = (g * 64) << CB;
    for (uint32_t f0 = 0; f0 < total; f0 += Q * 64) {
      int64_t pos[Q];
      uint32_t v[Q];
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const uint32_t f = f0 + (uint32_t)(q * 64 + lane);
        int lo = 0;
#pragma unroll
        for (int k = 1; k < 8; k++) lo += bp[k] <= f ? 8 : 0;
#pragma unroll
        for (int step = 4; step >= 1; step >>= 1)
          if ((uint32_t)__shfl((int)pre, lo + step) <= f) lo += step;
        const int32_t dl = __shfl(delta, lo);
        v[q] = (uint32_t)__shfl((int)val, lo);
        pos[q] = f < total ? gbase + ((int64_t)lo << CB) + (int64_t)f + dl : -1;
      }
      fn(pos, v);
    }
  }
}

// Region served by workgroup b when a grid of 8 ceil(B / 8) workgroups takes
// one region each: consecutive regions on one XCD (workgroup b runs on XCD
// b % 8), so lines shared by neighbouring regions' runs are read into one L2.
__device__ inline int region_of_block(unsigned b, int B) {
  const int per = (B + 7) / 8;
  return (int)(b % 8) * per + (int)(b / 8);
}

}  // namespace bale
