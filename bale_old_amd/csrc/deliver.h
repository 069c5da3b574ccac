// deliver.h -- stable delivery of items to destinations (one GPU).
//
// A chunk of items is cut into tiles; each wave ranks its consecutive items
// of a tile per destination with ballots, the waves are prefixed in input
// order, the tile is staged in LDS in destination order and each
// destination's run is written to consecutive output slots.  Per-(chunk,
// destination) counts from a count pass (rows) and the destinations' first
// slots (start) place every run, so the output is the input stably
// partitioned by destination.  Used by the device conveyor and the exstack
// services (destination = the PE number pushed with the item).
#pragma once
#include "bale_internal.h"

namespace bale {

constexpr int kMaxPes = 256;   // virtual PEs: 8 ballots rank a tile's items
constexpr int kDlvTile = 1024;    // items ranked per step (one per thread)
constexpr int kDlvWaves = kDlvTile / 64;

// Item copies with the widest aligned vector the item size allows
// (16-byte pieces for 16/32-byte items, 8-byte for 8/24, 4-byte for 4/12).
template <int W>
__device__ inline void item_load(const uint32_t* __restrict__ src, size_t i, uint32_t* v) {
  if constexpr (W % 4 == 0) {
    const u32x4* p = (const u32x4*)(src + i * W);
#pragma unroll
    for (int k = 0; k < W / 4; k++) {
      const u32x4 x = p[k];
      v[4 * k] = x.x; v[4 * k + 1] = x.y; v[4 * k + 2] = x.z; v[4 * k + 3] = x.w;
    }
  } else if constexpr (W % 2 == 0) {
    const uint2* p = (const uint2*)(src + i * W);
#pragma unroll
    for (int k = 0; k < W / 2; k++) {
      const uint2 x = p[k];
      v[2 * k] = x.x; v[2 * k + 1] = x.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < W; k++) v[k] = src[i * W + k];
  }
}
template <int W>
__device__ inline void item_store(uint32_t* __restrict__ dst, size_t i, const uint32_t* v) {
  if constexpr (W % 4 == 0) {
    u32x4* p = (u32x4*)(dst + i * W);
#pragma unroll
    for (int k = 0; k < W / 4; k++) p[k] = u32x4{v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
  } else if constexpr (W % 2 == 0) {
    uint2* p = (uint2*)(dst + i * W);
#pragma unroll
    for (int k = 0; k < W / 2; k++) p[k] = make_uint2(v[2 * k], v[2 * k + 1]);
  } else {
#pragma unroll
    for (int k = 0; k < W; k++) dst[i * W + k] = v[k];
  }
}

// Items per thread per tile: as many as the LDS staging allows (4 for items
// up to 16 bytes, 2 above).
template <int W>
constexpr int deliver_u() { return W <= 4 ? 4 : 2; }

// LDS of k_convey_deliver<W>: per wave and PE a count, then an offset; PE
// totals, tile starts and output cursors; the staged items and their PEs.
template <int W>
inline size_t deliver_lds(int npe) {
  constexpr int U = deliver_u<W>();
  size_t words = (size_t)kDlvWaves * npe + 3 * (size_t)kMaxPes + 1;
  words = (words + 3) & ~(size_t)3;  // 16-byte alignment for the staged items
  return words * 4 + (size_t)U * kDlvTile * W * 4 + (size_t)U * kDlvTile * 2;
}

// Stable delivery of chunk blockIdx.x.  W: 32-bit words per item.  A tile is
// U x 1024 consecutive items; wave w owns items w * 64U .. w * 64U + 64U - 1
// of it, 64 per round (item u * 64 + lane of the wave in round u).  A wave
// ranks each round's items per PE with ballots on top of its running count
// per PE (LDS), and the waves are prefixed in order, so the ranks follow push
// order.  rows[c][p]: items of PE p in chunks before c; start[p]: PE p's
// first slot.
template <int W>
__global__ __launch_bounds__(kDlvTile) void k_convey_deliver(const uint32_t* __restrict__ items,
                                                          const int64_t* __restrict__ pes,
                                                          int64_t n, int npe, int nbits,
                                                          int64_t chunk,
                                                          const uint32_t* __restrict__ rows,
                                                          const uint32_t* __restrict__ start,
                                                          uint32_t* __restrict__ out) {
  constexpr int U = deliver_u<W>();
  extern __shared__ __align__(16) uint32_t lds[];
  uint32_t* wcnt = lds;                                  // [kDlvWaves * npe]: count, then offset
  uint32_t* tot = wcnt + (size_t)kDlvWaves * npe;        // [kMaxPes] tile items per PE
  uint32_t* tstart = tot + kMaxPes;                      // [kMaxPes + 1] first staged slot per PE
  uint32_t* cursor = tstart + kMaxPes + 1;               // [kMaxPes] next output slot per PE
  size_t head = ((size_t)kDlvWaves * npe + 3 * (size_t)kMaxPes + 1 + 3) & ~(size_t)3;
  uint32_t* stage = lds + head;                          // [U * kDlvTile * W]
  uint16_t* sdst = (uint16_t*)(stage + (size_t)U * kDlvTile * W);  // [U * kDlvTile]
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const int64_t c = blockIdx.x;
  const int64_t lo = c * chunk, hi = lo + chunk < n ? lo + chunk : n;
  for (int p = tid; p < npe; p += kDlvTile) cursor[p] = start[p] + rows[(size_t)c * npe + p];
  const uint64_t below = lane ? (~0ULL >> (64 - lane)) : 0ULL;  // lanes before this one
  const int64_t mine = (int64_t)wave * (64 * U) + lane;           // this lane's first item in a tile
  uint32_t* wc = wcnt + (size_t)wave * npe;
  // a tile's PE numbers and items are loaded while the previous tile is
  // ranked and written
  int64_t pn[U];
  uint32_t vn[U][W];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int64_t e = lo + mine + (int64_t)u * 64;
    pn[u] = -1;
    if (e < hi) {
      item_load<W>(items, (size_t)e, vn[u]);
      pn[u] = pes[e];
    }
  }
  for (int64_t base = lo; base < hi; base += (int64_t)U * kDlvTile) {
    int d[U];
    uint32_t v[U][W];
#pragma unroll
    for (int u = 0; u < U; u++) {
      // npe: nothing to deliver (past the end, or PE out of range)
      d[u] = (uint64_t)pn[u] < (uint64_t)npe ? (int)pn[u] : npe;
#pragma unroll
      for (int k = 0; k < W; k++) v[u][k] = vn[u][k];
      const int64_t e = base + (int64_t)U * kDlvTile + mine + (int64_t)u * 64;
      pn[u] = -1;
      if (e < hi) {
        item_load<W>(items, (size_t)e, vn[u]);
        pn[u] = pes[e];
      }
    }
    for (int k = tid; k < kDlvWaves * npe; k += kDlvTile) wcnt[k] = 0;
    __syncthreads();
    // round by round: lanes of this wave with the same PE (one ballot per bit
    // of d), ranked after the wave's earlier items of that PE
    uint32_t rank[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool has = d[u] < npe;
      uint64_t peers = __ballot(has);
      for (int bit = 0; bit < nbits; bit++) {
        const bool one = (d[u] >> bit) & 1;
        const uint64_t b = __ballot(one);
        peers &= one ? b : ~b;
      }
      const uint32_t r = (uint32_t)__popcll(peers & below);
      const uint32_t before = has ? wc[d[u]] : 0u;  // every lane reads before the leader adds
      rank[u] = before + r;
      if (has && r == 0) wc[d[u]] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // offsets of each wave's items within the PE's part of the tile
    for (int p = tid; p < npe; p += kDlvTile) {
      uint32_t run = 0;
#pragma unroll
      for (int w = 0; w < kDlvWaves; w++) {
        const uint32_t t = wcnt[w * npe + p];
        wcnt[w * npe + p] = run;
        run += t;
      }
      tot[p] = run;
    }
    __syncthreads();
    if (wave == 0) {
      // exclusive scan of tot over the PEs, 4 per lane
      uint32_t sv[4], sum = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int p = lane * 4 + k;
        sv[k] = p < npe ? tot[p] : 0u;
        sum += sv[k];
      }
      uint32_t incl = sum;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
      }
      uint32_t run = incl - sum;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int p = lane * 4 + k;
        if (p < npe) tstart[p] = run;
        run += sv[k];
      }
      if (lane == 63) tstart[npe] = incl;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++)
      if (d[u] < npe) {
        const uint32_t q = tstart[d[u]] + wc[d[u]] + rank[u];
        item_store<W>(stage, q, v[u]);
        sdst[q] = (uint16_t)d[u];
      }
    __syncthreads();
    // staged order is PE order, push order within a PE: each PE's run goes
    // to consecutive output slots
    const uint32_t ntile = tstart[npe];
    for (uint32_t q = tid; q < ntile; q += kDlvTile) {
      const int p = sdst[q];
      const size_t g = (size_t)cursor[p] + q - tstart[p];
      uint32_t w[W];
      item_load<W>(stage, (size_t)q, w);
      item_store<W>(out, g, w);
    }
    __syncthreads();
    for (int p = tid; p < npe; p += kDlvTile) cursor[p] += tot[p];
  }
}

template <int W>
[[maybe_unused]] static void launch_deliver(dim3 g, hipStream_t s, const void* items, const int64_t* pes, int64_t n,
                           int npe, int nbits, int64_t chunk, const uint32_t* rows,
                           const uint32_t* start, void* out) {
  const size_t lds = deliver_lds<W>(npe);
  allow_lds(k_convey_deliver<W>, lds);
  hipLaunchKernelGGL(k_convey_deliver<W>, g, dim3(kDlvTile), lds, s, (const uint32_t*)items, pes, n,
                     npe, nbits, chunk, rows, start, (uint32_t*)out);
}


}  // namespace bale
