// randperm.hip -- random permutation of 0..N-1 by dart throwing
// (spmat/spmat_agi.upc:56-107, rand_permp_agi), and by a Fisher-Yates pass
// (apps/randperm_src/alternates/randperm_agi_opt.upc, rand_permp_agi_opt).
//
// At one PE the reference is sequential: darts r = lrand48() % M (M = 2N) are
// thrown in turn, and the i-th dart that finds its slot empty places item i;
// the permutation lists the filled slots in order.  The outcome has a parallel
// statement: dart k hits iff no earlier dart chose its slot, and the item it
// places is the number of hits before it (placing stops at N items).  From one
// 2^15-slot region per CU up to 2^27 slots the darts are sorted by slot region
// and resolved in LDS (the k_rs_* kernels below, no device atomics and no
// random global accesses); otherwise over the whole slot array:
//   k_darts    each dart's slot, from the drand48 sequence (jump-ahead per thread)
//   k_first    first[slot] = smallest dart index choosing it (device atomic min)
//   k_hits     darts that are their slot's first, flagged (bit 31), counted per tile
//   k_place    first[slot] = the hit's item (or empty from item N on)
//   k_occ / k_compact   filled slots counted per tile, then written in slot order
// with k_scan_tiles turning per-tile counts into offsets.  Every step is exact,
// so the result equals the reference's run at one PE for the same seed.
#include <algorithm>
#include <cmath>

#include "bale_internal.h"
#include "runwalk.h"

namespace bale {

namespace rp {

constexpr uint64_t kA = 0x5DEECE66DULL, kC = 0xBULL, kMask48 = (1ULL << 48) - 1;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kThreads = 1024;
constexpr int kPer = 16;                     // elements per thread per tile
constexpr int kTile = kThreads * kPer;       // 16384 elements per tile
constexpr int kDartsPerThread = 16;          // consecutive darts per thread in k_darts

// X -> a X + c (mod 2^48); composing two such maps of the same generator.
struct Lcg {
  uint64_t a, c;
};
__host__ __device__ inline Lcg lcg_then(Lcg f, Lcg g) {  // f, then g
  return {(g.a * f.a) & kMask48, (g.a * f.c + g.c) & kMask48};
}
// The map of k steps of the generator.
__host__ __device__ inline Lcg lcg_steps(uint64_t k) {
  Lcg r{1, 0}, b{kA, kC};
  while (k) {
    if (k & 1) r = lcg_then(r, b);
    b = lcg_then(b, b);
    k >>= 1;
  }
  return r;
}

// slot[k] = (lrand48 value k) mod M: lrand48 returns the top 31 bits of the
// state after k + 1 steps from x0.
__global__ __launch_bounds__(256) void k_darts(uint64_t x0, uint64_t D, uint32_t M,
                                               uint32_t* __restrict__ slot) {
  const uint64_t k0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kDartsPerThread;
  if (k0 >= D) return;
  const Lcg j = lcg_steps(k0);
  uint64_t x = (j.a * x0 + j.c) & kMask48;
  uint32_t v[kDartsPerThread];
#pragma unroll
  for (int i = 0; i < kDartsPerThread; i++) {
    x = (kA * x + kC) & kMask48;
    v[i] = (uint32_t)(x >> 17) % M;
  }
  if (k0 + kDartsPerThread <= D) {
    u32x4* o = (u32x4*)(slot + k0);  // k0 is a multiple of 16: 64-byte aligned
#pragma unroll
    for (int q = 0; q < kDartsPerThread / 4; q++) o[q] = u32x4{v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
  } else {
    for (int i = 0; i < kDartsPerThread && k0 + i < D; i++) slot[k0 + i] = v[i];
  }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ __launch_bounds__(256) void k_first(const uint32_t* __restrict__ slot, uint64_t D,
                                               uint32_t* __restrict__ first) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < D;
       k += (uint64_t)gridDim.x * blockDim.x)
    __hip_atomic_fetch_min(first + slot[k], (uint32_t)k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of one value per thread over a workgroup of kThreads; every thread gets it.
__device__ inline uint32_t block_sum(uint32_t v, uint32_t* sh) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  uint32_t t = 0;
  for (int w = 0; w < kThreads / 64; w++) t += sh[w];
  return t;
}

// Hit darts of a tile (element e = tile*kTile + i*kThreads + tid): flag them in
// bit 31 of their slot word and count them.
__global__ __launch_bounds__(kThreads) void k_hits(uint32_t* __restrict__ slot, uint64_t D,
                                                   const uint32_t* __restrict__ first,
                                                   uint32_t* __restrict__ tile_count) {
  __shared__ uint32_t sh[kThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t n = 0;
#pragma unroll 4
  for (int i = 0; i < kPer; i++) {
    const uint64_t k = base + (uint64_t)i * kThreads + threadIdx.x;
    if (k < D) {
      const uint32_t s = slot[k];
      if (first[s] == (uint32_t)k) {
        slot[k] = s | 0x80000000u;
        n++;
      }
    }
  }
  const uint32_t t = block_sum(n, sh);
  if (threadIdx.x == 0) tile_count[blockIdx.x] = t;
}

// Exclusive prefix of tile_count[0..nt) into tile_off, total in tile_off[nt].
// One workgroup.
__global__ __launch_bounds__(kThreads) void k_scan_tiles(const uint32_t* __restrict__ tile_count,
                                                         uint32_t nt, uint32_t* __restrict__ tile_off) {
  __shared__ uint32_t part[kThreads];
  const int t = threadIdx.x;
  const uint32_t per = (nt + kThreads - 1) / kThreads, t0 = t * per;
  uint32_t sum = 0;
  for (uint32_t i = t0; i < t0 + per && i < nt; i++) sum += tile_count[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0u;
  for (uint32_t i = t0; i < t0 + per && i < nt; i++) {
    tile_off[i] = run;
    run += tile_count[i];
  }
  if (t == kThreads - 1) tile_off[nt] = part[kThreads - 1];
}

// Rank of this thread's flag among the tile's flags in element order (round i,
// then thread); `round_base` carries the tile's flags of earlier rounds.
__device__ inline uint32_t round_rank(bool f, uint32_t& round_base, uint32_t* sh) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint64_t bal = __ballot(f);
  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
  __syncthreads();
  if (lane == 0) sh[wave] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (int w = 0; w < kThreads / 64; w++) {
    const uint32_t c = sh[w];
    before += w < wave ? c : 0u;
    total += c;
  }
  const uint32_t r = round_base + before + below;
  round_base += total;
  return r;
}

// first[slot] = item of the slot's hit dart: its rank among all hits, or empty
// from item N on (the sequential process stops after N hits).
__global__ __launch_bounds__(kThreads) void k_place(const uint32_t* __restrict__ slot, uint64_t D,
                                                    const uint32_t* __restrict__ tile_off,
                                                    uint32_t N, uint32_t* __restrict__ first) {
  __shared__ uint32_t sh[kThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t rb = tile_off[blockIdx.x];
  for (int i = 0; i < kPer; i++) {
    const uint64_t k = base + (uint64_t)i * kThreads + threadIdx.x;
    const uint32_t s = k < D ? slot[k] : 0u;
    const bool hit = (s >> 31) != 0;
    const uint32_t item = round_rank(hit, rb, sh);
    if (hit) first[s & 0x7FFFFFFFu] = item < N ? item : kEmpty;
  }
}

// Filled slots per tile.
__global__ __launch_bounds__(kThreads) void k_occ(const uint32_t* __restrict__ first, uint64_t M,
                                                  uint32_t* __restrict__ tile_count) {
  __shared__ uint32_t sh[kThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t n = 0;
#pragma unroll 4
  for (int i = 0; i < kPer; i++) {
    const uint64_t s = base + (uint64_t)i * kThreads + threadIdx.x;
    if (s < M && first[s] != kEmpty) n++;
  }
  const uint32_t t = block_sum(n, sh);
  if (threadIdx.x == 0) tile_count[blockIdx.x] = t;
}

// perm[rank of filled slot s] = its item, in slot order.
__global__ __launch_bounds__(kThreads) void k_compact(const uint32_t* __restrict__ first, uint64_t M,
                                                      const uint32_t* __restrict__ tile_off,
                                                      int64_t* __restrict__ perm) {
  __shared__ uint32_t sh[kThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t rb = tile_off[blockIdx.x];
  for (int i = 0; i < kPer; i++) {
    const uint64_t s = base + (uint64_t)i * kThreads + threadIdx.x;
    const uint32_t v = s < M ? first[s] : kEmpty;
    const bool f = v != kEmpty;
    const uint32_t pos = round_rank(f, rb, sh);
    if (f) perm[pos] = (int64_t)v;
  }
}

// ---- the same darts resolved by slot region, in LDS -----------------------
// With one region of 2^15 slots per CU up to kRsMaxRegions of them (M <= 2^27)
// no dart meets a device atomic or a random global access:
//   k_rs_sort   one workgroup per chunk of 2^14 darts generates them and writes
//               them grouped by region as u32 entries (slot in region << 14 |
//               dart in chunk), with the chunk's run starts (one u16 row)
//   (the run starts transposed to one row per region)
//   k_rs_first  one workgroup per region walks its runs chunk by chunk: an LDS
//               atomic min per slot gives each slot's first dart, a second walk
//               flags the entries whose dart is that first one (bit 31)
//   k_rs_rank   one workgroup per chunk: the flagged darts' ranks in dart order
//               within the chunk (u16 per entry) and the chunk's count
//   k_rs_place  one workgroup per region: each slot's item (the chunk's offset
//               plus the rank, empty from item N on) in slot order, with the
//               filled slots counted per tile
// and k_scan_tiles + k_compact list the filled slots as before.
constexpr int kRsChunkBits = 14;
constexpr uint32_t kRsChunk = 1u << kRsChunkBits;  // darts per chunk (= kTile)
constexpr int kRsRegionBits = 15;
constexpr uint32_t kRsRegion = 1u << kRsRegionBits;  // slots per region (2 tiles)
constexpr int kRsMaxRegions = 4096;
constexpr uint16_t kNoRank = 0xFFFF;
constexpr uint32_t kRsFirst = 0x80000000u;  // entry flag: its dart is its slot's first
static_assert(kRsChunk == (uint32_t)kTile && kRsRegion == 2 * (uint32_t)kTile, "tiles and regions");

__global__ __launch_bounds__(kThreads) void k_rs_sort(uint64_t x0, uint64_t D, uint32_t M, int B,
                                                      uint32_t* __restrict__ ent,
                                                      uint16_t* __restrict__ starts) {
  extern __shared__ uint32_t sh[];
  uint32_t* cnt = sh;                     // [B] counts, then run starts
  uint32_t* stage = sh + ((B + 3) & ~3);  // [kRsChunk] entries grouped by region
  __shared__ uint32_t s_wave[kThreads / 64];
  __shared__ uint32_t s_valid;
  const int tid = threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * kRsChunk;
  for (int r = tid; r < B; r += kThreads) cnt[r] = 0;
  // darts base + 16 tid + i, as k_darts makes them
  const uint64_t k0 = base + (uint64_t)tid * kDartsPerThread;
  uint32_t e[kDartsPerThread], reg[kDartsPerThread], rank[kDartsPerThread];
  const Lcg j = lcg_steps(k0);
  uint64_t x = (j.a * x0 + j.c) & kMask48;
#pragma unroll
  for (int i = 0; i < kDartsPerThread; i++) {
    x = (kA * x + kC) & kMask48;
    const uint32_t s = (uint32_t)(x >> 17) % M;
    reg[i] = k0 + i < D ? s >> kRsRegionBits : ~0u;
    e[i] = ((s & (kRsRegion - 1)) << kRsChunkBits) | (uint32_t)(tid * kDartsPerThread + i);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kDartsPerThread; i++) rank[i] = reg[i] != ~0u ? atomicAdd(&cnt[reg[i]], 1u) : 0u;
  __syncthreads();
  // exclusive scan of cnt[0..B): each thread sums a run of up to 4 counters
  const int per = (B + kThreads - 1) / kThreads;
  const int r0 = tid * per;
  uint32_t mine = 0;
  for (int q = 0; q < per; q++)
    if (r0 + q < B) mine += cnt[r0 + q];
  uint32_t incl = mine;
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) s_wave[wv] = incl;
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int q = 0; q < kThreads / 64; q++) {
      const uint32_t t = s_wave[q];
      s_wave[q] = run;
      run += t;
    }
    s_valid = run;
  }
  __syncthreads();
  uint32_t run = s_wave[wv] + incl - mine;
  for (int q = 0; q < per; q++)
    if (r0 + q < B) {
      const uint32_t c = cnt[r0 + q];
      cnt[r0 + q] = run;
      run += c;
    }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kDartsPerThread; i++)
    if (reg[i] != ~0u) stage[cnt[reg[i]] + rank[i]] = e[i];
  const uint32_t valid = s_valid;
  uint16_t* row = starts + (size_t)blockIdx.x * (B + 1);
  for (int r = tid; r < B; r += kThreads) row[r] = (uint16_t)cnt[r];
  if (tid == 0) row[B] = (uint16_t)valid;
  __syncthreads();
  uint32_t* out = ent + base;
  for (uint32_t q = tid; q < valid / 4; q += kThreads) ((u32x4*)out)[q] = ((const u32x4*)stage)[q];
  for (uint32_t q = (valid & ~3u) + tid; q < valid; q += kThreads) out[q] = stage[q];
}

// First dart of each slot of the region; then each entry whose dart is its
// slot's first is flagged in place (bit 31, above the entry's 29 bits), so the
// flags land in the entry lines the walk has just read.
__global__ __launch_bounds__(kThreads) void k_rs_first(uint32_t* __restrict__ ent,
                                                       const uint16_t* __restrict__ offT, int64_t C, int B) {
  extern __shared__ uint32_t tbl[];  // [kRsRegion] smallest dart per slot
  const int r = region_of_block(blockIdx.x, B);
  if (r >= B) return;
  for (uint32_t i = threadIdx.x; i < kRsRegion; i += kThreads) tbl[i] = kEmpty;
  __syncthreads();
  constexpr int U = 4;
  auto none = [](int64_t) { return 0u; };
  region_walk<kThreads, kRsChunkBits, U>(offT, C, r, none, [&](const int64_t* pos, const uint32_t*) {
    uint32_t v[U];
#pragma unroll
    for (int q = 0; q < U; q++) v[q] = pos[q] >= 0 ? ent[pos[q]] : 0u;
#pragma unroll
    for (int q = 0; q < U; q++)
      if (pos[q] >= 0) {
        const uint32_t k = (uint32_t)(pos[q] & ~(int64_t)(kRsChunk - 1)) | (v[q] & (kRsChunk - 1));
        atomicMin(&tbl[v[q] >> kRsChunkBits], k);
      }
  });
  __syncthreads();  // every minimum is in place
  region_walk<kThreads, kRsChunkBits, U>(offT, C, r, none, [&](const int64_t* pos, const uint32_t*) {
    uint32_t v[U];
#pragma unroll
    for (int q = 0; q < U; q++) v[q] = pos[q] >= 0 ? ent[pos[q]] : 0u;
#pragma unroll
    for (int q = 0; q < U; q++)
      if (pos[q] >= 0) {
        const uint32_t k = (uint32_t)(pos[q] & ~(int64_t)(kRsChunk - 1)) | (v[q] & (kRsChunk - 1));
        if (tbl[v[q] >> kRsChunkBits] == k) ent[pos[q]] = v[q] | kRsFirst;
      }
  });
}

// Ranks of the chunk's first darts in dart order: per entry (kNoRank for the
// others) and the chunk's count.
__global__ __launch_bounds__(kThreads) void k_rs_rank(const uint32_t* __restrict__ ent,
                                                      const uint16_t* __restrict__ starts, int B,
                                                      uint16_t* __restrict__ lrank, uint32_t* __restrict__ cnt) {
  __shared__ __align__(16) uint8_t flag[kRsChunk];
  __shared__ __align__(16) uint16_t rk[kRsChunk];
  __shared__ uint32_t s_wave[kThreads / 64 + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * kRsChunk;
  const uint32_t valid = starts[(size_t)blockIdx.x * (B + 1) + B];
  ((uint32_t*)flag)[tid * 4 + 0] = 0;
  ((uint32_t*)flag)[tid * 4 + 1] = 0;
  ((uint32_t*)flag)[tid * 4 + 2] = 0;
  ((uint32_t*)flag)[tid * 4 + 3] = 0;
  __syncthreads();
  for (uint32_t q = tid; q < valid; q += kThreads) {
    const uint32_t v = ent[base + q];
    if (v & kRsFirst) flag[v & (kRsChunk - 1)] = 1;
  }
  __syncthreads();
  // thread t ranks darts 16 t .. 16 t + 15
  const u32x4 f4 = ((const u32x4*)flag)[tid];
  const uint32_t fw[4] = {f4.x, f4.y, f4.z, f4.w};
  uint32_t n = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) n += __popc(fw[i]);  // one flag byte holds 0 or 1
  uint32_t incl = n;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) s_wave[wv] = incl;
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int q = 0; q < kThreads / 64; q++) {
      const uint32_t t = s_wave[q];
      s_wave[q] = run;
      run += t;
    }
    s_wave[kThreads / 64] = run;
  }
  __syncthreads();
  uint32_t rr = s_wave[wv] + incl - n;
#pragma unroll
  for (int i = 0; i < kDartsPerThread; i++) {
    const bool h = (fw[i >> 2] >> (8 * (i & 3))) & 1u;
    rk[tid * kDartsPerThread + i] = h ? (uint16_t)rr : kNoRank;
    rr += h;
  }
  __syncthreads();
  for (uint32_t q = tid; q < valid; q += kThreads) lrank[base + q] = rk[ent[base + q] & (kRsChunk - 1)];
  if (tid == 0) cnt[blockIdx.x] = s_wave[kThreads / 64];
}

// Items of the region's slots in slot order (first[]), the filled slots per tile.
__global__ __launch_bounds__(kThreads) void k_rs_place(const uint32_t* __restrict__ ent,
                                                       const uint16_t* __restrict__ lrank,
                                                       const uint16_t* __restrict__ offT,
                                                       const uint32_t* __restrict__ coff, int64_t C, int B,
                                                       uint32_t M, uint32_t N, uint32_t* __restrict__ first,
                                                       uint32_t* __restrict__ tile_count) {
  extern __shared__ uint32_t tbl[];  // [kRsRegion] item per slot
  __shared__ uint32_t sh[kThreads / 64];
  const int r = region_of_block(blockIdx.x, B);
  if (r >= B) return;
  for (uint32_t i = threadIdx.x; i < kRsRegion; i += kThreads) tbl[i] = kEmpty;
  __syncthreads();
  constexpr int U = 4;
  region_walk<kThreads, kRsChunkBits, U>(offT, C, r, [&](int64_t c) { return c < C ? coff[c] : 0u; },
             [&](const int64_t* pos, const uint32_t* cbase) {
               uint32_t v[U];
               uint16_t k[U];
#pragma unroll
               for (int q = 0; q < U; q++) {
                 v[q] = pos[q] >= 0 ? ent[pos[q]] : 0u;
                 k[q] = pos[q] >= 0 ? lrank[pos[q]] : kNoRank;
               }
#pragma unroll
               for (int q = 0; q < U; q++)
                 if (k[q] != kNoRank) {
                   const uint32_t item = cbase[q] + k[q];
                   if (item < N) tbl[(v[q] >> kRsChunkBits) & (kRsRegion - 1)] = item;
                 }
             });
  __syncthreads();  // every item is in place
  const uint64_t s0 = (uint64_t)r * kRsRegion;
  uint32_t n0 = 0, n1 = 0;
  for (uint32_t i = threadIdx.x; i < kRsRegion; i += kThreads) {
    const uint64_t sl = s0 + i;
    if (sl < M) {
      const uint32_t v = tbl[i];
      first[sl] = v;
      if (v != kEmpty) {
        if (i < (uint32_t)kTile) n0++; else n1++;
      }
    }
  }
  const uint32_t t0 = block_sum(n0, sh);
  const uint32_t t1 = block_sum(n1, sh);
  const uint64_t nts = ((uint64_t)M + kTile - 1) / kTile;
  if (threadIdx.x == 0) {
    tile_count[2 * r] = t0;
    if ((uint64_t)2 * r + 1 < nts) tile_count[2 * r + 1] = t1;
  }
}

// perm holds 0..N-1 once each: seen[] marks values; any duplicate or value out
// of range sets *bad.
__global__ __launch_bounds__(256) void k_check_perm(const int64_t* __restrict__ perm, uint64_t N,
                                                    uint32_t* __restrict__ seen,
                                                    uint32_t* __restrict__ bad) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int64_t v = perm[i];
    if (v < 0 || (uint64_t)v >= N || atomicExch(seen + v, 1u) != 0u) atomicOr(bad, 1u);
  }
}

// ---- Fisher-Yates over 0..N-1 (rand_permp_agi_opt at one PE) -------------
// Step i swaps positions i and j_i >= i.  Position i is final after step i,
// and a position p >= i has been written only by steps k < i with j_k = p, the
// last of which left there what position k held when step k ran.  So with
//   pred(k) = the last step before k whose target is k (none: k untouched),
//   W(k)    = what position k holds when step k runs
//           = W(pred(k)), or k when there is no pred(k),
//   prev(i) = the last step before i with the same target j_i,
// the result is perm[i] = W(prev(i)), or j_i when there is no prev(i).  pred
// strictly decreases, so W is a finite chase.  The steps sharing a target are
// linked into a list per target (one atomic exchange each), and each step
// finds prev and pred by walking two such lists; a target's list holds about
// ln(N / (N - p)) steps, fewer than one on average.

// j[i] = i + (lrand48 value N + i) mod (N - i): the first N values choose the
// (only) PE and are skipped.
__global__ __launch_bounds__(256) void k_fy_targets(uint64_t x0, uint32_t N, uint32_t* __restrict__ j) {
  const uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kDartsPerThread;
  if (i0 >= N) return;
  const Lcg f = lcg_steps((uint64_t)N + i0);
  uint64_t x = (f.a * x0 + f.c) & kMask48;
  for (int t = 0; t < kDartsPerThread && i0 + t < N; t++) {
    x = (kA * x + kC) & kMask48;
    const uint32_t i = (uint32_t)(i0 + t);
    j[i] = i + (uint32_t)(x >> 17) % (N - i);
  }
}

// next[i] = the previous head of target j_i's list; head[j_i] = i.
__global__ __launch_bounds__(256) void k_fy_link(const uint32_t* __restrict__ j, uint32_t N,
                                                 uint32_t* __restrict__ head, uint32_t* __restrict__ next) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
    next[i] = __hip_atomic_exchange(head + j[i], i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The largest member below `below` of the list starting at q (kEmpty if none).
__device__ inline uint32_t fy_last_below(uint32_t q, uint32_t below, const uint32_t* __restrict__ next) {
  uint32_t best = kEmpty;
  while (q != kEmpty) {
    if (q < below && (best == kEmpty || q > best)) best = q;
    q = next[q];
  }
  return best;
}

__global__ __launch_bounds__(256) void k_fy_walk(const uint32_t* __restrict__ j, uint32_t N,
                                                 const uint32_t* __restrict__ head,
                                                 const uint32_t* __restrict__ next,
                                                 uint32_t* __restrict__ prev, uint32_t* __restrict__ pred) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    prev[i] = fy_last_below(head[j[i]], i, next);
    pred[i] = fy_last_below(head[i], i, next);
  }
}

__global__ __launch_bounds__(256) void k_fy_place(const uint32_t* __restrict__ j, uint32_t N,
                                                  const uint32_t* __restrict__ prev,
                                                  const uint32_t* __restrict__ pred,
                                                  int64_t* __restrict__ perm) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    uint32_t q = prev[i];
    if (q == kEmpty) {
      perm[i] = (int64_t)j[i];
      continue;
    }
    for (uint32_t r = pred[q]; r != kEmpty; r = pred[q]) q = r;
    perm[i] = (int64_t)q;
  }
}

}  // namespace rp

int launch_randperm_shuffle(int64_t* perm, int64_t N, int64_t seed, Device* d, hipStream_t s) {
  using namespace rp;
  if (N <= 0) return 0;
  const uint64_t n = (uint64_t)N, b = ((n * 4 + 255) & ~(uint64_t)255);
  char* ws = (char*)workspace(d, 5 * b, s);
  if (!ws) return (int)hipErrorOutOfMemory;
  uint32_t* j = (uint32_t*)ws;
  uint32_t* head = (uint32_t*)(ws + b);
  uint32_t* next = (uint32_t*)(ws + 2 * b);
  uint32_t* prev = (uint32_t*)(ws + 3 * b);
  uint32_t* pred = (uint32_t*)(ws + 4 * b);
  const uint64_t x0 = seed != 0 ? ((uint64_t)(uint32_t)seed << 16) | 0x330EULL : 0x1234ABCD330EULL;
  const unsigned g = (unsigned)std::min<uint64_t>((uint64_t)d->cus * 16, (n + 255) / 256);
  hipLaunchKernelGGL(k_fill_u32, dim3(std::min(g, 4096u)), dim3(256), 0, s, head, n, kEmpty);
  const uint64_t nthr = (n + kDartsPerThread - 1) / kDartsPerThread;
  hipLaunchKernelGGL(k_fy_targets, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, s, x0, (uint32_t)N, j);
  hipLaunchKernelGGL(k_fy_link, dim3(g), dim3(256), 0, s, (const uint32_t*)j, (uint32_t)N, head, next);
  hipLaunchKernelGGL(k_fy_walk, dim3(g), dim3(256), 0, s, (const uint32_t*)j, (uint32_t)N,
                     (const uint32_t*)head, (const uint32_t*)next, prev, pred);
  hipLaunchKernelGGL(k_fy_place, dim3(g), dim3(256), 0, s, (const uint32_t*)j, (uint32_t)N,
                     (const uint32_t*)prev, (const uint32_t*)pred, perm);
  return (int)hipGetLastError();
}

// Darts for N hits: the expected count is M ln(M / (M - N)) = 2N ln 2 for
// M = 2N, with a standard deviation below sqrt(N); eight of them plus a
// constant leave a shortfall practically impossible, and one is handled by
// starting again with twice the darts.
static uint64_t randperm_darts(uint64_t N) {
  return (uint64_t)(1.3862944 * (double)N + 8.0 * sqrt((double)N)) + 4096;
}

static uint64_t up256(uint64_t b) { return (b + 255) & ~(uint64_t)255; }

// The darts resolved by slot region (M <= kRsRegion * kRsMaxRegions).
static int randperm_by_region(int64_t* perm, uint64_t N, uint64_t M, uint64_t x0, Device* d, hipStream_t s,
                              int64_t* darts_used) {
  using namespace rp;
  const int B = (int)((M + kRsRegion - 1) / kRsRegion);
  const unsigned grid_r = (unsigned)(8 * ((B + 7) / 8));
  const size_t lds_sort = (size_t)(((B + 3) & ~3) + kRsChunk) * 4;
  const size_t lds_region = (size_t)kRsRegion * 4;
  for (uint64_t D = randperm_darts(N);; D *= 2) {
    if (D >= ((uint64_t)1 << 31)) return (int)hipErrorInvalidValue;
    const uint64_t C = (D + kRsChunk - 1) / kRsChunk, nts = (M + kTile - 1) / kTile;
    const uint64_t nt = C > nts ? C : nts;
    const uint64_t b_ent = up256(C * kRsChunk * 4), b_st = up256(C * (B + 1) * 2),
                   b_rank = up256(C * kRsChunk * 2), b_first = up256(M * 4), b_cnt = up256(nt * 4);
    char* ws = (char*)workspace(d, b_ent + 2 * b_st + b_rank + b_first + b_cnt + up256((nt + 1) * 4), s);
    if (!ws) return (int)hipErrorOutOfMemory;
    uint32_t* ent = (uint32_t*)ws;
    uint16_t* starts = (uint16_t*)(ws + b_ent);
    uint16_t* offT = (uint16_t*)(ws + b_ent + b_st);
    uint16_t* lrank = (uint16_t*)(ws + b_ent + 2 * b_st);
    uint32_t* first = (uint32_t*)(ws + b_ent + 2 * b_st + b_rank);
    uint32_t* cnt = (uint32_t*)(ws + b_ent + 2 * b_st + b_rank + b_first);
    uint32_t* off = (uint32_t*)(ws + b_ent + 2 * b_st + b_rank + b_first + b_cnt);
    allow_lds(k_rs_sort, lds_sort);
    hipLaunchKernelGGL(k_rs_sort, dim3((unsigned)C), dim3(kThreads), lds_sort, s, x0, D, (uint32_t)M, B, ent,
                       starts);
    int rc = launch_starts_transpose(starts, offT, B + 1, (int64_t)C, s);
    if (rc) return rc;
    allow_lds(k_rs_first, lds_region);
    hipLaunchKernelGGL(k_rs_first, dim3(grid_r), dim3(kThreads), lds_region, s, ent,
                       (const uint16_t*)offT, (int64_t)C, B);
    hipLaunchKernelGGL(k_rs_rank, dim3((unsigned)C), dim3(kThreads), 0, s, (const uint32_t*)ent,
                       (const uint16_t*)starts, B, lrank, cnt);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, s, (const uint32_t*)cnt, (uint32_t)C, off);
    uint32_t hits = 0;
    rc = (int)hipMemcpyAsync(&hits, off + C, 4, hipMemcpyDeviceToHost, s);
    if (!rc) rc = (int)hipStreamSynchronize(s);
    if (rc) return rc;
    if (hits < N) continue;  // too few darts: again with twice as many
    // off: the chunks' offsets, read by k_rs_place; cnt then holds tile counts
    allow_lds(k_rs_place, lds_region);
    hipLaunchKernelGGL(k_rs_place, dim3(grid_r), dim3(kThreads), lds_region, s, (const uint32_t*)ent,
                       (const uint16_t*)lrank, (const uint16_t*)offT, (const uint32_t*)off, (int64_t)C, B,
                       (uint32_t)M, (uint32_t)N, first, cnt);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, s, (const uint32_t*)cnt, (uint32_t)nts, off);
    hipLaunchKernelGGL(k_compact, dim3((unsigned)nts), dim3(kThreads), 0, s, (const uint32_t*)first, M,
                       (const uint32_t*)off, perm);
    if (darts_used) *darts_used = (int64_t)D;
    return (int)hipGetLastError();
  }
}

int launch_randperm(int64_t* perm, int64_t N, int64_t seed, Device* d, hipStream_t s,
                    int64_t* darts_used) {
  using namespace rp;
  if (N <= 0) return 0;
  const uint64_t M = 2 * (uint64_t)N;
  // srand48(seed) when seed != 0, else drand48's initial state
  const uint64_t x0 = seed != 0 ? ((uint64_t)(uint32_t)seed << 16) | 0x330EULL : 0x1234ABCD330EULL;
  // By region when there are regions for every CU (a region's walk takes the
  // same time at any size, so with fewer regions than CUs the atomics win:
  // measured on MI355X, 2^20 elements 0.17 against 0.41 ms, 2^21 0.26 against
  // 0.27, 2^22 0.46 against 0.21) and at most kRsMaxRegions of them.
  const uint64_t nreg = (M + kRsRegion - 1) / kRsRegion;
  if (nreg <= (uint64_t)kRsMaxRegions && nreg >= (uint64_t)d->cus)
    return randperm_by_region(perm, (uint64_t)N, M, x0, d, s, darts_used);
  // otherwise first darts by device atomic min over all slots
  for (uint64_t D = randperm_darts((uint64_t)N);; D *= 2) {
    if (D >= ((uint64_t)1 << 31)) return (int)hipErrorInvalidValue;
    const uint64_t ntd = (D + kTile - 1) / kTile, nts = (M + kTile - 1) / kTile;
    const uint64_t nt = ntd > nts ? ntd : nts;
    auto up = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
    const uint64_t b_slot = up(D * 4), b_first = up(M * 4), b_cnt = up(nt * 4), b_off = up((nt + 1) * 4);
    char* ws = (char*)workspace(d, b_slot + b_first + b_cnt + b_off + 256, s);
    if (!ws) return (int)hipErrorOutOfMemory;
    uint32_t* slot = (uint32_t*)ws;
    uint32_t* first = (uint32_t*)(ws + b_slot);
    uint32_t* cnt = (uint32_t*)(ws + b_slot + b_first);
    uint32_t* off = (uint32_t*)(ws + b_slot + b_first + b_cnt);
    const unsigned g_fill = (unsigned)std::min<uint64_t>(4096, (M + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(g_fill), dim3(256), 0, s, first, M, kEmpty);
    const uint64_t nthr_d = (D + kDartsPerThread - 1) / kDartsPerThread;
    hipLaunchKernelGGL(k_darts, dim3((unsigned)((nthr_d + 255) / 256)), dim3(256), 0, s, x0, D,
                       (uint32_t)M, slot);
    const unsigned g_d = (unsigned)std::min<uint64_t>((uint64_t)d->cus * 32, (D + 255) / 256);
    hipLaunchKernelGGL(k_first, dim3(g_d), dim3(256), 0, s, (const uint32_t*)slot, D, first);
    hipLaunchKernelGGL(k_hits, dim3((unsigned)ntd), dim3(kThreads), 0, s, slot, D,
                       (const uint32_t*)first, cnt);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, s, (const uint32_t*)cnt,
                       (uint32_t)ntd, off);
    uint32_t hits = 0;
    int rc = (int)hipMemcpyAsync(&hits, off + ntd, 4, hipMemcpyDeviceToHost, s);
    if (!rc) rc = (int)hipStreamSynchronize(s);
    if (rc) return rc;
    if (hits < (uint64_t)N) continue;  // too few darts: again with twice as many
    hipLaunchKernelGGL(k_place, dim3((unsigned)ntd), dim3(kThreads), 0, s, (const uint32_t*)slot, D,
                       (const uint32_t*)off, (uint32_t)N, first);
    hipLaunchKernelGGL(k_occ, dim3((unsigned)nts), dim3(kThreads), 0, s, (const uint32_t*)first, M,
                       cnt);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, s, (const uint32_t*)cnt,
                       (uint32_t)nts, off);
    hipLaunchKernelGGL(k_compact, dim3((unsigned)nts), dim3(kThreads), 0, s, (const uint32_t*)first,
                       M, (const uint32_t*)off, perm);
    if (darts_used) *darts_used = (int64_t)D;
    return (int)hipGetLastError();
  }
}

int launch_check_perm(const int64_t* perm, int64_t N, Device* d, hipStream_t s, int* ok) {
  using namespace rp;
  *ok = 0;
  if (N < 0) return (int)hipErrorInvalidValue;
  if (N == 0) {
    *ok = 1;
    return 0;
  }
  const uint64_t bytes = (((uint64_t)N * 4 + 255) & ~(uint64_t)255) + 256;
  char* ws = (char*)workspace(d, bytes, s);
  if (!ws) return (int)hipErrorOutOfMemory;
  uint32_t* seen = (uint32_t*)ws;
  uint32_t* bad = (uint32_t*)(ws + bytes - 256);
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<uint64_t>(4096, ((uint64_t)N + 255) / 256)),
                     dim3(256), 0, s, seen, (uint64_t)N, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, bad, (uint64_t)1, 0u);
  hipLaunchKernelGGL(k_check_perm, dim3((unsigned)std::min<uint64_t>((uint64_t)d->cus * 32, ((uint64_t)N + 255) / 256)),
                     dim3(256), 0, s, perm, (uint64_t)N, seen, bad);
  uint32_t b = 1;
  int rc = (int)hipMemcpyAsync(&b, bad, 4, hipMemcpyDeviceToHost, s);
  if (!rc) rc = (int)hipStreamSynchronize(s);
  if (rc) return rc;
  *ok = b == 0;
  return 0;
}

}  // namespace bale
