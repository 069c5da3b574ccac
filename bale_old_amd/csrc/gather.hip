// gather.hip -- index gather tgt[i] = table[index[i]] on gfx950.
//
//   DIRECT  every lane issues its U table loads back to back so each wave
//           keeps many independent random reads in flight (memory-level
//           parallelism is what bounds a random gather); index and tgt move
//           as 16-byte non-temporal streams so they do not evict table lines.
//           The AGI model (apps/ig_src/ig_agi.upc:61-64).
//   BUCKET  requests sorted by table region within chunks, each region served
//           from one XCD's L2 (or, for regions of at most 2^14 entries, from a
//           workgroup's LDS copy of it), replies put back in order through LDS
//           (below) --
//           the request/reply exchange of apps/ig_src/ig_exstack.upc:69-94 and
//           ig_conveyor.upc:78-101.
#include "bale_internal.h"
#include "partition.h"
#include "runwalk.h"

namespace bale {

typedef i64x2 v2i64;

// VEC: index and tgt are 16-byte aligned, so pairs move as one 16-byte access.
// The adaptive plan's verdict (written on the device by the sample): the
// requests repeat table entries (hot entries or clustered runs), which the
// caches serve best straight from the table, so DIRECT runs; otherwise the
// BUCKET passes run.  Every kernel of both reads it and the other one exits.
__device__ inline bool gather_skewed(const uint32_t* flags) {
  return (*flags & (kFlagHotBins | kFlagRunsSeen)) != 0;
}

template <int U, bool PACKED, bool VEC>
__global__ __launch_bounds__(1024) void k_gather_direct(int64_t* __restrict__ tgt,
                                                        const int64_t* __restrict__ idx, int64_t n,
                                                        const int64_t* __restrict__ table,
                                                        uint64_t tlen, ErrWords* err,
                                                        const uint32_t* __restrict__ flags) {
  static_assert(U % 2 == 0, "U counts int64 elements, loaded in pairs");
  if (flags && !gather_skewed(flags)) return;  // the adaptive plan took BUCKET
  const int B = blockDim.x;
  // lane owns pairs (2*(k*B + tid), +1) of the block's span of B*U elements
  const int64_t base = (int64_t)blockIdx.x * B * U;
  v2i64 raw[U / 2];
#pragma unroll
  for (int k = 0; k < U / 2; k++) {
    int64_t e = base + 2 * ((int64_t)k * B + threadIdx.x);
    if (VEC && e + 1 < n) {
      raw[k] = __builtin_nontemporal_load((const v2i64*)(idx + e));
    } else if (e + 1 < n) {
      raw[k].x = __builtin_nontemporal_load(idx + e);
      raw[k].y = __builtin_nontemporal_load(idx + e + 1);
    } else {
      raw[k].x = e < n ? idx[e] : 0;
      raw[k].y = 0;
    }
  }
  unsigned bad_r = 0, bad_p = 0;
  int64_t val[U];
#pragma unroll
  for (int k = 0; k < U / 2; k++) {
    int64_t e = base + 2 * ((int64_t)k * B + threadIdx.x);
    uint64_t j0, j1;
    bool ok0 = e < n && decode_index<PACKED>(raw[k].x, tlen, j0, bad_r, bad_p);
    bool ok1 = e + 1 < n && decode_index<PACKED>(raw[k].y, tlen, j1, bad_r, bad_p);
    val[2 * k] = ok0 ? table[j0] : 0;
    val[2 * k + 1] = ok1 ? table[j1] : 0;
  }
#pragma unroll
  for (int k = 0; k < U / 2; k++) {
    int64_t e = base + 2 * ((int64_t)k * B + threadIdx.x);
    if (VEC && e + 1 < n) {
      v2i64 o;
      o.x = val[2 * k];
      o.y = val[2 * k + 1];
      __builtin_nontemporal_store(o, (v2i64*)(tgt + e));
    } else if (e + 1 < n) {
      __builtin_nontemporal_store(val[2 * k], tgt + e);
      __builtin_nontemporal_store(val[2 * k + 1], tgt + e + 1);
    } else if (e < n) {
      tgt[e] = val[2 * k];
    }
  }
  flush_errors(err, bad_r, bad_p);
}

// ---- BUCKET: request/reply by table region --------------------------------
// Three passes over requests grouped in chunks of kGChunk:
//   sort       one workgroup per chunk counts its requests per table region
//              in LDS and writes them back grouped by region, as u32 entries
//              (offset in the region << kGSlotBits | slot in the chunk), with
//              the chunk's run starts (u16, one row per chunk);
//   region     the run starts, transposed to one row per region, let a
//              workgroup walk one region's runs over a block of chunks; each
//              XCD serves its own block of regions in order, so a region's
//              table lines are read into one L2 once and answer every request
//              for that region from there; each reply goes to its entry's
//              position, so a chunk's replies fill one contiguous block;
//   unpermute  one workgroup per chunk reads its entries and replies (two
//              contiguous blocks) and places each reply at its slot in an LDS
//              copy of the chunk's tgt, then writes it out as one stream.
// Measured on MI355X at 2^28 requests into a 2^28-entry table: sort 0.55 ms,
// transpose 0.03, region 2.9, unpermute 0.9.  Writing the replies region by
// region instead (one contiguous stream, positions from scans of the run
// lengths) moved 0.55 ms from the region pass to the unpermute, which then
// reads one short range per region: no gain, so the simpler layout stays.
// The request/reply exchange of apps/ig_src/ig_exstack.upc:69-94 and
// ig_conveyor.upc:78-101, with the table region as the destination.
constexpr int kGSlotBits = 14;
constexpr uint32_t kGChunk = 1u << kGSlotBits;  // requests per chunk
constexpr int kGSortThreads = 1024;
constexpr int kGPer = kGChunk / kGSortThreads;  // requests per sort thread
constexpr int kGRegionMaxBits = 32 - kGSlotBits;
constexpr int kXcds = 8;  // workgroup b runs on XCD b % 8 (round-robin dispatch)

template <bool PACKED>
__global__ __launch_bounds__(kGSortThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_gsort(const int64_t* __restrict__ idx, int64_t n,
                                                        uint64_t tlen, int w, int B,
                                                        uint32_t* __restrict__ sorted,
                                                        uint16_t* __restrict__ starts,
                                                        ErrWords* err,
                                                        const uint32_t* __restrict__ flags) {
  if (flags && gather_skewed(flags)) return;  // the adaptive plan took DIRECT
  extern __shared__ uint32_t sh[];
  uint32_t* cnt = sh;                     // [B] counts, then run cursors
  uint32_t* stage = sh + ((B + 3) & ~3);  // [kGChunk] entries grouped by region
  __shared__ uint32_t s_wave[kGSortThreads / 64];
  __shared__ uint32_t s_valid;
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kGChunk;
  const uint32_t len = (uint32_t)min((int64_t)kGChunk, n - base);
  for (int r = tid; r < B; r += kGSortThreads) cnt[r] = 0;
  __syncthreads();
  // element k of this thread: local position 2 * (p * 1024 + tid) + (k & 1), p = k / 2;
  // only its table index (below 2^30: at most 4096 regions of 2^18) is kept
  // across the count, so the kernel fits 64 VGPRs and two workgroups per CU
  // (0.55 ms at 2^28 requests, against 0.73 with each request's rank kept
  // from the count and one workgroup per CU).  A request's place in its
  // region's run comes from an atomic on the run's cursor, so a run holds its
  // requests in no set order (each entry carries its slot).
  uint32_t jj[kGPer];
  unsigned bad_r = 0, bad_p = 0;
  auto decode = [&](int k, uint32_t l, int64_t raw) {
    uint64_t j = 0;
    const bool ok = l < len && decode_index<PACKED>(raw, tlen, j, bad_r, bad_p);
    jj[k] = ok ? (uint32_t)j : ~0u;
  };
  if (len == kGChunk && (((uintptr_t)(idx + base)) & 15) == 0) {
    // a whole aligned chunk (every chunk but the last, in practice): 16-byte
    // loads only
#pragma unroll
    for (int p = 0; p < kGPer / 2; p++) {
      const uint32_t l0 = 2u * ((uint32_t)p * kGSortThreads + tid);
      const i64x2 v = __builtin_nontemporal_load((const i64x2*)(idx + base + l0));
      decode(2 * p, l0, v.x);
      decode(2 * p + 1, l0 + 1, v.y);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kGPer; k++) {
      const uint32_t l = 2u * ((uint32_t)(k / 2) * kGSortThreads + tid) + (k & 1);
      decode(k, l, l < len ? idx[base + l] : 0);
    }
  }
#pragma unroll
  for (int k = 0; k < kGPer; k++)
    if (jj[k] != ~0u) atomicAdd(&cnt[jj[k] >> w], 1u);
  __syncthreads();
  // exclusive scan of cnt[0..B): each thread sums a run of up to 4 counters
  const int per = (B + kGSortThreads - 1) / kGSortThreads;
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
    for (int q = 0; q < kGSortThreads / 64; q++) {
      const uint32_t t = s_wave[q];
      s_wave[q] = run;
      run += t;
    }
    s_valid = run;
  }
  __syncthreads();
  uint16_t* row = starts + (size_t)blockIdx.x * (B + 1);
  uint32_t run = s_wave[wv] + incl - mine;
  for (int q = 0; q < per; q++)
    if (r0 + q < B) {
      const uint32_t c = cnt[r0 + q];
      cnt[r0 + q] = run;  // the run's cursor
      row[r0 + q] = (uint16_t)run;
      run += c;
    }
  const uint32_t valid = s_valid;
  if (tid == 0) row[B] = (uint16_t)valid;
  __syncthreads();
  const uint32_t omask = (1u << w) - 1;
#pragma unroll
  for (int k = 0; k < kGPer; k++)
    if (jj[k] != ~0u) {
      const uint32_t l = 2u * ((uint32_t)(k / 2) * kGSortThreads + tid) + (k & 1);
      stage[atomicAdd(&cnt[jj[k] >> w], 1u)] = ((jj[k] & omask) << kGSlotBits) | l;
    }
  __syncthreads();
  uint32_t* out = sorted + base;
  for (uint32_t q = tid; q < valid / 4; q += kGSortThreads)
    __builtin_nontemporal_store(((const u32x4*)stage)[q], (u32x4*)out + q);
  for (uint32_t q = (valid & ~3u) + tid; q < valid; q += kGSortThreads) out[q] = stage[q];
  flush_errors(err, bad_r, bad_p);
}

// offT[r][c] = starts[c][r]: the run starts as one row per region (u16 tiles
// of 64 x 64 through LDS).
__global__ __launch_bounds__(256) void k_gtranspose(const uint16_t* __restrict__ starts,
                                                    uint16_t* __restrict__ offT, int rows, int64_t C,
                                                    const uint32_t* __restrict__ flags) {
  if (flags && gather_skewed(flags)) return;
  __shared__ uint16_t t[64][65];
  const int r0 = blockIdx.x * 64;
  const int64_t c0 = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int y = ty; y < 64; y += 4) {
    const int64_t c = c0 + y;
    const int r = r0 + tx;
    t[y][tx] = (c < C && r < rows) ? starts[c * rows + r] : (uint16_t)0;
  }
  __syncthreads();
  for (int y = ty; y < 64; y += 4) {
    const int r = r0 + y;
    const int64_t c = c0 + tx;
    if (r < rows && c < C) offT[(int64_t)r * C + c] = t[tx][y];
  }
}

// Persistent.  Segment s of XCD x's queue is block s % nblk of region
// rb(x) + s / nblk, where XCD x owns regions [x B / 8, (x + 1) B / 8); a
// workgroup serves its own XCD's queue first, then helps the others.
constexpr int kGRegThreads = 512;
constexpr int kGSegChunks = 256;  // chunks per segment (at most kGRegThreads)
__global__ __launch_bounds__(kGRegThreads) void k_gregion(
    const uint32_t* __restrict__ sorted, const uint16_t* __restrict__ offT, int B, int64_t C,
    int w, int seg_chunks, const int64_t* __restrict__ table, int64_t* __restrict__ reply,
    uint32_t* __restrict__ heads, const uint32_t* __restrict__ flags) {
  if (flags && gather_skewed(flags)) return;
  __shared__ uint32_t s_a[kGSegChunks], s_pre[kGSegChunks + 1];
  __shared__ uint32_t s_seg;
  __shared__ uint32_t s_wave[kGRegThreads / 64];
  const int tid = threadIdx.x;
  const uint32_t nblk = (uint32_t)((C + seg_chunks - 1) / seg_chunks);
  const int x0 = blockIdx.x % kXcds;
  for (int qi = 0; qi < kXcds; qi++) {
    const int x = (x0 + qi) % kXcds;
    const int rlo = (int)((int64_t)x * B / kXcds), rhi = (int)((int64_t)(x + 1) * B / kXcds);
    const uint32_t nseg = (uint32_t)(rhi - rlo) * nblk;
    while (true) {
      if (tid == 0) s_seg = atomicAdd(&heads[x], 1u);
      __syncthreads();
      const uint32_t sg = s_seg;
      __syncthreads();
      if (sg >= nseg) break;
      const int r = rlo + (int)(sg / nblk);
      const int64_t c0 = (int64_t)(sg % nblk) * seg_chunks;
      const int nc = (int)min((int64_t)seg_chunks, C - c0);
      // run k of the segment: chunk c0 + k, entries [a, b) of its sorted block
      uint32_t len = 0;
      if (tid < nc) {
        const uint32_t a = offT[(int64_t)r * C + c0 + tid];
        const uint32_t b = offT[(int64_t)(r + 1) * C + c0 + tid];
        s_a[tid] = a;
        len = b - a;
      }
      uint32_t incl = len;
      const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
      }
      if (lane == 63) s_wave[wv] = incl;
      __syncthreads();
      if (tid == 0) {
        uint32_t acc = 0;
        for (int q = 0; q < kGRegThreads / 64; q++) {
          const uint32_t t = s_wave[q];
          s_wave[q] = acc;
          acc += t;
        }
        s_pre[seg_chunks] = acc;
      }
      __syncthreads();
      if (tid < nc) s_pre[tid] = s_wave[wv] + incl - len;
      if (tid == 0 && nc < seg_chunks) s_pre[nc] = s_pre[seg_chunks];
      __syncthreads();
      const uint32_t total = s_pre[seg_chunks];
      const int64_t* region = table + ((int64_t)r << w);
      // consecutive lanes take consecutive entries of the segment's runs
      constexpr int Q = 4;
      for (uint32_t f0 = tid; f0 < total; f0 += Q * kGRegThreads) {
        int64_t e[Q];  // entry position in the sorted blocks (-1: none)
#pragma unroll
        for (int q = 0; q < Q; q++) {
          const uint32_t f = f0 + q * kGRegThreads;
          e[q] = -1;
          if (f < total) {
            int lo = 0, hi = nc - 1;  // last run whose prefix is <= f
            while (lo < hi) {
              const int mid = (lo + hi + 1) >> 1;
              if (s_pre[mid] <= f) lo = mid; else hi = mid - 1;
            }
            e[q] = (c0 + lo) * (int64_t)kGChunk + s_a[lo] + (f - s_pre[lo]);
          }
        }
        // plain loads: a run shares its 128-byte line with the next region's
        // run, which this XCD serves next
        uint32_t o[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) o[q] = e[q] >= 0 ? sorted[e[q]] : 0u;
        int64_t v[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) v[q] = e[q] >= 0 ? region[o[q] >> kGSlotBits] : 0;
        // plain stores: a run's replies are a partial line that the L2 merges
        // with the neighbouring runs' (measured faster than non-temporal ones)
#pragma unroll
        for (int q = 0; q < Q; q++)
          if (e[q] >= 0) reply[e[q]] = v[q];
      }
    }
  }
}

// Regions of at most 2^14 entries (the plan's choice for tables of up to 2^25
// entries) are served from LDS instead: one workgroup per region copies its slice of the
// table (at most 128 KiB) into LDS with coalesced loads, then walks the
// region's runs over all chunks and answers each request with a random LDS
// read, which the CU serves many times faster than random L2 reads.  The
// reply still goes to the entry's position.
constexpr int kGLdsMaxBits = 14;
__global__ __launch_bounds__(1024) void k_gregion_lds(const uint32_t* __restrict__ sorted,
                                                      const uint16_t* __restrict__ offT, int B, int64_t C,
                                                      int w, const int64_t* __restrict__ table,
                                                      int64_t tlen, int64_t* __restrict__ reply,
                                                      const uint32_t* __restrict__ flags) {
  if (flags && gather_skewed(flags)) return;
  extern __shared__ int64_t tbl[];  // [1 << w]
  const int r = region_of_block(blockIdx.x, B);
  if (r >= B) return;
  const int64_t t0 = (int64_t)r << w;
  const int64_t tn = min((int64_t)1 << w, tlen - t0);
  for (int64_t i = threadIdx.x; i < tn; i += 1024) tbl[i] = table[t0 + i];
  __syncthreads();
  constexpr int Q = 4;
  auto none = [](int64_t) { return 0u; };
  region_walk<1024, kGSlotBits, Q>(offT, C, r, none, [&](const int64_t* pos, const uint32_t*) {
    uint32_t o[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) o[q] = pos[q] >= 0 ? sorted[pos[q]] : 0u;
#pragma unroll
    for (int q = 0; q < Q; q++)
      if (pos[q] >= 0) reply[pos[q]] = tbl[o[q] >> kGSlotBits];
  });
}

// One workgroup per chunk: tgt of the chunk from its entries and replies.
__global__ __launch_bounds__(1024) void k_gunpermute(const uint32_t* __restrict__ sorted,
                                                     const int64_t* __restrict__ reply,
                                                     const uint16_t* __restrict__ starts, int B,
                                                     int64_t n, int64_t* __restrict__ tgt,
                                                     const uint32_t* __restrict__ flags) {
  if (flags && gather_skewed(flags)) return;
  extern __shared__ int64_t buf[];  // [kGChunk]
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kGChunk;
  const uint32_t len = (uint32_t)min((int64_t)kGChunk, n - base);
  const uint32_t valid = starts[(size_t)blockIdx.x * (B + 1) + B];
  if (valid < len)  // skipped requests read 0
    for (uint32_t k = tid; k < len; k += 1024) buf[k] = 0;
  __syncthreads();
  const uint32_t* ent = sorted + base;
  const int64_t* rep = reply + base;
  for (uint32_t q = tid; q < valid / 4; q += 1024) {
    const u32x4 o = __builtin_nontemporal_load((const u32x4*)ent + q);
    const i64x2 a = __builtin_nontemporal_load((const i64x2*)rep + 2 * q);
    const i64x2 b = __builtin_nontemporal_load((const i64x2*)rep + 2 * q + 1);
    buf[o.x & (kGChunk - 1)] = a.x;
    buf[o.y & (kGChunk - 1)] = a.y;
    buf[o.z & (kGChunk - 1)] = b.x;
    buf[o.w & (kGChunk - 1)] = b.y;
  }
  for (uint32_t q = (valid & ~3u) + tid; q < valid; q += 1024) buf[ent[q] & (kGChunk - 1)] = rep[q];
  __syncthreads();
  int64_t* out = tgt + base;
  if ((((uintptr_t)out) & 15) == 0) {
    for (uint32_t q = tid; q < len / 2; q += 1024) {
      i64x2 v;
      v.x = buf[2 * q];
      v.y = buf[2 * q + 1];
      __builtin_nontemporal_store(v, (i64x2*)out + q);
    }
    if ((len & 1) && tid == 0) out[len - 1] = buf[len - 1];
  } else {
    for (uint32_t q = tid; q < len; q += 1024) out[q] = buf[q];
  }
}

int launch_starts_transpose(const uint16_t* starts, uint16_t* offT, int rows, int64_t C, hipStream_t s) {
  if (C <= 0 || rows <= 0) return 0;
  if (ceil_div(C, 64) > 65535) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gtranspose, dim3((unsigned)ceil_div(rows, 64), (unsigned)ceil_div(C, 64)), dim3(256), 0, s,
                     starts, offT, rows, C, (const uint32_t*)nullptr);
  return (int)hipGetLastError();
}

template <bool PACKED, bool VEC>
static void go_direct(int U, dim3 g, dim3 b, hipStream_t s, int64_t* tgt, const int64_t* idx,
                      int64_t n, const int64_t* table, uint64_t tlen, ErrWords* err,
                      const uint32_t* flags) {
  if (U == 16)
    hipLaunchKernelGGL((k_gather_direct<16, PACKED, VEC>), g, b, 0, s, tgt, idx, n, table, tlen, err, flags);
  else if (U == 8)
    hipLaunchKernelGGL((k_gather_direct<8, PACKED, VEC>), g, b, 0, s, tgt, idx, n, table, tlen, err, flags);
  else
    hipLaunchKernelGGL((k_gather_direct<4, PACKED, VEC>), g, b, 0, s, tgt, idx, n, table, tlen, err, flags);
}

// flags (optional): run only when the adaptive plan's sample chose DIRECT.
static int direct_launch(int64_t* tgt, const int64_t* idx, int64_t n, const int64_t* table,
                         int64_t table_len, bool packed, const bale_config& c, ErrWords* err,
                         const uint32_t* flags, hipStream_t s) {
  if (n <= 0) return 0;
  int block = (c.block == 256 || c.block == 512 || c.block == 1024) ? c.block : 256;
  int U = (c.unroll == 4 || c.unroll == 8 || c.unroll == 16) ? c.unroll : 16;
  int64_t grid = ceil_div(n, (int64_t)block * U);
  if (grid > 0x7fffffff) return (int)hipErrorInvalidValue;
  dim3 g((unsigned)grid), b(block);
  bool vec = (((uintptr_t)tgt | (uintptr_t)idx) & 15) == 0;
  if (packed) {
    if (vec) go_direct<true, true>(U, g, b, s, tgt, idx, n, table, (uint64_t)table_len, err, flags);
    else go_direct<true, false>(U, g, b, s, tgt, idx, n, table, (uint64_t)table_len, err, flags);
  } else {
    if (vec) go_direct<false, true>(U, g, b, s, tgt, idx, n, table, (uint64_t)table_len, err, flags);
    else go_direct<false, false>(U, g, b, s, tgt, idx, n, table, (uint64_t)table_len, err, flags);
  }
  return (int)hipGetLastError();
}

int launch_gather_direct(int64_t* tgt, const int64_t* idx, int64_t n, const int64_t* table,
                         int64_t table_len, bool packed, const bale_config& c, ErrWords* err,
                         hipStream_t s) {
  return direct_launch(tgt, idx, n, table, table_len, packed, c, err, nullptr, s);
}

constexpr int kMaxRegions = 4096;

int launch_gather_bucket(int64_t* tgt, const int64_t* idx, int64_t n, const int64_t* table,
                         int64_t table_len, bool packed, const bale_config& c, Device* d,
                         hipStream_t s) {
  if (n <= 0) return 0;
  if (n >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;  // caller splits
  // regions of 2^tile_log2 entries, 2^18 (2 MiB, half an XCD's L2) unless
  // set; those of at most 2^14 entries are served from LDS; an entry holds the
  // offset in the region beside the slot in the chunk
  int w = c.tile_log2 > 0 ? std::min(c.tile_log2, kGRegionMaxBits) : 18;
  const bool lds_regions = w <= kGLdsMaxBits;
  int64_t B = ceil_div(table_len, (int64_t)1 << w);
  if (B > kMaxRegions)
    return bad_config("BUCKET gather: a table of %lld entries needs more than %d regions of at most "
                      "2^%d entries; use DIRECT", (long long)table_len, kMaxRegions, w);
  const int64_t C = ceil_div(n, (int64_t)kGChunk);
  const size_t sorted_b = ((size_t)C * kGChunk * 4 + 255) & ~(size_t)255;
  const size_t reply_b = (size_t)C * kGChunk * 8;
  const size_t starts_b = ((size_t)C * (B + 1) * 2 + 255) & ~(size_t)255;
  // hot = 2 (adaptive): a sample of the requests picks DIRECT or these passes
  // on the device; its hot table and scratch follow the queue heads
  const bool adaptive = c.hot == 2;
  const size_t sketch_b = adaptive ? (size_t)(kHotSlots + kSketchScratch) * 4 : 0;
  char* ws = (char*)workspace(d, sorted_b + reply_b + 2 * starts_b + 256 + sketch_b, s);
  if (!ws) return (int)hipErrorOutOfMemory;
  uint32_t* sorted = (uint32_t*)ws;
  int64_t* reply = (int64_t*)(ws + sorted_b);
  uint16_t* starts = (uint16_t*)(ws + sorted_b + reply_b);
  uint16_t* offT = (uint16_t*)(ws + sorted_b + reply_b + starts_b);
  uint32_t* heads = (uint32_t*)(ws + sorted_b + reply_b + 2 * starts_b);  // [8], flag word at [63]
  uint32_t* flags = adaptive ? heads + 63 : nullptr;
  ErrWords* err = stream_err(d, s);
  int rc = launch_zero_bytes(heads, 256, s);
  if (rc) return rc;
  if (adaptive) {
    uint32_t* hot = heads + 64;
    if (packed)
      launch_hot_sketch<true>(idx, n, (uint64_t)table_len, hot, flags, hot + kHotSlots, s);
    else
      launch_hot_sketch<false>(idx, n, (uint64_t)table_len, hot, flags, hot + kHotSlots, s);
    if ((rc = direct_launch(tgt, idx, n, table, table_len, packed, c, err, flags, s))) return rc;
  }
  const size_t lds_sort = (size_t)(((B + 3) & ~3) + kGChunk) * 4;
  if (packed) {
    allow_lds(k_gsort<true>, lds_sort);
    hipLaunchKernelGGL(k_gsort<true>, dim3((unsigned)C), dim3(kGSortThreads), lds_sort, s, idx, n,
                       (uint64_t)table_len, w, (int)B, sorted, starts, err, (const uint32_t*)flags);
  } else {
    allow_lds(k_gsort<false>, lds_sort);
    hipLaunchKernelGGL(k_gsort<false>, dim3((unsigned)C), dim3(kGSortThreads), lds_sort, s, idx, n,
                       (uint64_t)table_len, w, (int)B, sorted, starts, err, (const uint32_t*)flags);
  }
  hipLaunchKernelGGL(k_gtranspose, dim3((unsigned)ceil_div(B + 1, 64), (unsigned)ceil_div(C, 64)),
                     dim3(256), 0, s, (const uint16_t*)starts, offT, (int)(B + 1), C,
                     (const uint32_t*)flags);
  if (lds_regions) {
    const size_t lds_t = (size_t)8 << w;
    allow_lds(k_gregion_lds, lds_t);
    hipLaunchKernelGGL(k_gregion_lds, dim3((unsigned)(8 * ceil_div(B, 8))), dim3(1024), lds_t, s,
                       (const uint32_t*)sorted, (const uint16_t*)offT, (int)B, C, w, table, table_len, reply,
                       (const uint32_t*)flags);
  } else {
    // segments of about 2^task_log2 requests (a chunk holds kGChunk / B of a
    // region's on average), 2^11 by default: measured on MI355X at 2^28
    // requests into a 2^28-entry table against 2^10 and 2^12 (4.64 against
    // 4.89 and 4.85 ms), and 4 workgroups per CU against 2 and 8 (5.25, 4.73 ms)
    const int seg_log2 = c.task_log2 > 0 ? c.task_log2 : 11;
    const int seg_chunks =
        (int)std::max<int64_t>(16, std::min<int64_t>(kGSegChunks, ((int64_t)B << seg_log2) >> kGSlotBits));
    int gr = c.groups_per_cu > 0 ? c.groups_per_cu : 4;
    hipLaunchKernelGGL(k_gregion, dim3((unsigned)(d->cus * gr)), dim3(kGRegThreads), 0, s,
                       (const uint32_t*)sorted, (const uint16_t*)offT, (int)B, C, w, seg_chunks, table,
                       reply, heads, (const uint32_t*)flags);
  }
  const size_t lds_un = (size_t)kGChunk * 8;
  allow_lds(k_gunpermute, lds_un);
  hipLaunchKernelGGL(k_gunpermute, dim3((unsigned)C), dim3(1024), lds_un, s, (const uint32_t*)sorted,
                     (const int64_t*)reply, (const uint16_t*)starts, (int)B, n, tgt,
                     (const uint32_t*)flags);
  return (int)hipGetLastError();
}

}  // namespace bale
