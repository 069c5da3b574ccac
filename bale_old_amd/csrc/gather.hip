// gather.hip -- index gather tgt[i] = table[index[i]] on gfx950.
//
//   DIRECT  every lane issues its U table loads back to back so each wave
//           keeps many independent random reads in flight (memory-level
//           parallelism is what bounds a random gather); index and tgt move
//           as 16-byte non-temporal streams so they do not evict table lines.
//           The AGI model (apps/ig_src/ig_agi.upc:61-64).
//   BUCKET  requests are partitioned by table region through LDS send buffers
//           (partition.h); each apply workgroup then serves one region, whose
//           lines stay in its L2 while the region's requests are answered, and
//           writes each reply straight into its request slot -- the
//           request/reply exchange of apps/ig_src/ig_exstack.upc:69-94 and
//           ig_conveyor.upc:78-101.
#include "bale_internal.h"
#include "partition.h"

namespace bale {

typedef i64x2 v2i64;

// VEC: index and tgt are 16-byte aligned, so pairs move as one 16-byte access.
template <int U, bool PACKED, bool VEC>
__global__ __launch_bounds__(1024) void k_gather_direct(int64_t* __restrict__ tgt,
                                                        const int64_t* __restrict__ idx, int64_t n,
                                                        const int64_t* __restrict__ table,
                                                        uint64_t tlen, ErrWords* err) {
  static_assert(U % 2 == 0, "U counts int64 elements, loaded in pairs");
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

template <bool PACKED>
__global__ __launch_bounds__(1024) void k_gather_scatter(const int64_t* __restrict__ idx, int64_t n,
                                                         uint64_t tlen, int shift, int B, int S,
                                                         int64_t chunk,
                                                         const uint32_t* __restrict__ rows,
                                                         const uint32_t* __restrict__ start,
                                                         uint32_t* __restrict__ off_out,
                                                         uint32_t* __restrict__ pos_of) {
  // a shipped entry (slot << 32 | offset) leaves its offset in partition order
  // and its partition position in pos_of[slot]; a skipped request gets ~0
  part_scatter_body<PACKED, false, false, uint64_t>(
      blockIdx.x, idx, n, tlen, shift, B, S, chunk, rows, start, nullptr,
      [off_out, pos_of](uint32_t pos, uint64_t ent) {
        off_out[pos] = (uint32_t)ent;
        pos_of[ent >> 32] = pos;
      },
      [](int64_t e, uint64_t, uint64_t low) { return ((uint64_t)e << 32) | low; },
      [](uint32_t, uint32_t) {},
      [pos_of](int64_t e) { pos_of[e] = ~0u; });
}

constexpr int kXcds = 8;  // workgroup b runs on XCD b % 8 (round-robin dispatch)

// Per-XCD segment lists: XCD x serves the regions r with r % 8 == x, in order.
// xseg[r] = segments of XCD x's regions before r; misc[2 + x] = XCD x's total;
// misc[10 + x] = XCD x's queue head.  One workgroup, a thread per XCD.
__global__ __launch_bounds__(64) void k_gather_xseg(const uint32_t* __restrict__ total, int B,
                                                    uint32_t seg_len, uint32_t* __restrict__ xseg,
                                                    uint32_t* __restrict__ misc) {
  const int x = threadIdx.x;
  if (x >= kXcds) return;
  uint32_t run = 0;
  for (int r = x; r < B; r += kXcds) {
    xseg[r] = run;
    run += (total[r] + seg_len - 1) / seg_len;
  }
  misc[2 + x] = run;
  misc[10 + x] = 0;
}

// Persistent, XCD-aware: a workgroup takes segments only from its own XCD's
// queue, so one region's table lines are read into one XCD's L2 and served
// from there while that XCD answers the region's requests in partition order.
__global__ __launch_bounds__(1024) void k_gather_region(const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ xseg,
                                                        uint32_t* __restrict__ misc, int B,
                                                        uint32_t seg_len, int shift,
                                                        const int64_t* __restrict__ table,
                                                        int64_t* __restrict__ res) {
  __shared__ uint32_t s_work[2];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int x = blockIdx.x % kXcds;
  const uint32_t total = misc[2 + x];
  const int nr = (B - x + kXcds - 1) / kXcds;  // regions x, x+8, ...
  while (true) {
    if (tid == 0) {
      uint32_t w = atomicAdd(&misc[10 + x], 1u);
      uint32_t r = 0;
      if (w < total) {
        int lo = 0, hi = nr - 1;  // last region index k with xseg[x + 8k] <= w
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (xseg[x + kXcds * mid] <= w) lo = mid; else hi = mid - 1;
        }
        r = (uint32_t)(x + kXcds * lo);
      }
      s_work[0] = w;
      s_work[1] = r;
    }
    __syncthreads();
    const uint32_t w = s_work[0], r = s_work[1];
    __syncthreads();
    if (w >= total) break;
    const uint32_t e0 = start[r] + (w - xseg[r]) * seg_len;
    const uint32_t e1 = min(start[r + 1], e0 + seg_len);
    const int64_t* region = table + ((int64_t)r << shift);
    constexpr int Q = 8;
    for (uint32_t k = e0 + tid; k < e1; k += Q * nthr) {
      uint32_t o[Q];
      int64_t v[Q];
#pragma unroll
      for (int q = 0; q < Q; q++) {
        uint32_t kk = k + q * nthr;
        o[q] = kk < e1 ? __builtin_nontemporal_load(off + kk) : 0u;
      }
#pragma unroll
      for (int q = 0; q < Q; q++) v[q] = region[o[q]];
#pragma unroll
      for (int q = 0; q < Q; q++) {
        uint32_t kk = k + q * nthr;
        if (kk < e1) __builtin_nontemporal_store(v[q], res + kk);
      }
    }
  }
}

// tgt[i] = res[pos_of[i]]: a request's reply sits where its request was staged.
// One workgroup streams one partition chunk: that chunk's replies fill one
// range per region, each consumed front to back, so the reads stay in cache.
// A skipped request (~0) gets 0.
__global__ __launch_bounds__(1024) void k_gather_unpermute(const uint32_t* __restrict__ pos_of,
                                                           const int64_t* __restrict__ res,
                                                           const uint32_t* __restrict__ start,
                                                           int B, int64_t n, int64_t chunk,
                                                           int64_t* __restrict__ tgt) {
  const uint32_t valid = start[B];
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = min(n, lo + chunk);
  constexpr int Q = 4;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)blockDim.x * Q) {
    uint32_t p[Q];
    int64_t v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
      int64_t i = i0 + (int64_t)q * blockDim.x;
      p[q] = i < hi ? __builtin_nontemporal_load(pos_of + i) : ~0u;
    }
#pragma unroll
    for (int q = 0; q < Q; q++) v[q] = p[q] < valid ? res[p[q]] : 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      int64_t i = i0 + (int64_t)q * blockDim.x;
      if (i < hi) __builtin_nontemporal_store(v[q], tgt + i);
    }
  }
}

template <bool PACKED, bool VEC>
static void go_direct(int U, dim3 g, dim3 b, hipStream_t s, int64_t* tgt, const int64_t* idx,
                      int64_t n, const int64_t* table, uint64_t tlen, ErrWords* err) {
  if (U == 16)
    hipLaunchKernelGGL((k_gather_direct<16, PACKED, VEC>), g, b, 0, s, tgt, idx, n, table, tlen, err);
  else if (U == 8)
    hipLaunchKernelGGL((k_gather_direct<8, PACKED, VEC>), g, b, 0, s, tgt, idx, n, table, tlen, err);
  else
    hipLaunchKernelGGL((k_gather_direct<4, PACKED, VEC>), g, b, 0, s, tgt, idx, n, table, tlen, err);
}

int launch_gather_direct(int64_t* tgt, const int64_t* idx, int64_t n, const int64_t* table,
                         int64_t table_len, bool packed, const bale_config& c, ErrWords* err,
                         hipStream_t s) {
  if (n <= 0) return 0;
  int block = (c.block == 256 || c.block == 512 || c.block == 1024) ? c.block : 256;
  int U = (c.unroll == 4 || c.unroll == 8 || c.unroll == 16) ? c.unroll : 16;
  int64_t grid = ceil_div(n, (int64_t)block * U);
  if (grid > 0x7fffffff) return (int)hipErrorInvalidValue;
  dim3 g((unsigned)grid), b(block);
  bool vec = (((uintptr_t)tgt | (uintptr_t)idx) & 15) == 0;
  if (packed) {
    if (vec) go_direct<true, true>(U, g, b, s, tgt, idx, n, table, (uint64_t)table_len, err);
    else go_direct<true, false>(U, g, b, s, tgt, idx, n, table, (uint64_t)table_len, err);
  } else {
    if (vec) go_direct<false, true>(U, g, b, s, tgt, idx, n, table, (uint64_t)table_len, err);
    else go_direct<false, false>(U, g, b, s, tgt, idx, n, table, (uint64_t)table_len, err);
  }
  return (int)hipGetLastError();
}

constexpr int kMaxRegions = 4096;

int launch_gather_bucket(int64_t* tgt, const int64_t* idx, int64_t n, const int64_t* table,
                         int64_t table_len, bool packed, const bale_config& c, Device* d,
                         hipStream_t s) {
  if (n <= 0) return 0;
  if (n >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;  // caller splits
  // default region: 2^18 entries (2 MiB), half an XCD's L2
  int shift = c.tile_log2 > 0 ? c.tile_log2 : 18;
  if (shift > 30) shift = 30;
  int64_t B = ceil_div(table_len, (int64_t)1 << shift);
  const size_t budget = (size_t)d->lds_per_block;
  // widen regions until there are at most kMaxRegions and each has an LDS send
  // buffer of at least 8 entries
  while (shift < 30 &&
         (B > kMaxRegions || part_pick_depth((int)B, sizeof(uint64_t), budget, false) < 8)) {
    shift++;
    B = ceil_div(table_len, (int64_t)1 << shift);
  }
  int S = part_pick_depth((int)B, sizeof(uint64_t), budget, false);
  if (S == 0) return bad_config("BUCKET gather: %lld table regions do not fit in LDS", (long long)B);
  int gpc = c.groups_per_cu > 0 ? c.groups_per_cu : 1;
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * gpc, ceil_div(n, 4096)));
  int64_t chunk = ceil_div(n, C);
  C = ceil_div(n, chunk);
  // staging: offsets in partition order (u32), each request's partition
  // position (u32), and the replies in partition order (i64)
  const size_t half = (((size_t)n * 4) + 255) & ~(size_t)255;
  void* ws = workspace(d, part_bytes(C, B, 0, 1) + 2 * half + (size_t)n * 8 + 256, s);
  if (!ws) return (int)hipErrorOutOfMemory;
  PartLayout L = part_carve(ws, C, B);
  uint32_t* off = (uint32_t*)L.staged;
  uint32_t* pos_of = (uint32_t*)((char*)L.staged + half);
  int64_t* res = (int64_t*)((char*)L.staged + 2 * half);
  size_t lds_count = (size_t)B * 4;
  if (packed)
    hipLaunchKernelGGL((k_part_count<true, false, false>), dim3((unsigned)C), dim3(1024), lds_count, s, idx,
                       n, (uint64_t)table_len, shift, (int)B, chunk, L.chunk, nullptr, stream_err(d, s),
                       (const uint32_t*)nullptr);
  else
    hipLaunchKernelGGL((k_part_count<false, false, false>), dim3((unsigned)C), dim3(1024), lds_count, s, idx,
                       n, (uint64_t)table_len, shift, (int)B, chunk, L.chunk, nullptr, stream_err(d, s),
                       (const uint32_t*)nullptr);
  hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)B), dim3(256), 0, s, L.chunk,
                     (int)C, (int)B, L.total);
  // segments: one region's requests split over all of an XCD's workgroups,
  // so they work on the same (L2-resident) region at a time
  int gr = c.groups_per_cu > 0 ? c.groups_per_cu : 2;
  const int64_t per_xcd = std::max<int64_t>(1, (int64_t)d->cus * gr / kXcds);
  int64_t seg = std::max<int64_t>(2048, ceil_div(ceil_div(n, B), per_xcd));
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, (int)B, (uint32_t)seg, 0u,
                     (const uint32_t*)nullptr, L.start, L.segstart, L.misc);
  uint32_t* xseg = L.segstart;  // reused: per-XCD segment starts
  hipLaunchKernelGGL(k_gather_xseg, dim3(1), dim3(64), 0, s, L.total, (int)B, (uint32_t)seg, xseg,
                     L.misc);
  size_t lds = part_scatter_lds((int)B, S, sizeof(uint64_t), false);
  if (packed) {
    allow_lds(k_gather_scatter<true>, lds);
    hipLaunchKernelGGL(k_gather_scatter<true>, dim3((unsigned)C), dim3(1024), lds, s, idx, n,
                       (uint64_t)table_len, shift, (int)B, S, chunk, L.chunk, L.start, off, pos_of);
  } else {
    allow_lds(k_gather_scatter<false>, lds);
    hipLaunchKernelGGL(k_gather_scatter<false>, dim3((unsigned)C), dim3(1024), lds, s, idx, n,
                       (uint64_t)table_len, shift, (int)B, S, chunk, L.chunk, L.start, off, pos_of);
  }
  hipLaunchKernelGGL(k_gather_region, dim3((unsigned)(d->cus * gr)), dim3(1024), 0, s, off, L.start,
                     xseg, L.misc, (int)B, (uint32_t)seg, shift, table, res);
  hipLaunchKernelGGL(k_gather_unpermute, dim3((unsigned)C), dim3(1024), 0, s, pos_of, res, L.start,
                     (int)B, n, chunk, tgt);
  return (int)hipGetLastError();
}

}  // namespace bale
