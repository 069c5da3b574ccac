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
                                                         uint64_t* __restrict__ staged) {
  part_scatter_body<PACKED, false, false, uint64_t>(
      blockIdx.x, idx, n, tlen, shift, B, S, chunk, rows, start, nullptr, staged,
      [](int64_t e, uint64_t, uint64_t low) { return ((uint64_t)e << 32) | low; },
      [](uint32_t, uint32_t) {});
}

// Persistent: a workgroup takes a segment of one region's requests and answers
// them.  Entries are (request slot << 32) | offset in region.
__global__ __launch_bounds__(1024) void k_gather_apply(const uint64_t* __restrict__ staged,
                                                       const uint32_t* __restrict__ start,
                                                       const uint32_t* __restrict__ segstart,
                                                       uint32_t* __restrict__ misc, int B,
                                                       uint32_t seg_len, int shift,
                                                       const int64_t* __restrict__ table,
                                                       int64_t* __restrict__ tgt) {
  __shared__ uint32_t s_work[2];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const uint32_t total = misc[0];
  while (true) {
    if (tid == 0) {
      uint32_t w = atomicAdd(&misc[1], 1u);
      uint32_t b = 0;
      if (w < total) {
        int lo = 0, hi = B - 1;
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (segstart[mid] <= w) lo = mid; else hi = mid - 1;
        }
        b = (uint32_t)lo;
      }
      s_work[0] = w;
      s_work[1] = b;
    }
    __syncthreads();
    const uint32_t w = s_work[0], b = s_work[1];
    if (w >= total) break;
    const uint32_t s = w - segstart[b];
    const uint32_t e0 = start[b] + s * seg_len;
    const uint32_t e1 = min(start[b + 1], e0 + seg_len);
    const int64_t* region = table + ((int64_t)b << shift);
    for (uint32_t k = e0 + tid; k < e1; k += 4 * nthr) {
      uint64_t ent[4];
      int64_t v[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        uint32_t kk = k + q * nthr;
        ent[q] = kk < e1 ? __builtin_nontemporal_load(staged + kk) : ~0ULL;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] = ent[q] != ~0ULL ? region[(uint32_t)ent[q]] : 0;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (ent[q] != ~0ULL) tgt[ent[q] >> 32] = v[q];
    }
    __syncthreads();
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
  if (S == 0) return (int)hipErrorInvalidValue;
  int gpc = c.groups_per_cu > 0 ? c.groups_per_cu : 1;
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * gpc, ceil_div(n, 4096)));
  int64_t chunk = ceil_div(n, C);
  C = ceil_div(n, chunk);
  void* ws = workspace(d, part_bytes(C, B, n, sizeof(uint64_t)), s);
  if (!ws) return (int)hipErrorOutOfMemory;
  PartLayout L = part_carve(ws, C, B);
  size_t lds_count = (size_t)B * 4;
  if (packed)
    hipLaunchKernelGGL((k_part_count<true, false, false>), dim3((unsigned)C), dim3(1024), lds_count, s, idx,
                       n, (uint64_t)table_len, shift, (int)B, chunk, L.chunk, nullptr, d->err);
  else
    hipLaunchKernelGGL((k_part_count<false, false, false>), dim3((unsigned)C), dim3(1024), lds_count, s, idx,
                       n, (uint64_t)table_len, shift, (int)B, chunk, L.chunk, nullptr, d->err);
  hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)ceil_div(B, 256)), dim3(256), 0, s, L.chunk,
                     (int)C, (int)B, L.total);
  int64_t seg = std::max<int64_t>(4096, ceil_div(n, (int64_t)d->cus * 4));
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, (int)B, (uint32_t)seg,
                     L.start, L.segstart, L.misc);
  size_t lds = part_scatter_lds((int)B, S, sizeof(uint64_t), false);
  if (packed) {
    allow_lds(k_gather_scatter<true>, lds);
    hipLaunchKernelGGL(k_gather_scatter<true>, dim3((unsigned)C), dim3(1024), lds, s, idx, n,
                       (uint64_t)table_len, shift, (int)B, S, chunk, L.chunk, L.start,
                       (uint64_t*)L.staged);
  } else {
    allow_lds(k_gather_scatter<false>, lds);
    hipLaunchKernelGGL(k_gather_scatter<false>, dim3((unsigned)C), dim3(1024), lds, s, idx, n,
                       (uint64_t)table_len, shift, (int)B, S, chunk, L.chunk, L.start,
                       (uint64_t*)L.staged);
  }
  hipLaunchKernelGGL(k_gather_apply, dim3((unsigned)(d->cus * 2)), dim3(1024), 0, s,
                     (const uint64_t*)L.staged, L.start, L.segstart, L.misc, (int)B, (uint32_t)seg,
                     shift, table, tgt);
  return (int)hipGetLastError();
}

}  // namespace bale
