// batch.hip -- thousands of uneven histogram / gather instances in one graph
// launch, the exhaustive per-instance autotuner, and the plain one-launch-per-
// instance baseline.
//
// A batch is planned on the host into task lists, one per phase:
//   A  (persistent, 512 threads, 32 KiB LDS)  gather (straight from the table,
//      or from an LDS copy of a table of at most 4096 entries) / direct /
//      direct+hot / LDS-tile tasks of every instance, and the count pass of
//      bucketed ones
//   B  (one workgroup per bucketed instance)   chunk offsets + bucket scan
//   C  (persistent, 1024 threads)               bucket scatter, one task per chunk
//   D  (persistent, 1024 threads, 128 KiB LDS)  bucket apply, one task per segment slot
// Each persistent kernel takes tasks from a device queue (one atomic per task),
// tasks sorted largest first.  The whole sequence -- queue reset, hot-set
// sketches, A, B, C, D -- is captured once into a HIP graph and replayed by
// bale_batch_launch.  Queue heads are cleared by a kernel (no memset node).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bale_internal.h"
#include "partition.h"

namespace bale {
namespace batchdev {

enum : uint32_t {
  OP_GATHER = 0,
  OP_HDIRECT = 1,
  OP_HDIRECT_RUNS = 2,
  OP_HHOT = 3,
  OP_HLDS = 4,
  OP_HLDS_RUNS = 5,
  OP_BCOUNT = 6,
  OP_GLDS = 7,
};
constexpr uint32_t kFlagHot = 1, kFlagRuns = 2;
// the instance's counts overlap another histogram instance's counts in the
// same batch: phase D flushes even a single-segment bucket with atomics
constexpr uint32_t kFlagShared = 4;
constexpr int kBlockA = 512;
constexpr int kLdsAWords = 8192;  // 32 KiB: LDS tiles of up to 2^13 bins
constexpr int kLdsTileBits = 13;
constexpr int kBucketBits = 15;   // apply tiles of 2^15 bins (128 KiB of LDS)
constexpr int kXcds = 8;          // MI355X: 8 XCDs, each with its own L2

struct BInst {
  void* table;          // counts (histogram) or table (gather)
  const int64_t* idx;
  int64_t* tgt;
  int64_t table_len;
  int64_t n;
  uint32_t* hot;        // hot-bin hash table or null
  uint32_t* rows;       // bucket: [C*B] chunk counts, then chunk offsets
  uint32_t* start;      // [B+1]
  uint32_t* total;      // [B]
  uint32_t* segstart;   // [B+1]
  uint32_t* misc;       // [4]
  uint16_t* staged;     // [n]
  int64_t chunk;        // bucket chunk length
  int32_t kind, dtype;
  int32_t tile_bits;    // LDS tile (OP_HLDS) or bucket width (bucketed)
  int32_t B, C, S;
  uint32_t flags;
  uint32_t seg_len;
};

struct BTask {
  uint32_t inst, begin, end, op;  // op: low 8 bits opcode, high bits tile / chunk
};

__device__ inline void add_count(const BInst& I, uint64_t j, uint32_t c) {
  if (I.dtype == BALE_I64) atomic_add_dev((int64_t*)I.table + j, (int64_t)c);
  else atomic_add_dev((int32_t*)I.table + j, (int32_t)c);
}

template <int MERGE>
__device__ inline uint32_t wave_merge(uint64_t key, bool ok) {
  if (MERGE == 0) return ok ? 1u : 0u;
  const int lane = __lane_id();
  uint64_t prev = __shfl_up(key, 1);
  bool change = (lane == 0) || (prev != key);
  uint64_t chg = __ballot(change);
  uint32_t mine = 0;
  if (ok && change) {
    uint64_t later = chg & ~((lane == 63) ? ~0ULL : ((2ULL << lane) - 1));
    int end = later ? __builtin_ctzll(later) : 64;
    mine = (uint32_t)(end - lane);
  }
  return mine;
}

// Gather from an LDS copy of the table (at most kLdsAWords / 2 entries): the
// table is read once per task, then every lookup is an LDS read.
__device__ inline void task_glds(uint32_t* sh, const BInst& I, int64_t b, int64_t e, ErrWords* err) {
  constexpr int U = 8;
  const int T = blockDim.x, t = threadIdx.x;
  int64_t* st = (int64_t*)sh;
  const int64_t* __restrict__ table = (const int64_t*)I.table;
  for (int64_t k = t; k < I.table_len; k += T) st[k] = table[k];
  __syncthreads();
  const int64_t* __restrict__ idx = I.idx;
  int64_t* __restrict__ tgt = I.tgt;
  unsigned br = 0, bp = 0;
  for (int64_t base = b; base < e; base += (int64_t)T * U) {
    int64_t raw[U], v[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      raw[k] = i < e ? __builtin_nontemporal_load(idx + i) : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      uint64_t j;
      v[k] = (i < e && decode_index<false>(raw[k], (uint64_t)I.table_len, j, br, bp)) ? st[j] : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      if (i < e) __builtin_nontemporal_store(v[k], tgt + i);
    }
  }
  flush_errors(err, br, bp);
}

__device__ inline void task_gather(const BInst& I, int64_t b, int64_t e, ErrWords* err) {
  constexpr int U = 8;
  const int T = blockDim.x, t = threadIdx.x;
  const int64_t* __restrict__ idx = I.idx;
  const int64_t* __restrict__ table = (const int64_t*)I.table;
  int64_t* __restrict__ tgt = I.tgt;
  unsigned br = 0, bp = 0;
  for (int64_t base = b; base < e; base += (int64_t)T * U) {
    int64_t raw[U], v[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      raw[k] = i < e ? __builtin_nontemporal_load(idx + i) : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      uint64_t j;
      v[k] = (i < e && decode_index<false>(raw[k], (uint64_t)I.table_len, j, br, bp)) ? table[j] : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      if (i < e) __builtin_nontemporal_store(v[k], tgt + i);
    }
  }
  flush_errors(err, br, bp);
}

template <int MERGE>
__device__ inline void task_hdirect(const BInst& I, int64_t b, int64_t e, ErrWords* err) {
  constexpr int U = 8;
  const int T = blockDim.x, t = threadIdx.x;
  unsigned br = 0, bp = 0;
  for (int64_t base = b; base < e; base += (int64_t)T * U) {
    int64_t raw[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      raw[k] = i < e ? __builtin_nontemporal_load(I.idx + i) : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      uint64_t j = ~0ULL;
      bool ok = i < e && decode_index<false>(raw[k], (uint64_t)I.table_len, j, br, bp);
      uint32_t c = wave_merge<MERGE>(ok ? j : ~0ULL, ok);
      if (c) add_count(I, j, c);
    }
  }
  flush_errors(err, br, bp);
}

__device__ inline void task_hhot(uint32_t* sh, const BInst& I, int64_t b, int64_t e, ErrWords* err) {
  constexpr int U = 8;
  const int T = blockDim.x, t = threadIdx.x;
  uint32_t* hkeys = sh;
  uint32_t* hcnt = sh + kHotSlots;
  for (int k = t; k < kHotSlots; k += T) {
    hkeys[k] = I.hot[k];
    hcnt[k] = 0;
  }
  __syncthreads();
  unsigned br = 0, bp = 0;
  for (int64_t base = b; base < e; base += (int64_t)T * U) {
    int64_t raw[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      raw[k] = i < e ? __builtin_nontemporal_load(I.idx + i) : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      uint64_t j;
      if (i < e && decode_index<false>(raw[k], (uint64_t)I.table_len, j, br, bp)) {
        int hs = hot_find(hkeys, (uint32_t)j);
        if (hs >= 0) atomicAdd(&hcnt[hs], 1u);
        else add_count(I, j, 1u);
      }
    }
  }
  __syncthreads();
  for (int k = t; k < kHotSlots; k += T)
    if (hcnt[k]) add_count(I, hkeys[k], hcnt[k]);
  flush_errors(err, br, bp);
}

template <int MERGE>
__device__ inline void task_hlds(uint32_t* sh, const BInst& I, int64_t b, int64_t e, uint32_t tile,
                                 ErrWords* err) {
  constexpr int U = 8;
  const int T = blockDim.x, t = threadIdx.x;
  const uint64_t lo_bin = (uint64_t)tile << I.tile_bits;
  const uint64_t w64 = min((uint64_t)I.table_len - lo_bin, (uint64_t)1 << I.tile_bits);
  const uint32_t width = (uint32_t)w64;
  for (uint32_t k = t; k < width; k += T) sh[k] = 0;
  __syncthreads();
  unsigned br = 0, bp = 0;
  for (int64_t base = b; base < e; base += (int64_t)T * U) {
    int64_t raw[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      raw[k] = i < e ? __builtin_nontemporal_load(I.idx + i) : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t i = base + (int64_t)k * T + t;
      uint64_t j = ~0ULL;
      bool ok = i < e && decode_index<false>(raw[k], (uint64_t)I.table_len, j, br, bp);
      ok = ok && (j - lo_bin) < w64;
      uint32_t c = wave_merge<MERGE>(ok ? j : ~0ULL, ok);
      if (c) atomicAdd(&sh[j - lo_bin], c);
    }
  }
  __syncthreads();
  for (uint32_t k = t; k < width; k += T) {
    uint32_t c = sh[k];
    if (c) add_count(I, lo_bin + k, c);
  }
  if (tile == 0) flush_errors(err, br, bp);
}

// Phase A: a persistent loop over tasks.  Queue q holds tasks[qoff[q],
// qoff[q+1]); a workgroup takes from queue blockIdx % nq until it is empty,
// then from the next ones in turn, so the loop ends once all nq queues are
// drained (each head only grows).
__global__ __launch_bounds__(kBlockA) void k_batch_a(const BInst* __restrict__ insts,
                                                     const BTask* __restrict__ tasks,
                                                     const uint32_t* __restrict__ qoff, uint32_t nq,
                                                     uint32_t* __restrict__ head, ErrWords* err) {
  extern __shared__ uint32_t sh[];
  __shared__ BTask s_t;
  const uint32_t q0 = blockIdx.x % nq;
  uint32_t tried = 0;  // thread 0: queues found empty so far
  while (true) {
    if (threadIdx.x == 0) {
      BTask t{~0u, 0, 0, 0};
      while (tried < nq) {
        const uint32_t q = (q0 + tried) % nq;
        const uint32_t k = atomicAdd(head + q, 1u);
        if (k < qoff[q + 1] - qoff[q]) {
          t = tasks[qoff[q] + k];
          break;
        }
        tried++;
      }
      s_t = t;
    }
    __syncthreads();
    const BTask T = s_t;
    __syncthreads();
    if (T.inst == ~0u) break;
    const BInst& I = insts[T.inst];
    const uint32_t op = T.op & 0xff, arg = T.op >> 8;
    switch (op) {
      case OP_GATHER: task_gather(I, T.begin, T.end, err); break;
      case OP_GLDS: task_glds(sh, I, T.begin, T.end, err); break;
      case OP_HDIRECT: task_hdirect<0>(I, T.begin, T.end, err); break;
      case OP_HDIRECT_RUNS: task_hdirect<1>(I, T.begin, T.end, err); break;
      case OP_HHOT: task_hhot(sh, I, T.begin, T.end, err); break;
      case OP_HLDS: task_hlds<0>(sh, I, T.begin, T.end, arg, err); break;
      case OP_HLDS_RUNS: task_hlds<1>(sh, I, T.begin, T.end, arg, err); break;
      case OP_BCOUNT:
        switch (I.flags & 3) {
          case 0: part_count_body<false, false, false>(sh, arg, I.idx, I.n, I.table_len, I.tile_bits, I.B, I.chunk, I.rows, I.hot, err); break;
          case 1: part_count_body<false, true, false>(sh, arg, I.idx, I.n, I.table_len, I.tile_bits, I.B, I.chunk, I.rows, I.hot, err); break;
          case 2: part_count_body<false, false, true>(sh, arg, I.idx, I.n, I.table_len, I.tile_bits, I.B, I.chunk, I.rows, I.hot, err); break;
          default: part_count_body<false, true, true>(sh, arg, I.idx, I.n, I.table_len, I.tile_bits, I.B, I.chunk, I.rows, I.hot, err); break;
        }
        break;
      default: break;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void k_batch_reset(uint32_t* head, int k) {
  if ((int)threadIdx.x < k) head[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(1024) void k_batch_sketch(const BInst* __restrict__ insts,
                                                       const uint32_t* __restrict__ list) {
  const BInst& I = insts[list[blockIdx.x]];
  hot_sketch_body<false>(I.idx, I.n, (uint64_t)I.table_len, I.hot);
}

// Phase B: one workgroup per bucketed instance.
__global__ __launch_bounds__(256) void k_batch_b(const BInst* __restrict__ insts,
                                                 const uint32_t* __restrict__ list) {
  const BInst& I = insts[list[blockIdx.x]];
  for (int b = threadIdx.x; b < I.B; b += blockDim.x) {
    uint32_t run = 0;
    for (int c = 0; c < I.C; c++) {
      uint32_t v = I.rows[(size_t)c * I.B + b];
      I.rows[(size_t)c * I.B + b] = run;
      run += v;
    }
    I.total[b] = run;
  }
  __syncthreads();
  part_scan_body(I.total, I.B, I.seg_len, I.start, I.segstart, I.misc);
}

// Phase C: scatter tasks (inst, chunk).
__global__ __launch_bounds__(1024) void k_batch_c(const BInst* __restrict__ insts,
                                                  const BTask* __restrict__ tasks, uint32_t ntasks,
                                                  uint32_t* __restrict__ head) {
  __shared__ BTask s_t;
  while (true) {
    if (threadIdx.x == 0) {
      uint32_t k = atomicAdd(head, 1u);
      s_t = k < ntasks ? tasks[k] : BTask{~0u, 0, 0, 0};
    }
    __syncthreads();
    const BTask T = s_t;
    __syncthreads();
    if (T.inst == ~0u) break;
    const BInst& I = insts[T.inst];
    auto make = [](int64_t, uint64_t, uint64_t low) { return (uint16_t)low; };
    auto flush = [&I](uint32_t bin, uint32_t c) { add_count(I, bin, c); };
    uint16_t* staged = I.staged;
    auto store = [staged](uint32_t pos, uint16_t v) { staged[pos] = v; };
    auto skip = [](int64_t) {};
    switch (I.flags & 3) {
      case 0: part_scatter_body<false, false, false, uint16_t>(T.begin, I.idx, I.n, I.table_len, I.tile_bits, I.B, I.S, I.chunk, I.rows, I.start, I.hot, store, make, flush, skip); break;
      case 1: part_scatter_body<false, true, false, uint16_t>(T.begin, I.idx, I.n, I.table_len, I.tile_bits, I.B, I.S, I.chunk, I.rows, I.start, I.hot, store, make, flush, skip); break;
      case 2: part_scatter_body<false, false, true, uint16_t>(T.begin, I.idx, I.n, I.table_len, I.tile_bits, I.B, I.S, I.chunk, I.rows, I.start, I.hot, store, make, flush, skip); break;
      default: part_scatter_body<false, true, true, uint16_t>(T.begin, I.idx, I.n, I.table_len, I.tile_bits, I.B, I.S, I.chunk, I.rows, I.start, I.hot, store, make, flush, skip); break;
    }
    __syncthreads();
  }
}

// Phase D: apply tasks (inst, segment slot); slots past the instance's segment
// count are empty.
__global__ __launch_bounds__(1024) void k_batch_d(const BInst* __restrict__ insts,
                                                  const BTask* __restrict__ tasks, uint32_t ntasks,
                                                  uint32_t* __restrict__ head) {
  extern __shared__ uint32_t sh[];
  __shared__ BTask s_t;
  __shared__ uint32_t s_b;
  const int tid = threadIdx.x, T = blockDim.x;
  while (true) {
    if (tid == 0) {
      uint32_t k = atomicAdd(head, 1u);
      s_t = k < ntasks ? tasks[k] : BTask{~0u, 0, 0, 0};
    }
    __syncthreads();
    const BTask tk = s_t;
    __syncthreads();
    if (tk.inst == ~0u) break;
    const BInst& I = insts[tk.inst];
    const uint32_t w = tk.begin;
    if (w >= I.misc[0]) continue;  // uniform: every thread skips
    if (tid == 0) {
      int lo = 0, hi = I.B - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (I.segstart[mid] <= w) lo = mid; else hi = mid - 1;
      }
      s_b = (uint32_t)lo;
    }
    __syncthreads();
    const uint32_t b = s_b;
    const uint32_t nseg = I.segstart[b + 1] - I.segstart[b];
    const uint32_t e0 = I.start[b] + (w - I.segstart[b]) * I.seg_len;
    const uint32_t e1 = min(I.start[b + 1], e0 + I.seg_len);
    const uint64_t lo_bin = (uint64_t)b << I.tile_bits;
    const uint32_t width = (uint32_t)min((uint64_t)1 << I.tile_bits, (uint64_t)I.table_len - lo_bin);
    for (uint32_t k = tid; k < width; k += T) sh[k] = 0;
    __syncthreads();
    const bool runs = (I.flags & kFlagRuns) != 0;
    // each thread takes 8 consecutive entries; with runs, equal neighbours
    // within the group are added at once
    for (uint32_t g = e0 + 8u * tid; g < e1; g += 8u * T) {
      uint32_t m = min(8u, e1 - g);
      if (!runs) {
        for (uint32_t q = 0; q < m; q++) atomicAdd(&sh[I.staged[g + q]], 1u);
      } else {
        uint32_t cur = I.staged[g], cnt = 1;
        for (uint32_t q = 1; q < m; q++) {
          uint32_t v = I.staged[g + q];
          if (v == cur) {
            cnt++;
          } else {
            atomicAdd(&sh[cur], cnt);
            cur = v;
            cnt = 1;
          }
        }
        atomicAdd(&sh[cur], cnt);
      }
    }
    __syncthreads();
    if (nseg == 1 && !(I.flags & kFlagShared)) {
      // the bucket is this task's alone (no other instance writes these counts)
      for (uint32_t k = tid; k < width; k += T) {
        uint32_t c = sh[k];
        if (!c) continue;
        if (I.dtype == BALE_I64) ((int64_t*)I.table)[lo_bin + k] += (int64_t)c;
        else ((int32_t*)I.table)[lo_bin + k] += (int32_t)c;
      }
    } else {
      for (uint32_t k = tid; k < width; k += T) {
        uint32_t c = sh[k];
        if (c) add_count(I, lo_bin + k, c);
      }
    }
    __syncthreads();
  }
}

// Add the graph's error counts to a stream's record and clear them.
__global__ void k_batch_err_move(ErrWords* __restrict__ from, ErrWords* __restrict__ to) {
  if (threadIdx.x == 0) {
    const unsigned long long r = from->range, p = from->pe;
    if (r) atomicAdd(&to->range, r);
    if (p) atomicAdd(&to->pe, p);
    from->range = 0;
    from->pe = 0;
  }
}

}  // namespace batchdev
}  // namespace bale

// ============================== host side ===================================
namespace bale {
using namespace batchdev;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  ~DevBuf() { dev_free(p); }
  int ensure(size_t bytes) {
    if (cap >= bytes) return 0;
    dev_free(p);
    p = nullptr;
    cap = 0;
    if (!(p = BALE_DEV_ALLOC(bytes ? bytes : 256))) return (int)hipErrorOutOfMemory;
    cap = bytes ? bytes : 256;
    return 0;
  }
};

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
inline int ceil_log2(int64_t v) {
  int b = 0;
  while (((int64_t)1 << b) < v) b++;
  return b;
}

// Arena offsets of one instance's scratch, patched into pointers after allocation.
struct Offs {
  size_t hot = SIZE_MAX, rows = SIZE_MAX, start = 0, total = 0, segstart = 0, misc = 0, staged = 0;
};

struct Plan {
  std::vector<BInst> inst;
  std::vector<Offs> off;
  std::vector<BTask> a, c, d;
  std::vector<uint32_t> qoff;  // phase-A queue q (one per XCD) holds a[qoff[q], qoff[q+1])
  std::vector<uint32_t> hot_list, bucket_list;
  size_t arena = 0;
  size_t lds_c = 0;
};

struct Buffers {
  DevBuf arena, inst, a, qoff, c, d, hot_list, bucket_list, heads;
  DevBuf err;  // the captured graph's index errors, moved to the launching stream's record
};

// Default per-instance rule when cfg.strategy is AUTO.
void batch_default(const bale_instance& in, bale_config& c) {
  memset(&c, 0, sizeof c);
  if (in.kind == 1) {
    c.strategy = BALE_STRAT_DIRECT;
    return;
  }
  if (in.table_len <= ((int64_t)1 << kLdsTileBits) && in.n >= 2 * in.table_len) {
    c.strategy = BALE_STRAT_LDS;
  } else if (in.table_len >= ((int64_t)1 << 20) && in.n >= in.table_len / 8) {
    c.strategy = BALE_STRAT_BUCKET;
    c.hot = 1;
  } else {
    c.strategy = BALE_STRAT_DIRECT;
    c.hot = 1;
  }
}

int plan_add(Plan& P, const bale_instance& in, const Device* d) {
  const uint32_t id = (uint32_t)P.inst.size();
  BInst I;
  memset(&I, 0, sizeof I);
  I.table = in.table;
  I.idx = in.index;
  I.tgt = in.tgt;
  I.table_len = in.table_len;
  I.n = in.n;
  I.kind = in.kind;
  I.dtype = in.dtype;
  Offs o;
  bale_config c = in.cfg;
  if (c.strategy == BALE_STRAT_AUTO) batch_default(in, c);
  int tl = c.task_log2 > 0 ? std::min(22, std::max(12, (int)c.task_log2)) : 16;
  const int64_t task = (int64_t)1 << tl;
  auto add_range_tasks = [&](uint32_t op) {
    for (int64_t b = 0; b < in.n; b += task)
      P.a.push_back(BTask{id, (uint32_t)b, (uint32_t)std::min(in.n, b + task), op});
  };
  if (in.kind == 1) {
    const bool in_lds = c.strategy == BALE_STRAT_LDS && in.table_len <= kLdsAWords / 2;
    add_range_tasks(in_lds ? OP_GLDS : OP_GATHER);
  } else if (c.strategy == BALE_STRAT_LDS &&
             ceil_div(in.table_len, (int64_t)1 << kLdsTileBits) <= 64) {
    int tb = std::max(8, std::min(kLdsTileBits, ceil_log2(in.table_len)));
    I.tile_bits = tb;
    int64_t tiles = ceil_div(in.table_len, (int64_t)1 << tb);
    uint32_t op = c.merge > 0 ? OP_HLDS_RUNS : OP_HLDS;
    for (int64_t b = 0; b < in.n; b += task)
      for (int64_t t = 0; t < tiles; t++)
        P.a.push_back(BTask{id, (uint32_t)b, (uint32_t)std::min(in.n, b + task),
                            op | ((uint32_t)t << 8)});
  } else if (c.strategy == BALE_STRAT_BUCKET && in.table_len >= 2 &&
             ceil_div(in.table_len, (int64_t)1 << kBucketBits) <= 4096) {
    int tb = (c.tile_log2 >= 10 && c.tile_log2 <= kBucketBits) ? c.tile_log2 : kBucketBits;
    tb = std::min(tb, std::max(1, ceil_log2(in.table_len)));
    const int64_t B = ceil_div(in.table_len, (int64_t)1 << tb);
    const int64_t C = ceil_div(in.n, task);
    I.tile_bits = tb;
    I.B = (int32_t)B;
    I.C = (int32_t)C;
    I.chunk = task;
    I.flags = (c.hot ? kFlagHot : 0) | (c.merge > 0 ? kFlagRuns : 0);
    const size_t budget = (size_t)d->lds_per_block;
    int S = std::min(64, part_pick_depth((int)B, sizeof(uint16_t), budget, c.hot != 0));
    if (S < 4) return BALE_ERR_ARG;
    // Few buckets: deeper buffers (up to 8192 entries, 64 KiB of entries in
    // all).  A round places 8 entries per thread, so with one to four buckets
    // a 64-entry buffer overflows dozens of times per round, each overflow a
    // barrier and a ship.
    while (S < 8192 && B * (int64_t)S * 2 * (int64_t)sizeof(uint16_t) <= ((int64_t)64 << 10) &&
           part_scatter_lds((int)B, 2 * S, sizeof(uint16_t), c.hot != 0) <= budget)
      S *= 2;
    I.S = S;
    P.lds_c = std::max(P.lds_c, part_scatter_lds((int)B, S, sizeof(uint16_t), c.hot != 0));
    int64_t seg = std::max<int64_t>({(int64_t)4 << tb, 16384, ceil_div(in.n, 64)});
    I.seg_len = (uint32_t)seg;
    o.rows = P.arena;   P.arena = align256(P.arena + (size_t)C * B * 4);
    o.start = P.arena;  P.arena = align256(P.arena + (size_t)(B + 1) * 4);
    o.total = P.arena;  P.arena = align256(P.arena + (size_t)B * 4);
    o.segstart = P.arena; P.arena = align256(P.arena + (size_t)(B + 1) * 4);
    o.misc = P.arena;   P.arena = align256(P.arena + 16);
    o.staged = P.arena; P.arena = align256(P.arena + (size_t)in.n * 2);
    if (c.hot) {
      o.hot = P.arena;
      P.arena = align256(P.arena + kHotSlots * 4);
      P.hot_list.push_back(id);
    }
    P.bucket_list.push_back(id);
    for (int64_t k = 0; k < C; k++) {
      P.a.push_back(BTask{id, 0, (uint32_t)std::min(task, in.n - k * task), OP_BCOUNT | ((uint32_t)k << 8)});
      P.c.push_back(BTask{id, (uint32_t)k, 0, 0});
    }
    int64_t maxseg = ceil_div(in.n, seg) + B;
    for (int64_t w = 0; w < maxseg; w++) P.d.push_back(BTask{id, (uint32_t)w, 0, 0});
  } else if (c.hot && in.table_len < (int64_t)kHotEmpty) {
    o.hot = P.arena;
    P.arena = align256(P.arena + kHotSlots * 4);
    P.hot_list.push_back(id);
    add_range_tasks(OP_HHOT);
  } else {
    add_range_tasks(c.merge > 0 ? OP_HDIRECT_RUNS : OP_HDIRECT);
  }
  P.inst.push_back(I);
  P.off.push_back(o);
  return BALE_OK;
}

// Allocate, patch scratch pointers, order tasks and upload.
int plan_upload(Plan& P, Buffers& Bf, hipStream_t s) {
  int rc;
  if ((rc = Bf.arena.ensure(std::max<size_t>(P.arena, 256)))) return rc;
  char* base = (char*)Bf.arena.p;
  for (size_t i = 0; i < P.inst.size(); i++) {
    BInst& I = P.inst[i];
    const Offs& o = P.off[i];
    if (o.hot != SIZE_MAX) I.hot = (uint32_t*)(base + o.hot);
    if (o.rows != SIZE_MAX) {
      I.rows = (uint32_t*)(base + o.rows);
      I.start = (uint32_t*)(base + o.start);
      I.total = (uint32_t*)(base + o.total);
      I.segstart = (uint32_t*)(base + o.segstart);
      I.misc = (uint32_t*)(base + o.misc);
      I.staged = (uint16_t*)(base + o.staged);
    }
  }
  // Phase A runs from one queue per XCD (workgroup b is dispatched to XCD
  // b % kXcds and drains that queue first, then the others): whole instances
  // go to the least-loaded queue, heaviest first, and each queue lists the
  // bucket counts first (they gate phases B-D), then its instances' tasks
  // together, heaviest instance first.  An XCD's L2 then holds the tables of
  // a few instances at a time, not of every instance in flight.
  {
    const size_t ni = P.inst.size();
    std::vector<int64_t> w(ni, 0);
    for (const BTask& t : P.a) w[t.inst] += (int64_t)(t.end - t.begin);
    std::vector<uint32_t> ord(ni), q(ni, 0);
    for (size_t i = 0; i < ni; i++) ord[i] = (uint32_t)i;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return w[x] > w[y]; });
    int64_t load[kXcds] = {};
    for (uint32_t i : ord) {
      int best = 0;
      for (int k = 1; k < kXcds; k++)
        if (load[k] < load[best]) best = k;
      q[i] = (uint32_t)best;
      load[best] += w[i];
    }
    std::stable_sort(P.a.begin(), P.a.end(), [&](const BTask& x, const BTask& y) {
      if (q[x.inst] != q[y.inst]) return q[x.inst] < q[y.inst];
      bool cx = (x.op & 0xff) == OP_BCOUNT, cy = (y.op & 0xff) == OP_BCOUNT;
      if (cx != cy) return cx;
      if (w[x.inst] != w[y.inst]) return w[x.inst] > w[y.inst];
      return x.inst < y.inst;
    });
    P.qoff.assign(kXcds + 1, 0);
    for (const BTask& t : P.a) P.qoff[q[t.inst] + 1]++;
    for (int k = 0; k < kXcds; k++) P.qoff[k + 1] += P.qoff[k];
  }
  auto up = [&](DevBuf& b, const void* src, size_t bytes) -> int {
    int r = b.ensure(std::max<size_t>(bytes, 16));
    if (r) return r;
    if (bytes) return (int)hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s);
    return 0;
  };
  // on failure, copies already queued finish before the caller frees their
  // buffers and sources
  auto failed = [&](int r) {
    (void)hipStreamSynchronize(s);
    (void)hipGetLastError();
    return r;
  };
  if ((rc = up(Bf.inst, P.inst.data(), P.inst.size() * sizeof(BInst)))) return failed(rc);
  if ((rc = up(Bf.a, P.a.data(), P.a.size() * sizeof(BTask)))) return failed(rc);
  if ((rc = up(Bf.qoff, P.qoff.data(), P.qoff.size() * 4))) return failed(rc);
  if ((rc = up(Bf.c, P.c.data(), P.c.size() * sizeof(BTask)))) return failed(rc);
  if ((rc = up(Bf.d, P.d.data(), P.d.size() * sizeof(BTask)))) return failed(rc);
  if ((rc = up(Bf.hot_list, P.hot_list.data(), P.hot_list.size() * 4))) return failed(rc);
  if ((rc = up(Bf.bucket_list, P.bucket_list.data(), P.bucket_list.size() * 4))) return failed(rc);
  if ((rc = Bf.heads.ensure(64))) return failed(rc);
  if ((rc = Bf.err.ensure(sizeof(ErrWords)))) return failed(rc);
  if ((rc = (int)hipMemsetAsync(Bf.err.p, 0, sizeof(ErrWords), s))) return failed(rc);
  return (int)hipStreamSynchronize(s);
}

constexpr size_t kLdsD = (size_t)4 << kBucketBits;

void set_batch_lds(const Plan& P) {
  static std::once_flag once;
  std::call_once(once, [] {
    allow_lds(k_batch_a, (size_t)kLdsAWords * 4);
    allow_lds(k_batch_d, kLdsD);
  });
  allow_lds(k_batch_c, P.lds_c);
}

// Enqueue the phases on s (also used inside graph capture).
void plan_enqueue(const Plan& P, Buffers& Bf, Device* d, hipStream_t s, ErrWords* err) {
  const BInst* insts = (const BInst*)Bf.inst.p;
  uint32_t* heads = (uint32_t*)Bf.heads.p;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, heads, 4 + kXcds);
  if (!P.hot_list.empty())
    hipLaunchKernelGGL(k_batch_sketch, dim3((unsigned)P.hot_list.size()), dim3(1024), 0, s, insts,
                       (const uint32_t*)Bf.hot_list.p);
  if (!P.a.empty()) {
    int64_t g = std::min<int64_t>((int64_t)d->cus * 4, (int64_t)P.a.size());
    hipLaunchKernelGGL(k_batch_a, dim3((unsigned)g), dim3(kBlockA), (size_t)kLdsAWords * 4, s, insts,
                       (const BTask*)Bf.a.p, (const uint32_t*)Bf.qoff.p, (uint32_t)kXcds, heads + 4, err);
  }
  if (!P.bucket_list.empty()) {
    hipLaunchKernelGGL(k_batch_b, dim3((unsigned)P.bucket_list.size()), dim3(256), 0, s, insts,
                       (const uint32_t*)Bf.bucket_list.p);
    int64_t gc = std::min<int64_t>((int64_t)d->cus, (int64_t)P.c.size());
    hipLaunchKernelGGL(k_batch_c, dim3((unsigned)gc), dim3(1024), P.lds_c, s, insts,
                       (const BTask*)Bf.c.p, (uint32_t)P.c.size(), heads + 1);
    int64_t gd = std::min<int64_t>((int64_t)d->cus, (int64_t)P.d.size());
    hipLaunchKernelGGL(k_batch_d, dim3((unsigned)gd), dim3(1024), kLdsD, s, insts,
                       (const BTask*)Bf.d.p, (uint32_t)P.d.size(), heads + 2);
  }
}

// Histogram instances whose counts share bytes with another histogram
// instance's counts (many index streams into one table is a natural batch)
// get kFlagShared: their buckets' flushes race with the other instance's.
void mark_shared_counts(Plan& P) {
  struct Span { uintptr_t lo, hi; size_t id; };
  std::vector<Span> v;
  for (size_t i = 0; i < P.inst.size(); i++) {
    const BInst& I = P.inst[i];
    if (I.kind != 0 || I.n == 0) continue;
    const uintptr_t lo = (uintptr_t)I.table;
    v.push_back({lo, lo + (uintptr_t)I.table_len * (I.dtype == BALE_I64 ? 8u : 4u), i});
  }
  std::sort(v.begin(), v.end(), [](const Span& a, const Span& b) { return a.lo < b.lo; });
  // clusters of spans chained by overlaps: every member of a cluster of two
  // or more is marked (a plain flush must not meet an atomic one either)
  size_t first = 0;
  uintptr_t reach = 0;
  for (size_t k = 0; k <= v.size(); k++) {
    if (k < v.size() && k > first && v[k].lo < reach) {
      reach = std::max(reach, v[k].hi);
      continue;
    }
    if (k > first + 1)
      for (size_t m = first; m < k; m++) P.inst[v[m].id].flags |= kFlagShared;
    if (k < v.size()) {
      first = k;
      reach = v[k].hi;
    }
  }
}

int check_instance(const bale_instance& in) {
  if (in.kind != 0 && in.kind != 1) return BALE_ERR_ARG;
  if (in.n < 0 || in.n >= ((int64_t)1 << 32)) return BALE_ERR_ARG;
  if (in.n > 0 && (!in.index || !in.table || in.table_len <= 0)) return BALE_ERR_ARG;
  if (in.kind == 1 && in.n > 0 && !in.tgt) return BALE_ERR_ARG;
  if (in.kind == 0 && in.dtype != BALE_I64 && in.dtype != BALE_I32) return BALE_ERR_ARG;
  return BALE_OK;
}

}  // namespace
}  // namespace bale

struct bale_batch {
  bale::Plan plan;
  bale::Buffers bufs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture = nullptr;
  int device = -1;
};

using namespace bale;

extern "C" {

bale_batch* bale_batch_create(const bale_instance* inst, int64_t count, void* stream) {
  if (count < 0 || (count > 0 && !inst)) {
    bale_set_error(BALE_ERR_ARG, "bad instance array");
    return nullptr;
  }
  if (count >= ((int64_t)1 << 24)) {
    bale_set_error(BALE_ERR_ARG, "too many instances");
    return nullptr;
  }
  Device* d = device_state();
  if (!d) {
    bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
    return nullptr;
  }
  bale_batch* b = new bale_batch();
  b->device = d->id;
  for (int64_t i = 0; i < count; i++) {
    if (check_instance(inst[i]) != BALE_OK) {
      bale_set_error(BALE_ERR_ARG, "instance %lld: bad kind, size, pointer or dtype", (long long)i);
      delete b;
      return nullptr;
    }
    if (inst[i].n == 0) {
      bale_instance z = inst[i];
      z.cfg.strategy = BALE_STRAT_DIRECT;
      z.cfg.hot = 0;
      plan_add(b->plan, z, d);
      continue;
    }
    if (plan_add(b->plan, inst[i], d) != BALE_OK) {
      bale_set_error(BALE_ERR_ARG, "instance %lld: configuration does not fit", (long long)i);
      delete b;
      return nullptr;
    }
  }
  mark_shared_counts(b->plan);
  hipStream_t us = (hipStream_t)stream;
  int rc = plan_upload(b->plan, b->bufs, us);
  if (rc) {
    bale_set_error(rc == hipErrorOutOfMemory ? BALE_ERR_ALLOC : BALE_ERR_HIP, "batch upload: %s",
                   hipGetErrorString((hipError_t)rc));
    delete b;
    return nullptr;
  }
  set_batch_lds(b->plan);
  // capture on a private stream (the null stream cannot be captured)
  if (hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess ||
      hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    bale_set_error(BALE_ERR_HIP, "stream capture could not start");
    delete b;
    return nullptr;
  }
  plan_enqueue(b->plan, b->bufs, d, b->capture, (ErrWords*)b->bufs.err.p);
  if (hipStreamEndCapture(b->capture, &b->graph) != hipSuccess ||
      hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    bale_set_error(BALE_ERR_HIP, "graph capture failed");
    delete b;
    return nullptr;
  }
  return b;
}

int bale_batch_launch(bale_batch* b, void* stream) {
  if (!b || !b->exec) return bale_set_error(BALE_ERR_ARG, "null batch");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  hipStream_t s = (hipStream_t)stream;
  ErrWords* err = stream_err(d, s);
  if (!err) return BALE_ERR_ARG;
  int rc = (int)hipGraphLaunch(b->exec, s);
  if (rc) return bale_set_error(BALE_ERR_HIP, "graph launch: %s", hipGetErrorString((hipError_t)rc));
  // the graph counted into the batch's own record; report on this stream's
  hipLaunchKernelGGL(k_batch_err_move, dim3(1), dim3(64), 0, s, (ErrWords*)b->bufs.err.p, err);
  rc = (int)hipGetLastError();
  return rc ? bale_set_error(BALE_ERR_HIP, "batch error record: %s", hipGetErrorString((hipError_t)rc))
            : BALE_OK;
}

int bale_batch_launch_direct(bale_batch* b, void* stream) {
  if (!b) return bale_set_error(BALE_ERR_ARG, "null batch");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  ErrWords* err = stream_err(d, (hipStream_t)stream);
  if (!err) return BALE_ERR_ARG;
  plan_enqueue(b->plan, b->bufs, d, (hipStream_t)stream, err);
  int rc = (int)hipGetLastError();
  return rc ? bale_set_error(BALE_ERR_HIP, "batch launch: %s", hipGetErrorString((hipError_t)rc))
            : BALE_OK;
}

int bale_batch_stats(const bale_batch* b, int64_t* out, int nout) {
  if (!b || !out) return bale_set_error(BALE_ERR_ARG, "null argument");
  int64_t v[6] = {(int64_t)b->plan.a.size(), (int64_t)b->plan.c.size(), (int64_t)b->plan.d.size(),
                  (int64_t)b->plan.hot_list.size(), (int64_t)b->plan.bucket_list.size(),
                  (int64_t)b->plan.arena};
  for (int i = 0; i < nout && i < 6; i++) out[i] = v[i];
  return BALE_OK;
}

void bale_batch_destroy(bale_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  delete b;
}

int bale_batch_launch_each(const bale_instance* inst, int64_t count, void* stream) {
  if (count < 0 || (count > 0 && !inst)) return bale_set_error(BALE_ERR_ARG, "bad instance array");
  for (int64_t i = 0; i < count; i++) {
    const bale_instance& in = inst[i];
    if (check_instance(in) != BALE_OK)
      return bale_set_error(BALE_ERR_ARG, "instance %lld: bad kind, size, pointer or dtype", (long long)i);
    bale_config c = in.cfg;
    int rc;
    if (in.kind == 1) {
      // a gather from an LDS copy of the table exists only inside batches
      if (c.strategy == BALE_STRAT_LDS) c.strategy = BALE_STRAT_DIRECT;
      rc = bale_gather(in.tgt, in.index, in.n, (const int64_t*)in.table, in.table_len,
                       BALE_INDEX_PLAIN, &c, stream);
    } else {
      rc = bale_histo(in.index, in.n, in.table, in.table_len, in.dtype, BALE_INDEX_PLAIN, &c, stream);
    }
    if (rc) return rc;
  }
  return BALE_OK;
}

int bale_batch_launch_plain(const bale_instance* inst, int64_t count, void* stream) {
  if (count < 0 || (count > 0 && !inst)) return bale_set_error(BALE_ERR_ARG, "bad instance array");
  bale_config c;
  memset(&c, 0, sizeof c);
  c.strategy = BALE_STRAT_DIRECT;
  c.block = 256;
  c.unroll = 8;
  for (int64_t i = 0; i < count; i++) {
    const bale_instance& in = inst[i];
    if (check_instance(in) != BALE_OK)
      return bale_set_error(BALE_ERR_ARG, "instance %lld: bad kind, size, pointer or dtype", (long long)i);
    int rc = in.kind == 1
                 ? bale_gather(in.tgt, in.index, in.n, (const int64_t*)in.table, in.table_len,
                               BALE_INDEX_PLAIN, &c, stream)
                 : bale_histo(in.index, in.n, in.table, in.table_len, in.dtype, BALE_INDEX_PLAIN, &c,
                              stream);
    if (rc) return rc;
  }
  return BALE_OK;
}

int bale_tune_candidates(int kind, int64_t table_len, bale_config* out, int max) {
  if (kind != 0 && kind != 1) return bale_set_error(BALE_ERR_ARG, "bad instance kind %d", kind);
  if (table_len <= 0) return bale_set_error(BALE_ERR_ARG, "table_len must be positive");
  std::vector<bale_config> v;
  auto add = [&](int strat, int task_log2, int merge, int hot) {
    bale_config c;
    memset(&c, 0, sizeof c);
    c.strategy = strat;
    c.task_log2 = task_log2;
    c.merge = merge;
    c.hot = hot;
    v.push_back(c);
  };
  if (kind == 1) {
    for (int t : {14, 16, 18}) add(BALE_STRAT_DIRECT, t, 0, 0);
    if (table_len <= kLdsAWords / 2)
      for (int t : {14, 16, 18}) add(BALE_STRAT_LDS, t, 0, 0);
  } else {
    for (int t : {14, 16, 18}) add(BALE_STRAT_DIRECT, t, 0, 0);
    add(BALE_STRAT_DIRECT, 16, 1, 0);
    for (int t : {16, 18}) add(BALE_STRAT_DIRECT, t, 0, 1);
    if (table_len <= ((int64_t)1 << 15)) {
      for (int t : {14, 16, 18}) add(BALE_STRAT_LDS, t, 0, 0);
      add(BALE_STRAT_LDS, 16, 1, 0);
    }
    if (table_len >= ((int64_t)1 << 12))
      for (int t : {16, 18})
        for (int hot : {0, 1})
          for (int runs : {0, 1}) add(BALE_STRAT_BUCKET, t, runs, hot);
  }
  int n = (int)v.size();
  if (out)
    for (int i = 0; i < n && i < max; i++) out[i] = v[i];
  return n;
}

// Timed pairs per candidate (the median is kept; the reference's tuner also
// averages repeated runs, convey/tune.pl:33-100).
constexpr int kTunePairs = 3;

int bale_tune(bale_instance* inst, int64_t count, int64_t min_ops, float* costs, int max_cand,
              void* stream) {
  if (count < 0 || (count > 0 && !inst)) return bale_set_error(BALE_ERR_ARG, "bad instance array");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  hipStream_t s = (hipStream_t)stream;
  ErrWords* err = stream_err(d, s);
  if (!err) return BALE_ERR_ARG;
  if (min_ops <= 0) min_ops = (int64_t)1 << 22;
  Buffers bufs, bufs2;
  DevBuf scratch;
  hipEvent_t ev[4 * kTunePairs];
  for (hipEvent_t& e : ev) (void)hipEventCreate(&e);
  std::vector<bale_config> cand(64);
  int rc = BALE_OK;
  for (int64_t i = 0; i < count && rc == BALE_OK; i++) {
    bale_instance& in = inst[i];
    if (check_instance(in) != BALE_OK) {
      rc = bale_set_error(BALE_ERR_ARG, "instance %lld: bad kind, size, pointer or dtype", (long long)i);
      break;
    }
    int nc = bale_tune_candidates(in.kind, in.table_len, cand.data(), 64);
    if (costs)
      for (int k = 0; k < max_cand; k++) costs[i * max_cand + k] = -1.f;
    if (in.n == 0) continue;
    // results go to scratch so the caller's data is untouched
    size_t sbytes = in.kind == 1 ? (size_t)in.n * 8
                                 : (size_t)in.table_len * (in.dtype == BALE_I64 ? 8 : 4);
    if (scratch.ensure(sbytes)) {
      rc = bale_set_error(BALE_ERR_ALLOC, "tuner scratch");
      break;
    }
    int64_t R = std::max<int64_t>(1, std::min<int64_t>(256, ceil_div(min_ops, in.n)));
    float best = 3.4e38f;
    int bestk = 0;
    // The marginal cost of candidate t: in a large batch a strategy's phases
    // run anyway (launches, sketches, the bucketed plan's scans), so the cost
    // an instance adds is the difference between R and 2R copies, not R copies
    // alone.  Both plans are built, uploaded and warmed up once, then timed in
    // kTunePairs interleaved pairs; the cost is the median of the pairs'
    // (t(2R) - t(R)) / (R n), each floored at a tenth of the pair's average so
    // noise cannot make it 0.  < 0 when the plan does not take the candidate,
    // -2 on a failed run.
    auto marginal = [&](const bale_instance& t) -> float {
      Plan P1, P2;
      for (int64_t r = 0; r < R; r++)
        if (plan_add(P1, t, d) != BALE_OK) return -1.f;
      for (int64_t r = 0; r < 2 * R; r++)
        if (plan_add(P2, t, d) != BALE_OK) return -1.f;
      if (plan_upload(P1, bufs, s) || plan_upload(P2, bufs2, s)) {
        rc = bale_set_error(BALE_ERR_ALLOC, "tuner upload");
        return -2.f;
      }
      set_batch_lds(P1);
      set_batch_lds(P2);
      plan_enqueue(P1, bufs, d, s, err);  // warm-ups
      plan_enqueue(P2, bufs2, d, s, err);
      for (int k = 0; k < kTunePairs; k++) {
        (void)hipEventRecord(ev[4 * k], s);
        plan_enqueue(P1, bufs, d, s, err);
        (void)hipEventRecord(ev[4 * k + 1], s);
        (void)hipEventRecord(ev[4 * k + 2], s);
        plan_enqueue(P2, bufs2, d, s, err);
        (void)hipEventRecord(ev[4 * k + 3], s);
      }
      if (hipEventSynchronize(ev[4 * kTunePairs - 1]) != hipSuccess) {
        rc = bale_set_error(BALE_ERR_HIP, "tuner run failed");
        return -2.f;
      }
      float c[kTunePairs];
      for (int k = 0; k < kTunePairs; k++) {
        float m1 = 0.f, m2 = 0.f;
        (void)hipEventElapsedTime(&m1, ev[4 * k], ev[4 * k + 1]);
        (void)hipEventElapsedTime(&m2, ev[4 * k + 2], ev[4 * k + 3]);
        const float t1 = m1 * 1e-3f, t2 = m2 * 1e-3f;
        const float avg = t2 / (float)(2 * R * in.n);
        c[k] = std::max((t2 - t1) / (float)(R * in.n), 0.1f * avg);
      }
      std::sort(c, c + kTunePairs);
      return c[kTunePairs / 2];
    };
    for (int k = 0; k < nc; k++) {
      bale_instance t = in;
      t.cfg = cand[k];
      if (in.kind == 1) t.tgt = (int64_t*)scratch.p;
      else t.table = scratch.p;
      const float cost = marginal(t);
      if (cost == -2.f) break;
      if (cost < 0.f) continue;
      if (costs && k < max_cand) costs[i * max_cand + k] = cost;
      if (cost < best) {
        best = cost;
        bestk = k;
      }
    }
    in.cfg = cand[bestk];
  }
  for (hipEvent_t& e : ev) (void)hipEventDestroy(e);
  // the tuner's runs counted into scratch: clear the device error record
  int64_t bad = 0;
  (void)bale_sync_errors(stream, &bad);
  return rc;
}

}  // extern "C"
