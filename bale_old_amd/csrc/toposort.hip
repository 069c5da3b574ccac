// toposort.hip -- toposort of a morally upper-triangular matrix on one GPU
// (apps/topo_src/toposort_agi.upc:53-166, toposort_matrix_agi at one PE).
//
// The reference keeps, per row, the count of its remaining nonzeros and the
// sum of their columns; rows with one nonzero are queued.  Level by level it
// pops the rows queued before the level began, in queue order: the k-th row
// popped overall takes position n-1-k in both permutations (its remaining
// column is its column sum), and every row of that column loses one nonzero;
// a row whose count drops from 2 to 1 joins the queue, so the next level's
// order is the order of those events.
//
// Row counts and column sums live in one 64-bit word per row (sum above the
// count) when the largest of each fit together, so one atomic add updates
// both; otherwise in two arrays.
//
// Here a level runs as five launches:
//   k_topo_pivot    one wave per queued row: claim the positions, then per
//                   row s of the column: count and column sum decremented
//                   (a device atomic per nonzero)
//   k_topo_mark     the same walk: for a row s whose count ended at 1, each
//                   event's key(k, j) = (k + 1) << 32 | j (j = place in the
//                   column) kept as s's largest (a device atomic per event on
//                   such rows only)
//   k_topo_count    a row s joins the next level iff its count ended at 1 and
//                   the event is its last decrement of the level -- the one
//                   that took the count from 2 to 1 in the sequential order --
//                   i.e. trig[s] == key(k, j); counted per queued row (one
//                   wave each) and per tile of 16 queued rows
//   k_topo_scan     one workgroup: tile offsets, and the next level's queue
//                   range (a ring of level descriptors on the device)
//   k_topo_enqueue  the joining rows written in (k, j) order (ballot ranks)
// so the queue, and with it both permutations, equal the reference's.  The
// host launches levels in batches and reads one descriptor per batch; levels
// past the last one find an empty range and return at once.
#include <stdlib.h>

#include <algorithm>

#include "bale_internal.h"

namespace bale {

namespace tp {

constexpr int kRing = 4;        // level descriptors (each level reads one, writes the next)
constexpr int kBatch = 16;      // levels launched between host checks
constexpr int kWG = 256;        // k_topo_pivot
constexpr int kCWG = 1024;      // k_topo_count / k_topo_enqueue: one wave per queued row
constexpr int kTileE = kCWG / 64;  // queued rows per tile
constexpr int kInitRows = 1024;  // rows per workgroup in k_topo_init
constexpr int kQ0 = 1024 * 16;   // rows per tile in the first queue's compaction

struct Level {
  int64_t qs, qe;  // queue[qs, qe) is the level's rows
};

// err[0]: a column outside [0, n) met; err[1]: a column claimed twice (the
// matrix is not morally triangular); err[2]: levels with queued rows;
// err[3], err[4]: the largest row count and column sum.
// When both fit one 64-bit word (cb count bits below the sum), cs[r] holds
// (rowsum << cb) | rowcnt and one atomic add updates both.
struct Ws {
  int32_t* rowcnt;
  int64_t* rowsum;
  unsigned long long* cs;
  int cb;
  unsigned long long* trig;
  int64_t* queue;
  int64_t* qcol;
  int32_t* tcount;  // per queued row of the level (triggered rows)
  int32_t* tsum;    // per tile, then exclusive offsets
  Level* desc;      // [kRing]
  unsigned long long* err;  // [5]
};

__device__ inline uint64_t trig_key(int64_t k, int64_t j) {
  return ((uint64_t)(k + 1) << 32) | (uint64_t)j;
}

// rowcnt, rowsum and trig of rows [r0, r0 + kInitRows): one workgroup, the
// rows' offsets in LDS, each nonzero added to its row's LDS sum.
__global__ __launch_bounds__(1024) void k_topo_init(const int64_t* __restrict__ offset,
                                                    const int64_t* __restrict__ nonzero, int64_t n,
                                                    Ws w) {
  __shared__ int64_t so[kInitRows + 1];
  __shared__ unsigned long long ssum[kInitRows];
  const int64_t r0 = (int64_t)blockIdx.x * kInitRows;
  const int R = (int)std::min<int64_t>(kInitRows, n - r0);
  for (int i = threadIdx.x; i <= R; i += 1024) so[i] = offset[r0 + i];
  for (int i = threadIdx.x; i < R; i += 1024) ssum[i] = 0;
  __syncthreads();
  const int64_t p0 = so[0], p1 = so[R];
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 1024) {
    int lo = 0, hi = R;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (so[mid] <= p) lo = mid;
      else hi = mid;
    }
    atomicAdd(&ssum[lo], (unsigned long long)nonzero[p]);
  }
  __syncthreads();
  unsigned long long mc = 0, ms = 0;
  for (int i = threadIdx.x; i < R; i += 1024) {
    const int64_t cnt = so[i + 1] - so[i];
    w.rowcnt[r0 + i] = (int32_t)cnt;
    w.rowsum[r0 + i] = (int64_t)ssum[i];
    w.trig[r0 + i] = 0;
    mc = mc > (unsigned long long)cnt ? mc : (unsigned long long)cnt;
    ms = ms > ssum[i] ? ms : ssum[i];
  }
  if (mc) atomicMax(&w.err[3], mc);
  if (ms) atomicMax(&w.err[4], ms);
}

__global__ __launch_bounds__(256) void k_topo_pack(int64_t n, Ws w) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x)
    w.cs[r] = ((unsigned long long)w.rowsum[r] << w.cb) | (unsigned long long)(uint32_t)w.rowcnt[r];
}

// Flag ranks within a workgroup of 1024 threads, round by round.
__device__ inline uint32_t q0_rank(bool f, uint32_t& base, uint32_t* sh) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint64_t bal = __ballot(f);
  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
  __syncthreads();
  if (lane == 0) sh[wave] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (int v = 0; v < 16; v++) {
    const uint32_t c = sh[v];
    before += v < wave ? c : 0u;
    total += c;
  }
  const uint32_t r = base + before + below;
  base += total;
  return r;
}

// First queue: rows with one nonzero, in row order.  PLACE false: count per
// tile of kQ0 rows; true: write them at the tile's offset.
template <bool PLACE>
__global__ __launch_bounds__(1024) void k_topo_q0(int64_t n, Ws w) {
  __shared__ uint32_t sh[16];
  const int64_t base = (int64_t)blockIdx.x * kQ0;
  uint32_t rb = PLACE ? (uint32_t)w.tsum[blockIdx.x] : 0u;
  for (int i = 0; i < kQ0 / 1024; i++) {
    const int64_t r = base + (int64_t)i * 1024 + threadIdx.x;
    const bool f = r < n && w.rowcnt[r] == 1;
    const uint32_t pos = q0_rank(f, rb, sh);
    if (PLACE && f) w.queue[pos] = r;
  }
  if (!PLACE && threadIdx.x == 0) w.tsum[blockIdx.x] = (int32_t)rb;
}

// Exclusive prefix of tsum[0..nt) in place (one workgroup); the total opens
// the level range `into` = [start, start + total).
__device__ inline void scan_tiles_into(int32_t* tsum, int64_t nt, int64_t start, Level* into) {
  __shared__ int32_t part[1024];
  __shared__ int64_t carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nt; b0 += 1024) {
    const int64_t i = b0 + threadIdx.x;
    const int32_t v = i < nt ? tsum[i] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int32_t o = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += o;
      __syncthreads();
    }
    const int64_t carry = carry_s;
    if (i < nt) tsum[i] = (int32_t)(carry + part[threadIdx.x] - v);
    __syncthreads();
    if (threadIdx.x == 1023) carry_s = carry + part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    into->qs = start;
    into->qe = start + carry_s;
  }
}

__global__ __launch_bounds__(1024) void k_topo_q0_scan(int64_t nt, Ws w) {
  scan_tiles_into(w.tsum, nt, 0, &w.desc[0]);
}

// One wave per queued row k of the level: positions n-1-k, then the column's
// rows lose one nonzero each.
template <bool PK>
__global__ __launch_bounds__(kWG) void k_topo_pivot(int64_t level, int64_t n,
                                                    const int64_t* __restrict__ t_offset,
                                                    const int64_t* __restrict__ t_nonzero,
                                                    int64_t* __restrict__ rperm,
                                                    int64_t* __restrict__ cperm, Ws w) {
  const Level L = w.desc[level % kRing];
  const int lane = __lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * kWG + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * kWG) >> 6;
  for (int64_t k = L.qs + wid; k < L.qe; k += nw) {
    const int64_t r = w.queue[k];
    const int64_t c = PK ? (int64_t)(w.cs[r] >> w.cb) : w.rowsum[r];
    if ((uint64_t)c >= (uint64_t)n) {
      if (lane == 0) {
        atomicAdd(&w.err[0], 1ull);
        w.qcol[k] = -1;
      }
      continue;
    }
    if (lane == 0) {
      rperm[r] = n - 1 - k;
      if (atomicExch((unsigned long long*)&cperm[c], (unsigned long long)(n - 1 - k)) !=
          ~0ull)
        atomicAdd(&w.err[1], 1ull);
      w.qcol[k] = c;
    }
    const int64_t j0 = t_offset[c], j1 = t_offset[c + 1];
    const unsigned long long dec = ((unsigned long long)c << w.cb) + 1ull;
    for (int64_t j = j0 + lane; j < j1; j += 64) {
      const int64_t s = t_nonzero[j];
      if ((uint64_t)s >= (uint64_t)n) {
        atomicAdd(&w.err[0], 1ull);
        continue;
      }
      if (PK) {
        atomicAdd(&w.cs[s], ~dec + 1ull);  // minus (c << cb) + 1
      } else {
        atomicSub(&w.rowcnt[s], 1);
        atomicAdd((unsigned long long*)&w.rowsum[s], (unsigned long long)(-c));
      }
    }
  }
}

// One wave per queued row k of the level, after the pivots: every event (k, j)
// on a row whose count ended the level at 1 offers its key to the row's
// largest, so only those rows (about one per row joining the next level) take
// the atomic; a row that does not end at 1 never joins, whatever its key.
template <bool PK>
__global__ __launch_bounds__(kCWG) void k_topo_mark(int64_t level, int64_t n,
                                                    const int64_t* __restrict__ t_offset,
                                                    const int64_t* __restrict__ t_nonzero, Ws w) {
  const Level L = w.desc[level % kRing];
  const int64_t m = L.qe - L.qs;
  const int lane = __lane_id();
  const int64_t nw = (int64_t)gridDim.x * (kCWG / 64);
  for (int64_t e = (int64_t)blockIdx.x * (kCWG / 64) + (threadIdx.x >> 6); e < m; e += nw) {
    const int64_t k = L.qs + e;
    const int64_t c = w.qcol[k];
    if (c < 0) continue;
    const int64_t j0 = t_offset[c], j1 = t_offset[c + 1];
    for (int64_t j = j0 + lane; j < j1; j += 64) {
      const int64_t s = t_nonzero[j];
      if ((uint64_t)s >= (uint64_t)n) continue;
      const bool one = PK ? (w.cs[s] & ((1ull << w.cb) - 1)) == 1ull : w.rowcnt[s] == 1;
      if (one) atomicMax(&w.trig[s], (unsigned long long)trig_key(k, j - j0));
    }
  }
}

// Rows of queued row k's column that join the next level (a bit per lane of
// the chunk starting at j).
template <bool PK>
__device__ inline uint64_t joining(const Ws& w, int64_t k, int64_t j0, int64_t j, int64_t j1,
                                   const int64_t* __restrict__ t_nonzero, int64_t n,
                                   int64_t* s_out) {
  bool f = false;
  int64_t s = -1;
  if (j < j1) {
    s = t_nonzero[j];
    if ((uint64_t)s < (uint64_t)n) {
      const bool one = PK ? (w.cs[s] & ((1ull << w.cb) - 1)) == 1ull : w.rowcnt[s] == 1;
      f = one && w.trig[s] == (unsigned long long)trig_key(k, j - j0);
    }
  }
  *s_out = s;
  return __ballot(f);
}

// Per queued row (one wave each): rows joining the next level; per tile of
// kTileE queued rows: their sum.
template <bool PK>
__global__ __launch_bounds__(kCWG) void k_topo_count(int64_t level, int64_t n,
                                                     const int64_t* __restrict__ t_offset,
                                                     const int64_t* __restrict__ t_nonzero, Ws w) {
  __shared__ int32_t wsum[kCWG / 64];
  const Level L = w.desc[level % kRing];
  const int64_t m = L.qe - L.qs, nt = (m + kTileE - 1) / kTileE;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  for (int64_t t = blockIdx.x; t < nt; t += gridDim.x) {
    const int64_t e = t * kTileE + wave;
    int32_t cnt = 0;
    if (e < m) {
      const int64_t k = L.qs + e;
      const int64_t c = w.qcol[k];
      if (c >= 0) {
        const int64_t j0 = t_offset[c], j1 = t_offset[c + 1];
        for (int64_t j = j0; j < j1; j += 64) {
          int64_t s;
          cnt += __popcll(joining<PK>(w, k, j0, j + lane, j1, t_nonzero, n, &s));
        }
      }
      if (lane == 0) w.tcount[e] = cnt;
    }
    if (lane == 0) wsum[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t sum = 0;
      for (int v = 0; v < kCWG / 64; v++) sum += wsum[v];
      w.tsum[t] = sum;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void k_topo_scan(int64_t level, Ws w) {
  const Level L = w.desc[level % kRing];
  const int64_t m = L.qe - L.qs, nt = (m + kTileE - 1) / kTileE;
  scan_tiles_into(w.tsum, nt, L.qe, &w.desc[(level + 1) % kRing]);
  if (m > 0 && threadIdx.x == 0) w.err[2] = (unsigned long long)(level + 1);
}

// The joining rows, in (k, j) order, at queue[qe + tile offset + ...].
template <bool PK>
__global__ __launch_bounds__(kCWG) void k_topo_enqueue(int64_t level, int64_t n,
                                                       const int64_t* __restrict__ t_offset,
                                                       const int64_t* __restrict__ t_nonzero, Ws w) {
  __shared__ int32_t eoff[kTileE];
  const Level L = w.desc[level % kRing];
  const int64_t m = L.qe - L.qs, nt = (m + kTileE - 1) / kTileE;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  for (int64_t t = blockIdx.x; t < nt; t += gridDim.x) {
    if (wave == 0) {  // exclusive prefix of the tile's counts
      const int64_t e = t * kTileE + lane;
      const int32_t v = lane < kTileE && e < m ? w.tcount[e] : 0;
      int32_t incl = v;
      for (int d = 1; d < 64; d <<= 1) {
        const int32_t o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
      }
      if (lane < kTileE) eoff[lane] = incl - v;
    }
    __syncthreads();
    const int64_t e = t * kTileE + wave;
    if (e < m) {
      const int64_t k = L.qs + e;
      const int64_t c = w.qcol[k];
      if (c >= 0) {
        int64_t pos = L.qe + w.tsum[t] + eoff[wave];
        const int64_t j0 = t_offset[c], j1 = t_offset[c + 1];
        for (int64_t j = j0; j < j1; j += 64) {
          int64_t s;
          const uint64_t bal = joining<PK>(w, k, j0, j + lane, j1, t_nonzero, n, &s);
          const uint32_t below = __builtin_amdgcn_mbcnt_hi(
              (uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
          if (((bal >> lane) & 1) && pos + below < n) w.queue[pos + below] = s;
          pos += __popcll(bal);
        }
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_topo_fill(int64_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = -1;
}

__global__ void k_topo_zero5(unsigned long long* p) {
  if (threadIdx.x < 5) p[threadIdx.x] = 0;
}

}  // namespace tp

// Workspace bytes for n rows.
static size_t topo_ws_bytes(int64_t n) {
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  const size_t tiles = std::max<size_t>(ceil_div((int64_t)nn, tp::kTileE), ceil_div((int64_t)nn, tp::kQ0)) + 1;
  return up(nn * 4) + up(nn * 8) * 5 + up(nn * 4) + up(tiles * 4) + up(sizeof(tp::Level) * tp::kRing) + 256;
}

int launch_toposort(int64_t n, const int64_t* offset, const int64_t* nonzero, const int64_t* t_offset,
                    const int64_t* t_nonzero, int64_t* rperm, int64_t* cperm, Device* d, hipStream_t s,
                    int64_t* pivots, int64_t* levels, int64_t* bad_col, int64_t* twice) {
  using namespace tp;
  *pivots = *levels = *bad_col = *twice = 0;
  if (n <= 0) return 0;
  char* p = (char*)workspace(d, topo_ws_bytes(n), s);
  if (!p) return (int)hipErrorOutOfMemory;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t nn = (size_t)n;
  const size_t tiles = std::max<size_t>(ceil_div(n, kTileE), ceil_div(n, kQ0)) + 1;
  Ws w;
  w.rowcnt = (int32_t*)p; p += up(nn * 4);
  w.rowsum = (int64_t*)p; p += up(nn * 8);
  w.cs = (unsigned long long*)p; p += up(nn * 8);
  w.cb = 0;
  w.trig = (unsigned long long*)p; p += up(nn * 8);
  w.queue = (int64_t*)p; p += up(nn * 8);
  w.qcol = (int64_t*)p; p += up(nn * 8);
  w.tcount = (int32_t*)p; p += up(nn * 4);
  w.tsum = (int32_t*)p; p += up(tiles * 4);
  w.desc = (Level*)p; p += up(sizeof(Level) * kRing);
  w.err = (unsigned long long*)p;
  hipLaunchKernelGGL(k_topo_zero5, dim3(1), dim3(64), 0, s, w.err);
  const unsigned gf = (unsigned)std::min<int64_t>((int64_t)d->cus * 16, ceil_div(n, 256));
  hipLaunchKernelGGL(k_topo_fill, dim3(gf), dim3(256), 0, s, rperm, n);
  hipLaunchKernelGGL(k_topo_fill, dim3(gf), dim3(256), 0, s, cperm, n);
  hipLaunchKernelGGL(k_topo_init, dim3((unsigned)ceil_div(n, kInitRows)), dim3(1024), 0, s, offset,
                     nonzero, n, w);
  const unsigned nq0 = (unsigned)ceil_div(n, kQ0);
  hipLaunchKernelGGL(k_topo_q0<false>, dim3(nq0), dim3(1024), 0, s, n, w);
  hipLaunchKernelGGL(k_topo_q0_scan, dim3(1), dim3(1024), 0, s, (int64_t)nq0, w);
  hipLaunchKernelGGL(k_topo_q0<true>, dim3(nq0), dim3(1024), 0, s, n, w);
  // counts and column sums share one word when both fit
  unsigned long long mx[2] = {0, 0};
  int rc = (int)hipMemcpyAsync(mx, w.err + 3, sizeof mx, hipMemcpyDeviceToHost, s);
  if (!rc) rc = (int)hipStreamSynchronize(s);
  if (rc) return rc;
  auto bits = [](unsigned long long x) { int b = 0; while (x) { b++; x >>= 1; } return b; };
  const int cb = bits(mx[0]) + 1, sb = bits(mx[1]);
  // BALE_TOPOSORT_SEPARATE=1 keeps them apart regardless (it checks that path)
  const char* sep = getenv("BALE_TOPOSORT_SEPARATE");
  const bool pk = cb + sb <= 64 && !(sep && sep[0] == '1');
  if (pk) {
    w.cb = cb;
    hipLaunchKernelGGL(k_topo_pack, dim3(gf), dim3(256), 0, s, n, w);
  }
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * 4, ceil_div(n, 4)));
  const unsigned gc = (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * 2, ceil_div(n, kTileE)));
  Level h{0, 0};
  for (int64_t L0 = 0;; L0 += kBatch) {
    // a level queues at least one row, so there are at most n of them
    if (L0 > n + kBatch) return bale_set_error(BALE_ERR_ARG, "toposort: more levels than rows");
    for (int64_t l = L0; l < L0 + kBatch; l++) {
      if (pk) {
        hipLaunchKernelGGL(k_topo_pivot<true>, dim3(g), dim3(kWG), 0, s, l, n, t_offset, t_nonzero, rperm, cperm, w);
        hipLaunchKernelGGL(k_topo_mark<true>, dim3(gc), dim3(kCWG), 0, s, l, n, t_offset, t_nonzero, w);
        hipLaunchKernelGGL(k_topo_count<true>, dim3(gc), dim3(kCWG), 0, s, l, n, t_offset, t_nonzero, w);
        hipLaunchKernelGGL(k_topo_scan, dim3(1), dim3(1024), 0, s, l, w);
        hipLaunchKernelGGL(k_topo_enqueue<true>, dim3(gc), dim3(kCWG), 0, s, l, n, t_offset, t_nonzero, w);
      } else {
        hipLaunchKernelGGL(k_topo_pivot<false>, dim3(g), dim3(kWG), 0, s, l, n, t_offset, t_nonzero, rperm, cperm, w);
        hipLaunchKernelGGL(k_topo_mark<false>, dim3(gc), dim3(kCWG), 0, s, l, n, t_offset, t_nonzero, w);
        hipLaunchKernelGGL(k_topo_count<false>, dim3(gc), dim3(kCWG), 0, s, l, n, t_offset, t_nonzero, w);
        hipLaunchKernelGGL(k_topo_scan, dim3(1), dim3(1024), 0, s, l, w);
        hipLaunchKernelGGL(k_topo_enqueue<false>, dim3(gc), dim3(kCWG), 0, s, l, n, t_offset, t_nonzero, w);
      }
    }
    rc = (int)hipGetLastError();
    if (rc) return rc;
    rc = (int)hipMemcpyAsync(&h, &w.desc[(L0 + kBatch) % kRing], sizeof h, hipMemcpyDeviceToHost, s);
    if (!rc) rc = (int)hipStreamSynchronize(s);
    if (rc) return rc;
    if (h.qs == h.qe) {  // an empty level: every later one is empty too
      *pivots = h.qe;
      break;
    }
  }
  unsigned long long e[3] = {0, 0, 0};
  rc = (int)hipMemcpy(e, w.err, sizeof e, hipMemcpyDeviceToHost);
  *bad_col = (int64_t)e[0];
  *twice = (int64_t)e[1];
  *levels = (int64_t)e[2];
  return rc;
}

}  // namespace bale
