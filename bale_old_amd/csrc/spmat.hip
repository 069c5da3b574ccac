// spmat.hip -- sparse matrices (CSR) on one GPU: Erdos-Renyi generation,
// transpose (spmat/spmat_agi.upc:176-236), row and column permutation
// (spmat/spmat_agi.upc:119-169) and the triangularity check
// (spmat/spmat_utils.upc:377-453).
//
// A matrix is offset[numrows + 1] and nonzero[nnz] (int64, as in sparsemat_t,
// spmat/spmat.h:73-86, at one PE): row i's column indices are
// nonzero[offset[i] .. offset[i+1]).
//
// Transpose.  At one PE the reference histograms the column counts, turns
// them into row starts of the transpose, and walks A's rows in order, putting
// row r at the next free slot of each of its columns, so every row of the
// transpose lists its entries in increasing original row.  That output is the
// stable sort of the (column, row) pairs by column, which is computed here as
// least-significant-digit radix passes of up to 8 bits over the column, each a
// stable partition by digit (deliver.h); the row starts come from the
// histogram of the columns (the library's histogram) and an int64 scan.
//
// Permute.  rperm is rperminv inverted; row i of the result is row rperm[i]
// of A with every column j replaced by cperminv[j] (a gather).
#include <algorithm>

#include "bale_internal.h"
#include "deliver.h"
#include "partition.h"

namespace bale {

namespace sm {

constexpr int kThreads = 1024;
constexpr int kPer = 8;                     // scan elements per thread
constexpr int kScanTile = kThreads * kPer;  // 8192
constexpr int kRowsPerGroup = 1024;         // rows per workgroup in row-tiled kernels

// ---- int64 inclusive scan, in place ---------------------------------------------
// Inclusive prefix of one value per thread over a workgroup of kThreads;
// *total gets the workgroup's sum.
__device__ inline int64_t block_incl_scan(int64_t v, int64_t* sh, int64_t* total) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t o = __shfl_up(v, d);
    if (lane >= d) v += o;
  }
  __syncthreads();
  if (lane == 63) sh[wave] = v;
  __syncthreads();
  int64_t before = 0, all = 0;
  for (int w = 0; w < kThreads / 64; w++) {
    const int64_t t = sh[w];
    before += w < wave ? t : 0;
    all += t;
  }
  *total = all;
  return v + before;
}

// sums[t] = sum of a[t * kScanTile ..) (thread tid owns kPer consecutive values)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scan_sums(const T* __restrict__ a, int64_t n,
                                                        int64_t* __restrict__ sums) {
  __shared__ int64_t sh[kThreads / 64];
  const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kPer;
  int64_t v = 0;
#pragma unroll
  for (int k = 0; k < kPer; k++)
    if (b + k < n) v += a[b + k];
  int64_t total;
  (void)block_incl_scan(v, sh, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// sums[0..nt) -> exclusive prefix, in place (one workgroup, rounds of kThreads)
__global__ __launch_bounds__(kThreads) void k_scan_top(int64_t* __restrict__ sums, int64_t nt) {
  __shared__ int64_t sh[kThreads / 64];
  int64_t carry = 0;
  for (int64_t base = 0; base < nt; base += kThreads) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < nt ? sums[i] : 0;
    int64_t total;
    const int64_t incl = block_incl_scan(v, sh, &total);
    if (i < nt) sums[i] = carry + incl - v;
    carry += total;
  }
}

// a[i] = sum of a[0..i] (inclusive), with tile t starting from sums[t]
__global__ __launch_bounds__(kThreads) void k_scan_apply(int64_t* __restrict__ a, int64_t n,
                                                         const int64_t* __restrict__ sums) {
  __shared__ int64_t sh[kThreads / 64];
  const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kPer;
  int64_t x[kPer], v = 0;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    x[k] = b + k < n ? a[b + k] : 0;
    v += x[k];
  }
  int64_t total;
  int64_t run = block_incl_scan(v, sh, &total) - v + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    run += x[k];
    if (b + k < n) a[b + k] = run;
  }
}

// out[i] = sum of in[0..i] (inclusive) for 32-bit counts, into int64
__global__ __launch_bounds__(kThreads) void k_scan_apply32(const int32_t* __restrict__ in, int64_t n,
                                                           const int64_t* __restrict__ sums,
                                                           int64_t* __restrict__ out) {
  __shared__ int64_t sh[kThreads / 64];
  const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kPer;
  int64_t x[kPer], v = 0;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    x[k] = b + k < n ? in[b + k] : 0;
    v += x[k];
  }
  int64_t total;
  int64_t run = block_incl_scan(v, sh, &total) - v + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    run += x[k];
    if (b + k < n) out[b + k] = run;
  }
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ __launch_bounds__(256) void k_fill_i64(int64_t* __restrict__ p, int64_t n, int64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- Erdos-Renyi rows -------------------------------------------------------------
// log(x), x in (0, 1]: x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
// s = (m - 1)/(m + 1), by its series to s^25; plus e ln 2.  Plain correctly
// rounded operations only (no contraction), so the host computes the same bits.
__device__ inline double det_log(double x) {
#pragma clang fp contract(off)
  uint64_t b = (uint64_t)__double_as_longlong(x);
  int64_t e = (int64_t)((b >> 52) & 0x7FF) - 1022;
  b = (b & 0x000FFFFFFFFFFFFFULL) | (0x3FEULL << 52);
  double m = __longlong_as_double((long long)b);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0), z = s * s;
  double t = 1.0 / 25.0;
  for (int k = 11; k >= 0; k--) t = t * z + 1.0 / (double)(2 * k + 1);
  return (double)e * 0.69314718055994530942 + 2.0 * s * t;
}

__device__ inline uint64_t er_key(uint64_t seed, int64_t row, int half) {
  return mix64(seed + kGolden + kGolden * (uint64_t)(2 * row + half + 1));
}

// Nonzeros of one row in columns [lo, hi): geometric gaps with log(1 - p) = D.
// out == nullptr counts only.
__device__ inline int64_t er_half(uint64_t key, int64_t lo, int64_t hi, double p, double D,
                                  int64_t* out) {
#pragma clang fp contract(off)
  if (lo >= hi || !(p > 0.0)) return 0;
  int64_t cnt = 0;
  if (p >= 1.0) {
    for (int64_t c = lo; c < hi; c++) {
      if (out) out[cnt] = c;
      cnt++;
    }
    return cnt;
  }
  int64_t col = lo - 1;
  for (uint64_t i = 0;; i++) {
    const uint64_t x = stream_at(key, i);
    const double u = (double)((x >> 11) + 1) * 0x1p-53;
    const double g = floor(det_log(u) / D);
    if (g >= (double)(hi - col)) break;
    col += 1 + (int64_t)g;
    if (col >= hi) break;
    if (out) out[cnt] = col;
    cnt++;
  }
  return cnt;
}

// FILL false: len[r] = row r's nonzeros (len = offset + 1); true: write them.
template <bool FILL>
__global__ __launch_bounds__(256) void k_er_rows(int64_t n, double p, int unit_diag, int mode,
                                                 uint64_t seed, int64_t* __restrict__ offset,
                                                 int64_t* __restrict__ nonzero) {
#pragma clang fp contract(off)
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double D = p > 0.0 && p < 1.0 ? det_log(1.0 - p) : -1.0;
  int64_t* out = FILL ? nonzero + offset[r] : nullptr;
  int64_t cnt = 0;
  if (mode == 1 || mode == 3) cnt += er_half(er_key(seed, r, 0), 0, r, p, D, out);
  if (unit_diag) {
    if (FILL) out[cnt] = r;
    cnt++;
  }
  if (mode == 2 || mode == 3) cnt += er_half(er_key(seed, r, 1), r + 1, n, p, D, FILL ? out + cnt : nullptr);
  if (!FILL) offset[r + 1] = cnt;
}

// ---- row-tiled expansion: one workgroup per kRowsPerGroup rows ----------------------
// Largest r in [0, R) with so[r] <= p (so ascending, so[0] <= p < so[R]).
__device__ inline int row_of(const int64_t* so, int R, int64_t p) {
  int lo = 0, hi = R;  // so[lo] <= p < so[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (so[mid] <= p) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---- transpose radix passes -----------------------------------------------------------
// A pass is a stable partition of (column, row) items by one digit of the
// column.  A workgroup walks its stream in tiles of kRdxTile items; each wave
// ranks 64 items per round with ballots (one per digit bit) on top of its
// running count per digit, one block-wide scan over the (digit, wave) counts
// gives every item its slot in the tile's digit order, the tile is staged in
// LDS in that order and each digit's run is written to consecutive output
// slots from the digit's cursor.  8192-item tiles give runs of 32 items
// (256 bytes) per digit on average.
//
// IN 0: items (column | row << 32); IN 1: the matrix's own arrays (column in
//   the low word of nonzero[]; rows from offset[]: each tile marks the first
//   position of every row it meets with an LDS max and a max-scan over the
//   tile's positions gives every position its row).
// OUT 0: items; OUT 1: the rows, as int64 (the transpose's nonzero array).
// CUR 0: the stream is chunk blockIdx.x of the input; cursors from a count
//   pass over sub-chunks of chunk / rstride items (start[p] + rows[c rstride][p]:
//   the items of digit p before the chunk's first sub-chunk).
// CUR 1 (any pass after the first): the input is sorted by the low `shift`
//   bits of the column, so its items with low bits x form group x, in row
//   order.  Workgroup c takes groups [x0, x1): their items are contiguous, and
//   its items of digit y are those whose low shift + nbits bits lie in
//   [y 2^shift + x0, y 2^shift + x1), which the stable pass places in order
//   from P[y 2^shift + x0], P[v] being the number of items whose low
//   shift + nbits bits are below v (P indexed up to plen, clamped).  So the
//   stream and the cursors follow from P with no count pass.  The last pass
//   uses the transpose's row starts as P; a middle pass a prefix derived from
//   them (k_low_prefix).
constexpr int kRdxTile = 8192;  // items per tile
// Workgroup shapes: 512 threads x 16 items (75 KiB of LDS, two workgroups per
// CU) when there are enough workgroups to fill the CUs twice, else 1024 x 8
// (a pass by column groups with fewer groups than that: at 2^24 columns the
// middle pass has 256).
template <int NT>
struct Rdx {
  static constexpr int U = kRdxTile / NT;
  static constexpr int NW = NT / 64;
  // words before the stage: wcnt, tstart, delta, cursor, wsum, srange (16-byte aligned)
  static constexpr int head = (NW * 256 + 257 + 256 + 256 + NW + 3 + 3) & ~3;
  static constexpr size_t lds = (size_t)head * 4 + (size_t)kRdxTile * 8;
};
constexpr int kRdxPerCu = 2;  // 512-thread delivery workgroups resident per CU
// count workgroups per delivery chunk
constexpr int kCountSplit = 2;

__device__ inline int64_t clamp_col(int64_t c, int64_t len) { return c < len ? c : len; }

// P[v], v in [0, M]: nonzeros whose column modulo M is below v, from the row
// starts of the transpose (M a power of two)
__global__ __launch_bounds__(256) void k_low_prefix(const int64_t* __restrict__ t_offset, int64_t numcols,
                                                    int64_t M, int64_t* __restrict__ P) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > M) return;
  int64_t sum = 0;
#pragma unroll 8
  for (int64_t c0 = 0; c0 < numcols; c0 += M)
    sum += t_offset[clamp_col(c0 + v, numcols)] - t_offset[c0];
  P[v] = sum;
}

template <int IN, int OUT, int CUR, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_rdx(const uint64_t* __restrict__ items,
                                                     const int64_t* __restrict__ offset, int64_t numrows,
                                                     int64_t n,
                                                     int npe, int nbits, int shift, int64_t chunk,
                                                     const uint32_t* __restrict__ rows, int rstride,
                                                     const uint32_t* __restrict__ start,
                                                     const int64_t* __restrict__ P, int64_t plen,
                                                     int64_t ngroups,
                                                     uint64_t* __restrict__ out,
                                                     int64_t* __restrict__ out_rows) {
  constexpr int NW = Rdx<NT>::NW, U = Rdx<NT>::U, T = kRdxTile, K = NW * 256 / NT;
  extern __shared__ __align__(16) uint32_t lds[];
  uint32_t* wcnt = lds;                 // [NW * 256] per wave and digit: count, then first staged slot
  uint32_t* tstart = wcnt + NW * 256;   // [257] first staged slot per digit
  uint32_t* delta = tstart + 257;       // [256] output slot of staged slot q is delta[p] + q
  uint32_t* cursor = delta + 256;       // [256] next output slot per digit
  uint32_t* wsum = cursor + 256;        // [NW] wave totals of the scan
  uint32_t* srange = wsum + NW;         // [3] stream bounds (CUR 1), first row (IN 1)
  uint64_t* stage = (uint64_t*)(lds + Rdx<NT>::head);  // [T]
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const uint32_t dmask = (uint32_t)(npe - 1);
  int64_t lo, hi;
  if constexpr (CUR == 0) {
    lo = (int64_t)blockIdx.x * chunk;
    hi = lo + chunk < n ? lo + chunk : n;
    for (int p = tid; p < npe; p += NT)
      cursor[p] = start[p] + rows[(size_t)blockIdx.x * rstride * npe + p];
  } else {
    const int64_t x0 = (int64_t)blockIdx.x * ngroups / gridDim.x;
    const int64_t x1 = ((int64_t)blockIdx.x + 1) * ngroups / gridDim.x;
    if (tid < 2) srange[tid] = 0;
    __syncthreads();
    for (int p = tid; p < npe; p += NT) {
      const int64_t c0 = (int64_t)p << shift;
      const int64_t base = P[clamp_col(c0, plen)];
      const int64_t a = P[clamp_col(c0 + x0, plen)], b = P[clamp_col(c0 + x1, plen)];
      cursor[p] = (uint32_t)a;
      atomicAdd(&srange[0], (uint32_t)(a - base));
      atomicAdd(&srange[1], (uint32_t)(b - base));
    }
    __syncthreads();
    lo = srange[0];
    hi = srange[1] < n ? (int64_t)srange[1] : n;
  }
  const uint64_t below = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  const int64_t mine = (int64_t)wave * (64 * U) + lane;
  uint32_t* wc = wcnt + wave * 256;
  auto load = [&](int64_t e) -> uint64_t {
    if constexpr (IN == 0) return items[e];
    else return (uint64_t)(uint32_t)items[e];  // the row is added below
  };
  // IN 1: the row holding the stream's first position, by a 64-way search of
  // offset[] (one wave: each step probes 64 rows and keeps the bracket)
  uint32_t ra = 0;
  if constexpr (IN == 1) {
    if (wave == 0) {
      int64_t a = 0, b = numrows;  // offset[a] <= lo < offset[b] (b = numrows: offset is nnz > lo)
      while (b - a > 1) {
        const int64_t r = a + (b - a) * (lane + 1) / 65;
        const bool le = r > a && offset[r] <= lo;
        const uint64_t m = __ballot(le), nm = __ballot(!le && r > a);
        const int64_t na = m ? __shfl(r, 63 - __clzll(m)) : a;
        const int64_t nb = nm ? __shfl(r, __builtin_ctzll(nm)) : b;
        a = na;
        b = nb;
      }
      if (lane == 0) srange[2] = (uint32_t)a;
    }
    __syncthreads();
    ra = srange[2];
  }
  for (int64_t base = lo; base < hi; base += T) {
    // this tile's items; the other workgroup on the CU hides their latency
    // (IN 1 loads them once the rows are known, so the row pass runs with
    // the registers free)
    uint64_t v[U];
    if constexpr (IN == 0) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t e = base + mine + (int64_t)u * 64;
        v[u] = e < hi ? load(e) : 0;
      }
    }
    if constexpr (IN == 1) {
      // rows of the tile's positions, in the stage area (free until staging)
      constexpr int EPT = T / NT;
      uint32_t* rid = (uint32_t*)stage;
      for (int k = tid; k < T / 4; k += NT) ((u32x4*)rid)[k] = u32x4{0u, 0u, 0u, 0u};
      __syncthreads();
      // every row that starts before the tile's end marks its first position
      // in the tile (position 0 for the row running into the tile); rows that
      // share a start (empty ones) leave the largest, the one holding it
      for (int64_t r = (int64_t)ra + tid; r < numrows; r += NT) {
        const int64_t o = offset[r];
        if (o >= base + T) break;
        atomicMax(&rid[o > base ? (int)(o - base) : 0], (uint32_t)r);
      }
      __syncthreads();
      // inclusive max-scan: thread t owns positions [t EPT, (t + 1) EPT)
      u32x4* rq = (u32x4*)(rid + tid * EPT);
      uint32_t m = 0;
#pragma unroll
      for (int k = 0; k < EPT / 4; k++) {
        const u32x4 x = rq[k];
        m = max(m, max(max(x.x, x.y), max(x.z, x.w)));
      }
      uint32_t incl = m;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl = max(incl, o);
      }
      if (lane == 63) wsum[wave] = incl;
      __syncthreads();
      uint32_t wpre = lane < wave && lane < NW ? wsum[lane] : 0u;
#pragma unroll
      for (int off = 1; off < NW; off <<= 1) wpre = max(wpre, __shfl_xor(wpre, off));
      wpre = __shfl(wpre, 0);
      const uint32_t up = __shfl_up(incl, 1);
      uint32_t run = max(wpre, lane ? up : 0u);
#pragma unroll
      for (int k = 0; k < EPT / 4; k++) {
        u32x4 x = rq[k];
        x.x = run = max(run, x.x);
        x.y = run = max(run, x.y);
        x.z = run = max(run, x.z);
        x.w = run = max(run, x.w);
        rq[k] = x;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t e = base + mine + (int64_t)u * 64;
        v[u] = e < hi ? load(e) | (uint64_t)rid[mine + u * 64] << 32 : 0;
      }
      ra = rid[T - 1];  // the next tile starts in this row or a later one
    }
    for (int k = tid; k < NW * 256; k += NT) wcnt[k] = 0;
    __syncthreads();
    // ranks within the wave: lanes with the same digit (one ballot per bit),
    // after the wave's earlier items of that digit
    uint32_t rank[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool has = base + mine + (int64_t)u * 64 < hi;
      const uint32_t du = ((uint32_t)v[u] >> shift) & dmask;
      uint64_t peers = __ballot(has);
      for (int bit = 0; bit < nbits; bit++) {
        const bool one = (du >> bit) & 1;
        const uint64_t b = __ballot(one);
        peers &= one ? b : ~b;
      }
      const uint32_t r = (uint32_t)__popcll(peers & below);
      const uint32_t before = has ? wc[du] : 0u;  // every lane reads before the leader adds
      rank[u] = before + r;
      if (has && r == 0) wc[du] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // exclusive scan of the counts in (digit, wave) order: thread t holds
    // entries 4t .. 4t+3 (digit e / NW, wave e % NW); wave scan, then the
    // wave totals (lanes 0..NW-1 reduce them, lane 0's value is broadcast)
    {
      uint32_t cc[K], s = 0;
      const int e0 = tid * K;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int e = e0 + k, p = e / NW, w = e % NW;
        cc[k] = p < npe ? wcnt[w * 256 + p] : 0u;
        s += cc[k];
      }
      uint32_t incl = s;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
      }
      if (lane == 63) wsum[wave] = incl;
      __syncthreads();
      const uint32_t wt = lane < NW ? wsum[lane] : 0u;
      uint32_t wpre = lane < wave ? wt : 0u, all = wt;
#pragma unroll
      for (int off = 1; off < NW; off <<= 1) {
        wpre += __shfl_xor(wpre, off);
        all += __shfl_xor(all, off);
      }
      wpre = __shfl(wpre, 0);
      all = __shfl(all, 0);
      uint32_t run = wpre + incl - s;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int e = e0 + k, p = e / NW, w = e % NW;
        if (p < npe) {
          if (w == 0) tstart[p] = run;
          wcnt[w * 256 + p] = run;
        }
        run += cc[k];
      }
      if (tid == 0) tstart[npe] = all;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++)
      if (base + mine + (int64_t)u * 64 < hi) stage[wc[((uint32_t)v[u] >> shift) & dmask] + rank[u]] = v[u];
    for (int p = tid; p < npe; p += NT) delta[p] = cursor[p] - tstart[p];
    __syncthreads();
    // staged order is digit order, input order within a digit
    const uint32_t ntile = tstart[npe];
    for (uint32_t q = tid; q < ntile; q += NT) {
      const uint64_t w = stage[q];
      const uint32_t g = delta[((uint32_t)w >> shift) & dmask] + q;
      if (g >= (uint64_t)n) continue;  // only an input the histogram rejected lands here
      if constexpr (OUT == 0) out[g] = w;
      else out_rows[g] = (int64_t)(w >> 32);
    }
    __syncthreads();
    for (int p = tid; p < npe; p += NT) cursor[p] = delta[p] + tstart[p + 1];
  }
}

// Per-chunk digit counts: rows[c * B + d] = items of chunk c with digit d
// ((word 0 >> shift) & (B - 1)); one LDS counter set per wave.
__global__ __launch_bounds__(kThreads) void k_rx_count(const uint32_t* __restrict__ items, int64_t n,
                                                       int shift, int B, int64_t chunk,
                                                       uint32_t* __restrict__ rows) {
  __shared__ uint32_t cnt[(kThreads / 64) * 256];
  const int wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < (kThreads / 64) * B; i += kThreads) cnt[i] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = std::min<int64_t>(lo + chunk, n);
  const uint32_t mask = (uint32_t)B - 1;
  uint32_t* mine = cnt + wave * B;
  // two items per 16-byte load (chunks start at even items)
  for (int64_t e = lo + 2 * (int64_t)threadIdx.x; e < hi; e += 2 * kThreads) {
    if (e + 1 < hi) {
      const u32x4 v = *(const u32x4*)(items + 2 * e);
      atomicAdd(&mine[(v.x >> shift) & mask], 1u);
      atomicAdd(&mine[(v.z >> shift) & mask], 1u);
    } else {
      atomicAdd(&mine[(items[2 * e] >> shift) & mask], 1u);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < B; d += kThreads) {
    uint32_t s = 0;
    for (int w = 0; w < kThreads / 64; w++) s += cnt[w * B + d];
    rows[(size_t)blockIdx.x * B + d] = s;
  }
}

// ---- permutation ---------------------------------------------------------------------
// A's row i becomes result row rperminv[i]: its length and start in A go there
// as one 16-byte pair (A's offsets are read in order, and the result rows are
// written at random only once); values outside [0, n) are range errors.
__global__ __launch_bounds__(256) void k_perm_pairs(const int64_t* __restrict__ offset,
                                                    const int64_t* __restrict__ rperminv, int64_t n,
                                                    i64x2* __restrict__ pairs, ErrWords* err) {
  unsigned bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rperminv[i];
    if ((uint64_t)r < (uint64_t)n) {
      const int64_t a = offset[i];
      i64x2 v;
      v.x = offset[i + 1] - a;
      v.y = a;
      pairs[r] = v;
    } else {
      bad++;
    }
  }
  flush_errors(err, bad, 0);
}

// The pairs split into the result's row lengths (scanned next) and the source
// starts the copy reads in order; a result row that no row of A named keeps
// the fill's start of -1: length 0 and a range error.
__global__ __launch_bounds__(256) void k_perm_split(const i64x2* __restrict__ pairs, int64_t n,
                                                    int64_t* __restrict__ len, int64_t* __restrict__ src,
                                                    ErrWords* err) {
  unsigned bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const i64x2 v = pairs[i];
    const bool named = v.y >= 0;
    len[i] = named ? v.x : 0;
    src[i] = named ? v.y : 0;
    bad += !named;
  }
  flush_errors(err, bad, 0);
}

// Row i of the result = row rperm[i] of A (starting at srcs[i] in A), columns
// mapped by cperminv (MAP), or A's own columns (the map then runs as the
// library's gather).
template <bool MAP>
__global__ __launch_bounds__(kThreads) void k_perm_fill(const int64_t* __restrict__ srcs,
                                                        const int64_t* __restrict__ nonzero,
                                                        const int64_t* __restrict__ cperminv,
                                                        int64_t numrows, int64_t numcols,
                                                        const int64_t* __restrict__ p_offset,
                                                        int64_t nnz, int64_t* __restrict__ p_nonzero,
                                                        ErrWords* err) {
  __shared__ int64_t so[kRowsPerGroup + 1];
  __shared__ int64_t src[kRowsPerGroup];
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerGroup;
  const int R = (int)std::min<int64_t>(kRowsPerGroup, numrows - r0);
  for (int i = threadIdx.x; i <= R; i += kThreads) so[i] = p_offset[r0 + i];
  for (int i = threadIdx.x; i < R; i += kThreads) src[i] = srcs[r0 + i];
  __syncthreads();
  const int64_t p0 = so[0], p1 = std::min<int64_t>(so[R], nnz);
  unsigned bad = 0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
    const int r = row_of(so, R, p);
    const int64_t c = nonzero[src[r] + (p - so[r])];
    if constexpr (!MAP) {
      p_nonzero[p] = c;  // the gather counts a column out of range
    } else if ((uint64_t)c < (uint64_t)numcols) {
      p_nonzero[p] = cperminv[c];
    } else {
      p_nonzero[p] = -1;
      bad++;
    }
  }
  if constexpr (MAP) flush_errors(err, bad, 0);
}

// ---- triangularity --------------------------------------------------------------------
// bad[0] += entries on the wrong side of the diagonal, bad[1] += rows without
// their diagonal entry (UPPER: columns below the row are wrong).
template <bool UPPER>
__global__ __launch_bounds__(256) void k_triangular(const int64_t* __restrict__ offset,
                                                    const int64_t* __restrict__ nonzero,
                                                    int64_t numrows,
                                                    unsigned long long* __restrict__ bad) {
  unsigned long long wrong = 0, nopivot = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < numrows;
       r += (int64_t)gridDim.x * blockDim.x) {
    bool pivot = false;
    for (int64_t j = offset[r]; j < offset[r + 1]; j++) {
      const int64_t c = nonzero[j];
      if (UPPER ? c < r : c > r) wrong++;
      else if (c == r) pivot = true;
    }
    nopivot += !pivot;
  }
  if (wrong) atomicAdd(bad, wrong);
  if (nopivot) atomicAdd(bad + 1, nopivot);
}

// ---- comparison ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_count_diff(const int64_t* __restrict__ a,
                                                    const int64_t* __restrict__ b, int64_t n,
                                                    unsigned long long* __restrict__ ndiff) {
  unsigned long long d = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d += a[i] != b[i];
  if (d) atomicAdd(ndiff, d);
}

inline unsigned grid_for(int64_t n, int threads, const Device* d) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * 16, ceil_div(n, threads)));
}

}  // namespace sm

// In-place inclusive scan of a[0..n); tmp holds ceil(n / 8192) values.
static int scan_incl(int64_t* a, int64_t n, int64_t* tmp, hipStream_t s) {
  using namespace sm;
  if (n <= 0) return 0;
  const int64_t nt = ceil_div(n, kScanTile);
  hipLaunchKernelGGL(k_scan_sums<int64_t>, dim3((unsigned)nt), dim3(kThreads), 0, s, (const int64_t*)a, n, tmp);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kThreads), 0, s, tmp, nt);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(kThreads), 0, s, a, n, (const int64_t*)tmp);
  return (int)hipGetLastError();
}
// out[i] = sum of in[0..i] for int32 counts; tmp as for scan_incl.
static int scan_incl32(const int32_t* in, int64_t n, int64_t* out, int64_t* tmp, hipStream_t s) {
  using namespace sm;
  if (n <= 0) return 0;
  const int64_t nt = ceil_div(n, kScanTile);
  hipLaunchKernelGGL(k_scan_sums<int32_t>, dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kThreads), 0, s, tmp, nt);
  hipLaunchKernelGGL(k_scan_apply32, dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, (const int64_t*)tmp, out);
  return (int)hipGetLastError();
}
static size_t scan_tmp_bytes(int64_t n) {
  return (size_t)((ceil_div(std::max<int64_t>(n, 1), sm::kScanTile) * 8 + 255) & ~(int64_t)255);
}

int launch_spmat_gen_er(int64_t n, double p, int unit_diag, int mode, uint64_t seed,
                        int64_t* offset, int64_t* nonzero, Device* d, hipStream_t s) {
  using namespace sm;
  if (n <= 0) {
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, offset, (int64_t)1, (int64_t)0);
    return (int)hipGetLastError();
  }
  const unsigned g = (unsigned)ceil_div(n, 256);
  if (!nonzero) {
    int64_t* tmp = (int64_t*)workspace(d, scan_tmp_bytes(n), s);
    if (!tmp) return (int)hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, offset, (int64_t)1, (int64_t)0);
    hipLaunchKernelGGL(k_er_rows<false>, dim3(g), dim3(256), 0, s, n, p, unit_diag, mode, seed,
                       offset, (int64_t*)nullptr);
    return scan_incl(offset + 1, n, tmp, s);
  }
  hipLaunchKernelGGL(k_er_rows<true>, dim3(g), dim3(256), 0, s, n, p, unit_diag, mode, seed,
                     offset, nonzero);
  return (int)hipGetLastError();
}

int launch_spmat_transpose(int64_t numrows, int64_t numcols, int64_t nnz, const int64_t* offset,
                           const int64_t* nonzero, int64_t* t_offset, int64_t* t_nonzero,
                           Device* d, hipStream_t s) {
  using namespace sm;
  // row starts of the transpose: column counts, then their prefix.  The
  // counts fit 32 bits (nnz < 2^31); they go to the front of t_nonzero, which
  // only the last radix pass writes, when they fit there (numcols <= 2 nnz)
  // and t_nonzero is not the matrix's own array (a caller may transpose a
  // square matrix's nonzeros in place), else straight into t_offset as int64.
  const bool overlap = (const void*)t_nonzero < (const void*)(nonzero + nnz) &&
                       (const void*)nonzero < (const void*)(t_nonzero + nnz);
  const bool c32 = nnz > 0 && numcols <= 2 * nnz && !overlap;
  int32_t* cnt32 = (int32_t*)t_nonzero;
  if (c32) {
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, t_offset, (int64_t)1, (int64_t)0);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(numcols, 256, d)), dim3(256), 0, s, cnt32, numcols);
  } else {
    hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(numcols + 1, 256, d)), dim3(256), 0, s, t_offset,
                       numcols + 1, (int64_t)0);
  }
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  if (nnz > 0) {
    rc = c32 ? bale_histo(nonzero, nnz, cnt32, numcols, BALE_I32, BALE_INDEX_PLAIN, nullptr, s)
             : bale_histo(nonzero, nnz, t_offset + 1, numcols, BALE_I64, BALE_INDEX_PLAIN, nullptr, s);
    if (rc) return rc;  // a BALE_ERR_* code with the error text set
  }
  // digits: LSD passes of at most 8 bits over the column
  int bits = 1;
  while (bits < 63 && ((int64_t)1 << bits) < numcols) bits++;
  const int passes = (bits + 7) / 8;
  const int64_t n = nnz;
  // Passes with a count: chunks of whole tiles, two delivery workgroups per CU
  // (the 512-thread shape, also for the first pass: it measured 6.5 against
  // 6.85 ms for the call with the first pass on one 1024-thread workgroup per CU).
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * kRdxPerCu,
                                                     ceil_div(std::max<int64_t>(n, 1), kRdxTile)));
  const int64_t chunk = ceil_div(ceil_div(std::max<int64_t>(n, 1), C), kRdxTile) * kRdxTile;
  C = ceil_div(std::max<int64_t>(n, 1), chunk);
  const int Bmax = 256;
  const size_t head = ((part_words(C * kCountSplit, Bmax) * 4) + 255) & ~(size_t)255;
  const size_t ib = ((size_t)std::max<int64_t>(n, 1) * 8 + 255) & ~(size_t)255;
  const size_t stb = scan_tmp_bytes(numcols);
  // the middle passes' prefixes: the largest is the second-to-last pass's
  const int wlast = bits / passes;  // the last (smallest) of the balanced widths
  const size_t pb = passes >= 3 ? ((((size_t)1 << (bits - wlast)) + 1) * 8 + 255) & ~(size_t)255 : 0;
  char* ws = (char*)workspace(d, head + 2 * ib + pb + stb, s);
  if (!ws) return (int)hipErrorOutOfMemory;
  int64_t* pbuf = (int64_t*)(ws + head + 2 * ib);
  int64_t* stmp = (int64_t*)(ws + head + 2 * ib + pb);
  rc = c32 ? scan_incl32(cnt32, numcols, t_offset + 1, stmp, s) : scan_incl(t_offset + 1, numcols, stmp, s);
  if (rc || n == 0) return rc;
  uint64_t* buf[2] = {(uint64_t*)(ws + head), (uint64_t*)(ws + head + ib)};
  int shift = 0;
  for (int k = 0; k < passes; k++) {
    const int w = (bits - shift + (passes - k) - 1) / (passes - k);  // balanced digit widths
    const int B = 1 << w;
    const bool first = k == 0, last = k == passes - 1;
    // the first pass reads the matrix itself; the last one places the rows
    // by column offsets when there are enough groups to spread
    const uint64_t* in = first ? (const uint64_t*)nonzero : buf[k & 1];
    uint64_t* out = buf[(k + 1) & 1];
    // after the first pass, a pass with at least a group per CU needs no count
    const int64_t ngroups = (int64_t)1 << shift;
    const bool by_groups = !first && ngroups >= (int64_t)d->cus;
    const int64_t* P = t_offset;
    int64_t plen = numcols;
    if (by_groups && !last) {
      const int64_t M = (int64_t)1 << (shift + w);
      hipLaunchKernelGGL(k_low_prefix, dim3((unsigned)ceil_div(M + 1, 256)), dim3(256), 0, s, t_offset,
                         numcols, M, pbuf);
      P = pbuf;
      plen = M;
    }
    PartLayout L = part_carve(ws, C * kCountSplit, B);
    if (!by_groups) {
      hipLaunchKernelGGL(k_rx_count, dim3((unsigned)(C * kCountSplit)), dim3(kThreads), 0, s,
                         (const uint32_t*)in, n, shift, B, chunk / kCountSplit, L.chunk);
      hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)B), dim3(256), 0, s, L.chunk,
                         (int)(C * kCountSplit), B, L.total);
      hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, B, 1u << 30, 0u,
                         (const uint32_t*)nullptr, L.start, L.segstart, L.misc);
    }
    const uint32_t* cr = L.chunk;
    const uint32_t* cs = L.start;
#define BALE_RDX(IN, OUT, CUR, G)                                                                       \
  do {                                                                                                  \
    if ((G) >= (int64_t)d->cus * kRdxPerCu) {                                                           \
      allow_lds(k_rdx<IN, OUT, CUR, 512>, Rdx<512>::lds);                                               \
      hipLaunchKernelGGL((k_rdx<IN, OUT, CUR, 512>), dim3((unsigned)(G)), dim3(512), Rdx<512>::lds, s,  \
                         in, offset, numrows, n, B, w, shift, chunk, cr, kCountSplit, cs, P, plen, ngroups, out,  \
                         t_nonzero);                                                                    \
    } else {                                                                                            \
      allow_lds(k_rdx<IN, OUT, CUR, 1024>, Rdx<1024>::lds);                                             \
      hipLaunchKernelGGL((k_rdx<IN, OUT, CUR, 1024>), dim3((unsigned)(G)), dim3(1024), Rdx<1024>::lds, s, \
                         in, offset, numrows, n, B, w, shift, chunk, cr, kCountSplit, cs, P, plen, ngroups, out,  \
                         t_nonzero);                                                                    \
    }                                                                                                   \
  } while (0)
    const int64_t G = std::min<int64_t>(ngroups, (int64_t)d->cus * kRdxPerCu);
    if (first && last) BALE_RDX(1, 1, 0, C);
    else if (first) BALE_RDX(1, 0, 0, C);
    else if (by_groups && last) BALE_RDX(0, 1, 1, G);
    else if (by_groups) BALE_RDX(0, 0, 1, G);
    else if (last) BALE_RDX(0, 1, 0, C);
    else BALE_RDX(0, 0, 0, C);
#undef BALE_RDX
    rc = (int)hipGetLastError();
    if (rc) return rc;
    shift += w;
  }
  return 0;
}

int launch_spmat_permute(int64_t numrows, int64_t numcols, int64_t nnz, const int64_t* offset,
                         const int64_t* nonzero, const int64_t* rperminv, const int64_t* cperminv,
                         int64_t* p_offset, int64_t* p_nonzero, Device* d, hipStream_t s) {
  using namespace sm;
  const size_t rb = ((size_t)std::max<int64_t>(numrows, 1) * 8 + 255) & ~(size_t)255;
  char* ws = (char*)workspace(d, 3 * rb + scan_tmp_bytes(numrows), s);
  if (!ws) return (int)hipErrorOutOfMemory;
  i64x2* pairs = (i64x2*)ws;            // per result row: source length and start
  int64_t* srcs = (int64_t*)(ws + 2 * rb);  // where each result row's source row starts in A
  int64_t* stmp = (int64_t*)(ws + 3 * rb);
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(2 * numrows, 256, d)), dim3(256), 0, s, (int64_t*)pairs,
                     2 * numrows, (int64_t)-1);
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, p_offset, (int64_t)1, (int64_t)0);
  if (numrows == 0) return (int)hipGetLastError();
  hipLaunchKernelGGL(k_perm_pairs, dim3(grid_for(numrows, 256, d)), dim3(256), 0, s, offset, rperminv,
                     numrows, pairs, stream_err(d, s));
  hipLaunchKernelGGL(k_perm_split, dim3(grid_for(numrows, 256, d)), dim3(256), 0, s, (const i64x2*)pairs,
                     numrows, p_offset + 1, srcs, stream_err(d, s));
  int rc = scan_incl(p_offset + 1, numrows, stmp, s);
  if (rc) return rc;
  // The column map is a gather from cperminv.  Large matrices copy the rows
  // first and run the map as the library's gather, in place (its table-region
  // plan beats one random read per nonzero there): 3.0e8 nonzeros into 2^24
  // columns on MI355X, 7.1 ms mapped in the copy against 1.64 ms of copy and
  // 4.07 ms of gather (6.8 ms for the call instead of 7.75).  Small ones map
  // in the copy.
  const bool by_gather = nnz >= ((int64_t)1 << 24) && numcols > 0;
  if (by_gather)
    hipLaunchKernelGGL(k_perm_fill<false>, dim3((unsigned)ceil_div(numrows, kRowsPerGroup)), dim3(kThreads),
                       0, s, (const int64_t*)srcs, nonzero, cperminv, numrows, numcols,
                       (const int64_t*)p_offset, nnz, p_nonzero, stream_err(d, s));
  else
    hipLaunchKernelGGL(k_perm_fill<true>, dim3((unsigned)ceil_div(numrows, kRowsPerGroup)), dim3(kThreads),
                       0, s, (const int64_t*)srcs, nonzero, cperminv, numrows, numcols,
                       (const int64_t*)p_offset, nnz, p_nonzero, stream_err(d, s));
  rc = (int)hipGetLastError();
  if (rc || !by_gather) return rc;
  // the gather reads every request before it writes any reply, so tgt may be
  // its own index array
  return bale_gather(p_nonzero, p_nonzero, nnz, cperminv, std::max<int64_t>(numcols, 1), BALE_INDEX_PLAIN,
                     nullptr, s);
}

int launch_spmat_triangular(const int64_t* offset, const int64_t* nonzero, int64_t numrows, int upper,
                            Device* d, hipStream_t s, int64_t* wrong, int64_t* nopivot) {
  using namespace sm;
  unsigned long long* bad = (unsigned long long*)workspace(d, 256, s);
  if (!bad) return (int)hipErrorOutOfMemory;
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, (int64_t*)bad, (int64_t)2, (int64_t)0);
  const unsigned g = grid_for(numrows, 256, d);
  if (upper) hipLaunchKernelGGL(k_triangular<true>, dim3(g), dim3(256), 0, s, offset, nonzero, numrows, bad);
  else hipLaunchKernelGGL(k_triangular<false>, dim3(g), dim3(256), 0, s, offset, nonzero, numrows, bad);
  unsigned long long h[2] = {0, 0};
  int rc = (int)hipMemcpyAsync(h, bad, sizeof h, hipMemcpyDeviceToHost, s);
  if (!rc) rc = (int)hipStreamSynchronize(s);
  *wrong = (int64_t)h[0];
  *nopivot = (int64_t)h[1];
  return rc;
}

int launch_count_diff(const int64_t* a, const int64_t* b, int64_t n, Device* d, hipStream_t s,
                      int64_t* ndiff) {
  using namespace sm;
  unsigned long long* cnt = (unsigned long long*)workspace(d, 256, s);
  if (!cnt) return (int)hipErrorOutOfMemory;
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, (int64_t*)cnt, (int64_t)1, (int64_t)0);
  if (n > 0)
    hipLaunchKernelGGL(k_count_diff, dim3(grid_for(n, 256, d)), dim3(256), 0, s, a, b, n, cnt);
  unsigned long long h = 0;
  int rc = (int)hipMemcpyAsync(&h, cnt, sizeof h, hipMemcpyDeviceToHost, s);
  if (!rc) rc = (int)hipStreamSynchronize(s);
  *ndiff = (int64_t)h;
  return rc;
}

}  // namespace bale
