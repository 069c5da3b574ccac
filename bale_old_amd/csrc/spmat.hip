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
__global__ __launch_bounds__(kThreads) void k_scan_sums(const int64_t* __restrict__ a, int64_t n,
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

// items[j] = (column of nonzero j) | (its row << 32)
__global__ __launch_bounds__(kThreads) void k_expand(const int64_t* __restrict__ offset,
                                                     const int64_t* __restrict__ nonzero,
                                                     int64_t numrows, uint64_t* __restrict__ items) {
  __shared__ int64_t so[kRowsPerGroup + 1];
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerGroup;
  const int R = (int)std::min<int64_t>(kRowsPerGroup, numrows - r0);
  for (int i = threadIdx.x; i <= R; i += kThreads) so[i] = offset[r0 + i];
  __syncthreads();
  const int64_t p0 = so[0], p1 = so[R];
  for (int64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
    const int r = row_of(so, R, p);
    items[p] = (uint64_t)(uint32_t)nonzero[p] | ((uint64_t)(r0 + r) << 32);
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
// rperm[rperminv[i]] = i; entries outside [0, n) are counted as range errors.
__global__ __launch_bounds__(256) void k_invert(const int64_t* __restrict__ rperminv, int64_t n,
                                                int64_t* __restrict__ rperm, ErrWords* err) {
  unsigned bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rperminv[i];
    if ((uint64_t)r < (uint64_t)n) rperm[r] = i;
    else bad++;
  }
  flush_errors(err, bad, 0);
}

// len[i] = length of row rperm[i] of A (0 and a range error for a row that
// rperminv did not name)
__global__ __launch_bounds__(256) void k_perm_rowlen(const int64_t* __restrict__ offset,
                                                     const int64_t* __restrict__ rperm, int64_t n,
                                                     int64_t* __restrict__ len, ErrWords* err) {
  unsigned bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = rperm[i];
    if ((uint64_t)r < (uint64_t)n) {
      len[i] = offset[r + 1] - offset[r];
    } else {
      len[i] = 0;
      bad++;
    }
  }
  flush_errors(err, bad, 0);
}

// Row i of the result = row rperm[i] of A, columns mapped by cperminv.
__global__ __launch_bounds__(kThreads) void k_perm_fill(const int64_t* __restrict__ offset,
                                                        const int64_t* __restrict__ nonzero,
                                                        const int64_t* __restrict__ rperm,
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
  for (int i = threadIdx.x; i < R; i += kThreads) {
    const int64_t r = rperm[r0 + i];
    src[i] = (uint64_t)r < (uint64_t)numrows ? offset[r] : 0;
  }
  __syncthreads();
  const int64_t p0 = so[0], p1 = std::min<int64_t>(so[R], nnz);
  unsigned bad = 0;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
    const int r = row_of(so, R, p);
    const int64_t c = nonzero[src[r] + (p - so[r])];
    if ((uint64_t)c < (uint64_t)numcols) {
      p_nonzero[p] = cperminv[c];
    } else {
      p_nonzero[p] = -1;
      bad++;
    }
  }
  flush_errors(err, bad, 0);
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
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nt), dim3(kThreads), 0, s, (const int64_t*)a, n, tmp);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kThreads), 0, s, tmp, nt);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(kThreads), 0, s, a, n, (const int64_t*)tmp);
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
  // row starts of the transpose: column counts, then their prefix
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(numcols + 1, 256, d)), dim3(256), 0, s, t_offset,
                     numcols + 1, (int64_t)0);
  int rc = (int)hipGetLastError();
  if (rc) return rc;
  if (nnz > 0) {
    rc = bale_histo(nonzero, nnz, t_offset + 1, numcols, BALE_I64, BALE_INDEX_PLAIN, nullptr, s);
    if (rc) return rc;  // a BALE_ERR_* code with the error text set
  }
  // digits: LSD passes of at most 8 bits over the column
  int bits = 1;
  while (bits < 63 && ((int64_t)1 << bits) < numcols) bits++;
  const int passes = (bits + 7) / 8;
  // chunks of the item array (one workgroup each, like the conveyor)
  const int64_t n = nnz;
  // two chunks per CU: two delivery workgroups fit a CU's LDS, and the
  // second hides the first one's barriers and loads
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * 2, ceil_div(std::max<int64_t>(n, 1), kChunkAlign)));
  const int64_t chunk = ceil_div(ceil_div(std::max<int64_t>(n, 1), C), kChunkAlign) * kChunkAlign;
  C = ceil_div(std::max<int64_t>(n, 1), chunk);
  const int Bmax = 256;
  const size_t head = ((part_words(C, Bmax) * 4) + 255) & ~(size_t)255;
  const size_t ib = ((size_t)std::max<int64_t>(n, 1) * 8 + 255) & ~(size_t)255;
  const size_t stb = scan_tmp_bytes(numcols);
  char* ws = (char*)workspace(d, head + 2 * ib + stb, s);
  if (!ws) return (int)hipErrorOutOfMemory;
  int64_t* stmp = (int64_t*)(ws + head + 2 * ib);
  rc = scan_incl(t_offset + 1, numcols, stmp, s);
  if (rc || n == 0) return rc;
  uint64_t* buf[2] = {(uint64_t*)(ws + head), (uint64_t*)(ws + head + ib)};
  hipLaunchKernelGGL(k_expand, dim3((unsigned)ceil_div(numrows, kRowsPerGroup)), dim3(kThreads), 0, s,
                     offset, nonzero, numrows, buf[0]);
  int shift = 0;
  for (int k = 0; k < passes; k++) {
    const int w = (bits - shift + (passes - k) - 1) / (passes - k);  // balanced digit widths
    const int B = 1 << w;
    PartLayout L = part_carve(ws, C, B);
    const uint32_t* in = (const uint32_t*)buf[k & 1];
    hipLaunchKernelGGL(k_rx_count, dim3((unsigned)C), dim3(kThreads), 0, s, in, n, shift, B, chunk,
                       L.chunk);
    hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)B), dim3(256), 0, s, L.chunk, (int)C, B,
                       L.total);
    hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, B, 1u << 30, 0u,
                       (const uint32_t*)nullptr, L.start, L.segstart, L.misc);
    const bool last = k == passes - 1;
    void* out = last ? (void*)t_nonzero : (void*)buf[(k + 1) & 1];
    const size_t lds = deliver_lds<2>(B);
    if (last) {
      allow_lds(k_convey_deliver<2, 2>, lds);
      hipLaunchKernelGGL((k_convey_deliver<2, 2>), dim3((unsigned)C), dim3(kDlvTile), lds, s, in,
                         (const int64_t*)nullptr, n, B, w, chunk, (const uint32_t*)L.chunk,
                         (const uint32_t*)L.start, (uint32_t*)out, shift);
    } else {
      allow_lds(k_convey_deliver<2, 1>, lds);
      hipLaunchKernelGGL((k_convey_deliver<2, 1>), dim3((unsigned)C), dim3(kDlvTile), lds, s, in,
                         (const int64_t*)nullptr, n, B, w, chunk, (const uint32_t*)L.chunk,
                         (const uint32_t*)L.start, (uint32_t*)out, shift);
    }
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
  char* ws = (char*)workspace(d, rb + scan_tmp_bytes(numrows), s);
  if (!ws) return (int)hipErrorOutOfMemory;
  int64_t* rperm = (int64_t*)ws;
  int64_t* stmp = (int64_t*)(ws + rb);
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(numrows, 256, d)), dim3(256), 0, s, rperm, numrows,
                     (int64_t)-1);
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, p_offset, (int64_t)1, (int64_t)0);
  if (numrows == 0) return (int)hipGetLastError();
  hipLaunchKernelGGL(k_invert, dim3(grid_for(numrows, 256, d)), dim3(256), 0, s, rperminv, numrows,
                     rperm, stream_err(d, s));
  hipLaunchKernelGGL(k_perm_rowlen, dim3(grid_for(numrows, 256, d)), dim3(256), 0, s, offset,
                     (const int64_t*)rperm, numrows, p_offset + 1, stream_err(d, s));
  int rc = scan_incl(p_offset + 1, numrows, stmp, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_perm_fill, dim3((unsigned)ceil_div(numrows, kRowsPerGroup)), dim3(kThreads), 0,
                     s, offset, nonzero, (const int64_t*)rperm, cperminv, numrows, numcols,
                     (const int64_t*)p_offset, nnz, p_nonzero, stream_err(d, s));
  return (int)hipGetLastError();
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
