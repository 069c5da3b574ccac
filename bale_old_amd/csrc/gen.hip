// gen.hip -- seeded input generation on the GPU, plus small utility kernels.
//
// Every element depends only on (seed, i), so the arrays match a sequential
// host generator element for element.  Streams: element i of a uniform index
// array is reduce(splitmix64 stream(seed)[i], nbins).  Zipf ranks come from
// rejection-inversion (Hormann & Derflinger 1996) with the element's draw as
// the seed of its own sub-stream, and rank k is mapped to bin perm(k-1), a
// keyed 4-round Feistel permutation with cycle walking.  Clustered streams
// repeat the value drawn at the start of each run; element i starts a run
// when i % 4096 == 0 or its flag draw hits probability 1/run_mean.
#include <math.h>

#include "bale_internal.h"

namespace bale {

constexpr uint64_t kPermSalt = 0x5045524D5045524DULL;
constexpr uint64_t kFlagSalt = 0x52554E5352554E53ULL;
constexpr int64_t kRunCap = 4096;

struct Zipf {
  double s, hx1, hn, sc;
  int64_t n;
};

__host__ __device__ inline double z_helper1(double t) {
#pragma clang fp contract(off)
  return fabs(t) > 1e-8 ? log1p(t) / t : 1.0 - t * (0.5 - t / 3.0);
}
__host__ __device__ inline double z_helper2(double t) {
#pragma clang fp contract(off)
  return fabs(t) > 1e-8 ? expm1(t) / t : 1.0 + t * 0.5 * (1.0 + t / 3.0);
}
__host__ __device__ inline double z_H(const Zipf& z, double x) {
#pragma clang fp contract(off)
  double lx = log(x);
  return z_helper2((1.0 - z.s) * lx) * lx;
}
__host__ __device__ inline double z_h(const Zipf& z, double x) {
#pragma clang fp contract(off)
  return exp(-z.s * log(x));
}
__host__ __device__ inline double z_Hinv(const Zipf& z, double y) {
#pragma clang fp contract(off)
  double t = y * (1.0 - z.s);
  if (t < -1.0) t = -1.0;
  return exp(z_helper1(t) * y);
}

static Zipf zipf_make(int64_t n, double s) {
  Zipf z;
  z.n = n;
  z.s = s;
  z.hx1 = z_H(z, 1.5) - 1.0;
  z.hn = z_H(z, (double)n + 0.5);
  z.sc = 2.0 - z_Hinv(z, z_H(z, 2.5) - z_h(z, 2.0));
  return z;
}

__device__ inline int64_t zipf_draw(const Zipf& z, uint64_t subseed) {
#pragma clang fp contract(off)
  for (uint64_t k = 0;; k++) {
    uint64_t r = stream_at(subseed, k);
    double u01 = (double)(r >> 11) * 0x1.0p-53;
    double u = z.hn + u01 * (z.hx1 - z.hn);
    double x = z_Hinv(z, u);
    int64_t kk = (int64_t)(x + 0.5);
    if (kk < 1) kk = 1;
    if (kk > z.n) kk = z.n;
    if ((double)kk - x <= z.sc || u >= z_H(z, (double)kk + 0.5) - z_h(z, (double)kk)) return kk;
  }
}

__device__ inline uint64_t perm_keyed(uint64_t x, uint64_t n, uint64_t key) {
  if (n <= 1) return 0;
  int bits = 64 - __builtin_clzll(n - 1);
  int half = (bits + 1) / 2;
  if (half < 1) half = 1;
  uint64_t mask = (1ULL << half) - 1;
  uint64_t rk[4];
#pragma unroll
  for (int r = 0; r < 4; r++) rk[r] = stream_at(key, (uint64_t)r);
  do {
    uint64_t L = x >> half, R = x & mask;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      uint64_t f = mix64(R ^ rk[r]) & mask;
      uint64_t t = R;
      R = L ^ f;
      L = t;
    }
    x = (L << half) | R;
  } while (x >= n);
  return x;
}

__device__ inline int64_t draw_value(int64_t i, int64_t nbins, int dist, const Zipf& z,
                                     uint64_t seed, uint64_t pkey) {
  uint64_t r = stream_at(seed, (uint64_t)i);
  if (dist == 0) return (int64_t)reduce_range(r, (uint64_t)nbins);
  int64_t rank = zipf_draw(z, r);
  return (int64_t)perm_keyed((uint64_t)(rank - 1), (uint64_t)nbins, pkey);
}

__device__ inline bool run_start(int64_t i, int64_t run_mean, uint64_t seed) {
  if (i % kRunCap == 0) return true;
  return reduce_range(stream_at(seed ^ kFlagSalt, (uint64_t)i), (uint64_t)run_mean) == 0;
}

__global__ __launch_bounds__(256) void k_gen_index(int64_t* __restrict__ out, int64_t n,
                                                   int64_t nbins, int dist, Zipf z,
                                                   int64_t run_mean, uint64_t seed,
                                                   uint64_t pkey) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    int64_t src = i;
    if (run_mean > 1)
      while (!run_start(src, run_mean, seed)) src--;
    out[i] = draw_value(src, nbins, dist, z, seed, pkey);
  }
}

__global__ __launch_bounds__(256) void k_fill_splitmix(int64_t* __restrict__ t, int64_t n) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
    t[j] = (int64_t)mix64((uint64_t)j + kGolden);
}

__global__ __launch_bounds__(256) void k_fill_neg(int64_t* __restrict__ t, int64_t n) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
    t[j] = -(j + 1);
}

__global__ __launch_bounds__(256) void k_pack(int64_t* __restrict__ pck,
                                              const int64_t* __restrict__ idx, int64_t n) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    pck[i] = idx[i] << 16;
}

// Zero `words` 32-bit words (a kernel rather than hipMemsetAsync so the
// sequence can be captured in a HIP graph).
__global__ __launch_bounds__(256) void k_zero(uint32_t* __restrict__ p, int64_t words) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
    p[i] = 0u;
}

// out2[0] = min, out2[1] = max of the decoded indices (out2 preset by caller to
// {INT64_MAX, INT64_MIN}).
__global__ __launch_bounds__(256) void k_minmax(const int64_t* __restrict__ idx, int64_t n,
                                                int packed, int64_t* out2) {
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    int64_t v = idx[i];
    if (packed) v = v >> 16;
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int off = 32; off > 0; off >>= 1) {
    int64_t a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin((long long*)&out2[0], (long long)lo);
    atomicMax((long long*)&out2[1], (long long)hi);
  }
}

static int grid_for(int64_t n, int block, int cap = 65536) {
  int64_t g = ceil_div(n > 0 ? n : 1, block);
  return (int)(g < cap ? g : cap);
}

int launch_fill_splitmix(int64_t* t, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_splitmix, dim3(grid_for(n, 256)), dim3(256), 0, s, t, n);
  return (int)hipGetLastError();
}
int launch_fill_neg(int64_t* t, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_neg, dim3(grid_for(n, 256)), dim3(256), 0, s, t, n);
  return (int)hipGetLastError();
}
int launch_gen_index(int64_t* out, int64_t n, int64_t nbins, int dist, double zs,
                     int64_t run_mean, uint64_t seed, hipStream_t s) {
  if (n <= 0) return 0;
  Zipf z = zipf_make(nbins, dist == 1 ? zs : 1.0);
  uint64_t pkey = mix64(seed ^ kPermSalt);
  hipLaunchKernelGGL(k_gen_index, dim3(grid_for(n, 256)), dim3(256), 0, s, out, n, nbins, dist, z,
                     run_mean, seed, pkey);
  return (int)hipGetLastError();
}
int launch_pack(int64_t* pck, const int64_t* idx, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_pack, dim3(grid_for(n, 256)), dim3(256), 0, s, pck, idx, n);
  return (int)hipGetLastError();
}
int launch_zero_bytes(void* p, size_t bytes, hipStream_t s) {
  int64_t words = (int64_t)((bytes + 3) / 4);
  if (words <= 0) return 0;
  hipLaunchKernelGGL(k_zero, dim3(grid_for(words, 256, 8192)), dim3(256), 0, s, (uint32_t*)p,
                     words);
  return (int)hipGetLastError();
}
int launch_minmax(const int64_t* idx, int64_t n, int packed, int64_t* out2, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_minmax, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, idx, n, packed,
                     out2);
  return (int)hipGetLastError();
}

}  // namespace bale
