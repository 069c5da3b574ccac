// probe.hip -- roofline probes that define the denominators for the path:
//   k_probe_copy    streaming 16-byte copy (HBM stream bandwidth)
//   k_probe_random  random 8-byte reads at hashed addresses with U reads in
//                   flight per lane and no index stream (random-access ceiling)
#include "bale_internal.h"

namespace bale {

__global__ __launch_bounds__(256) void k_probe_copy(u32x4* __restrict__ dst,
                                                    const u32x4* __restrict__ src, int64_t nvec) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

template <int U>
__global__ __launch_bounds__(256) void k_probe_random(const int64_t* __restrict__ table, uint64_t n,
                                                      int64_t nreads,
                                                      unsigned long long* __restrict__ out) {
  int64_t base = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * U;
  int64_t v[U];
#pragma unroll
  for (int k = 0; k < U; k++) {
    int64_t r = base + k;
    v[k] = r < nreads ? table[reduce_range(mix64((uint64_t)r), n)] : 0;
  }
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < U; k++) acc += (uint64_t)v[k];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)acc);
}

int launch_probe_copy(void* dst, const void* src, int64_t bytes, hipStream_t s) {
  int64_t nvec = bytes / 16;
  if (nvec <= 0) return 0;
  int64_t grid = std::min<int64_t>(ceil_div(nvec, 256), 65536);
  hipLaunchKernelGGL(k_probe_copy, dim3((unsigned)grid), dim3(256), 0, s, (u32x4*)dst,
                     (const u32x4*)src, nvec);
  return (int)hipGetLastError();
}

int launch_probe_random(const int64_t* table, int64_t n, int64_t nreads, int64_t* out,
                        hipStream_t s) {
  if (nreads <= 0 || n <= 0) return 0;
  constexpr int U = 16;
  int64_t grid = ceil_div(nreads, (int64_t)256 * U);
  if (grid > 0x7fffffff) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_probe_random<U>, dim3((unsigned)grid), dim3(256), 0, s, table,
                     (uint64_t)n, nreads, (unsigned long long*)out);
  return (int)hipGetLastError();
}

}  // namespace bale
