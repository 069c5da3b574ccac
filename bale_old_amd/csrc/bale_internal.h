// bale_internal.h -- shared device helpers and launcher declarations for
// libbale_hip.so (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bale_hip.h"

namespace bale {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

// 16-byte vector types the non-temporal builtins accept
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef long long i64x2 __attribute__((ext_vector_type(2)));

// ---- splitmix64 (Steele, Lea, Flood 2014) ---------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
// i-th output (0-based) of the stream seeded with `seed`
__host__ __device__ inline uint64_t stream_at(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * kGolden);
}
__device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }
// [0,n) from a 64-bit draw: top bits for powers of two, else multiply-high
__device__ inline uint64_t reduce_range(uint64_t x, uint64_t n) {
  if (n <= 1) return 0;
  if ((n & (n - 1)) == 0) return x >> (64 - __builtin_ctzll(n));
  return mulhi64(x, n);
}

// ---- per-device error record ---------------------------------------------
// err[0]: indices outside the table (skipped); err[1]: packed PE != 0.
struct ErrWords {
  unsigned long long range;
  unsigned long long pe;
};

// Decode one index. Returns false (and records why) when it must be skipped.
template <bool PACKED>
__device__ inline bool decode_index(int64_t raw, uint64_t limit, uint64_t& j,
                                    unsigned& nbad_range, unsigned& nbad_pe) {
  int64_t v = raw;
  if (PACKED) {
    if ((raw & 0xffff) != 0) { nbad_pe++; return false; }
    v = raw >> 16;
  }
  j = (uint64_t)v;
  if (j >= limit) { nbad_range++; return false; }
  return true;
}

__device__ inline void flush_errors(ErrWords* err, unsigned nbad_range, unsigned nbad_pe) {
  if (nbad_range) atomicAdd(&err->range, (unsigned long long)nbad_range);
  if (nbad_pe) atomicAdd(&err->pe, (unsigned long long)nbad_pe);
}

// device-scope relaxed add without return (global_atomic_add_x2 / _u32)
__device__ inline void atomic_add_dev(int64_t* p, int64_t v) {
  __hip_atomic_fetch_add((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void atomic_add_dev(int32_t* p, int32_t v) {
  __hip_atomic_fetch_add((unsigned*)p, (unsigned)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Raise a kernel's dynamic-LDS limit before a launch that needs more than the
// default 64 KiB.
template <typename K>
inline void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

// ---- per-device and per-stream state (capi.hip) ----------------------------
// Each stream a caller uses gets its own error record and workspace: a
// stream's calls are ordered by the stream, so its workspace can grow (after
// waiting for that stream alone) without freeing memory another stream's
// kernels still use.
struct StreamCtx {
  bool used = false;
  hipStream_t s = nullptr;
  ErrWords* err = nullptr;  // index errors of the calls on this stream
  void* ws = nullptr;       // grow-only workspace of the calls on this stream
  size_t ws_bytes = 0;
};
constexpr int kMaxStreamCtx = 256;
struct Device {
  int id = -1;
  int cus = 0;
  int lds_per_block = 0;
  // this device's address of the no-allocation scratch array (capi.hip),
  // resolved once when the device's state is created; nullptr if that failed
  void* scratch = nullptr;
  StreamCtx ctx[kMaxStreamCtx];
};
Device* device_state();  // current device's state (created on first use; allocates nothing)
// The stream's error record and workspace holder, created on the stream's
// first use (allocates the 16-byte record; refused while the stream is being
// captured).  nullptr on failure, with the error text set.
StreamCtx* stream_ctx(Device* d, hipStream_t s);
// The error record kernels launched on s report to (nullptr if stream_ctx failed).
ErrWords* stream_err(Device* d, hipStream_t s);
// Record the calling thread's error text (bale_hip_last_error) and return code.
int bale_set_error(int code, const char* fmt, ...);
// Workspace of at least `bytes` for kernels on s, 256-byte aligned; nullptr on
// failure.
void* workspace(Device* d, size_t bytes, hipStream_t s);
// Device memory for the library's own use: through the allocator installed
// with bale_hip_set_alc8r (hipMalloc by default), tagged with the calling
// file and line; nullptr on failure (no error text is set).  dev_free returns
// a pointer to the allocator that made it (nullptr is ignored).
void* dev_alloc(size_t bytes, const char* tag, uint64_t line);
void dev_free(void* p);
#define BALE_DEV_ALLOC(bytes) ::bale::dev_alloc((bytes), __FILE_NAME__, __LINE__)
// BALE_ERR_HIP with the runtime's text for rc (bale_hip_last_error).
int bale_fail_hip(int rc, const char* what);
// What a launcher returns for a configuration that cannot run on this shape
// (the reason is already in the thread's error text); the API turns it into
// BALE_ERR_ARG.  No HIP error code has this value.
constexpr int kBadConfig = 0x7fff0001;
int bad_config(const char* fmt, ...);
// The calling thread's bale_hip_last_timing values.
void set_last_timing(double gpu_s, double h2d_s, double d2h_s);
double last_gpu_seconds();
// Whether p is device (or managed) memory.
bool is_device_ptr(const void* p);
// Seconds between two recorded events.
double elapsed(hipEvent_t a, hipEvent_t b);

// ---- launchers (return hipError_t-like int, 0 = ok) -----------------------
int launch_fill_splitmix(int64_t* t, int64_t n, hipStream_t s);
int launch_fill_neg(int64_t* t, int64_t n, hipStream_t s);
int launch_gen_index(int64_t* out, int64_t n, int64_t nbins, int dist, double zs,
                     int64_t run_mean, uint64_t seed, hipStream_t s);
int launch_pack(int64_t* pck, const int64_t* idx, int64_t n, hipStream_t s);
int launch_zero_bytes(void* p, size_t bytes, hipStream_t s);
int launch_minmax(const int64_t* idx, int64_t n, int packed, int64_t* out2, hipStream_t s);

int launch_gather_direct(int64_t* tgt, const int64_t* idx, int64_t n, const int64_t* table,
                         int64_t table_len, bool packed, const bale_config& c, ErrWords* err,
                         hipStream_t s);
int launch_gather_bucket(int64_t* tgt, const int64_t* idx, int64_t n, const int64_t* table,
                         int64_t table_len, bool packed, const bale_config& c, Device* d,
                         hipStream_t s);

int launch_histo_direct(const int64_t* idx, int64_t n, void* counts, int64_t nbins, int dtype,
                        bool packed, const bale_config& c, Device* d, hipStream_t s);
int launch_histo_lds(const int64_t* idx, int64_t n, void* counts, int64_t nbins, int dtype,
                     bool packed, const bale_config& c, Device* d, hipStream_t s);
int launch_histo_bucket(const int64_t* idx, int64_t n, void* counts, int64_t nbins, int dtype,
                        bool packed, const bale_config& c, Device* d, hipStream_t s);

// offT[r][c] = starts[c][r] for u16 rows of `rows` entries, C rows (gather.hip)
int launch_starts_transpose(const uint16_t* starts, uint16_t* offT, int rows, int64_t C, hipStream_t s);

// random permutation by dart throwing (randperm.hip); darts_used (optional)
// receives the darts generated
int launch_randperm(int64_t* perm, int64_t N, int64_t seed, Device* d, hipStream_t s,
                    int64_t* darts_used);
// random permutation by a Fisher-Yates pass resolved in parallel (randperm.hip)
int launch_randperm_shuffle(int64_t* perm, int64_t N, int64_t seed, Device* d, hipStream_t s);
// *ok = perm[0..N) holds 0..N-1 once each (waits for the stream)
int launch_check_perm(const int64_t* perm, int64_t N, Device* d, hipStream_t s, int* ok);

// sparse matrices (spmat.hip)
int launch_spmat_gen_er(int64_t n, double p, int unit_diag, int mode, uint64_t seed,
                        int64_t* offset, int64_t* nonzero, Device* d, hipStream_t s);
int launch_spmat_transpose(int64_t numrows, int64_t numcols, int64_t nnz, const int64_t* offset,
                           const int64_t* nonzero, int64_t* t_offset, int64_t* t_nonzero,
                           Device* d, hipStream_t s);
int launch_spmat_permute(int64_t numrows, int64_t numcols, int64_t nnz, const int64_t* offset,
                         const int64_t* nonzero, const int64_t* rperminv, const int64_t* cperminv,
                         int64_t* p_offset, int64_t* p_nonzero, Device* d, hipStream_t s);
int launch_spmat_triangular(const int64_t* offset, const int64_t* nonzero, int64_t numrows, int upper,
                            Device* d, hipStream_t s, int64_t* wrong, int64_t* nopivot);
// toposort (toposort.hip): waits for the stream; pivots found, levels, the
// count of column indices out of range and of columns claimed twice
int launch_toposort(int64_t n, const int64_t* offset, const int64_t* nonzero, const int64_t* t_offset,
                    const int64_t* t_nonzero, int64_t* rperm, int64_t* cperm, Device* d, hipStream_t s,
                    int64_t* pivots, int64_t* levels, int64_t* bad_col, int64_t* twice);
// *ndiff = entries where a and b differ (waits for the stream)
int launch_count_diff(const int64_t* a, const int64_t* b, int64_t n, Device* d, hipStream_t s,
                      int64_t* ndiff);

int launch_probe_copy(void* dst, const void* src, int64_t bytes, hipStream_t s);
int launch_probe_random(const int64_t* table, int64_t n, int64_t nreads, int64_t* out,
                        hipStream_t s);

}  // namespace bale
