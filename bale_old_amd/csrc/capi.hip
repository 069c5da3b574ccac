// capi.hip -- the C-ABI of libbale_hip.so: per-device state, the native
// bale_* API and the reference drop-ins (histo_*, ig_*).  The random
// permutation, sparse-matrix and toposort entry points are in capi_spmat.hip.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <unordered_map>

#include "bale_internal.h"

namespace bale {

static thread_local std::string t_err;
static thread_local int t_code = BALE_OK;  // code of the last recorded error
static thread_local int t_strategy = BALE_STRAT_AUTO;  // strategy of the last histo / gather
static thread_local double t_gpu_s = 0.0, t_h2d_s = 0.0, t_d2h_s = 0.0;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  t_code = code;
  return code;
}
int bale_set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  t_code = code;
  return code;
}

int bad_config(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  t_code = BALE_ERR_ARG;
  return kBadConfig;
}

static int fail_hip(int rc, const char* what) {
  return fail(BALE_ERR_HIP, "%s: %s", what, hipGetErrorString((hipError_t)rc));
}
int bale_fail_hip(int rc, const char* what) { return fail_hip(rc, what); }
void set_last_timing(double gpu_s, double h2d_s, double d2h_s) {
  t_gpu_s = gpu_s;
  t_h2d_s = h2d_s;
  t_d2h_s = d2h_s;
}
double last_gpu_seconds() { return t_gpu_s; }

// ---- device allocations ---------------------------------------------------
// The installed allocator (a null grab means hipMalloc / hipFree), and which
// allocator made each pointer that a caller's allocator handed out, so that
// every pointer goes back to the allocator that made it even after another
// one is installed.
struct Alc8r {
  void* ctx = nullptr;
  void* (*grab)(void*, size_t, const char*, uint64_t) = nullptr;
  void (*release)(void*, void*) = nullptr;
};
static Alc8r g_alc8r;
static std::mutex g_alc8r_mu;
static std::unordered_map<void*, Alc8r>* g_ledger = nullptr;  // never destroyed

void* dev_alloc(size_t bytes, const char* tag, uint64_t line) {
  Alc8r a;
  {
    std::lock_guard<std::mutex> lk(g_alc8r_mu);
    a = g_alc8r;
  }
  if (!bytes) bytes = 256;
  void* p = nullptr;
  if (!a.grab) {
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
    }
    return p;
  }
  p = a.grab(a.ctx, bytes, tag, line);
  if (p) {
    std::lock_guard<std::mutex> lk(g_alc8r_mu);
    if (!g_ledger) g_ledger = new std::unordered_map<void*, Alc8r>();
    (*g_ledger)[p] = a;
  }
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  Alc8r a;
  bool theirs = false;
  {
    std::lock_guard<std::mutex> lk(g_alc8r_mu);
    if (g_ledger) {
      auto it = g_ledger->find(p);
      if (it != g_ledger->end()) {
        a = it->second;
        g_ledger->erase(it);
        theirs = true;
      }
    }
  }
  if (theirs) a.release(a.ctx, p);
  else (void)hipFree(p);
}

constexpr int kMaxDevices = 64;
static Device g_dev[kMaxDevices];
static std::mutex g_mu;

// ---- the no-allocation entry point's scratch ------------------------------
// One fixed array per device (each GPU the library runs on holds its own copy
// until the process ends).  In host code `scratch` names a host-side stand-in,
// not a device address: host code uses only Device::scratch, the address
// hipGetSymbolAddress gives for the current device.  Layout: 256 bytes of
// error words for calls that pass no error counters, then the workspace.
constexpr size_t kScratchBytes = (size_t)BALE_NOALLOC_SCRATCH_BYTES;
constexpr size_t kScratchErrBytes = 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// While a no-allocation call runs on this thread, workspace() and stream_err()
// hand out the scratch array instead of the stream's state (which may need
// allocating); `wanted` records the largest workspace asked for.
struct NoAllocScope {
  void* ws;
  size_t ws_bytes;
  ErrWords* err;
  size_t wanted = 0;
};
static thread_local NoAllocScope* t_noalloc = nullptr;

Device* device_state() {
  int id = 0;
  if (hipGetDevice(&id) != hipSuccess || id < 0 || id >= kMaxDevices) return nullptr;
  Device* d = &g_dev[id];
  if (d->id == id) return d;
  std::lock_guard<std::mutex> lk(g_mu);
  if (d->id == id) return d;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, id) != hipSuccess) return nullptr;
  d->cus = p.multiProcessorCount;
  int lds = (int)p.sharedMemPerBlock;
  int optin = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, id) == hipSuccess &&
      optin > lds)
    lds = optin;
  (void)hipGetLastError();
  d->lds_per_block = lds;
  // the scratch array's address on this device (the current one)
  void* sp = nullptr;
  if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) == hipSuccess && sp) {
    d->scratch = sp;
  } else {
    (void)hipGetLastError();
    d->scratch = nullptr;
  }
  d->id = id;
  return d;
}

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return st != hipStreamCaptureStatusNone;
}

StreamCtx* stream_ctx(Device* d, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  StreamCtx* free_slot = nullptr;
  for (StreamCtx& c : d->ctx) {
    if (c.used && c.s == s) return &c;
    if (!c.used && !free_slot) free_slot = &c;
  }
  if (!free_slot) {
    fail(BALE_ERR_ARG, "more than %d streams in use on device %d (bale_hip_release_stream frees one)",
         kMaxStreamCtx, d->id);
    return nullptr;
  }
  if (capturing(s)) {
    fail(BALE_ERR_ARG, "first call on a stream while it is being captured: make one call on it "
                       "outside capture first (or use bale_histo_noalloc)");
    return nullptr;
  }
  ErrWords* err = (ErrWords*)BALE_DEV_ALLOC(sizeof(ErrWords));
  if (!err) {
    fail(BALE_ERR_ALLOC, "error record allocation failed");
    return nullptr;
  }
  if (hipMemsetAsync(err, 0, sizeof(ErrWords), s) != hipSuccess) {
    (void)hipGetLastError();
    dev_free(err);
    fail(BALE_ERR_HIP, "clearing the error record failed");
    return nullptr;
  }
  free_slot->s = s;
  free_slot->err = err;
  free_slot->ws = nullptr;
  free_slot->ws_bytes = 0;
  free_slot->used = true;
  return free_slot;
}

ErrWords* stream_err(Device* d, hipStream_t s) {
  if (t_noalloc) return t_noalloc->err;
  StreamCtx* c = stream_ctx(d, s);
  return c ? c->err : nullptr;
}

// BALE_OK once the stream has its error record, else the error stream_ctx gave.
static int stream_ready(Device* d, hipStream_t s) { return stream_ctx(d, s) ? BALE_OK : t_code; }

// Launcher result to an API code: 0, a configuration the shape cannot take
// (reason already recorded), allocation failure, or another HIP error.
static int launch_result(int rc, const char* what) {
  if (rc == 0) return BALE_OK;
  if (rc == kBadConfig) return BALE_ERR_ARG;
  if (rc == hipErrorOutOfMemory) return t_code == BALE_ERR_ALLOC ? BALE_ERR_ALLOC : fail(BALE_ERR_ALLOC, "%s: workspace allocation failed", what);
  return fail_hip(rc, what);
}

// Fields of a caller's configuration that no shape can take.
static int check_config(const bale_config& c, bool histogram) {
  if (c.strategy < BALE_STRAT_AUTO || c.strategy > BALE_STRAT_BUCKET ||
      (!histogram && c.strategy == BALE_STRAT_LDS))
    return fail(BALE_ERR_ARG, "bad %s strategy %d", histogram ? "histogram" : "gather", c.strategy);
  if (c.block != 0 && c.block != 256 && c.block != 512 && c.block != 1024)
    return fail(BALE_ERR_ARG, "block %d: 0, 256, 512 or 1024 threads", c.block);
  if (c.unroll != 0 && c.unroll != 4 && c.unroll != 8 && c.unroll != 16)
    return fail(BALE_ERR_ARG, "unroll %d: 0, 4, 8 or 16", c.unroll);
  if (c.tile_log2 < 0 || c.tile_log2 > 30)
    return fail(BALE_ERR_ARG, "tile_log2 %d: 0 (default) or 1..30", c.tile_log2);
  if (c.merge < -1 || c.merge > BALE_MERGE_LEADER)
    return fail(BALE_ERR_ARG, "merge %d: -1 (default), 0, 1 or 2", c.merge);
  if (c.groups_per_cu < 0 || c.groups_per_cu > 16)
    return fail(BALE_ERR_ARG, "groups_per_cu %d: 0 (default) or 1..16", c.groups_per_cu);
  if (c.hot < 0 || c.hot > 2 || (c.hot == 2 && c.strategy != BALE_STRAT_BUCKET && c.strategy != BALE_STRAT_AUTO))
    return fail(BALE_ERR_ARG, "hot %d: 0, 1, or 2 with the BUCKET strategy", c.hot);
  if (c.task_log2 != 0 && (c.task_log2 < 8 || c.task_log2 > 30))
    return fail(BALE_ERR_ARG, "task_log2 %d: 0 (default) or 8..30", c.task_log2);
  return BALE_OK;
}

void* workspace(Device* d, size_t bytes, hipStream_t s) {
  if (t_noalloc) {
    t_noalloc->wanted = std::max(t_noalloc->wanted, bytes);
    return bytes <= t_noalloc->ws_bytes ? t_noalloc->ws : nullptr;
  }
  StreamCtx* c = stream_ctx(d, s);
  if (!c) return nullptr;
  if (c->ws_bytes >= bytes) return c->ws;
  if (capturing(s)) {
    fail(BALE_ERR_ARG, "the stream's workspace must grow to %zu bytes, which cannot happen during "
                       "capture: make the same call once outside capture first", bytes);
    return nullptr;
  }
  if (c->ws) {
    // only this stream's kernels use it
    if (hipStreamSynchronize(s) != hipSuccess) return nullptr;
    dev_free(c->ws);
    c->ws = nullptr;
    c->ws_bytes = 0;
  }
  size_t want = bytes + (bytes >> 3);  // grow with headroom
  c->ws = BALE_DEV_ALLOC(want);
  if (!c->ws) {
    fail(BALE_ERR_ALLOC, "workspace of %zu bytes could not be allocated", want);
    return nullptr;
  }
  c->ws_bytes = want;
  return c->ws;
}

static int max_tile_bits_for(const Device* d) {
  int tb = 15;
  while (tb > 8 && ((size_t)4 << tb) + 1024 > (size_t)d->lds_per_block) tb--;
  return tb;
}

// Default rule when no tuned configuration is given.
static void plan_histo(const Device* d, int64_t n, int64_t nbins, int dtype, bale_config* c) {
  memset(c, 0, sizeof *c);
  c->merge = BALE_MERGE_NONE;
  int tb = max_tile_bits_for(d);
  int64_t bytes = nbins * (dtype == BALE_I64 ? 8 : 4);
  (void)bytes;
  if (nbins <= ((int64_t)1 << tb) && n >= 8 * nbins) {
    c->strategy = BALE_STRAT_LDS;
    c->block = 1024;
    c->unroll = 8;
  } else if (n >= ((int64_t)1 << 20) && ceil_div(nbins, (int64_t)1 << tb) <= 4096 &&
             n >= nbins / 4) {
    // partition through LDS; the sample decides about hot bins and runs
    c->strategy = BALE_STRAT_BUCKET;
    c->block = 1024;
    c->hot = 2;
    // bucket width: about 256 buckets (one per CU in the apply pass), within
    // 2^10 bins and the packed 16-bit tile's 2^16; measured on MI355X against
    // the earlier fixed 2^14 (up to 2^22 bins) and 2^15/2^16 widths: 1.0-1.7x
    // on uniform inputs, 0.85-1.2x on Zipf(1.1) ones
    int lb = 0;
    while (((int64_t)1 << lb) < nbins) lb++;
    c->tile_log2 = std::max(10, std::min(16, lb - 8));
  } else {
    c->strategy = BALE_STRAT_DIRECT;
    c->block = 256;
    c->unroll = 8;
  }
}
static void plan_gather(const Device*, int64_t n, int64_t table_len, bale_config* c) {
  memset(c, 0, sizeof *c);
  c->strategy = BALE_STRAT_DIRECT;
  c->block = 256;
  // 8 table loads in flight per lane: 6.62 ms for 2^28 requests into a 2 GiB
  // table on MI355X, against 6.79 ms with 16 and 6.73 ms with 4
  c->unroll = 8;
  c->merge = BALE_MERGE_NONE;
  // Large uniform gathers go by table region (BUCKET); a sample taken on the
  // device sends inputs that repeat entries (hot entries, clustered runs) to
  // DIRECT, which the caches serve better.  Measured on MI355X, 2^28 uniform
  // requests: 56.6-58.9 against 40.9 GUPS into a 2^28-entry table, 67.6 against
  // 54.0 into 2^22 entries; Zipf(1.1) into 2^28 entries: DIRECT 123 GUPS,
  // BUCKET 68.  BUCKET reads every region's table lines once however few
  // requests there are: even at 2^26 requests into 2^28 entries, and behind
  // DIRECT at 2^24, hence at least one request per two entries.
  if (n >= ((int64_t)1 << 24) && 2 * n >= table_len && table_len >= ((int64_t)1 << 22) &&
      table_len <= ((int64_t)1 << 30)) {
    c->strategy = BALE_STRAT_BUCKET;
    c->hot = 2;
    // Regions of 2^18 entries served from an XCD's L2, or, up to 2^25 entries,
    // of 2^13-2^14 entries served from LDS (about 512-2048 regions).  2^28
    // requests on MI355X, L2 regions against LDS regions: 2^22 entries 72.9
    // against 98.5 GUPS (2^13), 2^23 72.0 / 85.2 (2^13), 2^24 72.0 / 86.5
    // (2^14), 2^25 71.7 / 78.5 (2^14); 2^26 69.3 against 62.5 (runs of 4).
    int lg = 0;
    while (((int64_t)2 << lg) <= table_len) lg++;  // floor(log2(table_len))
    c->tile_log2 = lg <= 25 ? std::max(13, std::min(14, lg - 10)) : 18;
    c->task_log2 = 11;
    c->groups_per_cu = 4;
  }
}

bool is_device_ptr(const void* p) {
  hipPointerAttribute_t a;
  memset(&a, 0, sizeof a);
  hipError_t rc = hipPointerGetAttributes(&a, p);
  if (rc != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Entries of int64 addressable from p to the end of its device allocation.
static int64_t device_extent(const void* p, size_t elem) {
  void* base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (void*)p) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  size_t off = (const char*)p - (const char*)base;
  return off > size ? 0 : (int64_t)((size - off) / elem);
}

}  // namespace bale

using namespace bale;

extern "C" {

int bale_hip_init(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(BALE_ERR_NODEV, "no HIP device");
  }
  if (device < 0 || device >= n) return fail(BALE_ERR_ARG, "device %d out of range [0,%d)", device, n);
  int rc = (int)hipSetDevice(device);
  if (rc) return fail_hip(rc, "hipSetDevice");
  if (!device_state()) return fail(BALE_ERR_HIP, "device state setup failed");
  return BALE_OK;
}

const char* bale_hip_last_error(void) { return t_err.c_str(); }

int bale_hip_last_strategy(void) { return t_strategy; }

void bale_hip_last_timing(double* gpu_s, double* h2d_s, double* d2h_s) {
  if (gpu_s) *gpu_s = t_gpu_s;
  if (h2d_s) *h2d_s = t_h2d_s;
  if (d2h_s) *d2h_s = t_d2h_s;
}

// cus, lds bytes per workgroup
int bale_hip_device_info(int* cus, int* lds_per_block) {
  Device* d = device_state();
  if (!d) return fail(BALE_ERR_NODEV, "no usable HIP device");
  if (cus) *cus = d->cus;
  if (lds_per_block) *lds_per_block = d->lds_per_block;
  return BALE_OK;
}

int bale_histo_plan(int64_t n, int64_t nbins, int dtype, bale_config* out) {
  Device* d = device_state();
  if (!d) return fail(BALE_ERR_NODEV, "no usable HIP device");
  if (!out) return fail(BALE_ERR_ARG, "null config");
  plan_histo(d, n, nbins, dtype, out);
  return BALE_OK;
}

int bale_gather_plan(int64_t n, int64_t table_len, bale_config* out) {
  Device* d = device_state();
  if (!d) return fail(BALE_ERR_NODEV, "no usable HIP device");
  if (!out) return fail(BALE_ERR_ARG, "null config");
  plan_gather(d, n, table_len, out);
  return BALE_OK;
}

// Arguments of bale_histo / bale_histo_noalloc that no device is needed to refuse.
static int check_histo_args(const int64_t* index, int64_t n, void* counts, int64_t nbins,
                            int dtype, int index_format, const bale_config* cfg) {
  if (n < 0) return fail(BALE_ERR_ARG, "negative update count %lld", (long long)n);
  if (n == 0) return BALE_OK;
  if (!index || !counts) return fail(BALE_ERR_ARG, "null index or counts");
  if (nbins <= 0) return fail(BALE_ERR_ARG, "nbins must be positive");
  if (dtype != BALE_I64 && dtype != BALE_I32) return fail(BALE_ERR_ARG, "bad dtype %d", dtype);
  if (index_format != BALE_INDEX_PLAIN && index_format != BALE_INDEX_PACKED)
    return fail(BALE_ERR_ARG, "bad index format %d", index_format);
  if (cfg) return check_config(*cfg, true);
  return BALE_OK;
}

// The histogram's launches (shared by bale_histo and bale_histo_noalloc).
static int histo_run(Device* d, hipStream_t s, const int64_t* index, int64_t n, void* counts,
                     int64_t nbins, int dtype, int index_format, const bale_config* cfg) {
  bale_config c;
  if (cfg && cfg->strategy != BALE_STRAT_AUTO) c = *cfg;
  else plan_histo(d, n, nbins, dtype, &c);
  t_strategy = c.strategy;
  bool packed = index_format == BALE_INDEX_PACKED;
  // the bucketed pass keeps 32-bit positions: run it in slices
  const int64_t slice = c.strategy == BALE_STRAT_BUCKET ? ((int64_t)1 << 30) : n;
  for (int64_t off = 0; off < n; off += slice) {
    int64_t m = std::min(slice, n - off);
    int rc;
    switch (c.strategy) {
      case BALE_STRAT_DIRECT:
        rc = launch_histo_direct(index + off, m, counts, nbins, dtype, packed, c, d, s);
        break;
      case BALE_STRAT_LDS:
        rc = launch_histo_lds(index + off, m, counts, nbins, dtype, packed, c, d, s);
        break;
      default:
        rc = launch_histo_bucket(index + off, m, counts, nbins, dtype, packed, c, d, s);
        break;
    }
    if (rc == hipErrorOutOfMemory && t_noalloc)
      return fail(BALE_ERR_ARG, "this histogram needs %zu bytes of scratch, more than the %zu the "
                                "no-allocation entry point has (use bale_histo)",
                  t_noalloc->wanted, t_noalloc->ws_bytes);
    if ((rc = launch_result(rc, "histogram launch"))) return rc;
  }
  return BALE_OK;
}

int bale_histo(const int64_t* index, int64_t n, void* counts, int64_t nbins, int dtype,
               int index_format, const bale_config* cfg, void* stream) {
  int e = check_histo_args(index, n, counts, nbins, dtype, index_format, cfg);
  if (e || n == 0) return e;
  Device* d = device_state();
  if (!d) return fail(BALE_ERR_NODEV, "no usable HIP device");
  hipStream_t s = (hipStream_t)stream;
  if ((e = stream_ready(d, s))) return e;
  return histo_run(d, s, index, n, counts, nbins, dtype, index_format, cfg);
}

int bale_histo_noalloc(const int64_t* index, int64_t n, void* counts, int64_t nbins, int dtype,
                       int index_format, const bale_config* cfg, int64_t* errors, void* stream) {
  int e = check_histo_args(index, n, counts, nbins, dtype, index_format, cfg);
  if (e || n == 0) return e;
  Device* d = device_state();
  if (!d) return fail(BALE_ERR_NODEV, "no usable HIP device");
  if (!d->scratch) return fail(BALE_ERR_HIP, "the scratch array's device address is unknown");
  NoAllocScope scope{(char*)d->scratch + kScratchErrBytes, kScratchBytes - kScratchErrBytes,
                     errors ? (ErrWords*)errors : (ErrWords*)d->scratch};
  t_noalloc = &scope;
  e = histo_run(d, (hipStream_t)stream, index, n, counts, nbins, dtype, index_format, cfg);
  t_noalloc = nullptr;
  return e;
}

int bale_gather(int64_t* tgt, const int64_t* index, int64_t n, const int64_t* table,
                int64_t table_len, int index_format, const bale_config* cfg, void* stream) {
  if (n < 0) return fail(BALE_ERR_ARG, "negative request count %lld", (long long)n);
  if (n == 0) return BALE_OK;
  if (!tgt || !index || !table) return fail(BALE_ERR_ARG, "null tgt, index or table");
  if (table_len <= 0) return fail(BALE_ERR_ARG, "table_len must be positive");
  if (index_format != BALE_INDEX_PLAIN && index_format != BALE_INDEX_PACKED)
    return fail(BALE_ERR_ARG, "bad index format %d", index_format);
  if (cfg) {
    int e = check_config(*cfg, false);
    if (e) return e;
  }
  Device* d = device_state();
  if (!d) return fail(BALE_ERR_NODEV, "no usable HIP device");
  hipStream_t s = (hipStream_t)stream;
  int e = stream_ready(d, s);
  if (e) return e;
  bale_config c;
  if (cfg && cfg->strategy != BALE_STRAT_AUTO) c = *cfg;
  else plan_gather(d, n, table_len, &c);
  t_strategy = c.strategy;
  bool packed = index_format == BALE_INDEX_PACKED;
  const int64_t slice = c.strategy == BALE_STRAT_BUCKET ? ((int64_t)1 << 30) : n;
  for (int64_t off = 0; off < n; off += slice) {
    int64_t m = std::min(slice, n - off);
    int rc;
    if (c.strategy == BALE_STRAT_BUCKET)
      rc = launch_gather_bucket(tgt + off, index + off, m, table, table_len, packed, c, d, s);
    else
      rc = launch_gather_direct(tgt + off, index + off, m, table, table_len, packed, c,
                                stream_err(d, s), s);
    if ((rc = launch_result(rc, "gather launch"))) return rc;
  }
  return BALE_OK;
}

int bale_sync_errors(void* stream, int64_t* bad) {
  Device* d = device_state();
  if (!d) return fail(BALE_ERR_NODEV, "no usable HIP device");
  hipStream_t s = (hipStream_t)stream;
  int rc = (int)hipStreamSynchronize(s);
  if (rc) return fail_hip(rc, "hipStreamSynchronize");
  if (bad) *bad = 0;
  ErrWords* err = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (StreamCtx& c : d->ctx)
      if (c.used && c.s == s) err = c.err;
  }
  if (!err) return BALE_OK;  // no call was made on this stream
  ErrWords h{0, 0}, z{0, 0};
  rc = (int)hipMemcpy(&h, err, sizeof h, hipMemcpyDeviceToHost);
  if (rc) return fail_hip(rc, "reading the error record");
  if (h.range || h.pe) {
    rc = (int)hipMemcpy(err, &z, sizeof z, hipMemcpyHostToDevice);
    if (rc) return fail_hip(rc, "clearing the error record");
  }
  if (bad) *bad = (int64_t)(h.range + h.pe);
  if (h.pe) return fail(BALE_ERR_PE, "%llu packed indices name a PE other than 0 (one GPU is PE 0)",
                        (unsigned long long)h.pe);
  if (h.range) return fail(BALE_ERR_RANGE, "%llu indices outside the table were skipped",
                           (unsigned long long)h.range);
  return BALE_OK;
}

int bale_hip_set_alc8r(const bale_alc8r_t* a) {
  if (a && (!a->grab || !a->free)) return fail(BALE_ERR_ARG, "an allocator needs both grab and free");
  std::lock_guard<std::mutex> lk(g_alc8r_mu);
  g_alc8r = a ? Alc8r{a->alc8r, a->grab, a->free} : Alc8r();
  return BALE_OK;
}

int bale_hip_release_stream(void* stream) {
  Device* d = device_state();
  if (!d) return fail(BALE_ERR_NODEV, "no usable HIP device");
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_mu);
  for (StreamCtx& c : d->ctx) {
    if (!c.used || c.s != s) continue;
    int rc = (int)hipStreamSynchronize(s);
    if (rc) return fail_hip(rc, "hipStreamSynchronize");
    dev_free(c.ws);
    dev_free(c.err);
    c = StreamCtx();
  }
  return BALE_OK;
}

void* bale_all_alloc(int64_t nblocks, int64_t blocksize) {
  if (nblocks < 0 || blocksize <= 0) {
    fail(BALE_ERR_ARG, "bad allocation size");
    return nullptr;
  }
  if (!device_state()) {
    fail(BALE_ERR_NODEV, "no usable HIP device");
    return nullptr;
  }
  size_t bytes = (size_t)(nblocks ? nblocks : 1) * (size_t)blocksize;
  void* p = BALE_DEV_ALLOC(bytes);
  if (!p) {
    fail(BALE_ERR_ALLOC, "device allocation of %zu bytes failed", bytes);
    return nullptr;
  }
  if (hipMemset(p, 0, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    (void)hipGetLastError();
    dev_free(p);
    fail(BALE_ERR_HIP, "zeroing the allocation failed");
    return nullptr;
  }
  return p;
}

void bale_all_free(void* p) { dev_free(p); }

int bale_memput(void* dst_dev, const void* src_host, int64_t bytes) {
  if (bytes < 0 || (bytes && (!dst_dev || !src_host))) return fail(BALE_ERR_ARG, "bad copy");
  int rc = (int)hipMemcpy(dst_dev, src_host, (size_t)bytes, hipMemcpyHostToDevice);
  return rc ? fail_hip(rc, "memput") : BALE_OK;
}

int bale_memget(void* dst_host, const void* src_dev, int64_t bytes) {
  if (bytes < 0 || (bytes && (!dst_host || !src_dev))) return fail(BALE_ERR_ARG, "bad copy");
  int rc = (int)hipMemcpy(dst_host, src_dev, (size_t)bytes, hipMemcpyDeviceToHost);
  return rc ? fail_hip(rc, "memget") : BALE_OK;
}

int bale_fill_splitmix(int64_t* table, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && !table)) return fail(BALE_ERR_ARG, "bad table");
  int rc = launch_fill_splitmix(table, n, (hipStream_t)stream);
  return rc ? fail_hip(rc, "fill") : BALE_OK;
}

int bale_fill_neg(int64_t* table, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && !table)) return fail(BALE_ERR_ARG, "bad table");
  int rc = launch_fill_neg(table, n, (hipStream_t)stream);
  return rc ? fail_hip(rc, "fill") : BALE_OK;
}

int bale_gen_index(int64_t* out, int64_t n, int64_t nbins, int dist, double s, int64_t run_mean,
                   uint64_t seed, void* stream) {
  if (n < 0 || (n > 0 && !out)) return fail(BALE_ERR_ARG, "bad output");
  if (nbins <= 0) return fail(BALE_ERR_ARG, "nbins must be positive");
  if (dist != 0 && dist != 1) return fail(BALE_ERR_ARG, "dist must be 0 (uniform) or 1 (zipf)");
  if (dist == 1 && !(s > 0.0 && s < 10.0)) return fail(BALE_ERR_ARG, "zipf exponent out of (0,10)");
  int rc = launch_gen_index(out, n, nbins, dist, s, run_mean, seed, (hipStream_t)stream);
  return rc ? fail_hip(rc, "gen_index") : BALE_OK;
}

int bale_pack(int64_t* pck, const int64_t* index, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && (!pck || !index))) return fail(BALE_ERR_ARG, "bad arrays");
  int rc = launch_pack(pck, index, n, (hipStream_t)stream);
  return rc ? fail_hip(rc, "pack") : BALE_OK;
}

int bale_probe_copy(void* dst, const void* src, int64_t bytes, void* stream) {
  if (bytes < 0 || (bytes % 16) || (bytes && (!dst || !src)))
    return fail(BALE_ERR_ARG, "copy size must be a multiple of 16");
  int rc = launch_probe_copy(dst, src, bytes, (hipStream_t)stream);
  return rc ? fail_hip(rc, "probe_copy") : BALE_OK;
}

int bale_probe_random_read(const int64_t* table, int64_t n, int64_t nreads, int64_t* out,
                           void* stream) {
  if (n <= 0 || nreads < 0 || !table || !out) return fail(BALE_ERR_ARG, "bad probe arguments");
  int rc = launch_probe_random(table, n, nreads, out, (hipStream_t)stream);
  return rc ? fail_hip(rc, "probe_random") : BALE_OK;
}

}  // extern "C"

// ======================= reference drop-ins ===============================
namespace bale {

struct Staged {
  void* p = nullptr;
  bool owned = false;
  ~Staged() {
    if (owned) dev_free(p);
  }
};

double elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
  return ms * 1e-3;
}

// Shared body of every drop-in.  kind 0: histogram (counts += ...),
// kind 1: gather (tgt = table[...]).
static double dropin(int kind, int64_t* out, int64_t* index, int64_t n, int64_t* table_or_counts,
                     bool packed, int strategy, int merge) {
  t_gpu_s = t_h2d_s = t_d2h_s = 0.0;
  if (n < 0) { fail(BALE_ERR_ARG, "negative length"); return -1.0; }
  if (n > 0 && (!index || !table_or_counts || (kind == 1 && !out))) {
    fail(BALE_ERR_ARG, "null array");
    return -1.0;
  }
  Device* d = device_state();
  if (!d) { fail(BALE_ERR_NODEV, "no usable HIP device"); return -1.0; }
  if (n == 0) return 0.0;
  hipStream_t s = 0;
  hipEvent_t e0, e1, e2, e3;
  hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&e2); hipEventCreate(&e3);
  auto cleanup = [&]() {
    hipEventDestroy(e0); hipEventDestroy(e1); hipEventDestroy(e2); hipEventDestroy(e3);
  };
  double result = -1.0;
  Staged sidx, stab, sout;
  int64_t tlen = -1;
  int rc;
  hipEventRecord(e0, s);
  // index
  if (is_device_ptr(index)) {
    sidx.p = index;
  } else {
    if (!(sidx.p = BALE_DEV_ALLOC((size_t)n * 8))) { fail(BALE_ERR_ALLOC, "staging index"); cleanup(); return -1.0; }
    sidx.owned = true;
    if ((rc = (int)hipMemcpy(sidx.p, index, (size_t)n * 8, hipMemcpyHostToDevice))) { fail_hip(rc, "index H2D"); cleanup(); return -1.0; }
  }
  // table / counts: its needed length is one past the largest index passed
  bool host_tab = !is_device_ptr(table_or_counts);
  {
    int64_t* mm = nullptr;
    int64_t init[2] = {INT64_MAX, INT64_MIN}, got[2];
    if (!(mm = (int64_t*)BALE_DEV_ALLOC(16))) { fail(BALE_ERR_ALLOC, "minmax"); cleanup(); return -1.0; }
    (void)hipMemcpy(mm, init, 16, hipMemcpyHostToDevice);
    launch_minmax((const int64_t*)sidx.p, n, packed ? 1 : 0, mm, s);
    rc = (int)hipMemcpy(got, mm, 16, hipMemcpyDeviceToHost);
    dev_free(mm);
    if (rc) { fail_hip(rc, "minmax"); cleanup(); return -1.0; }
    if (got[0] < 0) { fail(BALE_ERR_RANGE, "negative index %lld", (long long)got[0]); cleanup(); return -1.0; }
    tlen = got[1] + 1;
  }
  if (!host_tab) {
    stab.p = table_or_counts;
    int64_t ext = device_extent(table_or_counts, 8);
    if (ext >= 0 && tlen > ext) {
      fail(BALE_ERR_RANGE, "index %lld is past the end of the device allocation (%lld entries)",
           (long long)(tlen - 1), (long long)ext);
      cleanup();
      return -1.0;
    }
  } else {
    if (!(stab.p = BALE_DEV_ALLOC((size_t)tlen * 8))) { fail(BALE_ERR_ALLOC, "staging table"); cleanup(); return -1.0; }
    stab.owned = true;
    if ((rc = (int)hipMemcpy(stab.p, table_or_counts, (size_t)tlen * 8, hipMemcpyHostToDevice))) { fail_hip(rc, "table H2D"); cleanup(); return -1.0; }
  }
  // gather output
  bool host_out = kind == 1 && !is_device_ptr(out);
  if (kind == 1) {
    if (host_out) {
      if (!(sout.p = BALE_DEV_ALLOC((size_t)n * 8))) { fail(BALE_ERR_ALLOC, "staging tgt"); cleanup(); return -1.0; }
      sout.owned = true;
    } else {
      sout.p = out;
    }
  }
  hipEventRecord(e1, s);
  bale_config c;
  memset(&c, 0, sizeof c);
  int r;
  if (strategy == BALE_STRAT_AUTO) {
    c.strategy = BALE_STRAT_AUTO;
  } else {
    if (kind == 0) plan_histo(d, n, tlen, BALE_I64, &c);
    else plan_gather(d, n, tlen, &c);
    const int planned = c.strategy;
    const int64_t tiles = ceil_div(tlen, (int64_t)1 << max_tile_bits_for(d));
    int take = strategy;
    // beyond 4096 bucket tiles the bucketed histogram is not offered
    if (kind == 0 && strategy == BALE_STRAT_BUCKET && tiles > 4096) take = BALE_STRAT_DIRECT;
    // the LDS strategy reads the whole index once per bin tile: beyond one
    // tile per CU it is slower than the plan by the number of tiles (a 2^26-bin
    // table would read the index ~2048 times), so the plan runs instead
    if (kind == 0 && strategy == BALE_STRAT_LDS && tiles > d->cus) take = planned;
    // a model's own strategy runs without the plan's sample-driven extras
    if (take != planned) c.hot = 0;
    c.strategy = take;
    if (take == BALE_STRAT_LDS || take == BALE_STRAT_BUCKET) c.block = 1024;
    if (take == BALE_STRAT_LDS && c.unroll == 0) c.unroll = 8;
    c.merge = merge;
  }
  if (kind == 0)
    r = bale_histo((const int64_t*)sidx.p, n, stab.p, tlen, BALE_I64,
                   packed ? BALE_INDEX_PACKED : BALE_INDEX_PLAIN, &c, (void*)s);
  else
    r = bale_gather((int64_t*)sout.p, (const int64_t*)sidx.p, n, (const int64_t*)stab.p, tlen,
                    packed ? BALE_INDEX_PACKED : BALE_INDEX_PLAIN, &c, (void*)s);
  hipEventRecord(e2, s);
  if (r == BALE_OK) {
    if (kind == 0 && host_tab)
      rc = (int)hipMemcpy(table_or_counts, stab.p, (size_t)tlen * 8, hipMemcpyDeviceToHost);
    else if (kind == 1 && host_out)
      rc = (int)hipMemcpy(out, sout.p, (size_t)n * 8, hipMemcpyDeviceToHost);
    else
      rc = 0;
    hipEventRecord(e3, s);
    if (rc) r = fail_hip(rc, "result D2H");
  }
  if (r == BALE_OK) {
    int64_t bad = 0;
    r = bale_sync_errors((void*)s, &bad);
  } else {
    (void)hipDeviceSynchronize();
  }
  if (r == BALE_OK) {
    hipEventSynchronize(e3);
    t_h2d_s = elapsed(e0, e1);
    t_gpu_s = elapsed(e1, e2);
    t_d2h_s = elapsed(e2, e3);
    result = t_gpu_s;
  }
  cleanup();
  return result;
}

}  // namespace bale

extern "C" {

double histo_agi(int64_t* index, int64_t len, int64_t* counts) {
  return dropin(0, nullptr, index, len, counts, false, BALE_STRAT_DIRECT, BALE_MERGE_NONE);
}
double histo_exstack(int64_t* pckindx, int64_t len, int64_t* lcounts, int64_t) {
  return dropin(0, nullptr, pckindx, len, lcounts, true, BALE_STRAT_BUCKET, BALE_MERGE_NONE);
}
double histo_exstack2(int64_t* pckindx, int64_t len, int64_t* lcounts, int64_t) {
  return dropin(0, nullptr, pckindx, len, lcounts, true, BALE_STRAT_LDS, BALE_MERGE_NONE);
}
double histo_conveyor(int64_t* pckindx, int64_t len, int64_t* lcounts) {
  return dropin(0, nullptr, pckindx, len, lcounts, true, BALE_STRAT_AUTO, BALE_MERGE_NONE);
}
double histo_exstack2_goto(int64_t* pckindx, int64_t T, int64_t* lcounts, int64_t) {
  return dropin(0, nullptr, pckindx, T, lcounts, true, BALE_STRAT_BUCKET, BALE_MERGE_NONE);
}
double histo_exstack2_cyclic(int64_t* pckindx, int64_t T, int64_t* lcounts, int64_t) {
  return dropin(0, nullptr, pckindx, T, lcounts, true, BALE_STRAT_LDS, BALE_MERGE_RUNS);
}
double histo_exstack_function(int64_t* pckindx, int64_t T, int64_t* lcounts, int64_t) {
  return dropin(0, nullptr, pckindx, T, lcounts, true, BALE_STRAT_BUCKET, BALE_MERGE_NONE);
}
double histo_exstack2_function(int64_t* pckindx, int64_t T, int64_t* lcounts, int64_t) {
  return dropin(0, nullptr, pckindx, T, lcounts, true, BALE_STRAT_LDS, BALE_MERGE_NONE);
}

double ig_agi(int64_t* tgt, int64_t* index, int64_t l_num_req, int64_t* table) {
  return dropin(1, tgt, index, l_num_req, table, false, BALE_STRAT_DIRECT, BALE_MERGE_NONE);
}
double ig_exstack(int64_t* tgt, int64_t* pckindx, int64_t l_num_req, int64_t* ltable, int64_t) {
  return dropin(1, tgt, pckindx, l_num_req, ltable, true, BALE_STRAT_BUCKET, BALE_MERGE_NONE);
}
double ig_exstack2(int64_t* tgt, int64_t* pckindx, int64_t l_num_req, int64_t* ltable, int64_t) {
  return dropin(1, tgt, pckindx, l_num_req, ltable, true, BALE_STRAT_DIRECT, BALE_MERGE_NONE);
}
double ig_conveyor(int64_t* tgt, int64_t* pckindx, int64_t l_num_req, int64_t* ltable) {
  return dropin(1, tgt, pckindx, l_num_req, ltable, true, BALE_STRAT_AUTO, BALE_MERGE_NONE);
}
double ig_exstack2_cyclic(int64_t* tgt, int64_t* pckindx, int64_t l_num_req, int64_t* ltable, int64_t) {
  return dropin(1, tgt, pckindx, l_num_req, ltable, true, BALE_STRAT_DIRECT, BALE_MERGE_NONE);
}
double ig_exstack2_goto(int64_t* tgt, int64_t* pckindx, int64_t l_num_req, int64_t* ltable, int64_t) {
  return dropin(1, tgt, pckindx, l_num_req, ltable, true, BALE_STRAT_DIRECT, BALE_MERGE_NONE);
}
double ig_exstack_function(int64_t* tgt, int64_t* pckindx, int64_t l_num_req, int64_t* ltable, int64_t) {
  return dropin(1, tgt, pckindx, l_num_req, ltable, true, BALE_STRAT_BUCKET, BALE_MERGE_NONE);
}
double ig_exstack_pkg(int64_t* tgt, int64_t* pckindx, int64_t l_num_req, int64_t* ltable, int64_t) {
  return dropin(1, tgt, pckindx, l_num_req, ltable, true, BALE_STRAT_BUCKET, BALE_MERGE_NONE);
}

}  // extern "C"
