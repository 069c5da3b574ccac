// exstack and exstack2 on one GPU: bale's bulk-synchronous and asynchronous
// aggregation buffers (exstack/exstack.h:94-107, :201-210) as a bulk service
// over device arrays.
//
// The reference's PEs each push items for other PEs into per-destination
// buffers; exstack_exchange moves every buffer to its destination at once
// (exstack/exstack.upc), exstack2 sends a buffer when it fills and the
// receiver pops whatever has arrived (exstack/exstack2.upc).  A receiver can
// pop from any source (exstack_pop, with the source returned) or from one
// named source in that source's push order (exstack_pop_thread, which
// apps/ig_src/ig_exstack.upc:88-92 relies on to retrace its requests).
//
// Here the PEs are virtual PEs of one GPU (up to kMaxPes - 1).  A push hands over
// a batch of items with their source and destination PEs (device arrays); a
// round holds up to n_pe * buf_cnt items in all (the reference bounds each
// destination's buffer by buf_cnt; a caller that pushes until a push is short
// and then exchanges behaves the same).  The exchange delivers the round by
// two stable passes of the conveyor's delivery kernel -- by source, then by
// destination -- so every destination's items come out grouped by source,
// each source's in push order, as one device span per destination and a
// sub-span per (destination, source).
#include "bale_internal.h"
#include "deliver.h"
#include "partition.h"

#include <algorithm>
#include <vector>

namespace bale {

// The delivery keys of a round: an item whose source (0 when `src` is null)
// and destination are both PEs keeps them; any other item gets P for both,
// a bucket after every PE's, and is counted on the error record.  Also
// counts[dst * P + src] += 1 for the kept items.
__global__ __launch_bounds__(256) void k_ex_keys(const int64_t* __restrict__ dst,
                                                 const int64_t* __restrict__ src, int64_t n,
                                                 int64_t P, int64_t* __restrict__ kd,
                                                 int64_t* __restrict__ ks,
                                                 uint32_t* __restrict__ counts, ErrWords* err) {
  unsigned bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = dst[i], s = src ? src[i] : 0;
    const bool ok = d >= 0 && d < P && s >= 0 && s < P;
    kd[i] = ok ? d : P;
    ks[i] = ok ? s : P;
    if (ok) atomicAdd(&counts[d * P + s], 1u);
    else bad++;
  }
  flush_errors(err, bad, 0u);
}

}  // namespace bale

using namespace bale;

struct bale_exstack {
  int device = -1;
  int kind = 1;                 // 1 exstack, 2 exstack2
  int64_t npe = 0, buf_cnt = 0, capacity = 0;
  size_t pkg = 0;
  int words = 0;                // pkg / 4
  void* q_items = nullptr;      // this round's pushes
  int64_t* q_src = nullptr;
  int64_t* q_dst = nullptr;
  int64_t pending = 0;
  bool any_src = false;         // some push of the round named sources
  void* t_items = nullptr;      // after the pass by source
  int64_t* t_src = nullptr;
  int64_t* t_dst = nullptr;
  void* out_items = nullptr;    // delivered, by (destination, source, push order)
  int64_t* out_src = nullptr;
  int64_t* k_dst = nullptr;     // the round's delivery keys (P for items with a bad PE)
  int64_t* k_src = nullptr;
  uint32_t* pair_dev = nullptr; // [P * P] counts of the last delivery
  std::vector<uint32_t> pair;   // host copy
  std::vector<int64_t> pstart;  // [P * P + 1] sub-span starts
  std::vector<int64_t> taken;   // [P * P] items of each sub-span popped
  std::vector<uint8_t> whole;   // [P] the destination's whole span was popped
  int64_t unpopped = 0;
  bool done = false;            // exstack2: the caller is done pushing
};

namespace {

int ex_fail_hip(int rc, const char* what) {
  return bale_set_error(BALE_ERR_HIP, "%s: %s", what, hipGetErrorString((hipError_t)rc));
}

bool pkg_ok(size_t b) { return b == 4 || b == 8 || b == 12 || b == 16 || b == 24 || b == 32; }

void deliver_words(int W, dim3 g, hipStream_t s, const void* items, const int64_t* keys, int64_t n,
                   int npe, int nbits, int64_t chunk, const PartLayout& L, void* out) {
  switch (W) {
    case 1: launch_deliver<1>(g, s, items, keys, n, npe, nbits, chunk, L.chunk, L.start, out); break;
    case 2: launch_deliver<2>(g, s, items, keys, n, npe, nbits, chunk, L.chunk, L.start, out); break;
    case 3: launch_deliver<3>(g, s, items, keys, n, npe, nbits, chunk, L.chunk, L.start, out); break;
    case 4: launch_deliver<4>(g, s, items, keys, n, npe, nbits, chunk, L.chunk, L.start, out); break;
    case 6: launch_deliver<6>(g, s, items, keys, n, npe, nbits, chunk, L.chunk, L.start, out); break;
    default: launch_deliver<8>(g, s, items, keys, n, npe, nbits, chunk, L.chunk, L.start, out); break;
  }
}

// A stable pass keyed by `keys`: the layout (counts, chunk offsets, starts)
// once, then each array moved with it.  Items whose key is out of range are
// counted on the stream's error record and dropped.
struct Move {
  const void* src;
  void* dst;
  int words;
};
int stable_pass(Device* d, const int64_t* keys, int64_t n, int npe, const Move* mv, int nmv,
                hipStream_t s) {
  int nbits = 0;
  while ((1 << nbits) < npe) nbits++;
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * 2, ceil_div(n, kChunkAlign)));
  const int64_t chunk = ceil_div(ceil_div(n, C), kChunkAlign) * kChunkAlign;
  C = ceil_div(n, chunk);
  void* ws = workspace(d, part_bytes(C, npe, 0, 1), s);
  if (!ws) return BALE_ERR_ALLOC;
  PartLayout L = part_carve(ws, C, npe);
  hipLaunchKernelGGL((k_part_count<false, false, false>), dim3((unsigned)C), dim3(1024),
                     (size_t)npe * 4, s, keys, n, (uint64_t)npe, 0, npe, chunk, L.chunk,
                     (const uint32_t*)nullptr, stream_err(d, s), (const uint32_t*)nullptr);
  hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)npe), dim3(256), 0, s, L.chunk, (int)C, npe,
                     L.total);
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, npe, 1u << 30, 0u,
                     (const uint32_t*)nullptr, L.start, L.segstart, L.misc);
  for (int k = 0; k < nmv; k++)
    deliver_words(mv[k].words, dim3((unsigned)C), s, mv[k].src, keys, n, npe, nbits, chunk, L,
                  mv[k].dst);
  int rc = (int)hipGetLastError();
  return rc ? ex_fail_hip(rc, "exstack exchange") : BALE_OK;
}

// Deliver the round: by source (when sources were named), then by destination,
// with P + 1 keys (the last for items with a bad PE, which end up after every
// span); pair counts and sub-span starts come back to the host.
int ex_deliver(bale_exstack* x, hipStream_t s) {
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  ErrWords* err = stream_err(d, s);
  if (!err) return BALE_ERR_ARG;
  const int64_t n = x->pending;
  const int P = (int)x->npe;
  int rc = launch_zero_bytes(x->pair_dev, (size_t)P * P * 4, s);
  if (rc) return ex_fail_hip(rc, "exstack exchange");
  if (n > 0) {
    hipLaunchKernelGGL(k_ex_keys, dim3((unsigned)std::min<int64_t>(ceil_div(n, 256), 4096)), dim3(256), 0,
                       s, (const int64_t*)x->q_dst, (const int64_t*)(x->any_src ? x->q_src : nullptr), n,
                       (int64_t)P, x->k_dst, x->k_src, x->pair_dev, err);
    if (x->any_src) {
      const Move m1[3] = {{x->q_items, x->t_items, x->words}, {x->k_dst, x->t_dst, 2},
                          {x->k_src, x->t_src, 2}};
      if ((rc = stable_pass(d, x->k_src, n, P + 1, m1, 3, s))) return rc;
      const Move m2[2] = {{x->t_items, x->out_items, x->words}, {x->t_src, x->out_src, 2}};
      if ((rc = stable_pass(d, x->t_dst, n, P + 1, m2, 2, s))) return rc;
    } else {
      const Move m[2] = {{x->q_items, x->out_items, x->words}, {x->k_src, x->out_src, 2}};
      if ((rc = stable_pass(d, x->k_dst, n, P + 1, m, 2, s))) return rc;
    }
  }
  rc = (int)hipMemcpyAsync(x->pair.data(), x->pair_dev, (size_t)P * P * 4, hipMemcpyDeviceToHost, s);
  if (rc) return ex_fail_hip(rc, "exstack exchange: counts");
  int64_t bad = 0;
  rc = bale_sync_errors((void*)s, &bad);  // waits
  x->pending = 0;
  x->any_src = false;
  int64_t run = 0;
  for (int64_t k = 0; k < (int64_t)P * P; k++) {
    x->pstart[(size_t)k] = run;
    run += x->pair[(size_t)k];
  }
  x->pstart[(size_t)P * P] = run;
  std::fill(x->taken.begin(), x->taken.end(), 0);
  std::fill(x->whole.begin(), x->whole.end(), 0);
  x->unpopped = run;
  if (rc == BALE_ERR_RANGE)
    return bale_set_error(BALE_ERR_RANGE, "exstack exchange: %lld items named a PE outside 0..%d and "
                          "were dropped", (long long)bad, P - 1);
  return rc;
}

bale_exstack* ex_new(int kind, int64_t n_pe, int64_t buf_cnt, size_t pkg) {
  const char* name = kind == 1 ? "exstack_init" : "exstack2_init";
  if (n_pe < 1 || n_pe >= kMaxPes) {  // one delivery key is kept for bad PE numbers
    bale_set_error(BALE_ERR_ARG, "%s: %lld PEs (1..%d on one GPU)", name, (long long)n_pe, kMaxPes - 1);
    return nullptr;
  }
  if (buf_cnt < 1 || n_pe * buf_cnt >= ((int64_t)1 << 31)) {
    bale_set_error(BALE_ERR_ARG, "%s: buf_cnt %lld (n_pe * buf_cnt below 2^31)", name, (long long)buf_cnt);
    return nullptr;
  }
  if (!pkg_ok(pkg)) {
    bale_set_error(BALE_ERR_ARG, "%s: package size %zu (4, 8, 12, 16, 24 or 32 bytes)", name, pkg);
    return nullptr;
  }
  Device* d = device_state();
  if (!d) {
    bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
    return nullptr;
  }
  auto* x = new bale_exstack;
  x->device = d->id;
  x->kind = kind;
  x->npe = n_pe;
  x->buf_cnt = buf_cnt;
  x->capacity = n_pe * buf_cnt;
  x->pkg = pkg;
  x->words = (int)(pkg / 4);
  const size_t ib = (size_t)x->capacity * pkg, pb = (size_t)x->capacity * 8;
  void* p[11] = {};
  const size_t sz[11] = {ib, pb, pb, ib, pb, pb, ib, pb, (size_t)n_pe * n_pe * 4, pb, pb};
  bool ok = true;
  for (int k = 0; k < 11 && ok; k++)
    if (!(p[k] = BALE_DEV_ALLOC(std::max<size_t>(sz[k], 256)))) ok = false;
  if (!ok) {
    for (void* q : p) dev_free(q);
    delete x;
    bale_set_error(BALE_ERR_ALLOC, "%s: cannot allocate the buffers", name);
    return nullptr;
  }
  x->q_items = p[0];
  x->q_src = (int64_t*)p[1];
  x->q_dst = (int64_t*)p[2];
  x->t_items = p[3];
  x->t_src = (int64_t*)p[4];
  x->t_dst = (int64_t*)p[5];
  x->out_items = p[6];
  x->out_src = (int64_t*)p[7];
  x->pair_dev = (uint32_t*)p[8];
  x->k_dst = (int64_t*)p[9];
  x->k_src = (int64_t*)p[10];
  x->pair.assign((size_t)n_pe * n_pe, 0);
  x->pstart.assign((size_t)n_pe * n_pe + 1, 0);
  x->taken.assign((size_t)n_pe * n_pe, 0);
  x->whole.assign((size_t)n_pe, 0);
  return x;
}

// Accept the longest prefix of the batch that fits the round.
int64_t ex_push(bale_exstack* x, const void* items, const int64_t* from_pes, const int64_t* to_pes,
                int64_t n, hipStream_t s, const char* name) {
  if (n < 0 || (n > 0 && (!items || !to_pes)))
    return bale_set_error(BALE_ERR_ARG, "%s: bad batch", name);
  const int64_t k = std::min(n, x->capacity - x->pending);
  if (k <= 0) return 0;
  int rc = (int)hipMemcpyAsync((char*)x->q_items + (size_t)x->pending * x->pkg, items,
                               (size_t)k * x->pkg, hipMemcpyDeviceToDevice, s);
  if (!rc)
    rc = (int)hipMemcpyAsync(x->q_dst + x->pending, to_pes, (size_t)k * 8, hipMemcpyDeviceToDevice, s);
  if (!rc) {
    if (from_pes) {
      if (!x->any_src && x->pending > 0)  // earlier pushes of the round came from PE 0
        rc = launch_zero_bytes(x->q_src, (size_t)x->pending * 8, s);
      if (!rc)
        rc = (int)hipMemcpyAsync(x->q_src + x->pending, from_pes, (size_t)k * 8,
                                 hipMemcpyDeviceToDevice, s);
      x->any_src = true;
    } else if (x->any_src) {
      rc = launch_zero_bytes(x->q_src + x->pending, (size_t)k * 8, s);
    }
  }
  if (rc) return ex_fail_hip(rc, name);
  x->pending += k;
  return k;
}

int64_t ex_pop(bale_exstack* x, int64_t pe, const void** items, const int64_t** from_pes,
               const char* name) {
  const int64_t P = x->npe;
  if (pe < 0 || pe >= P) return bale_set_error(BALE_ERR_ARG, "%s: PE %lld", name, (long long)pe);
  const int64_t lo = x->pstart[(size_t)(pe * P)], hi = x->pstart[(size_t)((pe + 1) * P)];
  for (int64_t s = 0; s < P; s++)
    if (x->taken[(size_t)(pe * P + s)])
      return bale_set_error(BALE_ERR_ARG, "%s: PE %lld's items from PE %lld were popped one source "
                            "at a time this round", name, (long long)pe, (long long)s);
  if (x->whole[(size_t)pe]) return 0;
  x->whole[(size_t)pe] = 1;
  if (items) *items = (const char*)x->out_items + (size_t)lo * x->pkg;
  if (from_pes) *from_pes = x->out_src + lo;
  x->unpopped -= hi - lo;
  return hi - lo;
}

}  // namespace

extern "C" {

bale_exstack* bale_exstack_init(int64_t n_pe, int64_t buf_cnt, size_t pkg_size) {
  return ex_new(1, n_pe, buf_cnt, pkg_size);
}

int64_t bale_exstack_proceed(bale_exstack* x, int done) {
  if (!x || x->kind != 1) return bale_set_error(BALE_ERR_ARG, "exstack_proceed: not an exstack");
  return (done && x->pending == 0 && x->unpopped == 0) ? 0 : 1;
}

int64_t bale_exstack_push(bale_exstack* x, const void* items, const int64_t* from_pes,
                          const int64_t* to_pes, int64_t n, void* stream) {
  if (!x || x->kind != 1) return bale_set_error(BALE_ERR_ARG, "exstack_push: not an exstack");
  return ex_push(x, items, from_pes, to_pes, n, (hipStream_t)stream, "exstack_push");
}

int bale_exstack_exchange(bale_exstack* x, void* stream) {
  if (!x || x->kind != 1) return bale_set_error(BALE_ERR_ARG, "exstack_exchange: not an exstack");
  return ex_deliver(x, (hipStream_t)stream);
}

int64_t bale_exstack_pop(bale_exstack* x, int64_t pe, const void** items, const int64_t** from_pes) {
  if (!x || x->kind != 1) return bale_set_error(BALE_ERR_ARG, "exstack_pop: not an exstack");
  return ex_pop(x, pe, items, from_pes, "exstack_pop");
}

int64_t bale_exstack_pop_thread(bale_exstack* x, int64_t pe, int64_t from_pe, const void** items) {
  if (!x || x->kind != 1) return bale_set_error(BALE_ERR_ARG, "exstack_pop_thread: not an exstack");
  const int64_t P = x->npe;
  if (pe < 0 || pe >= P || from_pe < 0 || from_pe >= P)
    return bale_set_error(BALE_ERR_ARG, "exstack_pop_thread: PE %lld from %lld", (long long)pe,
                          (long long)from_pe);
  if (x->whole[(size_t)pe])
    return bale_set_error(BALE_ERR_ARG, "exstack_pop_thread: PE %lld's items were popped whole this "
                          "round", (long long)pe);
  const size_t k = (size_t)(pe * P + from_pe);
  const int64_t lo = x->pstart[k], hi = x->pstart[k + 1];
  if (x->taken[k]) return 0;
  x->taken[k] = 1;
  if (items) *items = (const char*)x->out_items + (size_t)lo * x->pkg;
  x->unpopped -= hi - lo;
  return hi - lo;
}

int bale_exstack_reset(bale_exstack* x) {
  if (!x) return bale_set_error(BALE_ERR_ARG, "exstack_reset: null exstack");
  x->pending = 0;
  x->any_src = false;
  x->unpopped = 0;
  x->done = false;
  std::fill(x->pstart.begin(), x->pstart.end(), 0);
  std::fill(x->taken.begin(), x->taken.end(), 0);
  std::fill(x->whole.begin(), x->whole.end(), 0);
  return BALE_OK;
}

void bale_exstack_clear(bale_exstack* x) {
  if (!x) return;
  void* p[11] = {x->q_items, x->q_src, x->q_dst, x->t_items, x->t_src, x->t_dst, x->out_items,
                 x->out_src, x->pair_dev, x->k_dst, x->k_src};
  for (void* q : p) dev_free(q);
  delete x;
}

bale_exstack* bale_exstack2_init(int64_t n_pe, int64_t buf_cnt, size_t pkg_size) {
  return ex_new(2, n_pe, buf_cnt, pkg_size);
}

int64_t bale_exstack2_push(bale_exstack* x, const void* items, const int64_t* from_pes,
                           const int64_t* to_pes, int64_t n, void* stream) {
  if (!x || x->kind != 2) return bale_set_error(BALE_ERR_ARG, "exstack2_push: not an exstack2");
  if (x->done) return bale_set_error(BALE_ERR_ARG, "exstack2_push: pushing after proceed was told done");
  return ex_push(x, items, from_pes, to_pes, n, (hipStream_t)stream, "exstack2_push");
}

int64_t bale_exstack2_proceed(bale_exstack* x, int done_pushing, void* stream) {
  if (!x || x->kind != 2) return bale_set_error(BALE_ERR_ARG, "exstack2_proceed: not an exstack2");
  if (done_pushing) x->done = true;
  // sends what was pushed once everything delivered before has been popped
  if (x->pending > 0 && x->unpopped == 0) {
    int rc = ex_deliver(x, (hipStream_t)stream);
    if (rc != BALE_OK) return rc;
  }
  return (x->done && x->pending == 0 && x->unpopped == 0) ? 0 : 1;
}

int64_t bale_exstack2_pop(bale_exstack* x, int64_t pe, const void** items, const int64_t** from_pes) {
  if (!x || x->kind != 2) return bale_set_error(BALE_ERR_ARG, "exstack2_pop: not an exstack2");
  return ex_pop(x, pe, items, from_pes, "exstack2_pop");
}

int bale_exstack2_reset(bale_exstack* x) {
  if (!x || x->kind != 2) return bale_set_error(BALE_ERR_ARG, "exstack2_reset: not an exstack2");
  return bale_exstack_reset(x);
}

void bale_exstack2_clear(bale_exstack* x) { bale_exstack_clear(x); }

}  // extern "C"
