// Device conveyor: bale's aggregate-by-destination service on one GPU.
//
// The reference's conveyors (convey/convey.h:184-338) move items pushed on
// each PE to the PE named with the item, and deliver the items one source
// sends one destination in the order they were pushed
// (convey/narrative.h:135-137).  Here the PEs are virtual destinations on one
// GPU (up to kMaxPes of them), items are pushed and pulled in bulk from device
// memory, and advance() delivers everything pushed since the last delivery:
//
//   push     copies a batch of items and their PE numbers into the
//            conveyor's queue (device to device, on the caller's stream)
//   advance  partitions the queue by PE: a count pass (per-chunk counts per
//            PE, partition.h), the chunk offsets and PE starts, then a
//            stable delivery pass (k_convey_deliver) that ranks each tile's
//            items per PE with wave ballots, stages them in LDS in delivery
//            order and writes each PE's run with coalesced stores
//   pull     hands out a PE's delivered items as one span (apull's bulk form)
//
// Order inside a PE's span is push order, as the reference guarantees for
// one source.  Every span is contiguous in one delivery buffer.

#include "bale_internal.h"
#include "deliver.h"
#include "partition.h"

#include <algorithm>
#include <vector>

namespace bale {

}  // namespace bale

using namespace bale;

struct bale_convey {
  int device = -1;
  int64_t npe = 0, capacity = 0;
  size_t item_bytes = 0, max_item_bytes = 0;
  int state = 0;                 // 0 dormant, 1 working, 2 done (advance returned DONE)
  void* queue = nullptr;         // capacity items pushed, not yet delivered
  int64_t* qpes = nullptr;       // their PE numbers
  void* out = nullptr;           // capacity items: the delivered spans
  int64_t pending = 0;           // items in the queue
  std::vector<uint32_t> starts;  // npe + 1: PE p's span is [starts[p], starts[p+1])
  std::vector<int64_t> pulled;   // items of each span already pulled
  std::vector<int64_t> last;     // size of each PE's last pull (for unpull)
  int64_t unpulled = 0;          // delivered items not pulled yet
};

static bool item_size_ok(size_t b) {
  return b == 4 || b == 8 || b == 12 || b == 16 || b == 24 || b == 32;
}

static int convey_fail_hip(int rc, const char* what) {
  return bale_set_error(BALE_ERR_HIP, "%s: %s", what, hipGetErrorString((hipError_t)rc));
}

// Deliver the queue: count, offsets and starts, stable delivery; then the
// spans' starts come back to the host.
static int convey_deliver(bale_convey* c, hipStream_t s) {
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  const int64_t n = c->pending;
  const int npe = (int)c->npe;
  int nbits = 0;
  while ((1 << nbits) < npe) nbits++;
  // two chunks per CU (two delivery workgroups fit a CU's LDS)
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * 2, ceil_div(n, kChunkAlign)));
  const int64_t chunk = ceil_div(ceil_div(n, C), kChunkAlign) * kChunkAlign;
  C = ceil_div(n, chunk);
  void* ws = workspace(d, part_bytes(C, npe, 0, 1), s);
  if (!ws) return bale_set_error(BALE_ERR_ALLOC, "convey_advance: workspace");
  PartLayout L = part_carve(ws, C, npe);
  hipLaunchKernelGGL((k_part_count<false, false, false>), dim3((unsigned)C), dim3(1024),
                     (size_t)npe * 4, s, (const int64_t*)c->qpes, n, (uint64_t)npe, 0, npe, chunk,
                     L.chunk, (const uint32_t*)nullptr, stream_err(d, s), (const uint32_t*)nullptr);
  hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)npe), dim3(256), 0, s, L.chunk, (int)C, npe,
                     L.total);
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, npe, 1u << 30, 0u,
                     (const uint32_t*)nullptr, L.start, L.segstart, L.misc);
  const dim3 g((unsigned)C);
  switch (c->item_bytes / 4) {
    case 1: launch_deliver<1>(g, s, c->queue, c->qpes, n, npe, nbits, chunk, L.chunk, L.start, c->out); break;
    case 2: launch_deliver<2>(g, s, c->queue, c->qpes, n, npe, nbits, chunk, L.chunk, L.start, c->out); break;
    case 3: launch_deliver<3>(g, s, c->queue, c->qpes, n, npe, nbits, chunk, L.chunk, L.start, c->out); break;
    case 4: launch_deliver<4>(g, s, c->queue, c->qpes, n, npe, nbits, chunk, L.chunk, L.start, c->out); break;
    case 6: launch_deliver<6>(g, s, c->queue, c->qpes, n, npe, nbits, chunk, L.chunk, L.start, c->out); break;
    default: launch_deliver<8>(g, s, c->queue, c->qpes, n, npe, nbits, chunk, L.chunk, L.start, c->out); break;
  }
  int rc = (int)hipGetLastError();
  if (rc) return convey_fail_hip(rc, "convey_advance: launch");
  rc = (int)hipMemcpyAsync(c->starts.data(), L.start, (size_t)(npe + 1) * 4, hipMemcpyDeviceToHost, s);
  if (rc) return convey_fail_hip(rc, "convey_advance: reading the spans");
  c->pending = 0;
  // waits for the delivery; reports (and clears) PE numbers out of range,
  // whose items were dropped
  int64_t bad = 0;
  rc = bale_sync_errors((void*)s, &bad);
  std::fill(c->pulled.begin(), c->pulled.end(), 0);
  std::fill(c->last.begin(), c->last.end(), 0);
  c->unpulled = c->starts[(size_t)npe];
  if (rc == BALE_ERR_RANGE)
    return bale_set_error(BALE_ERR_RANGE, "convey_advance: %lld items named a PE outside 0..%d "
                          "and were dropped", (long long)bad, npe - 1);
  return rc;
}

extern "C" {

bale_convey* bale_convey_new(int64_t n_pe, int64_t capacity, size_t max_item_bytes) {
  if (n_pe < 1 || n_pe > kMaxPes) {
    bale_set_error(BALE_ERR_ARG, "convey_new: %lld PEs (1..%d on one GPU)", (long long)n_pe, kMaxPes);
    return nullptr;
  }
  if (capacity < 1 || capacity >= ((int64_t)1 << 31)) {
    bale_set_error(BALE_ERR_ARG, "convey_new: capacity %lld (1..2^31-1 items)", (long long)capacity);
    return nullptr;
  }
  if (!item_size_ok(max_item_bytes)) {
    bale_set_error(BALE_ERR_ARG, "convey_new: item size %zu (4, 8, 12, 16, 24 or 32 bytes)",
                   max_item_bytes);
    return nullptr;
  }
  Device* d = device_state();
  if (!d) {
    bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
    return nullptr;
  }
  auto* c = new bale_convey;
  c->device = d->id;
  c->npe = n_pe;
  c->capacity = capacity;
  c->max_item_bytes = max_item_bytes;
  const size_t bytes = (size_t)capacity * max_item_bytes;
  c->queue = BALE_DEV_ALLOC(bytes);
  c->qpes = (int64_t*)BALE_DEV_ALLOC((size_t)capacity * sizeof(int64_t));
  c->out = BALE_DEV_ALLOC(bytes);
  if (!c->queue || !c->qpes || !c->out) {
    dev_free(c->queue);
    dev_free(c->qpes);
    dev_free(c->out);
    delete c;
    bale_set_error(BALE_ERR_ALLOC, "convey_new: cannot allocate %zu bytes",
                   2 * bytes + (size_t)capacity * 8);
    return nullptr;
  }
  c->starts.assign((size_t)n_pe + 1, 0);
  c->pulled.assign((size_t)n_pe, 0);
  c->last.assign((size_t)n_pe, 0);
  return c;
}

int bale_convey_begin(bale_convey* c, size_t item_bytes) {
  if (!c) return bale_set_error(BALE_ERR_ARG, "convey: null conveyor");
  if (c->state != 0) return bale_set_error(BALE_ERR_ARG, "convey_begin: conveyor is not dormant");
  if (!item_size_ok(item_bytes) || item_bytes > c->max_item_bytes)
    return bale_set_error(BALE_ERR_ARG, "convey_begin: item size %zu (4, 8, 12, 16, 24 or 32, "
                          "at most %zu)", item_bytes, c->max_item_bytes);
  c->item_bytes = item_bytes;
  c->state = 1;
  return BALE_CONVEY_OK;
}

int bale_convey_push(bale_convey* c, const void* items, const int64_t* pes, int64_t n,
                     void* stream) {
  if (!c) return bale_set_error(BALE_ERR_ARG, "convey: null conveyor");
  if (c->state != 1) return bale_set_error(BALE_ERR_ARG, "convey_push: conveyor is not working");
  if (n < 0 || (n > 0 && (!items || !pes))) return bale_set_error(BALE_ERR_ARG, "convey_push: bad batch");
  if (n == 0) return BALE_CONVEY_OK;
  if (c->pending + n > c->capacity) return BALE_CONVEY_FAIL;  // full until the next delivery
  hipStream_t s = (hipStream_t)stream;
  int rc = (int)hipMemcpyAsync((char*)c->queue + (size_t)c->pending * c->item_bytes, items,
                               (size_t)n * c->item_bytes, hipMemcpyDeviceToDevice, s);
  if (!rc)
    rc = (int)hipMemcpyAsync(c->qpes + c->pending, pes, (size_t)n * sizeof(int64_t),
                             hipMemcpyDeviceToDevice, s);
  if (rc) return convey_fail_hip(rc, "convey_push");
  c->pending += n;
  return BALE_CONVEY_OK;
}

int bale_convey_advance(bale_convey* c, int done, void* stream) {
  if (!c) return bale_set_error(BALE_ERR_ARG, "convey: null conveyor");
  if (c->state == 2) return BALE_CONVEY_DONE;
  if (c->state != 1) return bale_set_error(BALE_ERR_ARG, "convey_advance: conveyor is not working");
  // a new delivery replaces the spans, so it waits until they are pulled
  if (c->pending > 0 && c->unpulled == 0) {
    int rc = convey_deliver(c, (hipStream_t)stream);
    if (rc != BALE_OK) return rc;
  }
  if (done && c->pending == 0 && c->unpulled == 0) {
    c->state = 2;
    return BALE_CONVEY_DONE;
  }
  return BALE_CONVEY_OK;
}

int bale_convey_pull(bale_convey* c, int64_t pe, const void** items, int64_t* count,
                     int64_t* from) {
  if (!c) return bale_set_error(BALE_ERR_ARG, "convey: null conveyor");
  if (c->state == 0) return bale_set_error(BALE_ERR_ARG, "convey_pull: conveyor is dormant");
  if (pe < 0 || pe >= c->npe) return bale_set_error(BALE_ERR_ARG, "convey_pull: PE %lld", (long long)pe);
  const int64_t lo = (int64_t)c->starts[(size_t)pe] + c->pulled[(size_t)pe];
  const int64_t hi = (int64_t)c->starts[(size_t)pe + 1];
  if (items) *items = (const char*)c->out + (size_t)lo * c->item_bytes;
  if (count) *count = hi - lo;
  if (from) *from = 0;  // one GPU is the only source
  c->last[(size_t)pe] = hi - lo;
  c->pulled[(size_t)pe] += hi - lo;
  c->unpulled -= hi - lo;
  return hi > lo ? BALE_CONVEY_OK : BALE_CONVEY_FAIL;
}

int bale_convey_unpull(bale_convey* c, int64_t pe, int64_t count) {
  if (!c) return bale_set_error(BALE_ERR_ARG, "convey: null conveyor");
  if (c->state == 0) return bale_set_error(BALE_ERR_ARG, "convey_unpull: conveyor is dormant");
  if (pe < 0 || pe >= c->npe) return bale_set_error(BALE_ERR_ARG, "convey_unpull: PE %lld", (long long)pe);
  if (count < 0 || count > c->last[(size_t)pe]) return BALE_CONVEY_FAIL;
  c->pulled[(size_t)pe] -= count;
  c->last[(size_t)pe] -= count;
  c->unpulled += count;
  return BALE_CONVEY_OK;
}

int bale_convey_reset(bale_convey* c) {
  if (!c) return bale_set_error(BALE_ERR_ARG, "convey: null conveyor");
  c->state = 0;
  c->pending = 0;
  c->unpulled = 0;
  std::fill(c->starts.begin(), c->starts.end(), 0);
  std::fill(c->pulled.begin(), c->pulled.end(), 0);
  std::fill(c->last.begin(), c->last.end(), 0);
  return BALE_CONVEY_OK;
}

int bale_convey_free(bale_convey* c) {
  if (!c) return BALE_CONVEY_OK;
  dev_free(c->queue);
  dev_free(c->qpes);
  dev_free(c->out);
  delete c;
  return BALE_CONVEY_OK;
}

int bale_memcpy_d2d(void* dst, const void* src, int64_t bytes, void* stream) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return bale_set_error(BALE_ERR_ARG, "memcpy: bad range");
  if (bytes == 0) return BALE_OK;
  int rc = (int)hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
  return rc ? convey_fail_hip(rc, "memcpy") : BALE_OK;
}

}  // extern "C"
