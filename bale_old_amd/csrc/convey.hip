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
#include "partition.h"

#include <algorithm>
#include <vector>

namespace bale {

constexpr int kMaxPes = 256;   // virtual PEs: 8 ballots rank a tile's items
constexpr int kTile = 1024;    // items ranked per step (one per thread)
constexpr int kWaves = kTile / 64;

// Item copies with the widest aligned vector the item size allows
// (16-byte pieces for 16/32-byte items, 8-byte for 8/24, 4-byte for 4/12).
template <int W>
__device__ inline void item_load(const uint32_t* __restrict__ src, size_t i, uint32_t* v) {
  if constexpr (W % 4 == 0) {
    const u32x4* p = (const u32x4*)(src + i * W);
#pragma unroll
    for (int k = 0; k < W / 4; k++) {
      const u32x4 x = p[k];
      v[4 * k] = x.x; v[4 * k + 1] = x.y; v[4 * k + 2] = x.z; v[4 * k + 3] = x.w;
    }
  } else if constexpr (W % 2 == 0) {
    const uint2* p = (const uint2*)(src + i * W);
#pragma unroll
    for (int k = 0; k < W / 2; k++) {
      const uint2 x = p[k];
      v[2 * k] = x.x; v[2 * k + 1] = x.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < W; k++) v[k] = src[i * W + k];
  }
}
template <int W>
__device__ inline void item_store(uint32_t* __restrict__ dst, size_t i, const uint32_t* v) {
  if constexpr (W % 4 == 0) {
    u32x4* p = (u32x4*)(dst + i * W);
#pragma unroll
    for (int k = 0; k < W / 4; k++) p[k] = u32x4{v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
  } else if constexpr (W % 2 == 0) {
    uint2* p = (uint2*)(dst + i * W);
#pragma unroll
    for (int k = 0; k < W / 2; k++) p[k] = make_uint2(v[2 * k], v[2 * k + 1]);
  } else {
#pragma unroll
    for (int k = 0; k < W; k++) dst[i * W + k] = v[k];
  }
}

// Items per thread per tile: as many as the LDS staging allows (4 for items
// up to 16 bytes, 2 above).
template <int W>
constexpr int deliver_u() { return W <= 4 ? 4 : 2; }

// LDS of k_convey_deliver<W>: per virtual wave (U x 16 of them) and PE a
// count, then an offset; PE totals, tile starts and output cursors; the
// staged items and their PEs.
template <int W>
inline size_t deliver_lds(int npe) {
  constexpr int U = deliver_u<W>();
  size_t words = (size_t)U * kWaves * npe + 3 * (size_t)kMaxPes + 1;
  words = (words + 3) & ~(size_t)3;  // 16-byte alignment for the staged items
  return words * 4 + (size_t)U * kTile * W * 4 + (size_t)U * kTile * 2;
}

// Stable delivery of chunk blockIdx.x.  W: 32-bit words per item.  A tile is
// U x 1024 consecutive items, item base + u * 1024 + tid on thread tid; push
// order within the tile is (u, wave, lane), so a "virtual wave" u * 16 + wave
// ranks its items with ballots and the virtual waves are prefixed in order.
// rows[c][p]: items of PE p in chunks before c; start[p]: PE p's first slot.
template <int W>
__global__ __launch_bounds__(kTile) void k_convey_deliver(const uint32_t* __restrict__ items,
                                                          const int64_t* __restrict__ pes,
                                                          int64_t n, int npe, int nbits,
                                                          int64_t chunk,
                                                          const uint32_t* __restrict__ rows,
                                                          const uint32_t* __restrict__ start,
                                                          uint32_t* __restrict__ out) {
  constexpr int U = deliver_u<W>(), V = U * kWaves;
  extern __shared__ __align__(16) uint32_t lds[];
  uint32_t* wcnt = lds;                           // [V * npe]: count, then offset
  uint32_t* tot = wcnt + (size_t)V * npe;         // [kMaxPes] tile items per PE
  uint32_t* tstart = tot + kMaxPes;               // [kMaxPes + 1] first staged slot per PE
  uint32_t* cursor = tstart + kMaxPes + 1;        // [kMaxPes] next output slot per PE
  size_t head = ((size_t)V * npe + 3 * (size_t)kMaxPes + 1 + 3) & ~(size_t)3;
  uint32_t* stage = lds + head;                   // [U * kTile * W]
  uint16_t* sdst = (uint16_t*)(stage + (size_t)U * kTile * W);  // [U * kTile]
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const int64_t c = blockIdx.x;
  const int64_t lo = c * chunk, hi = lo + chunk < n ? lo + chunk : n;
  for (int p = tid; p < npe; p += kTile) cursor[p] = start[p] + rows[(size_t)c * npe + p];
  const uint64_t below = lane ? (~0ULL >> (64 - lane)) : 0ULL;  // lanes before this one
  // a tile's PE numbers and items are loaded while the previous tile is
  // ranked and written
  int64_t pn[U];
  uint32_t vn[U][W];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int64_t e = lo + (int64_t)u * kTile + tid;
    pn[u] = -1;
    if (e < hi) {
      pn[u] = pes[e];
      item_load<W>(items, (size_t)e, vn[u]);
    }
  }
  for (int64_t base = lo; base < hi; base += (int64_t)U * kTile) {
    int d[U];
    uint32_t v[U][W];
#pragma unroll
    for (int u = 0; u < U; u++) {
      // npe: nothing to deliver (past the end, or PE out of range)
      d[u] = (uint64_t)pn[u] < (uint64_t)npe ? (int)pn[u] : npe;
#pragma unroll
      for (int k = 0; k < W; k++) v[u][k] = vn[u][k];
      const int64_t e = base + (int64_t)(U + u) * kTile + tid;
      pn[u] = -1;
      if (e < hi) {
        pn[u] = pes[e];
        item_load<W>(items, (size_t)e, vn[u]);
      }
    }
    for (int k = tid; k < V * npe; k += kTile) wcnt[k] = 0;
    __syncthreads();
    // lanes of this wave with the same PE, by one ballot per bit of d
    uint32_t rank[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool has = d[u] < npe;
      uint64_t peers = __ballot(has);
      for (int bit = 0; bit < nbits; bit++) {
        const bool one = (d[u] >> bit) & 1;
        const uint64_t b = __ballot(one);
        peers &= one ? b : ~b;
      }
      rank[u] = (uint32_t)__popcll(peers & below);
      if (has && rank[u] == 0) wcnt[(u * kWaves + wave) * npe + d[u]] = (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // offsets of each virtual wave's items within the PE's part of the tile
    for (int p = tid; p < npe; p += kTile) {
      uint32_t run = 0;
      for (int w = 0; w < V; w++) {
        const uint32_t t = wcnt[w * npe + p];
        wcnt[w * npe + p] = run;
        run += t;
      }
      tot[p] = run;
    }
    __syncthreads();
    if (wave == 0) {
      // exclusive scan of tot over the PEs, 4 per lane
      uint32_t sv[4], sum = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int p = lane * 4 + k;
        sv[k] = p < npe ? tot[p] : 0u;
        sum += sv[k];
      }
      uint32_t incl = sum;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
      }
      uint32_t run = incl - sum;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int p = lane * 4 + k;
        if (p < npe) tstart[p] = run;
        run += sv[k];
      }
      if (lane == 63) tstart[npe] = incl;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++)
      if (d[u] < npe) {
        const uint32_t q = tstart[d[u]] + wcnt[(u * kWaves + wave) * npe + d[u]] + rank[u];
        item_store<W>(stage, q, v[u]);
        sdst[q] = (uint16_t)d[u];
      }
    __syncthreads();
    // staged order is PE order, push order within a PE: each PE's run goes
    // to consecutive output slots
    const uint32_t ntile = tstart[npe];
    for (uint32_t q = tid; q < ntile; q += kTile) {
      const int p = sdst[q];
      const size_t g = (size_t)cursor[p] + q - tstart[p];
      uint32_t w[W];
      item_load<W>(stage, (size_t)q, w);
      item_store<W>(out, g, w);
    }
    __syncthreads();
    for (int p = tid; p < npe; p += kTile) cursor[p] += tot[p];
  }
}

template <int W>
static void launch_deliver(dim3 g, hipStream_t s, const void* items, const int64_t* pes, int64_t n,
                           int npe, int nbits, int64_t chunk, const uint32_t* rows,
                           const uint32_t* start, void* out) {
  const size_t lds = deliver_lds<W>(npe);
  allow_lds(k_convey_deliver<W>, lds);
  hipLaunchKernelGGL(k_convey_deliver<W>, g, dim3(kTile), lds, s, (const uint32_t*)items, pes, n,
                     npe, nbits, chunk, rows, start, (uint32_t*)out);
}

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
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus, ceil_div(n, kChunkAlign)));
  const int64_t chunk = ceil_div(ceil_div(n, C), kChunkAlign) * kChunkAlign;
  C = ceil_div(n, chunk);
  void* ws = workspace(d, part_bytes(C, npe, 0, 1), s);
  if (!ws) return bale_set_error(BALE_ERR_ALLOC, "convey_advance: workspace");
  PartLayout L = part_carve(ws, C, npe);
  hipLaunchKernelGGL((k_part_count<false, false, false>), dim3((unsigned)C), dim3(1024),
                     (size_t)npe * 4, s, (const int64_t*)c->qpes, n, (uint64_t)npe, 0, npe, chunk,
                     L.chunk, (const uint32_t*)nullptr, d->err, (const uint32_t*)nullptr);
  hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)npe), dim3(256), 0, s, L.chunk, (int)C, npe,
                     L.total);
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, npe, 1u << 30, L.start,
                     L.segstart, L.misc);
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
  if (hipMalloc(&c->queue, bytes) != hipSuccess ||
      hipMalloc((void**)&c->qpes, (size_t)capacity * sizeof(int64_t)) != hipSuccess ||
      hipMalloc(&c->out, bytes) != hipSuccess) {
    (void)hipFree(c->queue);
    (void)hipFree(c->qpes);
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
  (void)hipFree(c->queue);
  (void)hipFree(c->qpes);
  (void)hipFree(c->out);
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
