// partition.h -- bucket-by-destination partitioning shared by the bucketed
// histogram and the bucketed gather.
//
// This is the on-chip form of bale's aggregate-by-destination idea (exstack /
// conveyor: push an item into the send buffer of its destination, ship the
// buffer when it is full).  Here a "destination" is a bucket of 2^shift
// consecutive bins or table entries.  Each workgroup keeps one small LDS send
// buffer per bucket; an item is appended to its bucket's buffer, and a full
// buffer is shipped to the bucket's region of a global staging array with one
// coalesced write.  Regions are sized exactly by a counting pass, so the
// shipping cursor (one device atomic per buffer) never overflows.
//
//   pass 1  k_part_count   per-bucket totals (LDS histogram, one atomic per bucket)
//   pass 2  k_part_scan    exclusive scan -> region starts, cursors, segment list
//   pass 3  k_part_scatter LDS send buffers, coalesced flushes
#pragma once
#include "bale_internal.h"

namespace bale {

// Workspace layout of one partition (all 32-bit words, then the staging array).
struct PartLayout {
  uint32_t* count;     // [B]   items per bucket
  uint32_t* start;     // [B+1] region start per bucket
  uint32_t* cursor;    // [B]   shipping cursor per bucket
  uint32_t* segstart;  // [B+1] first segment of each bucket
  uint32_t* misc;      // [4]   [0] total segments, [1] work-queue head
  void* staged;        // [n]   entries grouped by bucket
};

// Carve `ws` into a PartLayout for B buckets and n entries of `esize` bytes.
inline size_t part_bytes(int64_t B, int64_t n, size_t esize) {
  size_t words = (size_t)(B + (B + 1) + B + (B + 1) + 4);
  size_t head = ((words * 4) + 255) & ~(size_t)255;
  return head + (size_t)n * esize + 256;
}
inline PartLayout part_carve(void* ws, int64_t B) {
  PartLayout L;
  uint32_t* w = (uint32_t*)ws;
  L.count = w;
  L.start = L.count + B;
  L.cursor = L.start + (B + 1);
  L.segstart = L.cursor + B;
  L.misc = L.segstart + (B + 1);
  size_t words = (size_t)(B + (B + 1) + B + (B + 1) + 4);
  size_t head = ((words * 4) + 255) & ~(size_t)255;
  L.staged = (char*)ws + head;
  return L;
}
inline size_t part_zero_bytes(int64_t B) { return (size_t)(B + (B + 1) + B + (B + 1) + 4) * 4; }

// ---- pass 1 -----------------------------------------------------------------
template <bool PACKED>
__global__ __launch_bounds__(1024) void k_part_count(const int64_t* __restrict__ idx, int64_t n,
                                                     uint64_t limit, int shift, int B,
                                                     int64_t chunk, uint32_t* __restrict__ count,
                                                     ErrWords* err) {
  extern __shared__ uint32_t hist[];
  for (int b = threadIdx.x; b < B; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  int64_t lo = (int64_t)blockIdx.x * chunk;
  int64_t hi = lo + chunk < n ? lo + chunk : n;
  unsigned bad_r = 0, bad_p = 0;
  for (int64_t e = lo + threadIdx.x; e < hi; e += blockDim.x) {
    uint64_t j;
    if (decode_index<PACKED>(__builtin_nontemporal_load(idx + e), limit, j, bad_r, bad_p))
      atomicAdd(&hist[j >> shift], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < B; b += blockDim.x)
    if (hist[b]) atomicAdd(&count[b], hist[b]);
  flush_errors(err, bad_r, bad_p);
}

// ---- pass 2 (one workgroup of 1024) -------------------------------------------
static __global__ __launch_bounds__(1024) void k_part_scan(const uint32_t* __restrict__ count, int B,
                                                    uint32_t seg_len, uint32_t* __restrict__ start,
                                                    uint32_t* __restrict__ cursor,
                                                    uint32_t* __restrict__ segstart,
                                                    uint32_t* __restrict__ misc) {
  __shared__ uint32_t part_a[1024], part_b[1024];
  // each thread owns a contiguous run of buckets
  int per = (B + 1023) / 1024;
  int b0 = threadIdx.x * per;
  uint32_t sa = 0, sb = 0;
  for (int b = b0; b < b0 + per && b < B; b++) {
    sa += count[b];
    sb += (count[b] + seg_len - 1) / seg_len;
  }
  part_a[threadIdx.x] = sa;
  part_b[threadIdx.x] = sb;
  __syncthreads();
  // Hillis-Steele inclusive scan over the 1024 partials
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t va = threadIdx.x >= off ? part_a[threadIdx.x - off] : 0;
    uint32_t vb = threadIdx.x >= off ? part_b[threadIdx.x - off] : 0;
    __syncthreads();
    part_a[threadIdx.x] += va;
    part_b[threadIdx.x] += vb;
    __syncthreads();
  }
  uint32_t ra = threadIdx.x ? part_a[threadIdx.x - 1] : 0;
  uint32_t rb = threadIdx.x ? part_b[threadIdx.x - 1] : 0;
  for (int b = b0; b < b0 + per && b < B; b++) {
    start[b] = ra;
    cursor[b] = ra;
    segstart[b] = rb;
    ra += count[b];
    rb += (count[b] + seg_len - 1) / seg_len;
  }
  if (threadIdx.x == 1023) {
    start[B] = part_a[1023];
    segstart[B] = part_b[1023];
    misc[0] = part_b[1023];
    misc[1] = 0;
  }
}

// ---- pass 3 -----------------------------------------------------------------
// MakeEntry(e, j, lo_bits) -> entry value stored for element e with decoded
// index j (lo_bits = j & ((1<<shift)-1)).
template <bool PACKED, typename E, typename MakeEntry>
__device__ inline void part_scatter_body(const int64_t* __restrict__ idx, int64_t n,
                                         uint64_t limit, int shift, int B, int S, int64_t chunk,
                                         uint32_t* __restrict__ cursor, E* __restrict__ staged,
                                         MakeEntry make) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* fill = (uint32_t*)smem;               // [B]
  uint32_t* list = fill + B;                       // [B] buckets to ship this step
  uint32_t* nlist = list + B;                      // [1]
  E* buf = (E*)(smem + (((size_t)(2 * B + 1) * 4 + 15) & ~(size_t)15));  // [B*S]
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = __lane_id(), wave = tid >> 6, nwaves = nthr >> 6;
  for (int b = tid; b < B; b += nthr) fill[b] = 0;
  if (tid == 0) *nlist = 0;
  __syncthreads();
  const uint64_t mask = (1ULL << shift) - 1;
  int64_t lo = (int64_t)blockIdx.x * chunk;
  int64_t hi = lo + chunk < n ? lo + chunk : n;

  // ship the first `cnt` entries of bucket b with one wave (lanes stride the entries)
  auto ship = [&](int b, uint32_t cnt) {
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(&cursor[b], cnt);
    pos = __shfl(pos, 0);
    for (uint32_t k = lane; k < cnt; k += 64) staged[pos + k] = buf[(size_t)b * S + k];
  };

  for (int64_t base = lo; base < hi; base += nthr) {
    int64_t e = base + tid;
    bool pending = false;
    int b = 0;
    E ent{};
    if (e < hi) {
      uint64_t j;
      unsigned br = 0, bp = 0;
      if (decode_index<PACKED>(__builtin_nontemporal_load(idx + e), limit, j, br, bp)) {
        b = (int)(j >> shift);
        ent = make(e, j, j & mask);
        pending = true;
      }
    }
    while (__syncthreads_or(pending)) {
      if (pending) {
        uint32_t slot = atomicAdd(&fill[b], 1u);
        if (slot < (uint32_t)S) {
          buf[(size_t)b * S + slot] = ent;
          pending = false;
          if (slot == (uint32_t)S - 1) list[atomicAdd(nlist, 1u)] = (uint32_t)b;
        }
      }
      __syncthreads();
      uint32_t nl = *nlist;
      for (uint32_t k = wave; k < nl; k += nwaves) ship((int)list[k], (uint32_t)S);
      __syncthreads();
      for (uint32_t k = tid; k < nl; k += nthr) fill[list[k]] = 0;
      if (tid == 0) *nlist = 0;
      // the next __syncthreads_or orders these resets before any new append
    }
  }
  // ship what is left in every buffer
  __syncthreads();
  for (int b = wave; b < B; b += nwaves) {
    uint32_t f = fill[b];
    if (f) ship(b, f < (uint32_t)S ? f : (uint32_t)S);
  }
}

// LDS bytes the scatter pass needs for B buckets of S entries of `esize` bytes.
inline size_t part_scatter_lds(int B, int S, size_t esize) {
  return (((size_t)(2 * B + 1) * 4 + 15) & ~(size_t)15) + (size_t)B * S * esize;
}
// Largest power-of-two buffer depth S in [4, 512] that fits `budget` bytes.
inline int part_pick_depth(int B, size_t esize, size_t budget) {
  int S = 512;
  while (S > 4 && part_scatter_lds(B, S, esize) > budget) S >>= 1;
  return part_scatter_lds(B, S, esize) <= budget ? S : 0;
}

}  // namespace bale
