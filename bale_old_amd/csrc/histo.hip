// histo.hip -- histogram scatter-add counts[index[i]] += 1 on gfx950.
//
// Three strategies (the reference's models, re-thought for one chip):
//   DIRECT  one device-scope atomic per update, the AGI model
//           (apps/histo_src/histo_agi.upc:61-66, libgetput/libgetput.upc:48-60).
//           With `hot`, bins a sample sees twice are counted in an LDS hash
//           table per workgroup first.
//   LDS     per-workgroup sub-histograms of a bin tile in LDS, flushed with one
//           device atomic per nonzero bin (updates aggregated before they leave
//           the CU, as exstack2 aggregates before sending).
//   BUCKET  partition the updates by bin range through LDS send buffers
//           (partition.h), then one LDS histogram per bucket segment, as
//           exstack/conveyor route each update to its owner, which then applies
//           it locally (apps/histo_src/histo_exstack.upc:64-76,
//           histo_conveyor.upc:72-82).  The paged plan reads the index once:
//           each bucket's send buffer is a ring in LDS whose completed halves
//           are shipped as pages of the workgroup's own pool, and a small pass
//           lists each bucket's pages for the apply pass.
// All three add into the existing counts (the drivers accumulate across
// models, apps/histo_src/histo.upc:208-227).
#include <type_traits>

#include "bale_internal.h"
#include "partition.h"
#include "pk_count.h"

namespace bale {

// ---- wave-level merging of equal bins -------------------------------------
// key = decoded bin, or ~0 for a lane with nothing to add.  Returns the count
// this lane must add (0 when another lane adds for it).  Every lane of the
// wave must call it together.
template <int MERGE>
__device__ inline uint32_t wave_merge(uint64_t key, bool ok) {
  if (MERGE == 0) return ok ? 1u : 0u;
  const int lane = __lane_id();
  uint32_t mine = 0;
  if (MERGE == 1) {
    // runs of equal keys in adjacent lanes: the first lane of a run adds its length
    uint64_t prev = __shfl_up(key, 1);
    bool change = (lane == 0) || (prev != key);
    uint64_t chg = __ballot(change);
    if (ok && change) {
      uint64_t later = chg & ~((lane == 63) ? ~0ULL : ((2ULL << lane) - 1));
      int end = later ? __builtin_ctzll(later) : 64;
      mine = (uint32_t)(end - lane);
    }
    return mine;
  }
  // MERGE == 2: every distinct key in the wave is added once by a leader lane
  uint64_t active = __ballot(ok);
  while (active) {
    int leader = __builtin_ctzll(active);
    uint64_t lk = __shfl(key, leader);
    uint64_t m = __ballot(ok && key == lk) & active;
    if (lane == leader) mine = (uint32_t)__builtin_popcountll(m);
    active &= ~m;
  }
  return mine;
}

// ---- DIRECT ---------------------------------------------------------------
template <int U, typename CT, bool PACKED, int MERGE>
__global__ __launch_bounds__(1024) void k_histo_direct(const int64_t* __restrict__ idx, int64_t n,
                                                       CT* __restrict__ counts, uint64_t nbins,
                                                       ErrWords* err) {
  const int B = blockDim.x;
  int64_t base = (int64_t)blockIdx.x * B * U + threadIdx.x;
  int64_t raw[U];
#pragma unroll
  for (int k = 0; k < U; k++) {
    int64_t e = base + (int64_t)k * B;
    raw[k] = e < n ? __builtin_nontemporal_load(idx + e) : 0;
  }
  unsigned bad_r = 0, bad_p = 0;
#pragma unroll
  for (int k = 0; k < U; k++) {
    int64_t e = base + (int64_t)k * B;
    uint64_t j = ~0ULL;
    bool ok = false;
    if (e < n) ok = decode_index<PACKED>(raw[k], nbins, j, bad_r, bad_p);
    uint32_t c = wave_merge<MERGE>(ok ? j : ~0ULL, ok);
    if (c) atomic_add_dev(counts + j, (CT)c);
  }
  flush_errors(err, bad_r, bad_p);
}

// DIRECT with hot bins: a bounded grid walks the input; hot bins are counted in
// the workgroup's LDS hash table and flushed once at the end.
template <int U, typename CT, bool PACKED>
__global__ __launch_bounds__(1024) void k_histo_direct_hot(const int64_t* __restrict__ idx, int64_t n,
                                                           CT* __restrict__ counts, uint64_t nbins,
                                                           const uint32_t* __restrict__ hot_g,
                                                           ErrWords* err) {
  __shared__ uint32_t hkeys[kHotSlots];
  __shared__ uint32_t hcnt[kHotSlots];
  const int B = blockDim.x;
  for (int k = threadIdx.x; k < kHotSlots; k += B) {
    hkeys[k] = hot_g[k];
    hcnt[k] = 0;
  }
  __syncthreads();
  unsigned bad_r = 0, bad_p = 0;
  const int64_t step = (int64_t)gridDim.x * B * U;
  for (int64_t base = (int64_t)blockIdx.x * B * U + threadIdx.x; base < n; base += step) {
    int64_t raw[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t e = base + (int64_t)k * B;
      raw[k] = e < n ? __builtin_nontemporal_load(idx + e) : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t e = base + (int64_t)k * B;
      uint64_t j;
      if (e < n && decode_index<PACKED>(raw[k], nbins, j, bad_r, bad_p)) {
        int hs = hot_find(hkeys, (uint32_t)j);
        if (hs >= 0) atomicAdd(&hcnt[hs], 1u);
        else atomic_add_dev(counts + j, (CT)1);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kHotSlots; k += B)
    if (hcnt[k]) atomic_add_dev(counts + hkeys[k], (CT)hcnt[k]);
  flush_errors(err, bad_r, bad_p);
}

// ---- LDS ------------------------------------------------------------------
// grid = (chunks, tiles); tile t covers bins [t<<tile_bits, +width).
template <int U, typename CT, bool PACKED, int MERGE>
__global__ __launch_bounds__(1024) void k_histo_lds(const int64_t* __restrict__ idx, int64_t n,
                                                    CT* __restrict__ counts, uint64_t nbins,
                                                    int tile_bits, int64_t chunk, ErrWords* err) {
  extern __shared__ uint32_t sh[];
  const int B = blockDim.x;
  const uint64_t lo_bin = (uint64_t)blockIdx.y << tile_bits;
  const uint64_t width64 = nbins - lo_bin < (1ULL << tile_bits) ? nbins - lo_bin : (1ULL << tile_bits);
  const uint32_t width = (uint32_t)width64;
  for (uint32_t b = threadIdx.x; b < width; b += B) sh[b] = 0;
  __syncthreads();
  int64_t lo = (int64_t)blockIdx.x * chunk;
  int64_t hi = lo + chunk < n ? lo + chunk : n;
  unsigned bad_r = 0, bad_p = 0;
  // the loop bound is uniform across the workgroup, so every lane of a wave
  // reaches wave_merge together
  for (int64_t base = lo; base < hi; base += (int64_t)B * U) {
    int64_t raw[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t e = base + (int64_t)k * B + threadIdx.x;
      raw[k] = e < hi ? __builtin_nontemporal_load(idx + e) : 0;
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      int64_t e = base + (int64_t)k * B + threadIdx.x;
      uint64_t j = ~0ULL;
      bool ok = false;
      if (e < hi) ok = decode_index<PACKED>(raw[k], nbins, j, bad_r, bad_p);
      ok = ok && (j - lo_bin) < width64;
      uint32_t c = wave_merge<MERGE>(ok ? j : ~0ULL, ok);
      if (c) atomicAdd(&sh[j - lo_bin], c);
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < width; b += B) {
    uint32_t c = sh[b];
    if (c) atomic_add_dev(counts + lo_bin + b, (CT)c);
  }
  // every tile decodes every index; report bad ones once
  if (blockIdx.y == 0) flush_errors(err, bad_r, bad_p);
}

// ---- BUCKET ---------------------------------------------------------------
template <bool PACKED, bool HOT, bool RUNS, typename CT>
__global__ __launch_bounds__(1024) void k_histo_scatter(const int64_t* __restrict__ idx, int64_t n,
                                                        uint64_t nbins, int shift, int B, int S,
                                                        int64_t chunk,
                                                        const uint32_t* __restrict__ rows,
                                                        const uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ hot_g,
                                                        uint16_t* __restrict__ staged,
                                                        CT* __restrict__ counts,
                                                        const uint32_t* __restrict__ flags,
                                                        int S_cold) {
  part_scatter_adaptive<PACKED, HOT, RUNS, uint16_t>(
      blockIdx.x, idx, n, nbins, shift, B, S, S_cold, chunk, rows, start, hot_g,
      [staged](uint32_t pos, uint16_t v) { staged[pos] = v; },
      [](int64_t, uint64_t, uint64_t low) { return (uint16_t)low; },
      [counts](uint32_t bin, uint32_t c) { atomic_add_dev(counts + bin, (CT)c); },
      [](int64_t) {}, flags);
}

// Add the 8 16-bit bin offsets packed in x to the LDS histogram.  RUNS: equal
// neighbours (clustered inputs) are combined into one add.
template <bool RUNS>
__device__ inline void lds_add8(uint32_t* sh, u32x4 x) {
  uint32_t v[8] = {x.x & 0xffff, x.x >> 16, x.y & 0xffff, x.y >> 16,
                   x.z & 0xffff, x.z >> 16, x.w & 0xffff, x.w >> 16};
  if (!RUNS) {
#pragma unroll
    for (int i = 0; i < 8; i++) atomicAdd(&sh[v[i]], 1u);
    return;
  }
  uint32_t cur = v[0], cnt = 1;
#pragma unroll
  for (int i = 1; i < 8; i++) {
    if (v[i] == cur) {
      cnt++;
    } else {
      atomicAdd(&sh[cur], cnt);
      cur = v[i];
      cnt = 1;
    }
  }
  atomicAdd(&sh[cur], cnt);
}

// Packed tile for 2^16-bin buckets: two 16-bit counters per LDS word (bin k in
// half k & 1 of word k >> 1), 128 KiB for the whole bucket.  The words only
// ever receive atomic adds; each lane reads from the value its own add returned
// what that add owes the word's two bins when a half wrapped (pk_owed), so the
// counts are exact whatever order the adds land in.  The owed amounts (2^16
// per wrap, less one per carry the high half took from the low one) are
// counted in LDS: a bit per bin says the bin is owed something and a small
// hash table (bin -> owed, mod 2^32) holds the amounts.  A wrap takes 2^16
// adds to one bin, so a segment of m updates notes at most m / 2^15 bins and
// the table (kTallySlots) fills only for segments of more than 2^25 updates;
// then the amount goes to the device counter with an atomic, and `flag` makes
// the flush atomic as well.
constexpr int kTallyBits = 10;
constexpr uint32_t kTallySlots = 1u << kTallyBits;
struct PkTally {
  uint32_t* flag;  // some owed amount went to a device counter
  uint32_t* bits;  // [2^tb / 32]
  uint32_t* keys;  // [kTallySlots], kHotEmpty when free
  uint32_t* cnts;  // [kTallySlots] owed, mod 2^32
  uint32_t width;  // bins of this tile
};
__device__ inline uint32_t tally_hash(uint32_t bin) { return (bin * 0x9E3779B1u) >> (32 - kTallyBits); }
__device__ inline bool pk_tally_note(const PkTally& t, uint32_t bin, uint32_t amount) {
  uint32_t h = tally_hash(bin);
  for (uint32_t probe = 0; probe < kTallySlots; probe++) {
    const uint32_t prev = atomicCAS(&t.keys[h], kHotEmpty, bin);
    if (prev == kHotEmpty || prev == bin) {
      atomicAdd(&t.cnts[h], amount);
      atomicOr(&t.bits[bin >> 5], 1u << (bin & 31u));
      return true;
    }
    h = (h + 1u) & (kTallySlots - 1u);
  }
  return false;
}
// What a bin whose bit is set is owed (after the segment's barrier).
__device__ inline uint32_t pk_tally_of(const PkTally& t, uint32_t bin) {
  uint32_t h = tally_hash(bin);
  for (uint32_t probe = 0; probe < kTallySlots; probe++) {
    if (t.keys[h] == bin) return t.cnts[h];
    h = (h + 1u) & (kTallySlots - 1u);
  }
  return 0;
}
// LDS words after a packed tile of 2^tb bins: tally bits and table.
__host__ __device__ inline uint32_t pk_tally_words(int tb) { return (1u << tb) / 32u + 2u * kTallySlots; }

template <typename CT>
__device__ inline void pk_tally_add(const PkTally& t, uint32_t bin, uint32_t amount, CT* dst) {
  if (!pk_tally_note(t, bin, amount)) {
    atomic_add_dev(dst + bin, (CT)(int32_t)amount);
    *t.flag = 1u;
  }
}
// Account for an add of inc to bin that returned old (pk_owed).  The high half
// of a tile's last word has no bin when the width is odd: its carries are
// never read.
template <typename CT>
__device__ inline void pk_account(uint32_t bin, uint32_t inc, uint32_t old, CT* dst,
                                  const PkTally& t) {
  if (__builtin_expect(!pk_wrapped(old, bin & 1u, inc), 1)) return;  // one compare per add
  const PkOwed o = pk_owed(old, bin & 1u, inc);
  if (o.own) pk_tally_add(t, bin, o.own, dst);
  if (o.high && (bin | 1u) < t.width) pk_tally_add(t, bin | 1u, o.high, dst);
}
// Zero a packed tile's tally bits and table.
__device__ inline void pk_tally_clear(const PkTally& t, int tb, int tid, int nthr) {
  for (uint32_t k = tid; k < (1u << tb) / 32u; k += nthr) t.bits[k] = 0;
  for (uint32_t k = tid; k < kTallySlots; k += nthr) {
    t.keys[k] = kHotEmpty;
    t.cnts[k] = 0;
  }
}

// lds_add8 for the packed tile; all of a call's atomics are issued before
// their results are checked.
template <bool RUNS, typename CT>
__device__ inline void lds_add8_pk(uint32_t* sh, u32x4 x, CT* dst, const PkTally& tally) {
  uint32_t v[8] = {x.x & 0xffff, x.x >> 16, x.y & 0xffff, x.y >> 16,
                   x.z & 0xffff, x.z >> 16, x.w & 0xffff, x.w >> 16};
  uint32_t inc[8], old[8];
#pragma unroll
  for (int i = 0; i < 8; i++) inc[i] = 1;
  if (RUNS) {
    // a run's count goes on its first entry; the others add nothing
#pragma unroll
    for (int i = 6; i >= 0; i--)
      if (v[i] == v[i + 1]) {
        inc[i] += inc[i + 1];
        inc[i + 1] = 0;
      }
  }
#pragma unroll
  for (int i = 0; i < 8; i++)
    old[i] = inc[i] ? atomicAdd(&sh[v[i] >> 1], inc[i] << ((v[i] & 1u) * 16u)) : 0u;
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (inc[i]) pk_account(v[i], inc[i], old[i], dst, tally);
}

template <typename CT>
__device__ inline void lds_add1_pk(uint32_t* sh, uint32_t bin, CT* dst, const PkTally& tally) {
  const uint32_t old = atomicAdd(&sh[bin >> 1], 1u << ((bin & 1u) * 16u));
  pk_account(bin, 1u, old, dst, tally);
}

// Add an LDS tile of counts (32-bit, or PK: two 16-bit per word) into
// dst[0..width).  Not plain: a device atomic per nonzero count.  Plain (the
// workgroup owns these bins): read-modify-write of four consecutive counters
// per thread per step with four steps' loads in flight, so the counter traffic
// streams instead of waiting on one load at a time.
typedef int i32x4 __attribute__((ext_vector_type(4)));
template <bool PK>
__device__ inline uint32_t tile_count(const uint32_t* sh, uint32_t k, const PkTally& t) {
  if (!PK) return sh[k];
  const uint32_t owed = ((t.bits[k >> 5] >> (k & 31u)) & 1u) ? pk_tally_of(t, k) : 0u;
  return pk_read(sh[k >> 1], k & 1u, owed);
}
// four consecutive counters as 16-byte vectors
struct Cnt4_64 { i64x2 a, b; };
__device__ inline Cnt4_64 load4(const int64_t* p) { return {*(const i64x2*)p, *(const i64x2*)(p + 2)}; }
__device__ inline void store4(int64_t* p, Cnt4_64 v, const uint32_t* c) {
  v.a.x += c[0]; v.a.y += c[1]; v.b.x += c[2]; v.b.y += c[3];
  *(i64x2*)p = v.a;
  *(i64x2*)(p + 2) = v.b;
}
__device__ inline i32x4 load4(const int32_t* p) { return *(const i32x4*)p; }
__device__ inline void store4(int32_t* p, i32x4 v, const uint32_t* c) {
  v.x += (int)c[0]; v.y += (int)c[1]; v.z += (int)c[2]; v.w += (int)c[3];
  *(i32x4*)p = v;
}
template <typename CT, bool PK>
__device__ inline void flush_tile(const uint32_t* sh, uint32_t width, CT* dst, bool plain, int tid,
                                  int nthr, const PkTally& sp) {
  if (!plain) {
    for (uint32_t k = tid; k < width; k += nthr) {
      const uint32_t c = tile_count<PK>(sh, k, sp);
      if (c) atomic_add_dev(dst + k, (CT)c);
    }
    return;
  }
  constexpr int F = 4;
  const uint32_t w4 = (((uintptr_t)dst & 15u) == 0) ? (width & ~3u) : 0u;
  const uint32_t step = (uint32_t)nthr * 4u;
  for (uint32_t g0 = (uint32_t)tid * 4u; g0 < w4; g0 += step * F) {
    uint32_t c[F][4];
    decltype(load4(dst)) v[F];
    bool any[F];
#pragma unroll
    for (int f = 0; f < F; f++) {
      const uint32_t k = g0 + f * step;
#pragma unroll
      for (int i = 0; i < 4; i++) c[f][i] = k < w4 ? tile_count<PK>(sh, k + i, sp) : 0u;
      any[f] = (c[f][0] | c[f][1] | c[f][2] | c[f][3]) != 0;
    }
    // all loads first, then the stores
#pragma unroll
    for (int f = 0; f < F; f++)
      if (any[f]) v[f] = load4(dst + g0 + f * step);
#pragma unroll
    for (int f = 0; f < F; f++)
      if (any[f]) store4(dst + g0 + f * step, v[f], c[f]);
  }
  for (uint32_t k = w4 + tid; k < width; k += nthr) {
    const uint32_t c = tile_count<PK>(sh, k, sp);
    if (c) dst[k] += (CT)c;
  }
}

// Persistent: each workgroup takes segments (<= seg_len entries of one bucket)
// from a queue, histograms them in LDS and adds the tile into counts.  A bucket
// with one segment is owned by one workgroup, so its flush is a plain add.
// PK: the bucket's tile is packed 16-bit counters (2^16-bin buckets).
template <typename CT, bool RUNS, bool PK>
__global__ __launch_bounds__(1024) void k_histo_apply(const uint16_t* __restrict__ staged,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ segstart,
                                                      uint32_t* __restrict__ misc, int B,
                                                      int shift, uint64_t nbins,
                                                      CT* __restrict__ counts,
                                                      const uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t sh[];
  __shared__ uint32_t s_work[4];
  __shared__ uint32_t s_owed_dev;  // PK: some owed amount of this segment went to a device counter
  // PK: tally bits and table after the tile (2^(shift-1) words)
  uint32_t* const tally_base = sh + (PK ? 1u << (shift - 1) : 0u);
  const bool merge = RUNS && (!flags || (*flags & kFlagRunsSeen));
  const int tid = threadIdx.x, nthr = blockDim.x;
  const uint32_t total = misc[0], seg_len = misc[kSegWord];
  while (true) {
    if (tid == 0) {
      uint32_t w = atomicAdd(&misc[1], 1u);
      uint32_t b = 0;
      if (w < total) {
        // last bucket whose first segment is <= w
        int lo = 0, hi = B - 1;
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (segstart[mid] <= w) lo = mid; else hi = mid - 1;
        }
        b = (uint32_t)lo;
      }
      s_work[0] = w;
      s_work[1] = b;
      s_owed_dev = 0;
    }
    __syncthreads();
    const uint32_t w = s_work[0], b = s_work[1];
    if (w >= total) break;
    const uint32_t nseg = segstart[b + 1] - segstart[b];
    const uint32_t s = w - segstart[b];
    const uint32_t e0 = start[b] + s * seg_len;
    const uint32_t e1 = min(start[b + 1], e0 + seg_len);
    const uint64_t lo_bin = (uint64_t)b << shift;
    const uint32_t width = (uint32_t)min((uint64_t)1 << shift, nbins - lo_bin);
    const uint32_t words = PK ? (width + 1u) >> 1 : width;
    CT* dst = counts + lo_bin;
    for (uint32_t k = tid; k < words; k += nthr) sh[k] = 0;
    const PkTally sp{&s_owed_dev, tally_base, tally_base + (1u << shift) / 32u,
                     tally_base + (1u << shift) / 32u + kTallySlots, width};
    if (PK) pk_tally_clear(sp, shift, tid, nthr);
    __syncthreads();
    // unaligned head, then 8 entries per 16-byte load
    uint32_t a0 = min(e1, (e0 + 7u) & ~7u);
    for (uint32_t k = e0 + tid; k < a0; k += nthr) {
      if (PK) lds_add1_pk(sh, staged[k], dst, sp);
      else atomicAdd(&sh[staged[k]], 1u);
    }
    uint32_t nvec = (e1 - a0) / 8u;
    const u32x4* v = (const u32x4*)(staged + a0);
    for (uint32_t q = tid; q < nvec; q += nthr) {
      u32x4 x = __builtin_nontemporal_load(v + q);
      if (PK) {
        if (merge) lds_add8_pk<true>(sh, x, dst, sp);
        else lds_add8_pk<false>(sh, x, dst, sp);
      } else {
        if (merge) lds_add8<true>(sh, x);
        else lds_add8<false>(sh, x);
      }
    }
    for (uint32_t k = a0 + nvec * 8u + tid; k < e1; k += nthr) {
      if (PK) lds_add1_pk(sh, staged[k], dst, sp);
      else atomicAdd(&sh[staged[k]], 1u);
    }
    __syncthreads();
    // a bucket with one segment is this workgroup's alone: plain adds, unless
    // an owed amount already went to its counters atomically
    const bool plain = nseg == 1 && !(PK && s_owed_dev);
    flush_tile<CT, PK>(sh, width, dst, plain, tid, nthr, sp);
    __syncthreads();
  }
}

// ---- BUCKET, paged: the index is read once ---------------------------------
// The scatter pass ships each full LDS buffer as a page of its chunk's own
// pool (no count pass), a small pass lists every bucket's pages, and the apply
// pass walks those pages.  The page depth follows the sample's verdict the
// same way in all three kernels: S with the hot table on, S_cold without.
template <bool HOT>
__device__ inline int paged_depth(const uint32_t* flags, int S, int S_cold) {
  const uint32_t fl = flags ? *flags : (kFlagHotBins | kFlagRunsSeen);
  const bool h = HOT && (fl & kFlagHotBins);
  return (!h && HOT && S_cold > 0) ? S_cold : S;
}

// Paged scatter through ring buffers.  Each bucket has a ring of R = 2P
// 16-bit offsets in LDS.  Slot s (the s-th update of this chunk to the bucket,
// from an LDS counter) goes to ring position s mod R once s < sent + R (sent:
// the bucket's entries already shipped); the slot that completes a half
// (s mod P == P - 1) lists that half, and after a barrier the listed halves
// are shipped as pages.  A round of 8192 updates brings about 8192 / B updates
// to a bucket, so a slot rarely has to wait for a ship; one that does stays in
// its thread's registers and is placed after the next ship.  Written slots are
// always a prefix of each bucket's slots, so every ship frees the half that
// the oldest waiting slot needs.
// LDS: fill[B], sent[B], list[L], misc[8], sink[64], [hot keys, hot counts], ring[B][R].
// Pages a chunk of `len` updates can fill with pages of P entries: full pages
// plus one short page per bucket.
__host__ __device__ inline int64_t part_pool_pages(int64_t len, int B, int P) {
  return (len + P - 1) / P + B;
}
constexpr int kRingRound = kScatterU * 1024;  // updates per round (1024 threads)
constexpr int kRingMisc = 8;                  // [0,1] list lengths, [2,3] waiting flags
// Halves one phase can list: one per touched bucket, plus a second one only
// for a bucket that took more than P slots in the phase.
__host__ __device__ inline int ring_list_len(int B, int R) { return B + kRingRound / (R / 2); }
__host__ __device__ inline size_t ring_scatter_lds(int B, int R, bool hot) {
  size_t head = (size_t)(2 * B + ring_list_len(B, R) + kRingMisc + 64) * 4 +
                (hot ? (size_t)2 * kHotSlots * 4 : 0);
  return ((head + 15) & ~(size_t)15) + (size_t)B * R * 2;
}
// Deepest ring (R a power of two in [16, 256]) that fits `budget` bytes; 0 if none.
static int ring_pick(int B, size_t budget, bool hot) {
  int R = 256;
  while (R > 16 && ring_scatter_lds(B, R, hot) > budget) R >>= 1;
  return ring_scatter_lds(B, R, hot) <= budget ? R : 0;
}

template <bool PACKED, bool HOT, bool RUNS, typename CT>
__device__ __attribute__((always_inline)) inline void ring_scatter_body(
    int64_t c, const int64_t* __restrict__ idx, int64_t n, uint64_t limit, int shift, int B, int R,
    int64_t chunk, const uint32_t* __restrict__ hot_g, uint16_t* __restrict__ out,
    uint32_t* __restrict__ tags, uint32_t pool, uint32_t* __restrict__ rows,
    uint32_t* __restrict__ used, CT* __restrict__ counts, ErrWords* err) {
  constexpr int U = kScatterU;  // updates per thread per round, loaded as U/2 pairs
  constexpr int U2 = U / 2;
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = ring_list_len(B, R);
  uint32_t* fill = (uint32_t*)smem;
  uint32_t* sent = fill + B;
  uint32_t* list = sent + B;
  uint32_t* misc = list + L;
  uint32_t* sink = misc + kRingMisc;
  uint32_t* hkeys = sink + 64;
  uint32_t* hcnt = hkeys + kHotSlots;
  const size_t head = (size_t)(2 * B + L + kRingMisc + 64) * 4 + (HOT ? (size_t)2 * kHotSlots * 4 : 0);
  uint16_t* ring = (uint16_t*)(smem + ((head + 15) & ~(size_t)15));
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = __lane_id();
  const uint32_t P = (uint32_t)R >> 1, R1 = (uint32_t)R - 1u, P1 = P - 1u;
  const int rsh = __builtin_ctz((uint32_t)R), psh = rsh - 1;
  const uint32_t vpp = P / 8u;  // 16-byte pieces per page
  const int vsh = __builtin_ctz(vpp);
  for (int b = tid; b < B; b += nthr) {
    fill[b] = 0;
    sent[b] = 0;
  }
  if (HOT)
    for (int k = tid; k < kHotSlots; k += nthr) {
      hkeys[k] = hot_g[k];
      hcnt[k] = 0;
    }
  if (tid < kRingMisc) misc[tid] = 0;
  __syncthreads();
  const uint32_t mask = (1u << shift) - 1u;
  const int64_t lo = c * chunk, hi = lo + chunk < n ? lo + chunk : n;
  uint32_t bad_r = 0, bad_p = 0;
  uint32_t pages = 0;  // full pages shipped (the same in every thread)
  int p = 0;           // phase parity: list length misc[p], waiting flag misc[2 + p]

  // One update: its bin (bins fit 32 bits), and whether it counts at all.
  auto decode = [&](int64_t raw, bool valid, uint32_t& j) -> bool {
    uint64_t v = (uint64_t)raw;
    bool pe_ok = true;
    if (PACKED) {
      pe_ok = (v & 0xffffu) == 0;
      v = (uint64_t)(raw >> 16);
    }
    const bool in = v < limit;
    bad_p += (valid && !pe_ok) ? 1u : 0u;
    bad_r += (valid && pe_ok && !in) ? 1u : 0u;
    j = (uint32_t)v;
    return valid && pe_ok && in;
  };
  // Count a hot bin in LDS or a long run with one device atomic; otherwise the
  // update is placed (returns true) with its bucket and 16-bit entry.  RUNS:
  // one run detection on the bin (adjacent lanes with equal bins): a hot run
  // is counted at its head, a long cold run is one device atomic, and a short
  // cold run reserves its slots at its head (rh, rl).  Every lane of the wave
  // calls it together (wave_runs).
  auto classify = [&](bool ok, uint32_t j, uint32_t& bk, uint32_t& ent, int& rh,
                      uint32_t& rl) -> bool {
    int hs = -1;
    if (HOT && ok) hs = hot_find(hkeys, j);
    bool cd = ok && hs < 0;
    rh = lane;
    rl = 1;
    if (RUNS) {
      rh = wave_runs(ok ? j : kHotEmpty, rl);
      const bool head = lane == rh;
      if (HOT && hs >= 0) {
        if (head) atomicAdd(&hcnt[hs], rl);
      } else if (cd && rl >= kLongRun) {
        if (head) atomic_add_dev(counts + j, (CT)rl);
        cd = false;
      }
    } else if (HOT && hs >= 0) {
      atomicAdd(&hcnt[hs], 1u);
    }
    bk = cd ? j >> shift : 0u;
    ent = j & mask;
    return cd;
  };
  // Append the halves completed by this thread's slots (done bits) to the
  // list: one LDS atomic per wave; ballots are recomputed, not kept.
  auto list_append = [&](uint32_t done, const uint32_t* bk, const uint32_t* s) {
    uint32_t tot = 0;
#pragma unroll
    for (int u = 0; u < U; u++) tot += (uint32_t)__popcll(__ballot((done >> u) & 1u));
    if (tot == 0) return;  // the same in every lane
    uint32_t b0 = 0;
    if (lane == 0) b0 = atomicAdd(&misc[p], tot);
    b0 = __builtin_amdgcn_readfirstlane(b0);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t bal = __ballot((done >> u) & 1u);
      if ((done >> u) & 1u) {
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
        list[b0 + below] = bk[u] | (((s[u] >> psh) & 1u) << 16);
      }
      b0 += (uint32_t)__popcll(bal);
    }
  };
  // Barrier, ship the listed halves, barrier.  Returns whether some thread
  // still has waiting slots (`wait`: this thread has).
  auto ship_phase = [&](bool wait) -> bool {
    if (__ballot(wait) != 0 && lane == 0) misc[2 + p] = 1u;
    __syncthreads();
    const uint32_t nl = misc[p];
    const bool waiting = misc[2 + p] != 0;
    const uint32_t total = nl << vsh;
    for (uint32_t v = (uint32_t)tid; v < total; v += (uint32_t)nthr) {
      const uint32_t k = v >> vsh, q = v & (vpp - 1u);
      const uint32_t e = list[k], b = e & 0xffffu, h = e >> 16;
      const uint32_t page = pool + pages + k;
      const u32x4 x = ((const u32x4*)(ring + ((size_t)b << rsh) + ((size_t)h << psh)))[q];
      __builtin_nontemporal_store(x, (u32x4*)out + ((size_t)page << vsh) + q);
      if (q == 0) {
        tags[page] = b | (P << 16);
        atomicAdd(&sent[b], P);
      }
    }
    pages += nl;
    // the other parity's length and flag were last read a phase ago, before
    // the barrier that ended it, and are next written after the barrier below
    if (tid == 0) {
      misc[p ^ 1] = 0;
      misc[2 + (p ^ 1)] = 0;
    }
    p ^= 1;
    __syncthreads();
    return waiting;
  };
  // One round: thread tid holds updates base + 2 (k * 1024 + tid) + {0, 1}.
  // FULL: the whole round is inside the chunk (no bounds tests).
  auto round = [&](auto full_c, int64_t base, const i64x2* raw) {
    constexpr bool FULL = decltype(full_c)::value;
    uint32_t bk[U], ent[U], s[U], rlen[U];
    int rhead[U];
    uint32_t cold = 0;
#pragma unroll
    for (int k = 0; k < U2; k++)
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int u = 2 * k + h;
        const bool valid = FULL || base + 2 * (k * 1024 + tid) + h < hi;
        uint32_t j;
        const bool ok = decode(h ? raw[k].y : raw[k].x, valid, j);
        cold |= (classify(ok, j, bk[u], ent[u], rhead[u], rlen[u]) ? 1u : 0u) << u;
      }
    // reserve slots: one LDS atomic per update (per run with RUNS); lanes with
    // nothing to place add to their own sink word
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool cd = (cold >> u) & 1u;
      s[u] = atomicAdd((cd && lane == rhead[u]) ? &fill[bk[u]] : &sink[lane], rlen[u]);
    }
    uint32_t se[U];
#pragma unroll
    for (int u = 0; u < U; u++) se[u] = sent[bk[u]];  // bucket 0 for updates not placed
    uint32_t done = 0, wait = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (RUNS) s[u] = __shfl(s[u], rhead[u]) + (uint32_t)(lane - rhead[u]);
      const bool cd = (cold >> u) & 1u;
      const bool fits = cd && s[u] - se[u] <= R1;
      if (fits) ring[(bk[u] << rsh) | (s[u] & R1)] = (uint16_t)ent[u];
      done |= ((fits && (s[u] & P1) == P1) ? 1u : 0u) << u;
      wait |= ((cd && !fits) ? 1u : 0u) << u;
    }
    list_append(done, bk, s);
    bool waiting = ship_phase(wait != 0);
    while (waiting) {
      // slots whose half of the ring has been shipped since
      done = 0;
      uint32_t still = 0;
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool w = (wait >> u) & 1u;
        const bool fits = w && s[u] - sent[bk[u]] <= R1;
        if (fits) ring[(bk[u] << rsh) | (s[u] & R1)] = (uint16_t)ent[u];
        done |= ((fits && (s[u] & P1) == P1) ? 1u : 0u) << u;
        still |= ((w && !fits) ? 1u : 0u) << u;
      }
      wait = still;
      list_append(done, bk, s);
      waiting = ship_phase(wait != 0);
    }
  };

  // 16-byte index loads when the index is 16-byte aligned (rounds start at
  // multiples of 8192 updates)
  const bool vec = ((uintptr_t)idx & 15u) == 0;
  auto load_round = [&](int64_t base, i64x2* dst) {
    if (vec && base + kRingRound <= hi) {
      const i64x2* rp = (const i64x2*)(idx + base) + tid;
#pragma unroll
      for (int k = 0; k < U2; k++) dst[k] = __builtin_nontemporal_load(rp + k * 1024);
    } else {
#pragma unroll
      for (int k = 0; k < U2; k++) {
        const int64_t e = base + 2 * (k * 1024 + tid);
        dst[k].x = e < hi ? __builtin_nontemporal_load(idx + e) : 0;
        dst[k].y = e + 1 < hi ? __builtin_nontemporal_load(idx + e + 1) : 0;
      }
    }
  };
  i64x2 nxt[U2];
  if (lo < hi) load_round(lo, nxt);
  for (int64_t base = lo; base < hi; base += kRingRound) {
    i64x2 raw[U2];
#pragma unroll
    for (int k = 0; k < U2; k++) raw[k] = nxt[k];
    if (base + kRingRound < hi) load_round(base + kRingRound, nxt);
    if (base + kRingRound <= hi) round(std::true_type(), base, raw);
    else round(std::false_type(), base, raw);
  }
  // each bucket's unshipped slots (fewer than P, in one half) become short
  // page `pages + b`; the last ship phase ended with a barrier
  const uint32_t total = (uint32_t)B << vsh;
  for (uint32_t v = (uint32_t)tid; v < total; v += (uint32_t)nthr) {
    const uint32_t b = v >> vsh, q = v & (vpp - 1u);
    const uint32_t sd = sent[b], cnt = fill[b] - sd, page = pool + pages + b;
    if (q * 8u < cnt)
      __builtin_nontemporal_store(
          ((const u32x4*)(ring + ((size_t)b << rsh) + ((size_t)((sd >> psh) & 1u) << psh)))[q],
          (u32x4*)out + ((size_t)page << vsh) + q);
    if (q == 0) {
      tags[page] = b | (cnt << 16);
      rows[(size_t)c * B + b] = (sd >> psh) + (cnt ? 1u : 0u);
    }
  }
  if (HOT)
    for (int k = tid; k < kHotSlots; k += nthr)
      if (hcnt[k]) atomic_add_dev(counts + hkeys[k], (CT)hcnt[k]);
  if (tid == 0) used[c] = pages + (uint32_t)B;
  flush_errors(err, bad_r, bad_p);
}

// The sample's verdict (flags) picks the instantiation and the ring depth:
// R with the hot table on, R_cold (deeper: the hot table's LDS is free) off.
template <bool PACKED, bool HOT, bool RUNS, typename CT>
__global__ __launch_bounds__(1024) void k_histo_scatter_ring(
    const int64_t* __restrict__ idx, int64_t n, uint64_t nbins, int shift, int B, int R,
    int R_cold, int64_t chunk, const uint32_t* __restrict__ hot_g, uint16_t* __restrict__ out,
    uint32_t* __restrict__ tags, uint32_t* __restrict__ rows, uint32_t* __restrict__ used,
    CT* __restrict__ counts, ErrWords* err, const uint32_t* __restrict__ flags) {
  const uint32_t fl = flags ? *flags : (kFlagHotBins | kFlagRunsSeen);
  const bool h = HOT && (fl & kFlagHotBins), r = RUNS && (fl & kFlagRunsSeen);
  const int Re = (!h && HOT && R_cold > 0) ? R_cold : R;
  const uint32_t pool = (uint32_t)(blockIdx.x * part_pool_pages(chunk, B, Re / 2));
  if (h && r)
    ring_scatter_body<PACKED, HOT, RUNS, CT>(blockIdx.x, idx, n, nbins, shift, B, Re, chunk, hot_g,
                                             out, tags, pool, rows, used, counts, err);
  else if (h)
    ring_scatter_body<PACKED, HOT, false, CT>(blockIdx.x, idx, n, nbins, shift, B, Re, chunk, hot_g,
                                              out, tags, pool, rows, used, counts, err);
  else if (r)
    ring_scatter_body<PACKED, false, RUNS, CT>(blockIdx.x, idx, n, nbins, shift, B, Re, chunk,
                                               hot_g, out, tags, pool, rows, used, counts, err);
  else
    ring_scatter_body<PACKED, false, false, CT>(blockIdx.x, idx, n, nbins, shift, B, Re, chunk,
                                                hot_g, out, tags, pool, rows, used, counts, err);
}

// One workgroup per chunk: its nonempty pages, grouped by bucket.
// plist[start[b] + rows[c][b] + k] = k-th page of bucket b from chunk c, with
// bit 31 set on a short page.  Pages go through LDS a window at a time,
// counting-sorted by bucket, so each bucket's pages of a window are written as
// one contiguous run.
// LDS: cur[B] (next list position per bucket), cnt[B], lstart[B],
// wsum[32], sorted[kListWin], sbkt[kListWin] (u16).
constexpr int kListWin = 8192;
__host__ __device__ inline size_t page_list_lds(int B) {
  return (size_t)(3 * B + 32 + kListWin) * 4 + (size_t)kListWin * 2;
}
template <bool HOT>
__global__ __launch_bounds__(1024) void k_page_list(const uint32_t* __restrict__ tags,
                                                    const uint32_t* __restrict__ rows,
                                                    const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ used, int B,
                                                    int64_t chunk, int S, int S_cold,
                                                    const uint32_t* __restrict__ flags,
                                                    uint32_t* __restrict__ plist) {
  extern __shared__ uint32_t lds[];
  uint32_t* cur = lds;
  uint32_t* cnt = cur + B;
  uint32_t* lstart = cnt + B;
  uint32_t* wsum = lstart + B;
  uint32_t* sorted = wsum + 32;
  uint16_t* sbkt = (uint16_t*)(sorted + kListWin);
  constexpr int PER = kListWin / 1024;  // pages per thread per window
  const int c = blockIdx.x, tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const int Se = paged_depth<HOT>(flags, S, S_cold);
  const uint32_t pool = (uint32_t)(c * part_pool_pages(chunk, B, Se));
  for (int b = tid; b < B; b += 1024) cur[b] = start[b] + rows[(size_t)c * B + b];
  const uint32_t np = used[c];
  const int per_b = (B + 1023) / 1024;  // buckets per thread in the scan
  for (uint32_t w0 = 0; w0 < np; w0 += kListWin) {
    for (int b = tid; b < B; b += 1024) cnt[b] = 0;
    __syncthreads();
    // count and rank this window's pages per bucket
    uint32_t bk[PER], rk[PER], pg[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const uint32_t k = w0 + q * 1024 + tid;
      const uint32_t t = k < np ? tags[pool + k] : 0u;
      bk[q] = t & 0xffffu;
      pg[q] = (pool + k) | ((t >> 16) < (uint32_t)Se ? 0x80000000u : 0u);
      rk[q] = (t >> 16) ? atomicAdd(&cnt[bk[q]], 1u) : ~0u;  // ~0: empty or past the end
    }
    __syncthreads();
    // exclusive scan of cnt into lstart: per-thread sums of per_b buckets,
    // then across the workgroup (waves, then the 16 wave totals)
    uint32_t mine = 0;
    for (int i = 0; i < per_b; i++) {
      const int b = tid * per_b + i;
      if (b < B) mine += cnt[b];
    }
    uint32_t incl = mine;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(incl, d);
      if (lane >= d) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; w++) before += wsum[w];
    uint32_t run = before + incl - mine;
    for (int i = 0; i < per_b; i++) {
      const int b = tid * per_b + i;
      if (b < B) {
        lstart[b] = run;
        run += cnt[b];
      }
    }
    __syncthreads();
    // place into the sorted window
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (rk[q] != ~0u) {
        const uint32_t pos = lstart[bk[q]] + rk[q];
        sorted[pos] = pg[q];
        sbkt[pos] = (uint16_t)bk[q];
      }
    __syncthreads();
    // write: element i of the sorted window is page i - lstart[b] of bucket b here
    uint32_t total = 0;
    for (int w = 0; w < 16; w++) total += wsum[w];
    for (uint32_t i = tid; i < total; i += 1024) {
      const uint32_t b = sbkt[i];
      plist[cur[b] + (i - lstart[b])] = sorted[i];
    }
    __syncthreads();
    for (int b = tid; b < B; b += 1024) cur[b] += cnt[b];
    // the next window's first barrier orders these updates before their reads
  }
}

// lds_add8 / lds_add8_pk for the first `valid` (1..8) of the 8 offsets in x.
template <bool RUNS, bool PK, typename CT>
__device__ inline void lds_add_valid(uint32_t* sh, u32x4 x, uint32_t valid, CT* dst,
                                     const PkTally& tally) {
  uint32_t v[8] = {x.x & 0xffff, x.x >> 16, x.y & 0xffff, x.y >> 16,
                   x.z & 0xffff, x.z >> 16, x.w & 0xffff, x.w >> 16};
  uint32_t inc[8], old[8];
#pragma unroll
  for (int i = 0; i < 8; i++) inc[i] = (uint32_t)i < valid ? 1u : 0u;
  if (RUNS) {
    // a run's count goes on its first entry; the others add nothing
#pragma unroll
    for (int i = 6; i >= 0; i--)
      if (v[i] == v[i + 1]) {
        inc[i] += inc[i + 1];
        inc[i + 1] = 0;
      }
  }
  if (!PK) {
#pragma unroll
    for (int i = 0; i < 8; i++)
      if (inc[i]) atomicAdd(&sh[v[i]], inc[i]);
    return;
  }
#pragma unroll
  for (int i = 0; i < 8; i++)
    old[i] = inc[i] ? atomicAdd(&sh[v[i] >> 1], inc[i] << ((v[i] & 1u) * 16u)) : 0u;
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (inc[i]) pk_account(v[i], inc[i], old[i], dst, tally);
}

// Persistent: each workgroup takes segments (<= seg_pages pages of one bucket)
// from a queue, histograms their entries in LDS and adds the tile into counts;
// eight lanes read one 128-byte page of 64 entries.
template <typename CT, bool HOT, bool RUNS, bool PK>
__global__ __launch_bounds__(1024) void k_histo_apply_paged(
    const uint16_t* __restrict__ staged, const uint32_t* __restrict__ tags,
    const uint32_t* __restrict__ plist, const uint32_t* __restrict__ start,
    const uint32_t* __restrict__ segstart, uint32_t* __restrict__ misc, int B,
    int shift, uint64_t nbins, int S, int S_cold, CT* __restrict__ counts,
    const uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t sh[];
  __shared__ uint32_t s_work[2];
  __shared__ uint32_t s_owed_dev;  // PK: some owed amount of this segment went to a device counter
  // PK: tally bits and table after the tile (2^(shift-1) words)
  uint32_t* const tally_base = sh + (PK ? 1u << (shift - 1) : 0u);
  const bool merge = RUNS && (!flags || (*flags & kFlagRunsSeen));
  const uint32_t Se = (uint32_t)paged_depth<HOT>(flags, S, S_cold);
  const uint32_t vpp = Se / 8u;  // 16-byte vectors per page (a power of two)
  const int vsh = __builtin_ctz(vpp);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const uint32_t total = misc[0], seg_pages = misc[kSegWord];
  while (true) {
    if (tid == 0) {
      uint32_t w = atomicAdd(&misc[1], 1u);
      uint32_t b = 0;
      if (w < total) {
        int lo = 0, hi = B - 1;  // last bucket whose first segment is <= w
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (segstart[mid] <= w) lo = mid; else hi = mid - 1;
        }
        b = (uint32_t)lo;
      }
      s_work[0] = w;
      s_work[1] = b;
      s_owed_dev = 0;
    }
    __syncthreads();
    const uint32_t w = s_work[0], b = s_work[1];
    if (w >= total) break;
    const uint32_t nseg = segstart[b + 1] - segstart[b];
    const uint32_t g0 = start[b] + (w - segstart[b]) * seg_pages;
    const uint32_t g1 = min(start[b + 1], g0 + seg_pages);
    const uint64_t lo_bin = (uint64_t)b << shift;
    const uint32_t width = (uint32_t)min((uint64_t)1 << shift, nbins - lo_bin);
    const uint32_t words = PK ? (width + 1u) >> 1 : width;
    CT* dst = counts + lo_bin;
    for (uint32_t k = tid; k < words; k += nthr) sh[k] = 0;
    const PkTally sp{&s_owed_dev, tally_base, tally_base + (1u << shift) / 32u,
                     tally_base + (1u << shift) / 32u + kTallySlots, width};
    if (PK) pk_tally_clear(sp, shift, tid, nthr);
    __syncthreads();
    const uint32_t nv = (g1 - g0) << vsh;
    constexpr int Q = 4;  // vectors per thread in flight
    for (uint32_t v0 = tid; v0 < nv; v0 += Q * nthr) {
      uint32_t ent[Q], cnt[Q];
      u32x4 x[Q];
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const uint32_t v = v0 + q * nthr;
        ent[q] = v < nv ? plist[g0 + (v >> vsh)] : 0u;
      }
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const uint32_t v = v0 + q * nthr;
        const uint32_t page = ent[q] & 0x7fffffffu;
        cnt[q] = 0;
        if (v < nv) {
          cnt[q] = (ent[q] >> 31) ? tags[page] >> 16 : Se;
          x[q] = __builtin_nontemporal_load((const u32x4*)(staged + (size_t)page * Se) +
                                            (v & (vpp - 1u)));
        }
      }
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const uint32_t v = v0 + q * nthr;
        const uint32_t first = (v & (vpp - 1u)) * 8u;
        if (v < nv && cnt[q] > first) {
          const uint32_t valid = min(8u, cnt[q] - first);
          if (merge) lds_add_valid<true, PK>(sh, x[q], valid, dst, sp);
          else lds_add_valid<false, PK>(sh, x[q], valid, dst, sp);
        }
      }
    }
    __syncthreads();
    // a bucket with one segment is this workgroup's alone: plain adds, unless
    // an owed amount already went to its counters atomically
    const bool plain = nseg == 1 && !(PK && s_owed_dev);
    flush_tile<CT, PK>(sh, width, dst, plain, tid, nthr, sp);
    __syncthreads();
  }
}

// ---- launchers ------------------------------------------------------------
static int norm_block(int b) { return (b == 256 || b == 512 || b == 1024) ? b : 256; }
static int norm_unroll(int u) { return (u == 4 || u == 8 || u == 16) ? u : 8; }

// Largest bin tile (log2) whose 32-bit LDS counters leave 1 KiB of the
// per-workgroup LDS for static variables.
static int max_tile_bits(const Device* d) {
  int tb = 15;
  while (tb > 8 && ((size_t)4 << tb) + 1024 > (size_t)d->lds_per_block) tb--;
  return tb;
}

template <int U, typename CT, bool PACKED>
static void go_direct(int merge, dim3 g, dim3 b, hipStream_t s, const int64_t* idx, int64_t n,
                      CT* counts, uint64_t nbins, ErrWords* err) {
  if (merge == 1)
    hipLaunchKernelGGL((k_histo_direct<U, CT, PACKED, 1>), g, b, 0, s, idx, n, counts, nbins, err);
  else if (merge == 2)
    hipLaunchKernelGGL((k_histo_direct<U, CT, PACKED, 2>), g, b, 0, s, idx, n, counts, nbins, err);
  else
    hipLaunchKernelGGL((k_histo_direct<U, CT, PACKED, 0>), g, b, 0, s, idx, n, counts, nbins, err);
}

template <typename CT, bool PACKED>
static int direct_typed(const int64_t* idx, int64_t n, CT* counts, int64_t nbins,
                        const bale_config& c, Device* d, hipStream_t s) {
  int block = norm_block(c.block), U = norm_unroll(c.unroll);
  int merge = c.merge < 0 ? 0 : c.merge;
  int64_t per = (int64_t)block * U;
  int64_t grid = ceil_div(n, per);
  if (c.hot && nbins < (int64_t)kHotEmpty) {
    void* ws = workspace(d, (kHotSlots + kSketchScratch) * 4, s);
    if (!ws) return (int)hipErrorOutOfMemory;
    uint32_t* hot = (uint32_t*)ws;
    launch_hot_sketch<PACKED>(idx, n, (uint64_t)nbins, hot, nullptr, hot + kHotSlots, s);
    int gpc = c.groups_per_cu > 0 ? c.groups_per_cu : (block == 1024 ? 2 : 4);
    grid = std::min<int64_t>(grid, (int64_t)d->cus * gpc);
    dim3 g((unsigned)grid), b(block);
    if (U == 16)
      hipLaunchKernelGGL((k_histo_direct_hot<16, CT, PACKED>), g, b, 0, s, idx, n, counts, (uint64_t)nbins, hot, stream_err(d, s));
    else if (U == 8)
      hipLaunchKernelGGL((k_histo_direct_hot<8, CT, PACKED>), g, b, 0, s, idx, n, counts, (uint64_t)nbins, hot, stream_err(d, s));
    else
      hipLaunchKernelGGL((k_histo_direct_hot<4, CT, PACKED>), g, b, 0, s, idx, n, counts, (uint64_t)nbins, hot, stream_err(d, s));
    return (int)hipGetLastError();
  }
  if (grid > 0x7fffffff) return bad_config("DIRECT histogram: %lld workgroups exceed the grid limit", (long long)grid);
  dim3 g((unsigned)grid), b(block);
  if (U == 16) go_direct<16, CT, PACKED>(merge, g, b, s, idx, n, counts, nbins, stream_err(d, s));
  else if (U == 8) go_direct<8, CT, PACKED>(merge, g, b, s, idx, n, counts, nbins, stream_err(d, s));
  else go_direct<4, CT, PACKED>(merge, g, b, s, idx, n, counts, nbins, stream_err(d, s));
  return (int)hipGetLastError();
}

int launch_histo_direct(const int64_t* idx, int64_t n, void* counts, int64_t nbins, int dtype,
                        bool packed, const bale_config& c, Device* d, hipStream_t s) {
  if (n <= 0) return 0;
  if (dtype == BALE_I64)
    return packed ? direct_typed<int64_t, true>(idx, n, (int64_t*)counts, nbins, c, d, s)
                  : direct_typed<int64_t, false>(idx, n, (int64_t*)counts, nbins, c, d, s);
  return packed ? direct_typed<int32_t, true>(idx, n, (int32_t*)counts, nbins, c, d, s)
                : direct_typed<int32_t, false>(idx, n, (int32_t*)counts, nbins, c, d, s);
}

template <int U, typename CT, bool PACKED>
static void go_lds(int merge, dim3 g, dim3 b, size_t lds, hipStream_t s, const int64_t* idx,
                   int64_t n, CT* counts, uint64_t nbins, int tb, int64_t chunk, ErrWords* err) {
  if (merge == 1) {
    allow_lds(k_histo_lds<U, CT, PACKED, 1>, lds);
    hipLaunchKernelGGL((k_histo_lds<U, CT, PACKED, 1>), g, b, lds, s, idx, n, counts, nbins, tb, chunk, err);
  } else if (merge == 2) {
    allow_lds(k_histo_lds<U, CT, PACKED, 2>, lds);
    hipLaunchKernelGGL((k_histo_lds<U, CT, PACKED, 2>), g, b, lds, s, idx, n, counts, nbins, tb, chunk, err);
  } else {
    allow_lds(k_histo_lds<U, CT, PACKED, 0>, lds);
    hipLaunchKernelGGL((k_histo_lds<U, CT, PACKED, 0>), g, b, lds, s, idx, n, counts, nbins, tb, chunk, err);
  }
}

template <typename CT, bool PACKED>
static void go_lds_u(int U, int merge, dim3 g, dim3 b, size_t lds, hipStream_t s,
                     const int64_t* idx, int64_t n, CT* counts, uint64_t nbins, int tb,
                     int64_t chunk, ErrWords* err) {
  if (U == 16) go_lds<16, CT, PACKED>(merge, g, b, lds, s, idx, n, counts, nbins, tb, chunk, err);
  else if (U == 8) go_lds<8, CT, PACKED>(merge, g, b, lds, s, idx, n, counts, nbins, tb, chunk, err);
  else go_lds<4, CT, PACKED>(merge, g, b, lds, s, idx, n, counts, nbins, tb, chunk, err);
}

int launch_histo_lds(const int64_t* idx, int64_t n, void* counts, int64_t nbins, int dtype,
                     bool packed, const bale_config& c, Device* d, hipStream_t s) {
  if (n <= 0) return 0;
  int block = norm_block(c.block), U = norm_unroll(c.unroll);
  int merge = c.merge < 0 ? 0 : c.merge;
  int tb = c.tile_log2 > 0 ? c.tile_log2 : max_tile_bits(d);
  if (tb > max_tile_bits(d)) tb = max_tile_bits(d);
  int64_t width = std::min<int64_t>((int64_t)1 << tb, nbins);
  while (tb > 8 && ((int64_t)1 << (tb - 1)) >= width) tb--;  // shrink to the table
  int64_t tiles = ceil_div(nbins, (int64_t)1 << tb);
  if (tiles > 65535)
    return bad_config("LDS histogram: %lld bin tiles of 2^%d bins exceed 65535; use BUCKET or DIRECT "
                      "for a table of %lld bins", (long long)tiles, tb, (long long)nbins);
  int gpc = c.groups_per_cu > 0 ? c.groups_per_cu : 2;
  int64_t target = (int64_t)d->cus * gpc;
  int64_t chunks = std::max<int64_t>(1, target / tiles);
  // keep the flush (one atomic per bin per chunk) well under the update count
  int64_t cap = std::max<int64_t>(1, n / (4 * std::min<int64_t>(nbins, (int64_t)1 << tb)));
  chunks = std::min(chunks, cap);
  int64_t chunk = ceil_div(n, chunks);
  chunks = ceil_div(n, chunk);
  dim3 g((unsigned)chunks, (unsigned)tiles), b(block);
  size_t lds = (size_t)4 << tb;
  if (dtype == BALE_I64) {
    if (packed) go_lds_u<int64_t, true>(U, merge, g, b, lds, s, idx, n, (int64_t*)counts, nbins, tb, chunk, stream_err(d, s));
    else go_lds_u<int64_t, false>(U, merge, g, b, lds, s, idx, n, (int64_t*)counts, nbins, tb, chunk, stream_err(d, s));
  } else {
    if (packed) go_lds_u<int32_t, true>(U, merge, g, b, lds, s, idx, n, (int32_t*)counts, nbins, tb, chunk, stream_err(d, s));
    else go_lds_u<int32_t, false>(U, merge, g, b, lds, s, idx, n, (int32_t*)counts, nbins, tb, chunk, stream_err(d, s));
  }
  return (int)hipGetLastError();
}

constexpr int kMaxBuckets = 4096;

static int too_many_buckets(int64_t B, int tb, int64_t nbins) {
  int need = tb;
  while (need < 16 && ceil_div(nbins, (int64_t)1 << need) > kMaxBuckets) need++;
  if (ceil_div(nbins, (int64_t)1 << need) > kMaxBuckets)
    return bad_config("BUCKET histogram: %lld bins need more than %d buckets of at most 2^16 bins; "
                      "use DIRECT", (long long)nbins, kMaxBuckets);
  return bad_config("BUCKET histogram: buckets of 2^%d bins give %lld buckets, more than %d; use "
                    "tile_log2 >= %d", tb, (long long)B, kMaxBuckets, need);
}

// Updates per apply segment: long enough that a segment's flush (up to one add
// per bin of the tile) stays a fraction of its updates, short enough to spread
// the pass over the chip; cfg.task_log2 sets it explicitly.
static int64_t apply_segment(int64_t n, int tb, const bale_config& c, const Device* d) {
  if (c.task_log2 > 0) return std::max<int64_t>(4096, (int64_t)1 << std::min(30, (int)c.task_log2));
  return std::max<int64_t>({4096, ceil_div(n, (int64_t)d->cus * 2), (int64_t)4 << tb});
}

// Apply segment of the adaptive plan for inputs with hot bins: about half the
// average bucket (two segments for an average one, more for the heavy ones),
// at most the default and at least 16384 updates (shorter segments cost more
// in flushes than they gain in balance).
static int64_t skewed_segment(int64_t n, int64_t B, int tb, const bale_config& c, const Device* d) {
  const int64_t def = apply_segment(n, tb, c, d);
  if (c.task_log2 > 0) return def;
  return std::max<int64_t>(16384, std::min<int64_t>(def, ceil_div(n, 2 * B)));
}

// Apply segment for inputs the sketch finds without hot bins (taken on the
// device): twice the average bucket, rounded up to a power of two, so nearly
// every bucket is one segment.  0 (no alternative) when the caller fixed the
// segment length.
static int64_t uniform_segment(int64_t n, int64_t B, const bale_config& c) {
  if (c.task_log2 > 0) return 0;
  int64_t v = 4096;
  while (v < 2 * ceil_div(n, B) && v < ((int64_t)1 << 30)) v <<= 1;
  return v;
}

template <bool PACKED, bool HOT, bool RUNS, typename CT>
static int bucket_typed(const int64_t* idx, int64_t n, CT* counts, int64_t nbins,
                        const bale_config& c, Device* d, hipStream_t s) {
  int tb = c.tile_log2 > 0 ? c.tile_log2 : max_tile_bits(d);
  // entries are 16-bit bin offsets; a bucket one bit wider than a 32-bit
  // tile allows is applied through the packed 16-bit tile
  const int tb_max = std::min(16, max_tile_bits(d) + 1);
  if (tb > tb_max) tb = tb_max;
  const bool pk = tb > max_tile_bits(d);
  int64_t B = ceil_div(nbins, (int64_t)1 << tb);
  if (B > kMaxBuckets) return too_many_buckets(B, tb, nbins);
  // groups_per_cu scatter workgroups share a CU's LDS, so one waits on
  // memory while another places
  int gpc = c.groups_per_cu > 0 ? c.groups_per_cu : 1;
  const size_t budget = (size_t)d->lds_per_block / (size_t)gpc;
  int S = part_pick_depth((int)B, sizeof(uint16_t), budget, HOT);
  if (S == 0)
    return bad_config("BUCKET histogram: %lld buckets do not fit in the LDS of %d workgroups per CU "
                      "(use wider buckets or fewer groups_per_cu)", (long long)B, gpc);
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * gpc, ceil_div(n, 4096)));
  // chunks are whole multiples of kChunkAlign so both passes see the same
  // aligned 64-update groups
  int64_t chunk = ceil_div(ceil_div(n, C), kChunkAlign) * kChunkAlign;
  C = ceil_div(n, chunk);
  void* ws = workspace(d, part_bytes(C, B, n, sizeof(uint16_t)), s);
  if (!ws) return (int)hipErrorOutOfMemory;
  PartLayout L = part_carve(ws, C, B);
  // hot == 2: the sketch decides on the device whether hot bins / runs are
  // worth handling; every pass reads its verdict from the flag word
  uint32_t* flags = c.hot == 2 ? L.misc + (kMiscWords - 1) : nullptr;
  if (HOT)
    launch_hot_sketch<PACKED>(idx, n, (uint64_t)nbins, L.hot, flags, L.sketch, s);
  size_t lds_count = (size_t)(B + (HOT ? kHotSlots : 0)) * 4;
  hipLaunchKernelGGL((k_part_count<PACKED, HOT, RUNS>), dim3((unsigned)C), dim3(1024), lds_count, s, idx,
                     n, (uint64_t)nbins, tb, (int)B, chunk, L.chunk, L.hot, stream_err(d, s),
                     (const uint32_t*)flags);
  hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)B), dim3(256), 0, s, L.chunk,
                     (int)C, (int)B, L.total);
  int64_t seg = flags ? skewed_segment(n, B, tb, c, d) : apply_segment(n, tb, c, d);
  const int64_t seg_alt = uniform_segment(n, B, c);
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, (int)B, (uint32_t)seg,
                     (uint32_t)seg_alt, (const uint32_t*)flags, L.start, L.segstart, L.misc);
  size_t lds = part_scatter_lds((int)B, S, sizeof(uint16_t), HOT);
  // adaptive mode: when the sample finds no hot bins, the scatter pass puts
  // its buffers over the hot table, as deep as the no-hot layout allows
  int S_cold = 0;
  if (HOT && flags) {
    S_cold = part_pick_depth((int)B, sizeof(uint16_t), budget, false);
    lds = std::max(lds, part_scatter_lds((int)B, S_cold, sizeof(uint16_t), false));
  }
  allow_lds(k_histo_scatter<PACKED, HOT, RUNS, CT>, lds);
  hipLaunchKernelGGL((k_histo_scatter<PACKED, HOT, RUNS, CT>), dim3((unsigned)C), dim3(1024), lds, s, idx,
                     n, (uint64_t)nbins, tb, (int)B, S, chunk, L.chunk, L.start, L.hot,
                     (uint16_t*)L.staged, counts, (const uint32_t*)flags, S_cold);
  size_t lds_apply = ((size_t)(pk ? 2 : 4) << tb) + (pk ? (size_t)pk_tally_words(tb) * 4 : 0);
  int64_t groups = (int64_t)d->cus * (lds_apply <= 65536 ? 2 : 1);
  if (pk) {
    allow_lds(k_histo_apply<CT, RUNS, true>, lds_apply);
    hipLaunchKernelGGL((k_histo_apply<CT, RUNS, true>), dim3((unsigned)groups), dim3(1024),
                       lds_apply, s, (const uint16_t*)L.staged, L.start, L.segstart, L.misc, (int)B,
                       tb, (uint64_t)nbins, counts, (const uint32_t*)flags);
  } else {
    allow_lds(k_histo_apply<CT, RUNS, false>, lds_apply);
    hipLaunchKernelGGL((k_histo_apply<CT, RUNS, false>), dim3((unsigned)groups), dim3(1024),
                       lds_apply, s, (const uint16_t*)L.staged, L.start, L.segstart, L.misc, (int)B,
                       tb, (uint64_t)nbins, counts, (const uint32_t*)flags);
  }
  return (int)hipGetLastError();
}

// The paged plan: sketch, scatter into page pools, page counts prefixed over
// chunks and buckets, page list, apply.  Returns 1 when the buffers would be
// too shallow for 16-byte pages (the counted plan runs instead).
template <bool PACKED, bool HOT, bool RUNS, typename CT>
static int bucket_paged_typed(const int64_t* idx, int64_t n, CT* counts, int64_t nbins,
                              const bale_config& c, Device* d, hipStream_t s) {
  int tb = c.tile_log2 > 0 ? c.tile_log2 : max_tile_bits(d);
  const int tb_max = std::min(16, max_tile_bits(d) + 1);
  if (tb > tb_max) tb = tb_max;
  const bool pk = tb > max_tile_bits(d);
  int64_t B = ceil_div(nbins, (int64_t)1 << tb);
  if (B > kMaxBuckets) return too_many_buckets(B, tb, nbins);
  int gpc = c.groups_per_cu > 0 ? c.groups_per_cu : 1;
  const size_t budget = (size_t)d->lds_per_block / (size_t)gpc;
  // ring depth with the hot table's LDS taken, and without it (adaptive plan)
  const int R = ring_pick((int)B, budget, HOT);
  const bool adaptive = HOT && c.hot == 2;
  const int R_cold = adaptive ? ring_pick((int)B, budget, false) : 0;
  if (R < 16 || (adaptive && R_cold < 16)) return 1;  // pages would be under 16 bytes
  const int P = R / 2, P_cold = R_cold / 2;
  const int p_lo = adaptive ? std::min(P, P_cold) : P;
  int64_t C = std::max<int64_t>(1, std::min<int64_t>((int64_t)d->cus * gpc, ceil_div(n, 4096)));
  int64_t chunk = ceil_div(ceil_div(n, C), kChunkAlign) * kChunkAlign;
  C = ceil_div(n, chunk);
  // pool sizes for either depth: pages (tags, list) and entries (staging)
  const int64_t pages = C * part_pool_pages(chunk, (int)B, p_lo);
  const int64_t entries = std::max(C * part_pool_pages(chunk, (int)B, P) * P,
                                   adaptive ? C * part_pool_pages(chunk, (int)B, P_cold) * P_cold : 0);
  if (entries >= ((int64_t)1 << 32) || pages >= ((int64_t)1 << 31))
    return bad_config("BUCKET histogram: %lld staged entries exceed 32-bit positions", (long long)entries);
  const size_t head = ((part_words(C, B) * 4) + 255) & ~(size_t)255;
  const size_t tag_b = ((size_t)pages * 4 + 255) & ~(size_t)255;
  const size_t used_b = ((size_t)C * 4 + 255) & ~(size_t)255;
  const size_t stage_b = ((size_t)entries * 2 + 255) & ~(size_t)255;
  void* ws = workspace(d, head + 2 * tag_b + used_b + stage_b, s);
  if (!ws) return (int)hipErrorOutOfMemory;
  PartLayout L = part_carve(ws, C, B);
  uint32_t* tags = (uint32_t*)((char*)ws + head);
  uint32_t* plist = (uint32_t*)((char*)ws + head + tag_b);
  uint32_t* used = (uint32_t*)((char*)ws + head + 2 * tag_b);
  uint16_t* staged = (uint16_t*)((char*)ws + head + 2 * tag_b + used_b);
  uint32_t* flags = adaptive ? L.misc + (kMiscWords - 1) : nullptr;
  if (HOT)
    launch_hot_sketch<PACKED>(idx, n, (uint64_t)nbins, L.hot, flags, L.sketch, s);
  size_t lds = ring_scatter_lds((int)B, R, HOT);
  if (adaptive) lds = std::max(lds, ring_scatter_lds((int)B, R_cold, false));
  allow_lds(k_histo_scatter_ring<PACKED, HOT, RUNS, CT>, lds);
  hipLaunchKernelGGL((k_histo_scatter_ring<PACKED, HOT, RUNS, CT>), dim3((unsigned)C), dim3(1024),
                     lds, s, idx, n, (uint64_t)nbins, tb, (int)B, R, R_cold, chunk, L.hot, staged,
                     tags, L.chunk, used, counts, stream_err(d, s), (const uint32_t*)flags);
  hipLaunchKernelGGL(k_part_offsets, dim3((unsigned)B), dim3(256), 0, s, L.chunk, (int)C, (int)B,
                     L.total);
  // apply segments in pages of the hot-on depth
  const int64_t seg = adaptive ? skewed_segment(n, B, tb, c, d) : apply_segment(n, tb, c, d);
  const uint32_t seg_pages = (uint32_t)ceil_div(seg, P);
  // the uniform alternative is only taken without hot bins: cold pages
  const uint32_t seg_alt_pages =
      adaptive ? (uint32_t)ceil_div(uniform_segment(n, B, c), (int64_t)P_cold) : 0u;
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, s, L.total, (int)B, seg_pages, seg_alt_pages,
                     (const uint32_t*)flags, L.start, L.segstart, L.misc);
  allow_lds(k_page_list<HOT>, page_list_lds((int)B));
  hipLaunchKernelGGL(k_page_list<HOT>, dim3((unsigned)C), dim3(1024), page_list_lds((int)B), s,
                     (const uint32_t*)tags, (const uint32_t*)L.chunk, (const uint32_t*)L.start,
                     (const uint32_t*)used, (int)B, chunk, P, P_cold, (const uint32_t*)flags, plist);
  size_t lds_apply = ((size_t)(pk ? 2 : 4) << tb) + (pk ? (size_t)pk_tally_words(tb) * 4 : 0);
  int64_t groups = (int64_t)d->cus * (lds_apply <= 65536 ? 2 : 1);
  if (pk) {
    allow_lds(k_histo_apply_paged<CT, HOT, RUNS, true>, lds_apply);
    hipLaunchKernelGGL((k_histo_apply_paged<CT, HOT, RUNS, true>), dim3((unsigned)groups),
                       dim3(1024), lds_apply, s, (const uint16_t*)staged, (const uint32_t*)tags,
                       (const uint32_t*)plist, (const uint32_t*)L.start, (const uint32_t*)L.segstart,
                       L.misc, (int)B, tb, (uint64_t)nbins, P, P_cold, counts,
                       (const uint32_t*)flags);
  } else {
    allow_lds(k_histo_apply_paged<CT, HOT, RUNS, false>, lds_apply);
    hipLaunchKernelGGL((k_histo_apply_paged<CT, HOT, RUNS, false>), dim3((unsigned)groups),
                       dim3(1024), lds_apply, s, (const uint16_t*)staged, (const uint32_t*)tags,
                       (const uint32_t*)plist, (const uint32_t*)L.start, (const uint32_t*)L.segstart,
                       L.misc, (int)B, tb, (uint64_t)nbins, P, P_cold, counts,
                       (const uint32_t*)flags);
  }
  return (int)hipGetLastError();
}

// The paged plan reads the index once and is the faster one on every measured
// shape (MI355X, 2^24..2^28 updates, uniform and Zipf); the counted plan runs
// only where buffers are too shallow for pages.
template <bool PACKED, bool HOT, bool RUNS, typename CT>
static int bucket_any(const int64_t* idx, int64_t n, CT* counts, int64_t nbins,
                      const bale_config& c, Device* d, hipStream_t s) {
  int rc = bucket_paged_typed<PACKED, HOT, RUNS, CT>(idx, n, counts, nbins, c, d, s);
  if (rc != 1) return rc;
  return bucket_typed<PACKED, HOT, RUNS, CT>(idx, n, counts, nbins, c, d, s);
}

template <typename CT>
static int bucket_dispatch(const int64_t* idx, int64_t n, CT* counts, int64_t nbins, bool packed,
                           const bale_config& c, Device* d, hipStream_t s) {
  const bool hot = c.hot != 0, runs = c.merge > 0 || c.hot == 2;
  const int sel = (packed ? 4 : 0) | (hot ? 2 : 0) | (runs ? 1 : 0);
  switch (sel) {
    case 0: return bucket_any<false, false, false, CT>(idx, n, counts, nbins, c, d, s);
    case 1: return bucket_any<false, false, true, CT>(idx, n, counts, nbins, c, d, s);
    case 2: return bucket_any<false, true, false, CT>(idx, n, counts, nbins, c, d, s);
    case 3: return bucket_any<false, true, true, CT>(idx, n, counts, nbins, c, d, s);
    case 4: return bucket_any<true, false, false, CT>(idx, n, counts, nbins, c, d, s);
    case 5: return bucket_any<true, false, true, CT>(idx, n, counts, nbins, c, d, s);
    case 6: return bucket_any<true, true, false, CT>(idx, n, counts, nbins, c, d, s);
    default: return bucket_any<true, true, true, CT>(idx, n, counts, nbins, c, d, s);
  }
}

int launch_histo_bucket(const int64_t* idx, int64_t n, void* counts, int64_t nbins, int dtype,
                        bool packed, const bale_config& c, Device* d, hipStream_t s) {
  if (n <= 0) return 0;
  if (n >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;  // caller splits
  if (dtype == BALE_I64)
    return bucket_dispatch<int64_t>(idx, n, (int64_t*)counts, nbins, packed, c, d, s);
  return bucket_dispatch<int32_t>(idx, n, (int32_t*)counts, nbins, packed, c, d, s);
}

}  // namespace bale
