// partition.h -- bucket-by-destination partitioning shared by the bucketed
// histogram and the bucketed gather.
//
// This is the on-chip form of bale's aggregate-by-destination idea (exstack /
// conveyor: push an item into the send buffer of its destination, ship the
// buffer when it is full).  Here a "destination" is a bucket of 2^shift
// consecutive bins or table entries, and a "PE" is a workgroup working on a
// fixed chunk of the input.
//
//   k_part_count    per-(chunk, bucket) counts, one row per workgroup, no atomics
//   k_part_offsets  column prefix over chunks: each chunk's first slot per bucket
//   k_part_scan     prefix over buckets: bucket starts and the apply segment list
//   part_scatter    one LDS send buffer per bucket; a full buffer is shipped with
//                   one coalesced write to the chunk's own, exactly sized range of
//                   the bucket (LDS cursor, no device atomics)
//
// Hot bins (histogram only): a sample of the input is sorted in LDS and bins
// seen at least twice become "hot"; they go to an LDS hash table in every
// workgroup (counted there, flushed once per workgroup) and skip partitioning.
// Zipf-like inputs put most of their mass on such bins.
#pragma once
#include "bale_internal.h"

namespace bale {

constexpr int kHotSlots = 2048;          // LDS hash slots (power of two)
constexpr int kHotMax = 1024;            // at most this many hot bins
constexpr uint32_t kHotEmpty = 0xFFFFFFFFu;
constexpr int kSketch = 4096;            // sampled indices for the hot set
// In RUNS mode a run of at least this many equal bins within an aligned group
// of 64 consecutive updates is added with one device atomic.  Both passes see
// the same groups because chunks are multiples of 64 updates and each wave
// handles one aligned group per step.
constexpr uint32_t kLongRun = 4;
constexpr int64_t kChunkAlign = 4096;

// Hot-bin table: kHotSlots slots in aligned pairs.  A bin lives in one of the
// two slots of the pair its hash names, or is left out of the hot set (it is
// then counted like any other bin), so a lookup is two adjacent LDS reads and
// no loop.
__device__ inline uint32_t hot_pair(uint32_t key) { return ((key * 0x9E3779B1u) >> (32 - 10)) << 1; }
static_assert(kHotSlots == 2 << 10, "hot_pair names kHotSlots / 2 pairs");

// Slot of `key` in the hot table `keys`, or -1.
__device__ inline int hot_find(const uint32_t* keys, uint32_t key) {
  const uint32_t h = hot_pair(key);
  const uint32_t k0 = keys[h], k1 = keys[h + 1];
  return k0 == key ? (int)h : k1 == key ? (int)(h + 1) : -1;
}

// Workspace layout (32-bit words, then the staging array).  misc: [0] total
// apply segments, [1] apply queue head, [2..9] per-XCD segment totals,
// [10..17] per-XCD queue heads.
constexpr int kMiscWords = 32;
struct PartLayout {
  uint32_t* chunk;     // [C*B] per-chunk counts, then per-chunk first slots
  uint32_t* start;     // [B+1] bucket starts
  uint32_t* total;     // [B]   items per bucket
  uint32_t* segstart;  // [B+1] first apply segment of each bucket
  uint32_t* misc;      // [kMiscWords] queue totals and heads
  uint32_t* hot;       // [kHotSlots] hot-bin hash table (keys)
  uint32_t* sketch;    // [kSketchScratch] sampled keys and pair codes
  void* staged;        // [n]   entries grouped by bucket
};

// Scratch of the two-kernel sketch: the sampled keys, then one pair code per sample.
constexpr int kSketchScratch = 2 * kSketch;

inline size_t part_words(int64_t C, int64_t B) {
  return (size_t)(C * B + (B + 1) + B + (B + 1) + kMiscWords + kHotSlots + kSketchScratch);
}
inline size_t part_bytes(int64_t C, int64_t B, int64_t n, size_t esize) {
  size_t head = ((part_words(C, B) * 4) + 255) & ~(size_t)255;
  return head + (size_t)n * esize + 256;
}
inline PartLayout part_carve(void* ws, int64_t C, int64_t B) {
  PartLayout L;
  uint32_t* w = (uint32_t*)ws;
  L.chunk = w;
  L.start = L.chunk + C * B;
  L.total = L.start + (B + 1);
  L.segstart = L.total + B;
  L.misc = L.segstart + (B + 1);
  L.hot = L.misc + kMiscWords;
  L.sketch = L.hot + kHotSlots;
  size_t head = ((part_words(C, B) * 4) + 255) & ~(size_t)255;
  L.staged = (char*)ws + head;
  return L;
}

// ---- hot-set sketch (one workgroup of 1024) -----------------------------------
// Samples kSketch indices at a fixed stride, counts them in an LDS hash table
// (2 * kSketch slots, so probing stays short and always finds a slot) and
// writes the hash table of bins that occur at least twice (at most kHotMax).
// With `flags`, also writes the sample's verdict for adaptive passes:
// bit kFlagHotBins when the repeated bins carry enough of the sample to be
// worth a hot table, bit kFlagRunsSeen when at least 1/8 of the sampled
// positions equal their right neighbour (clustered runs).
constexpr uint32_t kFlagHotBins = 1, kFlagRunsSeen = 2;
constexpr int kSketchSlots = 2 * kSketch;

__device__ inline uint32_t sketch_hash(uint32_t v) {
  return (v * 0x9E3779B1u) >> (32 - 13);  // kSketchSlots = 2^13
}
static_assert(kSketchSlots == (1 << 13), "sketch_hash assumes 2^13 slots");

// Sample k of kSketch sits at position floor(k * n / kSketch).
__device__ inline int64_t sketch_pos(int k, int64_t n) {
  return (int64_t)(((unsigned __int128)k * (uint64_t)n) / kSketch);
}

// Key of one sample (kHotEmpty when it is out of range or cannot be a hot
// bin) and its pair code: bit 0 = it has a right neighbour, bit 1 = the
// neighbour names the same bin.
template <bool PACKED>
__device__ inline uint32_t sketch_key(int64_t pos, int64_t n, uint64_t limit, int64_t r0, int64_t r1,
                                      bool pairs, uint32_t& code) {
  uint64_t j = 0, j2 = 0;
  unsigned br = 0, bp = 0;
  code = 0;
  const bool ok = pos < n && decode_index<PACKED>(r0, limit, j, br, bp);
  if (pairs && ok && pos + 1 < n)
    code = 1u | ((decode_index<PACKED>(r1, limit, j2, br, bp) && j2 == j) ? 2u : 0u);
  return ok && j < kHotEmpty ? (uint32_t)j : kHotEmpty;
}

// Sum of v over the 64 lanes of the wave (every lane gets it).
__device__ inline uint32_t wave_add(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// From the kSketch sampled keys (kPer per thread of a 1024-thread workgroup,
// sample tid + q * 1024 in key[q]) and the workgroup's pair counts: the hash
// table of bins that occur at least twice and, with `flags`, the verdict.
__device__ inline void sketch_tables(const uint32_t* key, uint32_t eq, uint32_t pairs, uint64_t limit,
                                     uint32_t* __restrict__ hot_out, uint32_t* __restrict__ flags) {
  __shared__ uint32_t keys[kSketchSlots];
  __shared__ uint32_t cnt2[kSketchSlots / 2];  // two 16-bit counts per word
  __shared__ uint32_t table[kHotSlots];
  __shared__ uint32_t nhot, neq, npairs, ndup;
  constexpr int kPer = kSketch / 1024;
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int k = tid; k < kSketchSlots; k += nthr) keys[k] = kHotEmpty;
  for (int k = tid; k < kSketchSlots / 2; k += nthr) cnt2[k] = 0;
  for (int k = tid; k < kHotSlots; k += nthr) table[k] = kHotEmpty;
  if (tid == 0) nhot = neq = npairs = ndup = 0;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const uint32_t v = key[q];
    if (v != kHotEmpty) {
      // at most kSketch keys in 2 * kSketch slots: the probe always ends
      uint32_t h = sketch_hash(v);
      for (;;) {
        uint32_t prev = atomicCAS(&keys[h], kHotEmpty, v);
        if (prev == kHotEmpty || prev == v) break;
        h = (h + 1) & (kSketchSlots - 1);
      }
      atomicAdd(&cnt2[h >> 1], 1u << (16 * (h & 1)));  // counts stay below 2^16
    }
  }
  if (flags) {
    // one LDS atomic per wave: same-address atomics from every lane serialise
    eq = wave_add(eq);
    pairs = wave_add(pairs);
    if ((tid & 63) == 0) {
      atomicAdd(&neq, eq);
      atomicAdd(&npairs, pairs);
    }
  }
  __syncthreads();
  // a bin is hot when it occurs at least twice in the sample
  uint32_t dup = 0;
  for (int h = tid; h < kSketchSlots; h += nthr) {
    uint32_t c = (cnt2[h >> 1] >> (16 * (h & 1))) & 0xffffu;
    if (c >= 2) {
      dup += c;
      if (atomicAdd(&nhot, 1u) < (uint32_t)kHotMax) {
        // first free slot of the bin's pair; a full pair leaves the bin cold
        const uint32_t v = keys[h], t = hot_pair(v);
        if (atomicCAS(&table[t], kHotEmpty, v) != kHotEmpty) (void)atomicCAS(&table[t + 1], kHotEmpty, v);
      }
    }
  }
  if (flags) {
    dup = wave_add(dup);
    if ((tid & 63) == 0) atomicAdd(&ndup, dup);
  }
  __syncthreads();
  for (int k = tid; k < kHotSlots; k += nthr) hot_out[k] = table[k];
  if (flags && tid == 0) {
    // Worth the hot table when the samples that share their bin with another
    // sample are at least 1/32 of the sample and 4x what uniform updates over
    // `limit` bins would give by chance (k(k-1)/limit for a sample of k).
    const double k = kSketch, chance = k * (k - 1.0) / (double)(limit ? limit : 1);
    const bool skewed = ndup * 32u >= (uint32_t)kSketch && (double)ndup >= 4.0 * chance;
    *flags = (skewed ? kFlagHotBins : 0u) | (npairs && neq * 8 >= npairs ? kFlagRunsSeen : 0u);
  }
  __syncthreads();
}

// The sketch in one workgroup of 1024 (the batch runs one per instance, side
// by side): every sample (and its right neighbour) is loaded before any is used.
template <bool PACKED>
__device__ inline void hot_sketch_body(const int64_t* __restrict__ idx, int64_t n, uint64_t limit,
                                       uint32_t* __restrict__ hot_out,
                                       uint32_t* __restrict__ flags = nullptr) {
  constexpr int kPer = kSketch / 1024;
  const int tid = threadIdx.x;
  int64_t r0s[kPer], r1s[kPer];
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int64_t pos = sketch_pos(tid + q * 1024, n);
    r0s[q] = pos < n ? idx[pos] : 0;
    r1s[q] = flags && pos + 1 < n ? idx[pos + 1] : 0;
  }
  uint32_t key[kPer], eq = 0, pairs = 0;
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    uint32_t code;
    key[q] = sketch_key<PACKED>(sketch_pos(tid + q * 1024, n), n, limit, r0s[q], r1s[q], flags != nullptr, code);
    pairs += code & 1u;
    eq += code >> 1;
  }
  sketch_tables(key, eq, pairs, limit, hot_out, flags);
}

// A single workgroup's scattered loads are bound by the misses one CU keeps
// in flight (~20 us for the 4096 samples and their neighbours).  Standalone
// calls spread the loads instead: k_sketch_sample (kSketchGroups workgroups,
// one sample per thread) writes each sample's key and pair code to scratch,
// and k_sketch_build (one workgroup) reads them back as streams.
constexpr int kSketchGroups = kSketch / 256;

template <bool PACKED>
__global__ __launch_bounds__(256) void k_sketch_sample(const int64_t* __restrict__ idx, int64_t n,
                                                       uint64_t limit, int pairs,
                                                       uint32_t* __restrict__ scratch) {
  const int k = blockIdx.x * 256 + threadIdx.x;  // grid kSketchGroups x 256 covers every sample
  const int64_t pos = sketch_pos(k, n);
  const int64_t r0 = pos < n ? idx[pos] : 0;
  const int64_t r1 = pairs && pos + 1 < n ? idx[pos + 1] : 0;
  uint32_t code;
  scratch[k] = sketch_key<PACKED>(pos, n, limit, r0, r1, pairs != 0, code);
  scratch[kSketch + k] = code;
}

[[maybe_unused]] static __global__ __launch_bounds__(1024) void k_sketch_build(const uint32_t* __restrict__ scratch, uint64_t limit,
                                                       uint32_t* __restrict__ hot_out,
                                                       uint32_t* __restrict__ flags) {
  constexpr int kPer = kSketch / 1024;
  uint32_t key[kPer], eq = 0, pairs = 0;
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int k = threadIdx.x + q * 1024;
    key[q] = scratch[k];
    const uint32_t code = scratch[kSketch + k];
    pairs += code & 1u;
    eq += code >> 1;
  }
  sketch_tables(key, eq, pairs, limit, hot_out, flags);
}

// Enqueue the two-kernel sketch; scratch holds kSketchScratch words.
template <bool PACKED>
inline void launch_hot_sketch(const int64_t* idx, int64_t n, uint64_t limit, uint32_t* hot_out,
                              uint32_t* flags, uint32_t* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_sketch_sample<PACKED>, dim3(kSketchGroups), dim3(256), 0, s, idx, n, limit,
                     flags ? 1 : 0, scratch);
  hipLaunchKernelGGL(k_sketch_build, dim3(1), dim3(1024), 0, s, (const uint32_t*)scratch, limit, hot_out,
                     flags);
}

// Runs of equal keys in adjacent lanes of the wave.  Returns the lane that heads
// this lane's run and sets len to the run's length.  Every lane of the wave
// must call it together.
__device__ inline int wave_runs(uint32_t key, uint32_t& len) {
  const int lane = __lane_id();
  uint32_t prev = __shfl_up(key, 1);
  bool change = (lane == 0) || (prev != key);
  uint64_t chg = __ballot(change);
  uint64_t upto = lane == 63 ? ~0ULL : ((2ULL << lane) - 1);  // lanes <= this one
  int head = 63 - __builtin_clzll(chg & upto);
  uint64_t later = chg & ~upto;
  int end = later ? __builtin_ctzll(later) : 64;
  len = (uint32_t)(end - head);
  return head;
}

// ---- pass 1: per-chunk bucket counts ---------------------------------------
// Counts chunk c's items per bucket into row c of `rows`.  sh: LDS of
// (B + kHotSlots) words.  Called by a whole workgroup.
template <bool PACKED, bool HOT, bool RUNS>
__device__ inline void part_count_body(uint32_t* sh, int64_t c, const int64_t* __restrict__ idx,
                                       int64_t n, uint64_t limit, int shift, int B, int64_t chunk,
                                       uint32_t* __restrict__ rows,
                                       const uint32_t* __restrict__ hot_g, ErrWords* err) {
  constexpr bool use_hot = HOT, use_runs = RUNS;
  uint32_t* hist = sh;              // [B]
  uint32_t* hot = sh + B;           // [kHotSlots] when HOT
  for (int b = threadIdx.x; b < B; b += blockDim.x) hist[b] = 0;
  if (use_hot)
    for (int k = threadIdx.x; k < kHotSlots; k += blockDim.x) hot[k] = hot_g[k];
  __syncthreads();
  int64_t lo = c * chunk;
  int64_t hi = lo + chunk < n ? lo + chunk : n;
  unsigned bad_r = 0, bad_p = 0;
  constexpr int U = 4;
  const int nthr = blockDim.x;
  // uniform loop bound: all lanes of a wave reach wave_runs together
  for (int64_t base = lo; base < hi; base += (int64_t)nthr * U) {
    int64_t raw[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      int64_t e = base + (int64_t)u * nthr + threadIdx.x;
      raw[u] = e < hi ? __builtin_nontemporal_load(idx + e) : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      int64_t e = base + (int64_t)u * nthr + threadIdx.x;
      uint32_t key = kHotEmpty;
      uint64_t j;
      if (e < hi && decode_index<PACKED>(raw[u], limit, j, bad_r, bad_p))
        if (!use_hot || hot_find(hot, (uint32_t)j) < 0) key = (uint32_t)(j >> shift);
      if (use_runs) {
        // a run of >= kLongRun equal bins is added directly by the scatter
        // pass and is not partitioned (same aligned 64-element groups there)
        uint32_t bin = key != kHotEmpty ? (uint32_t)j : kHotEmpty, blen;
        wave_runs(bin, blen);
        if (bin != kHotEmpty && blen >= kLongRun) key = kHotEmpty;
        uint32_t len;
        int head = wave_runs(key, len);
        if (key != kHotEmpty && __lane_id() == head) atomicAdd(&hist[key], len);
      } else if (key != kHotEmpty) {
        atomicAdd(&hist[key], 1u);
      }
    }
  }
  __syncthreads();
  uint32_t* row = rows + (size_t)c * B;
  for (int b = threadIdx.x; b < B; b += blockDim.x) row[b] = hist[b];
  flush_errors(err, bad_r, bad_p);
}

// `flags` (optional, written by the sketch) switches the HOT / RUNS handling
// off when the sample saw no hot bins / no runs; the scatter pass reads the
// same word, so both passes make the same choice.
template <bool PACKED, bool HOT, bool RUNS>
__global__ __launch_bounds__(1024) void k_part_count(const int64_t* __restrict__ idx, int64_t n,
                                                     uint64_t limit, int shift, int B,
                                                     int64_t chunk, uint32_t* __restrict__ rows,
                                                     const uint32_t* __restrict__ hot_g,
                                                     ErrWords* err,
                                                     const uint32_t* __restrict__ flags) {
  extern __shared__ uint32_t sh[];
  const int64_t c = blockIdx.x;
  const uint32_t fl = flags ? *flags : (kFlagHotBins | kFlagRunsSeen);
  const bool h = HOT && (fl & kFlagHotBins), r = RUNS && (fl & kFlagRunsSeen);
  if (h && r)
    part_count_body<PACKED, HOT, RUNS>(sh, c, idx, n, limit, shift, B, chunk, rows, hot_g, err);
  else if (h)
    part_count_body<PACKED, HOT, false>(sh, c, idx, n, limit, shift, B, chunk, rows, hot_g, err);
  else if (r)
    part_count_body<PACKED, false, RUNS>(sh, c, idx, n, limit, shift, B, chunk, rows, hot_g, err);
  else
    part_count_body<PACKED, false, false>(sh, c, idx, n, limit, shift, B, chunk, rows, hot_g, err);
}

// ---- pass 2a: column prefix over chunks (a workgroup of 256 per bucket) -----------
// rows[c][b] becomes the items of bucket b in chunks before c; total[b] the sum.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_part_offsets(
    uint32_t* __restrict__ rows, int C, int B, uint32_t* __restrict__ total) {
  __shared__ uint32_t part[256];
  const int b = blockIdx.x, t = threadIdx.x, T = blockDim.x;
  const int per = (C + T - 1) / T, c0 = t * per;
  uint32_t sum = 0;
  for (int c = c0; c < c0 + per && c < C; c++) sum += rows[(size_t)c * B + b];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    uint32_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0;
  for (int c = c0; c < c0 + per && c < C; c++) {
    uint32_t v = rows[(size_t)c * B + b];
    rows[(size_t)c * B + b] = run;
    run += v;
  }
  if (t == T - 1) total[b] = part[T - 1];
}

// ---- pass 2b: prefix over buckets (one workgroup, any size up to 1024) ------------
// start[b] = items before bucket b; segstart[b] = apply segments before bucket
// b; misc[0] = total segments, misc[1] = 0 (the apply queue head).
__device__ inline void part_scan_body(const uint32_t* __restrict__ count, int B, uint32_t seg_len,
                                      uint32_t* __restrict__ start,
                                      uint32_t* __restrict__ segstart,
                                      uint32_t* __restrict__ misc) {
  __shared__ uint32_t part_a[1024], part_b[1024];
  const int T = blockDim.x, t = threadIdx.x;
  int per = (B + T - 1) / T;
  int b0 = t * per;
  uint32_t sa = 0, sb = 0;
  for (int b = b0; b < b0 + per && b < B; b++) {
    sa += count[b];
    sb += (count[b] + seg_len - 1) / seg_len;
  }
  part_a[t] = sa;
  part_b[t] = sb;
  __syncthreads();
  // Hillis-Steele inclusive scan over the T partials
  for (int off = 1; off < T; off <<= 1) {
    uint32_t va = t >= off ? part_a[t - off] : 0;
    uint32_t vb = t >= off ? part_b[t - off] : 0;
    __syncthreads();
    part_a[t] += va;
    part_b[t] += vb;
    __syncthreads();
  }
  uint32_t ra = t ? part_a[t - 1] : 0;
  uint32_t rb = t ? part_b[t - 1] : 0;
  for (int b = b0; b < b0 + per && b < B; b++) {
    start[b] = ra;
    segstart[b] = rb;
    ra += count[b];
    rb += (count[b] + seg_len - 1) / seg_len;
  }
  if (t == T - 1) {
    start[B] = part_a[T - 1];
    segstart[B] = part_b[T - 1];
    misc[0] = part_b[T - 1];
    misc[1] = 0;
  }
  __syncthreads();
}

// misc[kSegWord]: the apply segment length k_part_scan chose (the last word
// holds the sketch's flags).
constexpr int kSegWord = kMiscWords - 2;

// seg_alt, when nonzero, replaces seg_len if the sketch (flags) found no hot
// bins: near-uniform buckets are about equally full, so segments of twice the
// average leave every bucket to one workgroup (a plain-add flush), while
// skewed inputs keep the shorter segments that spread their heavy buckets.
[[maybe_unused]] static __global__ __launch_bounds__(1024) void k_part_scan(const uint32_t* __restrict__ count, int B,
                                                           uint32_t seg_len, uint32_t seg_alt,
                                                           const uint32_t* __restrict__ flags,
                                                           uint32_t* __restrict__ start,
                                                           uint32_t* __restrict__ segstart,
                                                           uint32_t* __restrict__ misc) {
  const uint32_t seg = seg_alt && flags && !(*flags & kFlagHotBins) ? seg_alt : seg_len;
  part_scan_body(count, B, seg, start, segstart, misc);
  if (threadIdx.x == 0) misc[kSegWord] = seg;
}

// ---- pass 3: scatter through LDS send buffers -------------------------------------
// LDS layout: cursor[B], fill[B], list[B], nlist[2], [hot keys, hot counts], buf[B*S].
inline size_t part_scatter_lds(int B, int S, size_t esize, bool hot) {
  size_t head = (size_t)(3 * B + 2 + 64) * 4 + (hot ? (size_t)2 * kHotSlots * 4 : 0);
  return ((head + 15) & ~(size_t)15) + (size_t)B * S * esize;
}
// Largest power-of-two buffer depth S in [4, 256] that fits `budget` bytes (0 if none).
inline int part_pick_depth(int B, size_t esize, size_t budget, bool hot) {
  int S = 256;
  while (S > 4 && part_scatter_lds(B, S, esize, hot) > budget) S >>= 1;
  return part_scatter_lds(B, S, esize, hot) <= budget ? S : 0;
}

constexpr int kScatterU = 8;  // elements per thread per round (loaded a round ahead)

// MakeEntry(e, j, low) -> entry stored for element e with decoded index j
// (low = j & ((1<<shift)-1)).  HotFlush(bin, count) adds count to a bin
// directly (hot counters at the end, long runs as they are met).
// RUNS: adjacent lanes with the same bucket (or hot bin) reserve their slots
// with one LDS atomic per run, and runs of >= kLongRun equal bins skip the
// partition (clustered inputs).  chunk must be a multiple of 64.
// Store(pos, entry) writes one shipped entry to staging position pos;
// Skip(e) is called for an element whose index is out of range.
template <bool PACKED, bool HOT, bool RUNS, typename E, typename MakeEntry, typename HotFlush,
          typename Store, typename Skip>
__device__ inline void part_scatter_body(int64_t c, const int64_t* __restrict__ idx, int64_t n,
                                         uint64_t limit, int shift, int B, int S, int64_t chunk,
                                         const uint32_t* __restrict__ rows,
                                         const uint32_t* __restrict__ start,
                                         const uint32_t* __restrict__ hot_g, Store store,
                                         MakeEntry make, HotFlush hot_flush, Skip skip) {
  constexpr bool use_hot = HOT, use_runs = RUNS;
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* cursor = (uint32_t*)smem;  // [B] next staging slot of this chunk per bucket
  uint32_t* fill = cursor + B;         // [B]
  uint32_t* list = fill + B;           // [B] buckets to ship this step
  uint32_t* nlist = list + B;          // [2] list lengths, alternating between steps
  uint32_t* sink = nlist + 2;          // [64] per-lane target of placements that place nothing
  uint32_t* hkeys = sink + 64;         // [kHotSlots]
  uint32_t* hcnt = hkeys + kHotSlots;  // [kHotSlots]
  size_t head = (size_t)(3 * B + 2 + 64) * 4 + (HOT ? (size_t)2 * kHotSlots * 4 : 0);
  E* buf = (E*)(smem + ((head + 15) & ~(size_t)15));  // [B*S]
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = __lane_id(), wave = tid >> 6, nwaves = nthr >> 6;
  const uint32_t* row = rows + (size_t)c * B;
  for (int b = tid; b < B; b += nthr) {
    cursor[b] = start[b] + row[b];
    fill[b] = 0;
  }
  if (use_hot)
    for (int k = tid; k < kHotSlots; k += nthr) {
      hkeys[k] = hot_g[k];
      hcnt[k] = 0;
    }
  if (tid == 0) nlist[0] = nlist[1] = 0;
  __syncthreads();
  const uint64_t mask = (1ULL << shift) - 1;
  int64_t lo = c * chunk;
  int64_t hi = lo + chunk < n ? lo + chunk : n;
  uint32_t p = 0;  // which list length this step appends to

  // Ship every full buffer listed this step: wave w takes list entries
  // w, w + nwaves, ...; its lanes reserve the entries' output ranges at once
  // (a bucket is listed at most once per step, so its cursor has one owner),
  // then copy 64/S buffers per instruction.
  auto ship_listed = [&](uint32_t nl) {
    for (uint32_t g0 = wave; g0 < nl; g0 += (uint32_t)nwaves * 64u) {
      uint32_t k = g0 + (uint32_t)lane * nwaves;
      uint32_t b_l = 0, pos_l = 0;
      if (k < nl) {
        b_l = list[k];
        pos_l = cursor[b_l];
        cursor[b_l] = pos_l + (uint32_t)S;
        fill[b_l] -= (uint32_t)S;  // entries past S (overflow) stay counted
      }
      const uint32_t cnt = min(64u, (nl - g0 + nwaves - 1) / (uint32_t)nwaves);
      if (S <= 64) {
        const uint32_t per = 64u / (uint32_t)S;
        const uint32_t q = (uint32_t)lane % (uint32_t)S;
        for (uint32_t m0 = 0; m0 < cnt; m0 += per) {
          uint32_t m = m0 + (uint32_t)lane / (uint32_t)S;
          uint32_t b = __shfl(b_l, (int)(m & 63u)), pos = __shfl(pos_l, (int)(m & 63u));
          if (m < cnt) store(pos + q, buf[(size_t)b * S + q]);
        }
      } else {
        for (uint32_t m = 0; m < cnt; m++) {
          uint32_t b = __shfl(b_l, (int)m), pos = __shfl(pos_l, (int)m);
          for (uint32_t q = lane; q < (uint32_t)S; q += 64) store(pos + q, buf[(size_t)b * S + q]);
        }
      }
    }
  };

  // one wave ships the first `cnt` entries of bucket b to the chunk's range
  auto ship = [&](int b, uint32_t cnt) {
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(&cursor[b], cnt);  // LDS atomic
    pos = __shfl(pos, 0);
    for (uint32_t k = lane; k < cnt; k += 64) store(pos + k, buf[(size_t)b * S + k]);
  };

  const int64_t step = (int64_t)nthr * kScatterU;
  // indices of the next round are loaded while this round is placed
  int64_t next[kScatterU];
#pragma unroll
  for (int u = 0; u < kScatterU; u++) {
    int64_t e = lo + (int64_t)u * nthr + tid;
    next[u] = e < hi ? __builtin_nontemporal_load(idx + e) : 0;
  }
  for (int64_t base = lo; base < hi; base += step) {
    bool pending[kScatterU];
    int bk[kScatterU];
    E ent[kScatterU];
    int64_t raw[kScatterU];
#pragma unroll
    for (int u = 0; u < kScatterU; u++) {
      raw[u] = next[u];
      int64_t e = base + step + (int64_t)u * nthr + tid;
      next[u] = e < hi ? __builtin_nontemporal_load(idx + e) : 0;
    }
#pragma unroll
    for (int u = 0; u < kScatterU; u++) {
      int64_t e = base + (int64_t)u * nthr + tid;
      pending[u] = false;
      bk[u] = 0;
      int hs = -1;
      uint64_t j = 0;
      bool ok = false;
      if (e < hi) {
        unsigned br = 0, bp = 0;
        ok = decode_index<PACKED>(raw[u], limit, j, br, bp);
        if (ok && use_hot) hs = hot_find(hkeys, (uint32_t)j);
        if (!ok) skip(e);
      }
      if (use_hot) {
        if (use_runs) {
          uint32_t len;
          int head = wave_runs(hs >= 0 ? (uint32_t)hs : kHotEmpty, len);
          if (hs >= 0 && lane == head) atomicAdd(&hcnt[hs], len);
        } else if (hs >= 0) {
          atomicAdd(&hcnt[hs], 1u);
        }
      }
      bool cold = ok && hs < 0;
      if (use_runs) {
        // long runs of one bin are added directly (the count pass skips them)
        uint32_t len;
        int head = wave_runs(cold ? (uint32_t)j : kHotEmpty, len);
        if (cold && len >= kLongRun) {
          if (lane == head) hot_flush((uint32_t)j, len);
          cold = false;
        }
      }
      if (cold) {
        bk[u] = (int)(j >> shift);
        ent[u] = make(e, j, j & mask);
        pending[u] = true;
      }
    }
    // Place: one slot per element.  A slot past the buffer (overflow) is
    // kept and used after the buffer is shipped.  All of a round's slot
    // atomics are issued before any result is used (lanes with nothing to
    // place add to their own sink word, so no lane branches around them).
    uint32_t slot[kScatterU];
    int rhead[kScatterU];
#pragma unroll
    for (int u = 0; u < kScatterU; u++) {
      uint32_t len = 1;
      rhead[u] = lane;
      if (use_runs) rhead[u] = wave_runs(pending[u] ? (uint32_t)bk[u] : kHotEmpty, len);
      const bool adds = pending[u] && lane == rhead[u];
      slot[u] = atomicAdd(adds ? &fill[bk[u]] : &sink[lane], len);
    }
    bool any = false;
#pragma unroll
    for (int u = 0; u < kScatterU; u++) {
      if (use_runs) slot[u] = __shfl(slot[u], rhead[u]) + (uint32_t)(lane - rhead[u]);
      if (pending[u] && slot[u] < (uint32_t)S) {
        buf[(size_t)bk[u] * S + slot[u]] = ent[u];
        pending[u] = false;
        if (slot[u] == (uint32_t)S - 1) list[atomicAdd(&nlist[p], 1u)] = (uint32_t)bk[u];
      }
      any = any || pending[u];
    }
    for (;;) {
      __syncthreads();
      // ships take S entries off each listed bucket; the other list length
      // (last used a round ago, whose readers are past the barrier above) is
      // cleared here
      ship_listed(nlist[p]);
      if (tid == 0) nlist[p ^ 1] = 0;
      p ^= 1;
      // ships have read their buffers before overflow entries (and the next
      // round's placements) are written into them
      if (!__syncthreads_or(any)) break;
      any = false;
#pragma unroll
      for (int u = 0; u < kScatterU; u++) {
        if (pending[u]) {
          slot[u] -= (uint32_t)S;
          if (slot[u] < (uint32_t)S) {
            buf[(size_t)bk[u] * S + slot[u]] = ent[u];
            pending[u] = false;
            if (slot[u] == (uint32_t)S - 1) list[atomicAdd(&nlist[p], 1u)] = (uint32_t)bk[u];
          }
        }
        any = any || pending[u];
      }
      // more than S entries past a full buffer: ship again (rare)
      if (!__syncthreads_or(any)) break;
    }
  }
  __syncthreads();
  for (int b = wave; b < B; b += nwaves) {
    uint32_t f = fill[b];
    if (f) ship(b, f);
  }
  if (use_hot)
    for (int k = tid; k < kHotSlots; k += nthr)
      if (hcnt[k]) hot_flush(hkeys[k], hcnt[k]);
}

// The scatter pass under the count pass's choice from `flags` (see
// k_part_count).  Without the hot table, its LDS goes to deeper buffers:
// S_cold entries per bucket (the caller sized the allocation for both).
template <bool PACKED, bool HOT, bool RUNS, typename E, typename MakeEntry, typename HotFlush,
          typename Store, typename Skip>
__device__ inline void part_scatter_adaptive(int64_t c, const int64_t* __restrict__ idx, int64_t n,
                                             uint64_t limit, int shift, int B, int S, int S_cold,
                                             int64_t chunk, const uint32_t* __restrict__ rows,
                                             const uint32_t* __restrict__ start,
                                             const uint32_t* __restrict__ hot_g, Store store,
                                             MakeEntry make, HotFlush hot_flush, Skip skip,
                                             const uint32_t* __restrict__ flags) {
  const uint32_t fl = flags ? *flags : (kFlagHotBins | kFlagRunsSeen);
  const bool h = HOT && (fl & kFlagHotBins), r = RUNS && (fl & kFlagRunsSeen);
  const int Sc = S_cold > 0 ? S_cold : S;
  if (h && r)
    part_scatter_body<PACKED, HOT, RUNS, E>(c, idx, n, limit, shift, B, S, chunk, rows, start,
                                            hot_g, store, make, hot_flush, skip);
  else if (h)
    part_scatter_body<PACKED, HOT, false, E>(c, idx, n, limit, shift, B, S, chunk, rows, start,
                                             hot_g, store, make, hot_flush, skip);
  else if (r)
    part_scatter_body<PACKED, false, RUNS, E>(c, idx, n, limit, shift, B, HOT ? Sc : S, chunk,
                                              rows, start, hot_g, store, make, hot_flush, skip);
  else
    part_scatter_body<PACKED, false, false, E>(c, idx, n, limit, shift, B, HOT ? Sc : S, chunk,
                                               rows, start, hot_g, store, make, hot_flush, skip);
}

}  // namespace bale
