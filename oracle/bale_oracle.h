/*
 * bale_oracle.h -- CPU restatement of bale's histogram and index-gather
 * hot path, used ONLY as test infrastructure.
 *
 * This library is the checker.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library
 * (bale_old_amd/libbale_hip.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the jdevinney/bale_old tree).  How each function is pinned
 * (known-answer tests, closed forms, brute force, invariants) is listed
 * in tests/test_oracle.py; no function here is "parity unpinned".
 */
#ifndef BALE_ORACLE_H
#define BALE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- input generation (driver side of the path) ---------- */
/* splitmix64 (Steele, Lea, Flood 2014): one step of the stream. */
uint64_t or_splitmix64_next(uint64_t *state);
/* i-th output (0-based) of the splitmix64 stream seeded with `seed`. */
uint64_t or_splitmix64_at(uint64_t seed, uint64_t i);
/* Stateless table value: first output of the stream whose state is j. */
uint64_t or_splitmix64(uint64_t j);
/* Map a 64-bit draw to [0,n): top bits when n is a power of two, else the
 * multiply-high reduction (Lemire 2019). */
uint64_t or_reduce(uint64_t x, uint64_t n);
/* Keyed pseudo-random permutation of [0,n) (4-round Feistel network over
 * the smallest even-bit power-of-two domain >= n, with cycle walking). */
uint64_t or_perm(uint64_t x, uint64_t n, uint64_t key);

/* index[i] = reduce(stream(seed)[i], range) */
void or_gen_uniform(int64_t *out, int64_t n, int64_t range, uint64_t seed);
/* Zipf(s) ranks over [1,nbins] by rejection-inversion (Hormann & Derflinger
 * 1996), rank k mapped to bin perm(k-1, nbins, perm_key).  With
 * run_mean > 0, clustered: element i repeats the value of the start of its
 * run; i starts a run when i % 4096 == 0 or its flag draw hits 1/run_mean.
 * dist: 0 uniform, 1 zipf. */
void or_gen_index(int64_t *out, int64_t n, int64_t nbins, int dist, double s,
                  int64_t run_mean, uint64_t seed);
/* Expected probability of rank k (1-based) under Zipf(s) over n ranks. */
double or_zipf_pmf(int64_t k, int64_t n, double s);

/* table[j] = splitmix64(j) */
void or_fill_splitmix(int64_t *table, int64_t n);
/* table[j] = -(j+1): the reference ig driver's fill at THREADS==1
 * (apps/ig_src/ig.upc:159-161). */
void or_fill_neg(int64_t *table, int64_t n);
/* pckindx = ((idx / P) << 16) | (idx % P)   (apps/histo_src/histo.upc:146-150) */
void or_pack(int64_t *pck, const int64_t *idx, int64_t n, int64_t P);
/* The reference drivers' own inputs: srand(seed0 + pe); rand() % range
 * (apps/histo_src/histo.upc:143-146 with seed0=120348; apps/ig_src/ig.upc:168-170
 * with seed0=5).  Uses the C library's rand(), as the reference does. */
void or_gen_reference_rand(int64_t *out, int64_t n, int64_t range, unsigned seed0);

/* ---------------- histogram ------------------------------------------- */
/* AGI model: counts[index[i]] += 1 one update at a time
 * (apps/histo_src/histo_agi.upc:61-66, libgetput/libgetput.upc:48-60). */
void or_histo_agi(const int64_t *index, int64_t n, int64_t *counts);
void or_histo_agi_i32(const int64_t *index, int64_t n, int32_t *counts);
/* Aggregate-by-owner model over P virtual PEs: decode col = p>>16, pe =
 * p&0xffff, bucket by pe in push order, then each PE applies its bucket to its
 * local slice (apps/histo_src/histo_exstack.upc:64-76,
 * histo_conveyor.upc:72-82).  lcounts is P slices of lnum entries, pe-major.
 * Returns -1 if a decoded pe is >= P. */
int or_histo_owner(const int64_t *pckindx, int64_t n, int64_t *lcounts,
                   int64_t lnum, int64_t P);
/* Timed CPU baseline: nthreads threads split the updates, relaxed atomic add
 * per update (the AGI model on shared memory).  Returns the average over
 * threads of per-thread wall seconds between barriers
 * (libgetput/libgetput.upc:395-401). */
double or_histo_agi_mt(const int64_t *index, int64_t n, int64_t *counts, int nthreads);
/* Aggregate-by-owner with nthreads PEs (bin j owned by PE j % P), bulk
 * synchronous rounds of at most 1024 items per destination as exstack runs them
 * (apps/histo_src/histo_exstack.upc:64-76, histo_conveyor.upc:72-82); average
 * per-PE seconds between barriers (libgetput/libgetput.upc:395-401). */
double or_histo_owner_mt(const int64_t *index, int64_t n, int64_t *counts, int nthreads);
/* The request/reply form (apps/ig_src/ig_exstack.upc:69-94): requests to the
 * owner, replies in place, retraced in order. */
double or_ig_owner_mt(int64_t *tgt, const int64_t *index, int64_t n, const int64_t *table,
                      int nthreads);

/* ---------------- index gather ---------------------------------------- */
/* AGI model: tgt[i] = table[index[i]] (apps/ig_src/ig_agi.upc:61-64). */
void or_ig_agi(int64_t *tgt, const int64_t *index, int64_t n, const int64_t *table);
/* exstack model over P virtual PEs: requests are raw local indices pushed to
 * their owner in order; each owner replaces them with ltable values; the
 * requester retraces its requests in order popping from each source's FIFO
 * (apps/ig_src/ig_exstack.upc:69-94).  ltables is P slices of lnum, pe-major. */
int or_ig_exstack(int64_t *tgt, const int64_t *pckindx, int64_t n,
                  const int64_t *ltables, int64_t lnum, int64_t P);
/* conveyor model: {i, lindx} requests, {i, value} replies, tgt[i] = value
 * (apps/ig_src/ig_conveyor.upc:78-101). */
int or_ig_conveyor(int64_t *tgt, const int64_t *pckindx, int64_t n,
                   const int64_t *ltables, int64_t lnum, int64_t P);
double or_ig_agi_mt(int64_t *tgt, const int64_t *index, int64_t n,
                    const int64_t *table, int nthreads);

/* ---------------- batched mix ----------------------------------------- */
/* One instance of the batched uneven mix (see or_batch_params). */
typedef struct or_instance {
  int32_t kind;       /* 0 histogram, 1 gather */
  int32_t dist;       /* 0 uniform, 1 zipf, 2 clustered (uniform runs) */
  int64_t table_len;  /* bins (histogram) or table entries (gather) */
  int64_t nops;       /* updates or requests */
  double  s;          /* zipf exponent when dist == 1 */
  uint64_t seed;      /* index stream seed */
} or_instance;
void or_batch_params(or_instance *out, int64_t count, uint64_t seed,
                     int lo_tab, int hi_tab, int lo_ops, int hi_ops);

/* ---------------- conveyor delivery ---------------------------------- */
/* One source pushing n items (item_bytes each) to P destinations: every
 * destination receives the items addressed to it in push order
 * (convey/narrative.h:135-137; convey_push / convey_pull,
 * convey/convey.h:237,265).  out = the destinations' items one after another
 * (destination 0 first), starts[p] = first item of destination p, starts[P] =
 * items delivered.  Items naming a destination outside [0,P) are dropped;
 * returns how many. */
int64_t or_convey_deliver(const void *items, const int64_t *pes, int64_t n, int64_t item_bytes,
                          int64_t P, void *out, int64_t *starts);

/* ---------------- random permutation (dart throwing) ------------------ */
/* drand48 generator state after srand48(seed): (seed mod 2^32) << 16 | 0x330E. */
uint64_t or_srand48_state(int64_t seed);
/* n successive lrand48 values from a 48-bit state, restated from POSIX. */
void or_lrand48_seq(uint64_t state, int64_t n, int64_t *out);
/* glibc's own srand48(seed) then n calls of lrand48. */
void or_glibc_lrand48(int64_t seed, int64_t n, int64_t *out);
/* rand_permp_agi at one PE (spmat/spmat_agi.upc:56-107) into perm[0..N);
 * returns the darts thrown (-1: out of memory). */
int64_t or_randperm_agi(int64_t *perm, int64_t N, int64_t seed);
/* rand_permp_agi_opt at one PE (apps/randperm_src/alternates/randperm_agi_opt.upc):
 * N draws for the PE choice, then Fisher-Yates over 0..N-1 into perm[0..N). */
void or_randperm_agi_opt(int64_t *perm, int64_t N, int64_t seed);
/* is_perm (spmat/spmat_utils.upc:508): 1 if perm holds 0..N-1 once each. */
int or_is_perm(const int64_t *perm, int64_t N);

/* ---------------- sparse matrices (CSR at one PE) ---------------------- */
/* A matrix is offset[numrows + 1] (offset[0] = 0) and nonzero[offset[numrows]]:
 * row i's column indices are nonzero[offset[i] .. offset[i+1]). */

/* Natural logarithm of x in (0, 1] by the atanh series, in plain IEEE double
 * operations (no fused multiply-add) so the GPU computes the same bits. */
double or_det_log(double x);
/* Erdos-Renyi matrix of n rows with edge probability p drawn by geometric gaps
 * between nonzeros (Batagelj-Brandes; the reference uses the same method,
 * spmat/spmat_utils.upc:612-700), with splitmix64 streams per row and half:
 * mode 1 lower triangle, 2 upper, 3 both (asymmetric); the diagonal when
 * unit_diag.  offset always filled; nonzero written when not NULL.  Returns
 * the number of nonzeros. */
int64_t or_gen_er(int64_t n, double p, int unit_diag, int mode, uint64_t seed, int64_t *offset,
                  int64_t *nonzero);
/* transpose_matrix_agi at one PE (spmat/spmat_agi.upc:176-236). */
void or_transpose_agi(int64_t numrows, int64_t numcols, const int64_t *offset,
                      const int64_t *nonzero, int64_t *t_offset, int64_t *t_nonzero);
/* permute_matrix_agi at one PE (spmat/spmat_agi.upc:119-169): row i of A goes
 * to row rperminv[i], column j to column cperminv[j]. */
void or_permute_agi(int64_t numrows, const int64_t *offset, const int64_t *nonzero,
                    const int64_t *rperminv, const int64_t *cperminv, int64_t *p_offset,
                    int64_t *p_nonzero);

/* toposort_matrix_agi at one PE (apps/topo_src/toposort_agi.upc:53-166): mat
 * a permuted upper-triangular n x n matrix, tmat its transpose (rows sorted).
 * Writes rperm[row] and cperm[col] (the permutations that make mat upper
 * triangular) and, when `levels` is not NULL, the number of levels.  Returns
 * the pivots found (n on success). */
int64_t or_toposort_agi(int64_t n, const int64_t *offset, const int64_t *nonzero,
                        const int64_t *t_offset, const int64_t *t_nonzero, int64_t *rperm,
                        int64_t *cperm, int64_t *levels);

/* Order-independent 64-bit digest of an int64 array (sum of mix(v_i ^ i)). */
uint64_t or_digest(const int64_t *v, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
