/*
 * bale_hip.h -- C-ABI of libbale_hip.so, the MI355X (gfx950) implementation
 * of bale's histogram scatter-add and index gather.
 *
 * Two groups of entry points:
 *
 *  1. Drop-ins with the reference's exact prototypes (SHARED expands to
 *     nothing, as in the reference's SHMEM build, libgetput/libgetput.h:151).
 *     They run on one GPU, which is one PE (THREADS == 1), so a packed
 *     index is (lindx << 16) | 0.  Pointers may be host or device memory:
 *     host arrays are staged to the GPU and back (histograms are added to,
 *     never overwritten; gathers overwrite tgt[0..n)).  Each returns the GPU
 *     seconds of the operation with data resident (HIP events), or -1.0 on
 *     error (bale_hip_last_error() says why), as the reference returns -1.0
 *     when setup fails (apps/histo_src/histo_exstack.upc:58).
 *
 *  2. The native API (bale_*): device pointers, explicit sizes, counter
 *     type, strategy and tile shape, asynchronous on a caller's HIP stream.
 *
 * Indices are bounds-checked on the GPU: an index outside the table is
 * skipped and counted, and the call reports BALE_ERR_RANGE (the reference
 * does no check; an out-of-range index there is undefined behaviour).
 */
#ifndef BALE_HIP_H
#define BALE_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================= 1. reference drop-ins ======================== */

/* apps/histo_src/histo.h:47 -- AGI model: counts[index[i]] += 1.
 * Runs the direct-atomic strategy. */
double histo_agi(int64_t *index, int64_t len, int64_t *counts);
/* apps/histo_src/histo.h:48 -- exstack model: packed indices, bulk-synchronous
 * aggregation.  Runs the bucket-by-destination strategy (partition by bin
 * range, then LDS histogram per bucket).  buf_cnt is accepted and ignored. */
double histo_exstack(int64_t *pckindx, int64_t len, int64_t *lcounts, int64_t buf_cnt);
/* apps/histo_src/histo.h:49 -- exstack2 model: runs the LDS-privatized
 * strategy (one pass, per-workgroup sub-histograms). */
double histo_exstack2(int64_t *pckindx, int64_t len, int64_t *lcounts, int64_t buf_cnt);
/* apps/histo_src/histo.h:50 -- conveyor model: the library picks the
 * strategy (autotuner cache, else its default rule). */
double histo_conveyor(int64_t *pckindx, int64_t len, int64_t *lcounts);
/* apps/histo_src/alternates/histo_alternates.h:43-46 -- same results; the
 * goto/cyclic/function variants differ only in control flow in the
 * reference.  They map to: goto -> bucket, cyclic -> LDS with adjacent-run
 * merging, exstack_function -> bucket, exstack2_function -> LDS. */
double histo_exstack2_goto(int64_t *pckindx, int64_t T, int64_t *lcounts, int64_t buf_cnt);
double histo_exstack2_cyclic(int64_t *pckindx, int64_t T, int64_t *lcounts, int64_t buf_cnt);
double histo_exstack_function(int64_t *pckindx, int64_t T, int64_t *lcounts, int64_t buf_cnt);
double histo_exstack2_function(int64_t *pckindx, int64_t T, int64_t *lcounts, int64_t buf_cnt);

/* apps/ig_src/ig.h:49 -- AGI model: tgt[i] = table[index[i]] (direct gather). */
double ig_agi(int64_t *tgt, int64_t *index, int64_t l_num_req, int64_t *table);
/* apps/ig_src/ig.h:50 -- exstack model: packed indices; runs the bucketed
 * gather (requests grouped by table region, replies written to their slot). */
double ig_exstack(int64_t *tgt, int64_t *pckindx, int64_t l_num_req, int64_t *ltable, int64_t buf_cnt);
/* apps/ig_src/ig.h:51 -- exstack2 model: packed indices, direct gather. */
double ig_exstack2(int64_t *tgt, int64_t *pckindx, int64_t l_num_req, int64_t *ltable, int64_t buf_cnt);
/* apps/ig_src/ig.h:52 -- conveyor model: library-picked strategy. */
double ig_conveyor(int64_t *tgt, int64_t *pckindx, int64_t l_num_req, int64_t *ltable);
/* apps/ig_src/alternates/ig_alternates.h:42-45 -- same results:
 * exstack2_cyclic / exstack2_goto -> direct, exstack_function / exstack_pkg
 * -> bucketed. */
double ig_exstack2_cyclic(int64_t *tgt, int64_t *pckindx, int64_t l_num_req, int64_t *ltable, int64_t buf_cnt);
double ig_exstack2_goto(int64_t *tgt, int64_t *pckindx, int64_t l_num_req, int64_t *ltable, int64_t buf_cnt);
double ig_exstack_function(int64_t *tgt, int64_t *pckindx, int64_t l_num_req, int64_t *ltable, int64_t buf_cnt);
double ig_exstack_pkg(int64_t *tgt, int64_t *pckindx, int64_t l_num_req, int64_t *ltable, int64_t buf_cnt);

/* spmat/spmat.h:115-118 -- a random permutation of 0..N-1 by dart throwing
 * (spmat/spmat_agi.upc:56-107).  Returns device memory from bale_all_alloc
 * (free with bale_all_free, as the reference frees with lgp_all_free), NULL on
 * error; GPU seconds in bale_hip_last_timing.  Every model gives the result of
 * the reference's rand_permp_agi at one PE for the same seed (the reference's
 * exstack / exstack2 / conveyor models re-throw rejected darts in batches and
 * so give other, equally valid, permutations).  N < 2^30; buf_cnt is ignored. */
int64_t *rand_permp_agi(int64_t N, int seed);
int64_t *rand_permp_exstack(int64_t N, int seed, int64_t buf_cnt);
int64_t *rand_permp_exstack2(int64_t N, int seed, int64_t buf_cnt);
int64_t *rand_permp_conveyor(int64_t N, int seed);
/* apps/randperm_src/randperm_alternates.h:42 -- the alternate model
 * (apps/randperm_src/alternates/randperm_agi_opt.upc): N generator draws for
 * the PE choice, then a Fisher-Yates pass over 0..N-1, resolved in parallel;
 * the reference's result at one PE for the same seed.  Device memory from
 * bale_all_alloc, NULL on error; N < 2^30. */
int64_t *rand_permp_agi_opt(int64_t N, int seed);
/* spmat/spmat.h:152 -- 1 if perm (device or host memory) holds 0..N-1 exactly once. */
int is_perm(int64_t *perm, int64_t N);
/* spmat/spmat_utils.upc:92-107 -- the reference's wrapper (its default model, AGI). */
int64_t *rand_permp(int64_t N, int seed);

/* Sparse matrices: the reference's sparsemat_t (spmat/spmat.h:73-86), field
 * for field.  At one PE the local and global views coincide: lnumrows ==
 * numrows, lnnz == nnz, loffset == offset, lnonzero == nonzero.  Matrices the
 * library makes hold device memory (bale_all_alloc); inputs may hold device or
 * host memory (host arrays are staged).  Row i's column indices are
 * nonzero[offset[i] .. offset[i+1]). */
typedef struct sparsemat_t {
  int64_t local;
  int64_t numrows;
  int64_t lnumrows;
  int64_t numcols;
  int64_t nnz;
  int64_t lnnz;
  int64_t *offset;
  int64_t *loffset;
  int64_t *nonzero;
  int64_t *lnonzero;
} sparsemat_t;

/* spmat/spmat.h:125-128 (spmat/spmat_agi.upc:176-236) -- the transpose; every
 * row of it lists its entries in increasing row of A, as the reference's AGI
 * model does at one PE.  The four models give the same matrix; buf_cnt is
 * ignored.  NULL on error (bale_hip_last_error).  Needs numrows, numcols below
 * 2^32 and nnz below 2^31. */
sparsemat_t *transpose_matrix_agi(sparsemat_t *A);
sparsemat_t *transpose_matrix_exstack(sparsemat_t *A, int64_t buf_cnt);
sparsemat_t *transpose_matrix_exstack2(sparsemat_t *A, int64_t buf_cnt);
sparsemat_t *transpose_matrix_conveyor(sparsemat_t *A);
/* spmat/spmat_utils.upc:109-127 -- the reference's wrapper. */
sparsemat_t *transpose_matrix(sparsemat_t *A);
/* spmat/spmat.h:120-123 (spmat/spmat_agi.upc:119-169) -- row i of A becomes
 * row rperminv[i], column j becomes column cperminv[j]; rows keep the order of
 * their entries.  rperminv / cperminv: device memory (as rand_permp returns). */
sparsemat_t *permute_matrix_agi(sparsemat_t *A, int64_t *rperminv, int64_t *cperminv);
sparsemat_t *permute_matrix_exstack(sparsemat_t *A, int64_t *rperminv, int64_t *cperminv,
                                    int64_t buf_cnt);
sparsemat_t *permute_matrix_exstack2(sparsemat_t *A, int64_t *rperminv, int64_t *cperminv,
                                     int64_t buf_cnt);
sparsemat_t *permute_matrix_conveyor(sparsemat_t *A, int64_t *rperminv, int64_t *cperminv);
/* spmat/spmat_utils.upc:92-107 -- the reference's wrapper. */
sparsemat_t *permute_matrix(sparsemat_t *A, int64_t *rperminv, int64_t *cperminv);
/* spmat/spmat_utils.upc:554-610 -- an Erdos-Renyi matrix of n rows, edge
 * probability p, by geometric gaps between nonzeros (the reference's method);
 * mode 1 lower triangle, 2 upper, 3 both halves (asymmetric); the diagonal
 * when unit_diag.  The gaps come from splitmix64 streams per (row, half), not
 * from the reference's rand() sequence, so a seed gives another matrix of the
 * same distribution.  Mode 0 (symmetric) is not provided.  Device memory. */
sparsemat_t *gen_erdos_renyi_graph_dist(int n, double p, int64_t unit_diag, int64_t mode,
                                        int64_t seed);
/* spmat/spmat_utils.upc:1153-1183 -- zeroed device arrays: numrows + 1
 * offsets, nnz nonzeros. */
sparsemat_t *init_matrix(int64_t numrows, int64_t numcols, int64_t nnz);
/* spmat/spmat_utils.upc:1211-1220 -- frees the arrays (not the struct). */
void clear_matrix(sparsemat_t *mat);
/* spmat/spmat_utils.upc:1070-1125 -- 0 when equal (sizes, offsets, nonzeros),
 * 1 otherwise; compared on the GPU. */
int compare_matrix(sparsemat_t *lmat, sparsemat_t *rmat);
/* spmat/spmat_utils.upc:377-453 -- 1 when no entry lies below (upper) / above
 * (lower) the diagonal and, with unit_diagonal, every row has its diagonal
 * entry; 0 otherwise. */
int is_upper_triangular(sparsemat_t *A, int64_t unit_diagonal);
int is_lower_triangular(sparsemat_t *A, int64_t unit_diagonal);

/* apps/topo_src/toposort.h:47-50 and alternates/toposort_alternates.h:42-44
 * (apps/topo_src/toposort_agi.upc:53-166) -- rperm[row], cperm[col] (device
 * or host memory, numrows entries each) such that permute_matrix(mat, rperm, cperm)
 * is upper triangular with a full diagonal; mat is a row and column
 * permutation of such a matrix and tmat its transpose.  Every model returns
 * the permutations toposort_matrix_agi finds at one PE (the other models
 * process each level's queue in another order and may find other valid
 * ones).  Returns GPU seconds, or -1.0 when mat is not morally triangular or
 * on error (bale_hip_last_error). */
double toposort_matrix_agi(int64_t *rperm, int64_t *cperm, sparsemat_t *mat, sparsemat_t *tmat);
double toposort_matrix_exstack(int64_t *rperm, int64_t *cperm, sparsemat_t *mat, sparsemat_t *tmat);
double toposort_matrix_exstack2(int64_t *rperm, int64_t *cperm, sparsemat_t *mat, sparsemat_t *tmat);
double toposort_matrix_convey(int64_t *rperm, int64_t *cperm, sparsemat_t *mat, sparsemat_t *tmat);
double toposort_matrix_oo(int64_t *rperm, int64_t *cperm, sparsemat_t *mat, sparsemat_t *tmat);
double toposort_matrix_cooler(int64_t *rperm, int64_t *cperm, sparsemat_t *mat, sparsemat_t *tmat);
double toposort_matrix_exstack_orig(int64_t *rperm, int64_t *cperm, sparsemat_t *mat,
                                    sparsemat_t *tmat);

/* ============================ 2. native API =========================== */

#define BALE_OK          0
#define BALE_ERR_ARG    -1  /* bad argument (null pointer, size, enum) */
#define BALE_ERR_RANGE  -2  /* some indices were outside [0, table_len): skipped */
#define BALE_ERR_HIP    -3  /* a HIP runtime call failed */
#define BALE_ERR_ALLOC  -4  /* device memory could not be allocated */
#define BALE_ERR_PE     -5  /* a packed index names a PE other than 0 */
#define BALE_ERR_NODEV  -6  /* no usable GPU */

/* strategies */
#define BALE_STRAT_AUTO    0
#define BALE_STRAT_DIRECT  1  /* histogram: one device atomic per update; gather: one load per request */
#define BALE_STRAT_LDS     2  /* histogram: LDS sub-histograms per workgroup, bin-range tiled */
#define BALE_STRAT_BUCKET  3  /* partition by destination range in LDS, coalesced flush, then apply */

/* counter types */
#define BALE_I64 0
#define BALE_I32 1

/* index formats */
#define BALE_INDEX_PLAIN  0  /* global index */
#define BALE_INDEX_PACKED 1  /* (lindx << 16) | pe, pe must be 0 */

/* wave-level merging of duplicate bins before the add */
#define BALE_MERGE_NONE   0
#define BALE_MERGE_RUNS   1  /* adjacent lanes with equal bins are combined */
#define BALE_MERGE_LEADER 2  /* plus up to 4 rounds of leader-key ballot merging */

typedef struct bale_config {
  int32_t strategy;   /* BALE_STRAT_* */
  int32_t block;      /* threads per workgroup: 256, 512 or 1024; 0 = default */
  int32_t unroll;     /* items per lane per step: 4, 8 or 16; 0 = default */
  int32_t tile_log2;  /* LDS bin tile / bucket width in log2 bins; 0 = default */
  int32_t merge;      /* BALE_MERGE_*; -1 = default */
  int32_t groups_per_cu; /* resident workgroups per CU for persistent passes; 0 = default */
  int32_t hot;        /* histogram DIRECT/BUCKET: 1 = count the bins a 4096-index sample
                         sees twice in a per-workgroup LDS hash table; 0 = off;
                         BUCKET only: 2 = let that sample decide on the device, per
                         call, whether to use the hot table and run merging, and (with
                         task_log2 = 0) the apply segment length */
  int32_t task_log2;  /* 0 or 8..30.  Batches: operations per task (log2, 12..22); 0 = 16.
                         BUCKET gathers: about 2^task_log2 requests per region segment.
                         BUCKET histograms: updates per apply segment (log2); 0 = by size
                         (with hot = 2: twice the average bucket when the sample finds
                         no hot bins, half of it, at least 16384, when it does) */
} bale_config;

/* Select the device for this thread's later calls and create its state.
 * Optional: the first call on a device does it. */
int bale_hip_init(int device);
/* Text of the last error on this thread ("" if none). */
const char *bale_hip_last_error(void);
/* GPU seconds, host-to-device and device-to-host seconds of the last
 * drop-in call on this thread. */
void bale_hip_last_timing(double *gpu_s, double *h2d_s, double *d2h_s);

/* counts[index[i]] += 1 for i < n (counts is int64_t or int32_t per dtype).
 * Asynchronous on `stream` (a hipStream_t; NULL = the null stream).  Range
 * and PE errors are accumulated on the device; read them with
 * bale_sync_errors(). */
int bale_histo(const int64_t *index, int64_t n, void *counts, int64_t nbins,
               int dtype, int index_format, const bale_config *cfg, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): bale_histo with the same results, whose workspace is one
 * fixed array of BALE_NOALLOC_SCRATCH_BYTES per device instead of memory the
 * library allocates; it calls no hipMalloc or hipFree.  A call whose plan
 * needs more scratch than that returns BALE_ERR_ARG, saying how much (about
 * 2 bytes per update for the bucketed plan, so up to ~2^24 updates; DIRECT
 * and LDS need none beyond the hot-bin sample, 40 KiB).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (on other streams or
 * threads) race, and the caller must order them.  errors: NULL, or device
 * memory for two int64 counters {indices outside the table, packed indices
 * naming another PE} that the call adds to (they are not reported through
 * bale_sync_errors). */
#define BALE_NOALLOC_SCRATCH_BYTES (64ll << 20)
int bale_histo_noalloc(const int64_t *index, int64_t n, void *counts, int64_t nbins,
                       int dtype, int index_format, const bale_config *cfg, int64_t *errors,
                       void *stream);
/* tgt[i] = table[index[i]] for i < n. */
int bale_gather(int64_t *tgt, const int64_t *index, int64_t n, const int64_t *table,
                int64_t table_len, int index_format, const bale_config *cfg, void *stream);
/* Wait for `stream`, then return BALE_OK, BALE_ERR_RANGE or BALE_ERR_PE for the
 * calls made on that stream since its last check, and clear its record.
 * *bad (optional) receives the number of skipped indices.
 *
 * Streams: every stream used with this library gets its own error record and
 * grow-only workspace on its first call (which must not be made while the
 * stream is being captured into a graph).  Calls on different streams may
 * overlap; calls on one stream from several host threads must be serialised
 * by the caller, as the stream serialises their kernels. */
int bale_sync_errors(void *stream, int64_t *bad);
/* Wait for `stream` and free its error record and workspace (for callers that
 * create and destroy many streams; at most 256 are held per device). */
int bale_hip_release_stream(void *stream);
/* convey/convey_alc8r.h:53-60 -- the reference's allocator object
 * (convey_alc8r_t), here for all the device memory the library allocates for
 * itself from the call on: stream workspaces and error records,
 * bale_all_alloc, staging copies, batches, conveyors, exstacks.  grab must
 * return device memory of at least `size` bytes (256-byte aligned) or NULL;
 * `tag` and `value` are the library's source file and line.  Each pointer is
 * returned to the allocator that grabbed it, even after another is
 * installed.  NULL restores hipMalloc / hipFree.  A failed grab makes the
 * call that needed it return BALE_ERR_ALLOC (the drop-ins -1.0 or NULL).
 * The reference passes an allocator to each conveyor; here it is one per
 * process. */
typedef struct bale_alc8r {
  void *alc8r;
  void *(*grab)(void *alc8r, size_t size, const char *tag, uint64_t value);
  void (*free)(void *alc8r, void *ptr);
} bale_alc8r_t;
int bale_hip_set_alc8r(const bale_alc8r_t *a);
/* The strategy (BALE_STRAT_DIRECT, _LDS or _BUCKET) the last bale_histo,
 * bale_gather or drop-in call on this thread ran; BALE_STRAT_AUTO before any. */
int bale_hip_last_strategy(void);
/* The configuration BALE_STRAT_AUTO would run for this shape. */
int bale_histo_plan(int64_t n, int64_t nbins, int dtype, bale_config *out);
int bale_gather_plan(int64_t n, int64_t table_len, bale_config *out);

/* perm[0..n) (device memory) = the random permutation rand_permp_agi gives
 * at one PE for `seed` (seed 0: drand48's initial state, i.e. no srand48).
 * n < 2^30.  Waits for `stream` once, to check that the darts sufficed. */
int bale_randperm(int64_t *perm, int64_t n, int64_t seed, void *stream);
/* perm[0..n) (device memory) = the permutation rand_permp_agi_opt gives at
 * one PE for `seed` (seed 0: drand48's initial state).  n < 2^30.  Only
 * enqueues on `stream`. */
int bale_randperm_shuffle(int64_t *perm, int64_t n, int64_t seed, void *stream);
/* 1 if perm[0..n) (device memory) holds 0..n-1 exactly once, 0 if not,
 * negative BALE_ERR_* on error; waits for `stream`. */
int bale_is_perm(const int64_t *perm, int64_t n, void *stream);

/* Sparse matrices as device arrays (offset[numrows + 1], nonzero[nnz]).  Each
 * call finishes on `stream` before it returns (it reports range errors).
 * Transpose: t_offset[numcols + 1], t_nonzero[nnz]; numrows, numcols < 2^32,
 * nnz < 2^31. */
int bale_spmat_transpose(int64_t numrows, int64_t numcols, int64_t nnz, const int64_t *offset,
                         const int64_t *nonzero, int64_t *t_offset, int64_t *t_nonzero,
                         void *stream);
/* Permute: p_offset[numrows + 1], p_nonzero[nnz]; rperminv[numrows],
 * cperminv[numcols] permutations. */
int bale_spmat_permute(int64_t numrows, int64_t numcols, int64_t nnz, const int64_t *offset,
                       const int64_t *nonzero, const int64_t *rperminv, const int64_t *cperminv,
                       int64_t *p_offset, int64_t *p_nonzero, void *stream);
/* Erdos-Renyi rows (as gen_erdos_renyi_graph_dist): offset[n + 1] is always
 * written and *nnz set; nonzero (capacity nnz_cap) is filled when not NULL and
 * large enough (else BALE_ERR_ARG). */
int bale_spmat_gen_er(int64_t n, double p, int unit_diag, int mode, uint64_t seed, int64_t *offset,
                      int64_t *nonzero, int64_t nnz_cap, int64_t *nnz, void *stream);
/* Toposort of an n x n morally upper-triangular matrix (offset, nonzero) given
 * its transpose (t_offset, t_nonzero): rperm[n], cperm[n] as
 * toposort_matrix_agi; *levels = the levels of the queue (may be NULL).
 * BALE_ERR_ARG when the matrix is not morally triangular. */
int bale_toposort(int64_t n, const int64_t *offset, const int64_t *nonzero, const int64_t *t_offset,
                  const int64_t *t_nonzero, int64_t *rperm, int64_t *cperm, int64_t *levels,
                  void *stream);
/* *wrong = entries on the wrong side of the diagonal, *nopivot = rows without
 * their diagonal entry. */
int bale_spmat_triangular(int64_t numrows, const int64_t *offset, const int64_t *nonzero,
                          int upper, int64_t *wrong, int64_t *nopivot, void *stream);

/* ---- device memory for C callers ------------------------------------- */
/* libgetput/libgetput.h:159 (lgp_all_alloc) at one PE: zeroed device memory
 * for nblocks * blocksize bytes; NULL on failure.  Free with bale_all_free
 * (libgetput.h:160, lgp_all_free). */
void *bale_all_alloc(int64_t nblocks, int64_t blocksize);
void bale_all_free(void *p);
/* Host-to-device and device-to-host copies of `bytes` (synchronous). */
int bale_memput(void *dst_dev, const void *src_host, int64_t bytes);
int bale_memget(void *dst_host, const void *src_dev, int64_t bytes);

/* ---- input generation on the GPU (seeded, reproducible) ------------- */
/* table[j] = splitmix64(j) */
int bale_fill_splitmix(int64_t *table, int64_t n, void *stream);
/* table[j] = -(j+1) (the reference ig driver's fill, apps/ig_src/ig.upc:161) */
int bale_fill_neg(int64_t *table, int64_t n, void *stream);
/* Index streams: dist 0 uniform over [0,nbins); 1 Zipf(s) ranks over a keyed
 * permutation of the bins; run_mean > 1 makes runs of a repeated value with
 * geometric lengths of that mean.  Element i depends only on (seed, i). */
int bale_gen_index(int64_t *out, int64_t n, int64_t nbins, int dist, double s,
                   int64_t run_mean, uint64_t seed, void *stream);
/* pck[i] = index[i] << 16: the packed form at one PE (apps/histo_src/histo.upc:148-150). */
int bale_pack(int64_t *pck, const int64_t *index, int64_t n, void *stream);

/* ---- batches of uneven instances ------------------------------------- */
/* No counterpart in the reference (its drivers run one instance per call).
 * One histogram or gather instance; all pointers are device memory. */
typedef struct bale_instance {
  int32_t kind;          /* 0 histogram (counts += ...), 1 gather (tgt = ...) */
  int32_t dtype;         /* histogram counters: BALE_I64 or BALE_I32 */
  int64_t n;             /* updates or requests, < 2^32 */
  int64_t table_len;     /* bins or table entries */
  const int64_t *index;  /* plain global indices */
  void *table;           /* counts (histogram) or table (gather) */
  int64_t *tgt;          /* gather output */
  bale_config cfg;       /* strategy and task size; BALE_STRAT_AUTO = default rule */
} bale_instance;

typedef struct bale_batch bale_batch;
/* Plan `count` instances into per-phase task lists and capture the launch
 * sequence in one HIP graph.  Instances keep their pointers; relaunching
 * repeats the work (histograms add again). NULL on error.
 * Histogram instances may share counts (several index streams into one
 * table): instances whose counts overlap are flushed with atomics only.  A
 * gather's table or tgt must not overlap any histogram's counts in the same
 * batch (the phases do not order them).  Index errors of a launch are
 * reported by bale_sync_errors on the launching stream. */
bale_batch *bale_batch_create(const bale_instance *inst, int64_t count, void *stream);
/* One graph launch of the whole batch on `stream`. */
int bale_batch_launch(bale_batch *b, void *stream);
/* The same kernels enqueued directly (no graph). */
int bale_batch_launch_direct(bale_batch *b, void *stream);
/* out[0..5] = phase-A tasks, scatter tasks, apply slots, hot-set sketches,
 * bucketed instances, scratch bytes. */
int bale_batch_stats(const bale_batch *b, int64_t *out, int nout);
void bale_batch_destroy(bale_batch *b);
/* Baseline: one plain DIRECT kernel launch per instance, in order, on `stream`. */
int bale_batch_launch_plain(const bale_instance *inst, int64_t count, void *stream);
/* Baseline: each instance with its own cfg as one bale_histo / bale_gather call,
 * in order, on `stream` (what tuning buys without the batched launch). */
int bale_batch_launch_each(const bale_instance *inst, int64_t count, void *stream);

/* Exhaustive autotuner: time every candidate configuration of each instance
 * (as batches of that instance repeated R and 2R times, R x ops about
 * min_ops, results in scratch so caller data is untouched) and store the one
 * with the lowest marginal cost in inst[i].cfg: the time the R extra copies
 * add per operation, which is what an instance costs inside a large batch,
 * as the median of three interleaved R / 2R pairs.
 * costs (optional): count x max_cand seconds per operation, -1 where unused. */
int bale_tune(bale_instance *inst, int64_t count, int64_t min_ops, float *costs, int max_cand,
              void *stream);
/* The candidate list the tuner tries for an instance of this kind and size;
 * returns its length (writes at most max entries). */
int bale_tune_candidates(int kind, int64_t table_len, bale_config *out, int max);

/* ---- roofline probes ------------------------------------------------- */
/* dst[i] = src[i] over `bytes` (multiple of 16): streaming copy rate. */
int bale_probe_copy(void *dst, const void *src, int64_t bytes, void *stream);
/* nreads random 8-byte reads of table[0..n) at addresses hashed from the
 * read number (no index stream); a digest goes to *out (one int64). */
int bale_probe_random_read(const int64_t *table, int64_t n, int64_t nreads,
                           int64_t *out, void *stream);

/* ---- exstack and exstack2 ---------------------------------------------
 * The aggregation buffers of exstack/exstack.h:94-107 (bulk synchronous) and
 * :201-210 (asynchronous) as a bulk service on one GPU: items pushed with a
 * source and a destination virtual PE (1..255 PEs, device arrays), delivered
 * so that each destination receives its items grouped by source, each
 * source's in push order.  A round holds up to n_pe * buf_cnt items in all
 * (the reference bounds each destination's buffer by buf_cnt; a caller that
 * pushes until a push is short and then exchanges behaves the same).  Spans
 * handed out by the pops are device memory, valid until the next delivery.
 * Negative returns are BALE_ERR_* codes (bale_hip_last_error() says why). */
typedef struct bale_exstack bale_exstack;
typedef struct bale_exstack bale_exstack2;
/* exstack.h:94 exstack_init -- n_pe virtual PEs, buf_cnt items per PE per
 * round, packages of pkg_size bytes (4, 8, 12, 16, 24 or 32).  NULL on error. */
bale_exstack *bale_exstack_init(int64_t n_pe, int64_t buf_cnt, size_t pkg_size);
/* exstack.h:95 exstack_proceed -- 0 once done is set, nothing is pushed and
 * everything delivered was popped; else 1. */
int64_t bale_exstack_proceed(bale_exstack *ex, int done);
/* exstack.h:96 exstack_push, in bulk -- pushes the longest prefix of the n
 * items (device arrays: packages, source PEs or NULL for PE 0, destination
 * PEs) that fits this round, on `stream`; returns how many it took. */
int64_t bale_exstack_push(bale_exstack *ex, const void *items, const int64_t *from_pes,
                          const int64_t *to_pes, int64_t n, void *stream);
/* exstack.h:97 exstack_exchange -- delivers the round (waits for `stream`);
 * items naming a PE outside 0..n_pe-1 are dropped and reported as
 * BALE_ERR_RANGE after the rest is delivered. */
int bale_exstack_exchange(bale_exstack *ex, void *stream);
/* exstack.h:100 exstack_pop, in bulk -- the items PE `pe` received this round
 * (*items) and their source PEs (*from_pes, int64), grouped by source in
 * increasing order; returns the count (0 once popped). */
int64_t bale_exstack_pop(bale_exstack *ex, int64_t pe, const void **items, const int64_t **from_pes);
/* exstack.h:98 exstack_pop_thread, in bulk -- what `from_pe` sent `pe` this
 * round, in push order (the retrace of apps/ig_src/ig_exstack.upc:88-92);
 * returns the count (0 once popped).  A PE's round is popped either whole or
 * one source at a time. */
int64_t bale_exstack_pop_thread(bale_exstack *ex, int64_t pe, int64_t from_pe, const void **items);
/* exstack.h:106 exstack_reset: forget the round; :105 exstack_clear: free. */
int bale_exstack_reset(bale_exstack *ex);
void bale_exstack_clear(bale_exstack *ex);
/* exstack.h:201 exstack2_init -- as bale_exstack_init. */
bale_exstack2 *bale_exstack2_init(int64_t n_pe, int64_t buf_cnt, size_t pkg_size);
/* exstack.h:203 exstack2_push, in bulk -- as bale_exstack_push. */
int64_t bale_exstack2_push(bale_exstack2 *x, const void *items, const int64_t *from_pes,
                           const int64_t *to_pes, int64_t n, void *stream);
/* exstack.h:202 exstack2_proceed -- sends what was pushed once everything
 * delivered before was popped (the asynchronous send of a full buffer), and
 * returns 0 once done_pushing was given, nothing waits and everything
 * delivered was popped; else 1. */
int64_t bale_exstack2_proceed(bale_exstack2 *x, int done_pushing, void *stream);
/* exstack.h:204 exstack2_pop, in bulk -- as bale_exstack_pop. */
int64_t bale_exstack2_pop(bale_exstack2 *x, int64_t pe, const void **items, const int64_t **from_pes);
/* exstack.h:209 exstack2_reset, :210 exstack2_clear. */
int bale_exstack2_reset(bale_exstack2 *x);
void bale_exstack2_clear(bale_exstack2 *x);

/* ---- learned selector ----------------------------------------------------
 * 12 sketch features per instance and an fp32 MLP 12 -> 64 -> 64 -> 10
 * (3 strategy, 4 variant, 3 task-size logits); all arrays device memory. */
int bale_hip_device_info(int *cus, int *lds_per_block);
int bale_sketch_feature_count(void);
int bale_mlp_param_count(void);
/* features[count][12] of the instances' index distributions. */
int bale_sketch(const bale_instance *inst, int64_t count, float *features, void *stream);
/* logits[count][10] and per-head argmax classes[count][3]. */
int bale_mlp_infer(const float *params, const float *features, int64_t count, float *logits,
                   int32_t *classes, void *stream);
/* Minibatch Adam on per-head cross-entropy against labels[count][3]
 * (adam_state: 2 x param_count zeros; loss[steps] per-step mean loss). */
int bale_mlp_train(float *params, float *adam_state, const float *features, const int32_t *labels,
                   int64_t count, int steps, int batch, float lr, uint64_t seed, float *loss,
                   void *stream);
/* Minibatch Adam on a cost-weighted choice loss: cands[count][maxc][4] holds each
 * timed candidate's (strategy, variant, task) classes and its time over the
 * instance's best (strategy < 0 ends a list); the loss is the cross-entropy of the
 * softmax over the candidates' summed head log-probabilities against targets
 * proportional to (best / time)^8.  Features are standardised on the
 * device (scratch: 2 x 12 + count x 12 floats) and the standardisation is folded
 * into the first layer, so the trained params take raw features. */
int bale_mlp_train_cost(float *params, float *adam_state, const float *features,
                        const float *cands, int maxc, int64_t count, int steps, int batch,
                        float lr, uint64_t seed, float *loss, float *scratch, void *stream);

/* ---- device conveyor ---------------------------------------------------
 * The conveyor interface (convey/convey.h:184-338) as a bulk service on one
 * GPU: items go to up to 256 virtual PEs, and each PE receives its items in
 * push order (the one-source FIFO guarantee, convey/narrative.h:135-137).
 * Items and PE numbers are device memory; a delivery happens in
 * bale_convey_advance, after which each PE's items are one contiguous device
 * span.  Return codes follow convey.h:48-53: */
#define BALE_CONVEY_DONE 0 /* advance: done, everything pulled */
#define BALE_CONVEY_FAIL 0 /* push: full until the next delivery; pull: nothing */
#define BALE_CONVEY_OK 1
/* negative: a BALE_ERR_* code (bale_hip_last_error() says why). */
typedef struct bale_convey bale_convey;
/* convey.h:184 convey_new -- n_pe virtual PEs (1..256), room for `capacity`
 * pushed items between deliveries, items of at most max_item_bytes (4, 8,
 * 12, 16, 24 or 32).  NULL on error. */
bale_convey *bale_convey_new(int64_t n_pe, int64_t capacity, size_t max_item_bytes);
/* convey.h:225 convey_begin -- dormant -> working, with this item size. */
int bale_convey_begin(bale_convey *c, size_t item_bytes);
/* convey.h:237 convey_push, in bulk -- copies n items and their PE numbers
 * (device arrays) into the conveyor on `stream`.  BALE_CONVEY_FAIL if they do
 * not fit before the next delivery (nothing is copied then). */
int bale_convey_push(bale_convey *c, const void *items, const int64_t *pes, int64_t n,
                     void *stream);
/* convey.h:318 convey_advance -- delivers everything pushed once the previous
 * delivery has been pulled (waits for the GPU).  Returns BALE_CONVEY_DONE when
 * `done` is set and nothing is left to push or pull, else BALE_CONVEY_OK;
 * BALE_ERR_RANGE when items named a PE outside 0..n_pe-1 (they are dropped,
 * the rest are delivered). */
int bale_convey_advance(bale_convey *c, int done, void *stream);
/* convey.h:278 convey_apull, in bulk -- *items = device pointer to the items
 * delivered to `pe` and not pulled yet, *count of them, *from = 0 (the only
 * source).  BALE_CONVEY_OK if count > 0, BALE_CONVEY_FAIL if none. */
int bale_convey_pull(bale_convey *c, int64_t pe, const void **items, int64_t *count,
                     int64_t *from);
/* convey.h:301 convey_unpull -- gives back the last `count` items of the
 * latest pull from `pe`. */
int bale_convey_unpull(bale_convey *c, int64_t pe, int64_t count);
/* convey.h:328 convey_reset -- back to dormant; queued items are dropped. */
int bale_convey_reset(bale_convey *c);
/* convey.h:338 convey_free. */
int bale_convey_free(bale_convey *c);
/* Device-to-device copy of `bytes` on `stream` (to take items out of a span). */
int bale_memcpy_d2d(void *dst, const void *src, int64_t bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif
