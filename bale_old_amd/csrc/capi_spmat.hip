// capi_spmat.hip -- the C-ABI of the adjacent kernels in libbale_hip.so: the
// random permutation (rand_permp_*, bale_randperm), sparse matrices
// (sparsemat_t drop-ins, bale_spmat_*) and toposort (toposort_matrix_*,
// bale_toposort).
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include "bale_internal.h"

using namespace bale;

// ======================= random permutation ================================
extern "C" {

int bale_randperm(int64_t* perm, int64_t n, int64_t seed, void* stream) {
  if (n < 0) return bale_set_error(BALE_ERR_ARG, "negative permutation length %lld", (long long)n);
  if (n == 0) return BALE_OK;
  if (!perm) return bale_set_error(BALE_ERR_ARG, "null permutation");
  if (n >= ((int64_t)1 << 30))
    return bale_set_error(BALE_ERR_ARG, "permutation length must be below 2^30 (slots 2n below 2^31)");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  int rc = launch_randperm(perm, n, seed, d, (hipStream_t)stream, nullptr);
  if (rc == hipErrorOutOfMemory) return bale_set_error(BALE_ERR_ALLOC, "workspace allocation failed");
  return rc ? bale_fail_hip(rc, "randperm") : BALE_OK;
}

int bale_randperm_shuffle(int64_t* perm, int64_t n, int64_t seed, void* stream) {
  if (n < 0) return bale_set_error(BALE_ERR_ARG, "negative permutation length %lld", (long long)n);
  if (n == 0) return BALE_OK;
  if (!perm) return bale_set_error(BALE_ERR_ARG, "null permutation");
  if (n >= ((int64_t)1 << 30)) return bale_set_error(BALE_ERR_ARG, "permutation length must be below 2^30");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  int rc = launch_randperm_shuffle(perm, n, seed, d, (hipStream_t)stream);
  if (rc == hipErrorOutOfMemory) return bale_set_error(BALE_ERR_ALLOC, "workspace allocation failed");
  return rc ? bale_fail_hip(rc, "randperm_shuffle") : BALE_OK;
}

int bale_is_perm(const int64_t* perm, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && !perm)) return bale_set_error(BALE_ERR_ARG, "bad permutation array");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  int ok = 0;
  int rc = launch_check_perm(perm, n, d, (hipStream_t)stream, &ok);
  if (rc == hipErrorOutOfMemory) return bale_set_error(BALE_ERR_ALLOC, "workspace allocation failed");
  if (rc) return bale_fail_hip(rc, "is_perm");
  return ok;
}

// The reference's models of rand_permp (spmat/spmat.h:115-118) all return a
// permutation in the shared heap; here it is device memory from
// bale_all_alloc, filled by the dart-throwing AGI model (the result the
// reference's rand_permp_agi gives at one PE), and the GPU seconds go to
// bale_hip_last_timing.  NULL on error.
static int64_t* randperm_dropin(int64_t N, int seed, bool shuffle = false) {
  set_last_timing(0.0, 0.0, 0.0);
  if (N < 0 || N >= ((int64_t)1 << 30)) {
    bale_set_error(BALE_ERR_ARG, "permutation length %lld outside [0, 2^30)", (long long)N);
    return nullptr;
  }
  int64_t* perm = (int64_t*)bale_all_alloc(N ? N : 1, sizeof(int64_t));
  if (!perm) return nullptr;
  Device* d = device_state();
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, 0);
  int rc = !N ? 0 : shuffle ? launch_randperm_shuffle(perm, N, seed, d, 0) : launch_randperm(perm, N, seed, d, 0, nullptr);
  hipEventRecord(e1, 0);
  if (!rc) rc = (int)hipEventSynchronize(e1);
  if (!rc) set_last_timing(elapsed(e0, e1), 0.0, 0.0);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) {
    bale_all_free(perm);
    if (rc == hipErrorOutOfMemory) bale_set_error(BALE_ERR_ALLOC, "workspace allocation failed");
    else bale_fail_hip(rc, "randperm");
    return nullptr;
  }
  return perm;
}

int64_t* rand_permp_agi(int64_t N, int seed) { return randperm_dropin(N, seed); }
int64_t* rand_permp_exstack(int64_t N, int seed, int64_t) { return randperm_dropin(N, seed); }
int64_t* rand_permp_exstack2(int64_t N, int seed, int64_t) { return randperm_dropin(N, seed); }
int64_t* rand_permp_conveyor(int64_t N, int seed) { return randperm_dropin(N, seed); }

int64_t* rand_permp(int64_t N, int seed) { return randperm_dropin(N, seed); }

// The alternate model: the Fisher-Yates pass, the result the reference's
// rand_permp_agi_opt gives at one PE.
int64_t* rand_permp_agi_opt(int64_t N, int seed) { return randperm_dropin(N, seed, true); }

// ---- sparse matrices: native API ------------------------------------------------

static int spmat_rc(int rc, const char* what) {
  if (rc < 0) return rc;  // a BALE_ERR_* code, text already set
  if (rc == hipErrorOutOfMemory) return bale_set_error(BALE_ERR_ALLOC, "%s: workspace allocation failed", what);
  return bale_fail_hip(rc, what);
}

int bale_spmat_transpose(int64_t numrows, int64_t numcols, int64_t nnz, const int64_t* offset,
                         const int64_t* nonzero, int64_t* t_offset, int64_t* t_nonzero,
                         void* stream) {
  if (numrows < 0 || numcols < 0 || nnz < 0) return bale_set_error(BALE_ERR_ARG, "negative matrix size");
  if (numrows >= ((int64_t)1 << 32) || numcols >= ((int64_t)1 << 32) || nnz >= ((int64_t)1 << 31))
    return bale_set_error(BALE_ERR_ARG, "transpose needs numrows, numcols < 2^32 and nnz < 2^31");
  if (!offset || !t_offset || (nnz && (!nonzero || !t_nonzero)))
    return bale_set_error(BALE_ERR_ARG, "null matrix array");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  int rc = launch_spmat_transpose(numrows, numcols, nnz, offset, nonzero, t_offset, t_nonzero, d,
                                  (hipStream_t)stream);
  if (rc) return spmat_rc(rc, "transpose");
  return bale_sync_errors(stream, nullptr);
}

int bale_spmat_permute(int64_t numrows, int64_t numcols, int64_t nnz, const int64_t* offset,
                       const int64_t* nonzero, const int64_t* rperminv, const int64_t* cperminv,
                       int64_t* p_offset, int64_t* p_nonzero, void* stream) {
  if (numrows < 0 || numcols < 0 || nnz < 0) return bale_set_error(BALE_ERR_ARG, "negative matrix size");
  if (!offset || !p_offset || (numrows && !rperminv) || (nnz && (!nonzero || !p_nonzero || !cperminv)))
    return bale_set_error(BALE_ERR_ARG, "null matrix or permutation array");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  int rc = launch_spmat_permute(numrows, numcols, nnz, offset, nonzero, rperminv, cperminv, p_offset,
                                p_nonzero, d, (hipStream_t)stream);
  if (rc) return spmat_rc(rc, "permute");
  return bale_sync_errors(stream, nullptr);
}

int bale_spmat_gen_er(int64_t n, double p, int unit_diag, int mode, uint64_t seed, int64_t* offset,
                      int64_t* nonzero, int64_t nnz_cap, int64_t* nnz, void* stream) {
  if (n < 0 || n >= ((int64_t)1 << 32)) return bale_set_error(BALE_ERR_ARG, "row count outside [0, 2^32)");
  if (mode < 1 || mode > 3) return bale_set_error(BALE_ERR_ARG, "mode %d (1 lower, 2 upper, 3 both)", mode);
  if (!(p >= 0.0 && p <= 1.0)) return bale_set_error(BALE_ERR_ARG, "edge probability outside [0, 1]");
  if (!offset || !nnz) return bale_set_error(BALE_ERR_ARG, "null offset or nnz");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_spmat_gen_er(n, p, unit_diag, mode, seed, offset, nullptr, d, s);
  if (rc) return spmat_rc(rc, "gen_er");
  int64_t total = 0;
  rc = (int)hipMemcpyAsync(&total, offset + n, 8, hipMemcpyDeviceToHost, s);
  if (!rc) rc = (int)hipStreamSynchronize(s);
  if (rc) return bale_fail_hip(rc, "gen_er: reading nnz");
  *nnz = total;
  if (!nonzero) return BALE_OK;
  if (total > nnz_cap)
    return bale_set_error(BALE_ERR_ARG, "gen_er: %lld nonzeros, capacity %lld", (long long)total, (long long)nnz_cap);
  rc = launch_spmat_gen_er(n, p, unit_diag, mode, seed, offset, nonzero, d, s);
  if (!rc) rc = (int)hipStreamSynchronize(s);
  return rc ? spmat_rc(rc, "gen_er") : BALE_OK;
}

int bale_toposort(int64_t n, const int64_t* offset, const int64_t* nonzero, const int64_t* t_offset,
                  const int64_t* t_nonzero, int64_t* rperm, int64_t* cperm, int64_t* levels,
                  void* stream) {
  if (n < 0 || n >= ((int64_t)1 << 31)) return bale_set_error(BALE_ERR_ARG, "toposort: n outside [0, 2^31)");
  if (n && (!offset || !nonzero || !t_offset || !t_nonzero || !rperm || !cperm))
    return bale_set_error(BALE_ERR_ARG, "toposort: null array");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  int64_t piv = 0, lev = 0, bad_col = 0, twice = 0;
  int rc = launch_toposort(n, offset, nonzero, t_offset, t_nonzero, rperm, cperm, d,
                           (hipStream_t)stream, &piv, &lev, &bad_col, &twice);
  if (rc) return spmat_rc(rc, "toposort");
  if (levels) *levels = lev;
  if (bad_col) return bale_set_error(BALE_ERR_RANGE, "toposort: %lld column indices outside [0, %lld)",
                           (long long)bad_col, (long long)n);
  if (twice || piv != n)
    return bale_set_error(BALE_ERR_ARG, "toposort: %lld pivots for %lld rows, %lld columns claimed twice: "
                "the matrix is not morally upper triangular", (long long)piv, (long long)n,
                (long long)twice);
  return BALE_OK;
}

int bale_spmat_triangular(int64_t numrows, const int64_t* offset, const int64_t* nonzero, int upper,
                          int64_t* wrong, int64_t* nopivot, void* stream) {
  if (numrows < 0 || !offset || !wrong || !nopivot) return bale_set_error(BALE_ERR_ARG, "bad arguments");
  Device* d = device_state();
  if (!d) return bale_set_error(BALE_ERR_NODEV, "no usable HIP device");
  int rc = launch_spmat_triangular(offset, nonzero, numrows, upper, d, (hipStream_t)stream, wrong,
                                   nopivot);
  return rc ? spmat_rc(rc, "triangular") : BALE_OK;
}

}  // extern "C"

// ---- sparse matrices: the reference's spmat interface ----------------------------
namespace bale {

// A device copy of n int64 values when p is host memory (freed with the
// object), else p itself.
struct StagedI64 {
  const int64_t* p = nullptr;
  int64_t* owned = nullptr;
  ~StagedI64() {
    dev_free(owned);
  }
  int stage(const int64_t* src, int64_t n) {
    if (!src || n <= 0 || is_device_ptr(src)) {
      p = src;
      return 0;
    }
    owned = (int64_t*)BALE_DEV_ALLOC((size_t)n * 8);
    if (!owned) return bale_set_error(BALE_ERR_ALLOC, "staging %lld values", (long long)n);
    int rc = (int)hipMemcpy(owned, src, (size_t)n * 8, hipMemcpyHostToDevice);
    if (rc) return bale_fail_hip(rc, "staging to the device");
    p = owned;
    return 0;
  }
};

// Device memory for an output of n int64 values: `dst` itself when it is
// device memory, else a device buffer copied back to `dst` by back().
struct StagedOut {
  int64_t* p = nullptr;
  int64_t* owned = nullptr;
  ~StagedOut() {
    dev_free(owned);
  }
  int stage(int64_t* dst, int64_t n) {
    if (n <= 0 || is_device_ptr(dst)) {
      p = dst;
      return 0;
    }
    owned = (int64_t*)BALE_DEV_ALLOC((size_t)n * 8);
    if (!owned) return bale_set_error(BALE_ERR_ALLOC, "staging %lld values", (long long)n);
    p = owned;
    return 0;
  }
  int back(int64_t* dst, int64_t n) {
    if (!owned) return 0;
    int rc = (int)hipMemcpy(dst, owned, (size_t)n * 8, hipMemcpyDeviceToHost);
    return rc ? bale_fail_hip(rc, "copying back to the host") : 0;
  }
};

static sparsemat_t* new_matrix(int64_t numrows, int64_t numcols, int64_t nnz) {
  sparsemat_t* m = (sparsemat_t*)calloc(1, sizeof(sparsemat_t));
  if (!m) {
    bale_set_error(BALE_ERR_ALLOC, "matrix struct");
    return nullptr;
  }
  m->numrows = m->lnumrows = numrows;
  m->numcols = numcols;
  m->nnz = m->lnnz = nnz;
  m->offset = m->loffset = (int64_t*)bale_all_alloc(numrows + 1, sizeof(int64_t));
  m->nonzero = m->lnonzero = (int64_t*)bale_all_alloc(nnz ? nnz : 1, sizeof(int64_t));
  if (!m->offset || !m->nonzero) {
    bale_all_free(m->offset);
    bale_all_free(m->nonzero);
    free(m);
    return nullptr;
  }
  return m;
}

static void drop_matrix(sparsemat_t* m) {
  if (!m) return;
  bale_all_free(m->offset);
  bale_all_free(m->nonzero);
  free(m);
}

static bool matrix_ok(const sparsemat_t* A) {
  if (!A || A->numrows < 0 || A->numcols < 0 || A->nnz < 0 || !A->offset ||
      (A->nnz && !A->nonzero)) {
    bale_set_error(BALE_ERR_ARG, "bad sparse matrix");
    return false;
  }
  return true;
}

// Times the GPU part of a spmat drop-in on the null stream.
struct DropinClock {
  hipEvent_t e0{}, e1{};
  DropinClock() {
    set_last_timing(0.0, 0.0, 0.0);
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, 0);
  }
  void stop() {
    hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) == hipSuccess) set_last_timing(elapsed(e0, e1), 0.0, 0.0);
  }
  ~DropinClock() {
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
};

static sparsemat_t* transpose_dropin(sparsemat_t* A) {
  if (!matrix_ok(A)) return nullptr;
  StagedI64 off, nz;
  if (off.stage(A->offset, A->numrows + 1) || nz.stage(A->nonzero, A->nnz)) return nullptr;
  sparsemat_t* At = new_matrix(A->numcols, A->numrows, A->nnz);
  if (!At) return nullptr;
  DropinClock clk;
  int rc = bale_spmat_transpose(A->numrows, A->numcols, A->nnz, off.p, nz.p, At->offset, At->nonzero,
                                nullptr);
  clk.stop();
  if (rc) {
    drop_matrix(At);
    return nullptr;
  }
  return At;
}

static sparsemat_t* permute_dropin(sparsemat_t* A, const int64_t* rperminv, const int64_t* cperminv) {
  if (!matrix_ok(A)) return nullptr;
  StagedI64 off, nz, rp, cp;
  if (off.stage(A->offset, A->numrows + 1) || nz.stage(A->nonzero, A->nnz) ||
      rp.stage(rperminv, A->numrows) || cp.stage(cperminv, A->numcols))
    return nullptr;
  sparsemat_t* Ap = new_matrix(A->numrows, A->numcols, A->nnz);
  if (!Ap) return nullptr;
  DropinClock clk;
  int rc = bale_spmat_permute(A->numrows, A->numcols, A->nnz, off.p, nz.p, rp.p, cp.p, Ap->offset,
                              Ap->nonzero, nullptr);
  clk.stop();
  if (rc) {
    drop_matrix(Ap);
    return nullptr;
  }
  return Ap;
}

static double toposort_dropin(int64_t* rperm, int64_t* cperm, sparsemat_t* mat, sparsemat_t* tmat) {
  if (!matrix_ok(mat) || !matrix_ok(tmat)) return -1.0;
  if (mat->numrows != mat->numcols || tmat->numrows != mat->numcols || tmat->nnz != mat->nnz) {
    bale_set_error(BALE_ERR_ARG, "toposort: mat must be square and tmat its transpose");
    return -1.0;
  }
  StagedI64 off, nz, toff, tnz;
  if (off.stage(mat->offset, mat->numrows + 1) || nz.stage(mat->nonzero, mat->nnz) ||
      toff.stage(tmat->offset, tmat->numrows + 1) || tnz.stage(tmat->nonzero, tmat->nnz))
    return -1.0;
  const int64_t n = mat->numrows;
  if (!rperm || !cperm) {
    bale_set_error(BALE_ERR_ARG, "toposort: null permutation");
    return -1.0;
  }
  // host arrays for the permutations get device copies, written back after
  StagedOut rp, cp;
  if (rp.stage(rperm, n) || cp.stage(cperm, n)) return -1.0;
  DropinClock clk;
  int rc = bale_toposort(n, off.p, nz.p, toff.p, tnz.p, rp.p, cp.p, nullptr, nullptr);
  clk.stop();
  if (!rc) rc = rp.back(rperm, n);
  if (!rc) rc = cp.back(cperm, n);
  return rc ? -1.0 : last_gpu_seconds();
}

static int is_perm_dropin(const int64_t* perm, int64_t N) {
  StagedI64 p;
  if (N < 0 || p.stage(perm, N)) return 0;
  return bale_is_perm(p.p, N, nullptr) > 0 ? 1 : 0;
}

static int triangular_dropin(sparsemat_t* A, int64_t unit_diagonal, int upper) {
  if (!matrix_ok(A)) return 0;
  StagedI64 off, nz;
  if (off.stage(A->offset, A->numrows + 1) || nz.stage(A->nonzero, A->nnz)) return 0;
  int64_t wrong = 0, nopivot = 0;
  if (bale_spmat_triangular(A->numrows, off.p, nz.p, upper, &wrong, &nopivot, nullptr)) return 0;
  return !(wrong || (unit_diagonal && nopivot));
}

}  // namespace bale

extern "C" {

sparsemat_t* transpose_matrix_agi(sparsemat_t* A) { return transpose_dropin(A); }
sparsemat_t* transpose_matrix_exstack(sparsemat_t* A, int64_t) { return transpose_dropin(A); }
sparsemat_t* transpose_matrix_exstack2(sparsemat_t* A, int64_t) { return transpose_dropin(A); }
sparsemat_t* transpose_matrix_conveyor(sparsemat_t* A) { return transpose_dropin(A); }
sparsemat_t* transpose_matrix(sparsemat_t* A) { return transpose_dropin(A); }

sparsemat_t* permute_matrix_agi(sparsemat_t* A, int64_t* rperminv, int64_t* cperminv) {
  return permute_dropin(A, rperminv, cperminv);
}
sparsemat_t* permute_matrix_exstack(sparsemat_t* A, int64_t* rperminv, int64_t* cperminv, int64_t) {
  return permute_dropin(A, rperminv, cperminv);
}
sparsemat_t* permute_matrix_exstack2(sparsemat_t* A, int64_t* rperminv, int64_t* cperminv, int64_t) {
  return permute_dropin(A, rperminv, cperminv);
}
sparsemat_t* permute_matrix_conveyor(sparsemat_t* A, int64_t* rperminv, int64_t* cperminv) {
  return permute_dropin(A, rperminv, cperminv);
}
sparsemat_t* permute_matrix(sparsemat_t* A, int64_t* rperminv, int64_t* cperminv) {
  return permute_dropin(A, rperminv, cperminv);
}

sparsemat_t* gen_erdos_renyi_graph_dist(int n, double p, int64_t unit_diag, int64_t mode, int64_t seed) {
  set_last_timing(0.0, 0.0, 0.0);
  if (n < 0 || mode < 1 || mode > 3) {
    bale_set_error(BALE_ERR_ARG, "gen_erdos_renyi_graph_dist: n %d, mode %lld (1, 2 or 3)", n, (long long)mode);
    return nullptr;
  }
  int64_t* off = (int64_t*)bale_all_alloc((int64_t)n + 1, sizeof(int64_t));
  if (!off) return nullptr;
  int64_t nnz = 0;
  if (bale_spmat_gen_er(n, p, (int)(unit_diag != 0), (int)mode, (uint64_t)seed, off, nullptr, 0, &nnz,
                        nullptr)) {
    bale_all_free(off);
    return nullptr;
  }
  sparsemat_t* m = (sparsemat_t*)calloc(1, sizeof(sparsemat_t));
  int64_t* nz = (int64_t*)bale_all_alloc(nnz ? nnz : 1, sizeof(int64_t));
  if (!m || !nz ||
      bale_spmat_gen_er(n, p, (int)(unit_diag != 0), (int)mode, (uint64_t)seed, off, nz, nnz, &nnz,
                        nullptr)) {
    free(m);
    bale_all_free(nz);
    bale_all_free(off);
    if (!m) bale_set_error(BALE_ERR_ALLOC, "matrix struct");
    return nullptr;
  }
  m->numrows = m->lnumrows = n;
  m->numcols = n;
  m->nnz = m->lnnz = nnz;
  m->offset = m->loffset = off;
  m->nonzero = m->lnonzero = nz;
  return m;
}

sparsemat_t* init_matrix(int64_t numrows, int64_t numcols, int64_t nnz) {
  if (numrows < 0 || numcols < 0 || nnz < 0) {
    bale_set_error(BALE_ERR_ARG, "init_matrix: negative size");
    return nullptr;
  }
  return new_matrix(numrows, numcols, nnz);
}

void clear_matrix(sparsemat_t* mat) {
  if (!mat) return;
  bale_all_free(mat->offset);
  bale_all_free(mat->nonzero);
  mat->offset = mat->loffset = mat->nonzero = mat->lnonzero = nullptr;
}

int compare_matrix(sparsemat_t* lmat, sparsemat_t* rmat) {
  if (!matrix_ok(lmat) || !matrix_ok(rmat)) return 1;
  if (lmat->numrows != rmat->numrows || lmat->numcols != rmat->numcols || lmat->nnz != rmat->nnz)
    return 1;
  Device* d = device_state();
  if (!d) return 1;
  StagedI64 lo, ro, ln, rn;
  if (lo.stage(lmat->offset, lmat->numrows + 1) || ro.stage(rmat->offset, rmat->numrows + 1) ||
      ln.stage(lmat->nonzero, lmat->nnz) || rn.stage(rmat->nonzero, rmat->nnz))
    return 1;
  int64_t d1 = 0, d2 = 0;
  if (launch_count_diff(lo.p, ro.p, lmat->numrows + 1, d, 0, &d1)) return 1;
  if (d1) return 1;
  if (launch_count_diff(ln.p, rn.p, lmat->nnz, d, 0, &d2)) return 1;
  return d2 ? 1 : 0;
}

double toposort_matrix_agi(int64_t* rperm, int64_t* cperm, sparsemat_t* mat, sparsemat_t* tmat) {
  return toposort_dropin(rperm, cperm, mat, tmat);
}
double toposort_matrix_exstack(int64_t* rperm, int64_t* cperm, sparsemat_t* mat, sparsemat_t* tmat) {
  return toposort_dropin(rperm, cperm, mat, tmat);
}
double toposort_matrix_exstack2(int64_t* rperm, int64_t* cperm, sparsemat_t* mat, sparsemat_t* tmat) {
  return toposort_dropin(rperm, cperm, mat, tmat);
}
double toposort_matrix_convey(int64_t* rperm, int64_t* cperm, sparsemat_t* mat, sparsemat_t* tmat) {
  return toposort_dropin(rperm, cperm, mat, tmat);
}
double toposort_matrix_oo(int64_t* rperm, int64_t* cperm, sparsemat_t* mat, sparsemat_t* tmat) {
  return toposort_dropin(rperm, cperm, mat, tmat);
}
double toposort_matrix_cooler(int64_t* rperm, int64_t* cperm, sparsemat_t* mat, sparsemat_t* tmat) {
  return toposort_dropin(rperm, cperm, mat, tmat);
}
double toposort_matrix_exstack_orig(int64_t* rperm, int64_t* cperm, sparsemat_t* mat,
                                    sparsemat_t* tmat) {
  return toposort_dropin(rperm, cperm, mat, tmat);
}

// spmat/spmat.h:152 -- 1 if perm holds 0..N-1 once each (checked on the GPU;
// host arrays are staged).
int is_perm(int64_t* perm, int64_t N) { return is_perm_dropin(perm, N); }

int is_upper_triangular(sparsemat_t* A, int64_t unit_diagonal) {
  return triangular_dropin(A, unit_diagonal, 1);
}
int is_lower_triangular(sparsemat_t* A, int64_t unit_diagonal) {
  return triangular_dropin(A, unit_diagonal, 0);
}

}  // extern "C"
