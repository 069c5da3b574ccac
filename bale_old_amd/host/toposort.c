/*
 * toposort.c -- the reference's toposort driver (apps/topo_src/toposort.upc:215-340)
 * on one MI355X through libbale_hip.so: the input from generate_toposort_input
 * (an upper-triangular Erdos-Renyi matrix with a full diagonal, rows and
 * columns permuted with rand_permp seeds 1230 and 45), its transpose, each
 * model's seconds, and check_is_triangle on every result.  With -f the
 * input is a Matrix Market file instead, read as the reference's
 * read_matrix_mm_to_dist does (spmat/spmat_utils.upc:217-365).
 */
#include <getopt.h>
#include <stdint.h>
#include <string.h>

#include "driver.h"
#include "mm_read.h"

static void usage(void) {
  fprintf(stderr,
          "topo: the row and column permutations that make a permuted upper-triangular matrix\n"
          "      upper triangular again, found on one MI355X.\n"
          "options (defaults in brackets):\n"
          "  -n N     rows per PE of the generated matrix [1000]\n"
          "  -Z N     mean nonzeros per row of the generated matrix [35] (sets the edge probability)\n"
          "  -e P     edge probability of the generated matrix (used when -Z is not given)\n"
          "  -f FILE  read the matrix from FILE (Matrix Market coordinate format) instead\n"
          "  -M MASK  models to run, a sum of 1 AGI, 2 exstack, 4 exstack2, 8 conveyor, 16 alternates [15]\n"
          "  -b N     buffer length of the buffered models (accepted; the GPU sizes its own)\n"
          "  -D       accepted, unused (as in the reference)\n"
          "  -h       this text\n");
  exit(0);
}

/* A Matrix Market file (mm_read.h) as a matrix in device memory; NULL (with
 * a message) on a bad file. */
static sparsemat_t *read_matrix_mm(const char *name) {
  mm_csr m;
  if (mm_read(name, &m)) return NULL;
  sparsemat_t *A = init_matrix(m.nr, m.nc, m.nnz);
  if (A && (bale_memput(A->offset, m.offset, (m.nr + 1) * 8) != BALE_OK ||
            (m.nnz && bale_memput(A->nonzero, m.nonzero, m.nnz * 8) != BALE_OK))) {
    clear_matrix(A);
    free(A);
    A = NULL;
  }
  if (!A) fprintf(stderr, "ERROR: read_matrix_mm: %s\n", bale_hip_last_error());
  mm_free(&m);
  return A;
}

/* toposort.upc:147-167 */
static int check_is_triangle(sparsemat_t *mat, int64_t *rperminv, int64_t *cperminv) {
  int rf = is_perm(rperminv, mat->numrows), cf = is_perm(cperminv, mat->numcols);
  if (!rf || !cf) {
    fprintf(stderr, "ERROR: check_is_triangle is_perm(rperminv2) = %d is_perm(cperminv2) = %d\n", rf, cf);
    return 1;
  }
  sparsemat_t *mat2 = permute_matrix(mat, rperminv, cperminv);
  int ret = 0;
  if (!mat2 || !is_upper_triangular(mat2, 1)) {
    fprintf(stderr, "ERROR: check_is_triangle fails\n");
    ret = 1;
  }
  if (mat2) {
    clear_matrix(mat2);
    free(mat2);
  }
  return ret;
}

/* toposort.upc:176-213 */
static sparsemat_t *generate_toposort_input(int64_t numrows, double prob, int64_t rand_seed) {
  fprintf(stderr, "Creating input matrix for toposort\n");
  double t = wall_seconds();
  sparsemat_t *omat = gen_erdos_renyi_graph_dist((int)numrows, prob, 1, 2, rand_seed);
  printf("generate ER graph time %lf\n", wall_seconds() - t);
  if (!omat || !is_upper_triangular(omat, 1)) exit(1);
  t = wall_seconds();
  int64_t *rperminv = rand_permp(numrows, 1230);
  int64_t *cperminv = rand_permp(numrows, 45);
  printf("generate perms time %lf\n", wall_seconds() - t);
  if (!rperminv || !cperminv) {
    printf("ERROR: topo_rand_permp returns NULL!\n");
    return NULL;
  }
  t = wall_seconds();
  sparsemat_t *mat = permute_matrix(omat, rperminv, cperminv);
  printf("permute matrix time %lf\n", wall_seconds() - t);
  if (!mat) {
    printf("ERROR: permute_matrix returned NULL");
    return NULL;
  }
  clear_matrix(omat);
  free(omat);
  bale_all_free(rperminv);
  bale_all_free(cperminv);
  return mat;
}

int main(int argc, char *argv[]) {
  int64_t l_numrows = 100000, buf_cnt = 1024, rand_seed = 5, models_mask = ALL_Models, cores = 1;
  double nz_per_row = 35, erdos_renyi_prob = 0.0;
  int printhelp = 0, read_graph = 0, opt;
  char filename[4096] = "";
  while ((opt = getopt(argc, argv, "hb:c:M:n:f:Z:e:D")) != -1) {
    switch (opt) {
      case 'h': printhelp = 1; break;
      case 'b': sscanf(optarg, "%ld", &buf_cnt); break;
      case 'c': sscanf(optarg, "%ld", &cores); break;
      case 'M': sscanf(optarg, "%ld", &models_mask); break;
      case 'n': sscanf(optarg, "%ld", &l_numrows); break;
      case 'f': read_graph = 1; snprintf(filename, sizeof filename, "%s", optarg); break;
      case 'Z': sscanf(optarg, "%lf", &nz_per_row); break;
      case 'e': sscanf(optarg, "%lf", &erdos_renyi_prob); break;
      case 'D': break;
      default: break;
    }
  }
  if (printhelp) usage();
  driver_start(argc, argv);
  sparsemat_t *mat = NULL;
  if (read_graph) {
    mat = read_matrix_mm(filename);
    if (!mat) exit(1);
    l_numrows = mat->numrows;
    if (mat->numcols != mat->numrows) {
      fprintf(stderr, "ERROR: %s is %ld x %ld, toposort needs a square matrix\n", filename,
              mat->numrows, mat->numcols);
      exit(1);
    }
  }
  const int64_t numrows = l_numrows, numcols = numrows;
  if (erdos_renyi_prob == 0.0) { /* toposort.upc:257-263 */
    erdos_renyi_prob = (2.0 * nz_per_row) / (numrows - 1);
    if (erdos_renyi_prob > 1.0) erdos_renyi_prob = 1.0;
  } else {
    nz_per_row = erdos_renyi_prob * numrows;
  }
  fprintf(stderr, "Running toposort on %d threads\n", 1);
  fprintf(stderr, "buf_cnt (stack size)           (-b)   %ld\n", buf_cnt);
  fprintf(stderr, "Number of rows per thread      (-n)   %ld\n", l_numrows);
  fprintf(stderr, "Avg # of nonzeros per row      (-Z)   %2.2lf\n", nz_per_row);
  fprintf(stderr, "Erdos-Renyi edge probability   (-e)   %lf\n", erdos_renyi_prob);
  fprintf(stderr, "task mask (M) = %ld (should be 1,2,4,8,16 for agi, exstack, exstack2, conveyors, alternates\n", models_mask);
  if (!mat) mat = generate_toposort_input(numrows, erdos_renyi_prob, rand_seed);
  if (!mat) {
    printf("ERROR: mat is NULL!\n");
    exit(1);
  }
  printf("Input matrix has %ld rows and %ld nonzeros\n", mat->numrows, mat->nnz);
  sparsemat_t *tmat = transpose_matrix(mat);
  if (!tmat) {
    printf("ERROR: tmat is NULL!\n");
    exit(1);
  }
  fprintf(stderr, "Run toposort on mat (and tmat) ...\n");
  int64_t *rperminv2 = must(bale_all_alloc(numrows, sizeof(int64_t)), "alloc");
  int64_t *cperminv2 = must(bale_all_alloc(numcols, sizeof(int64_t)), "alloc");
  int64_t error = 0;
  for (int64_t use_model = 1L; use_model < 32; use_model *= 2) {
    double laptime = 0.0;
    switch (use_model & models_mask) {
      case AGI_Model:
        fprintf(stderr, "      AGI: ");
        laptime = toposort_matrix_agi(rperminv2, cperminv2, mat, tmat);
        break;
      case EXSTACK_Model:
        fprintf(stderr, "  Exstack: ");
        laptime = toposort_matrix_exstack(rperminv2, cperminv2, mat, tmat);
        break;
      case EXSTACK2_Model:
        fprintf(stderr, " Exstack2: ");
        laptime = toposort_matrix_exstack2(rperminv2, cperminv2, mat, tmat);
        break;
      case CONVEYOR_Model:
        fprintf(stderr, " Conveyor: ");
        laptime = toposort_matrix_convey(rperminv2, cperminv2, mat, tmat);
        break;
      case ALTERNATE_Model:
        fprintf(stderr, " ALTERNATE: ");
        laptime = toposort_matrix_cooler(rperminv2, cperminv2, mat, tmat);
        break;
      default:
        continue;
    }
    if (laptime < 0) {
      fprintf(stderr, "ERROR: %s\n", bale_hip_last_error());
      error++;
      continue;
    }
    fprintf(stderr, "  %8.3lf seconds\n", laptime);
    if (check_is_triangle(mat, rperminv2, cperminv2)) {
      printf("\nERROR: After toposort_matrix_upc: mat2 is not upper-triangular!\n");
      error++;
    }
  }
  bale_all_free(rperminv2);
  bale_all_free(cperminv2);
  clear_matrix(tmat);
  free(tmat);
  clear_matrix(mat);
  free(mat);
  return (int)error;
}
