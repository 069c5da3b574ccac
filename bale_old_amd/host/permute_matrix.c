/*
 * permute_matrix.c -- the reference's permute_matrix driver
 * (apps/permute_matrix_src/permute_matrix.upc:88-190) on one MI355X through
 * libbale_hip.so: row and column permutations from rand_permp (seeds seed and
 * seed + 1), an Erdos-Renyi matrix (seed + 2), each model's seconds.  Beyond
 * the reference, every result is checked against the AGI model's.
 */
#include <getopt.h>
#include <stdint.h>

#include "driver.h"

static void usage(void) {
  fprintf(stderr,
          "permute_matrix: rows and columns of an Erdos-Renyi matrix permuted on one MI355X.\n"
          "options (defaults in brackets):\n"
          "  -n N     rows per PE [10000]\n"
          "  -Z N     mean nonzeros per row [10] (sets the edge probability)\n"
          "  -e P     edge probability (used when -Z is not given)\n"
          "  -M MASK  models to run, a sum of 1 AGI, 2 exstack, 4 exstack2, 8 conveyor [15]\n"
          "  -s SEED  seed of the matrix and the permutations [101892]\n"
          "  -b N     buffer length of the buffered models (accepted; the GPU sizes its own)\n"
          "  -h       this text\n");
  exit(0);
}

int main(int argc, char *argv[]) {
  int64_t models_mask = 0xF, buf_cnt = 1024, nz_per_row = -1, l_numrows = 10000, seed = 101892;
  int64_t cores = 1;
  double erdos_renyi_prob = 0.0;
  int printhelp = 0, opt;
  while ((opt = getopt(argc, argv, "e:hc:n:M:s:Z:b:")) != -1) {
    switch (opt) {
      case 'h': printhelp = 1; break;
      case 'b': sscanf(optarg, "%ld", &buf_cnt); break;
      case 'c': sscanf(optarg, "%ld", &cores); break;
      case 'e': sscanf(optarg, "%lf", &erdos_renyi_prob); break;
      case 'n': sscanf(optarg, "%ld", &l_numrows); break;
      case 'M': sscanf(optarg, "%ld", &models_mask); break;
      case 's': sscanf(optarg, "%ld", &seed); break;
      case 'Z': sscanf(optarg, "%ld", &nz_per_row); break;
      default: break;
    }
  }
  if (printhelp) usage();
  driver_start(argc, argv);
  const int64_t numrows = l_numrows;
  /* permute_matrix.upc:121-129 */
  if (nz_per_row == -1 && erdos_renyi_prob == 0.0) nz_per_row = 10;
  else if (nz_per_row == -1) nz_per_row = (int64_t)(erdos_renyi_prob * numrows);
  erdos_renyi_prob = (2.0 * (nz_per_row - 1)) / numrows;
  if (erdos_renyi_prob > 1.0) erdos_renyi_prob = 1.0;
  fprintf(stderr, "Running permute_matrix on %d threads\n", 1);
  fprintf(stderr, "buf_cnt (stack size)         (-b)= %ld\n", buf_cnt);
  fprintf(stderr, "Erdos-Renyi edge probability (-e)= %lf\n", erdos_renyi_prob);
  fprintf(stderr, "rows per thread (-n)             = %ld\n", l_numrows);
  fprintf(stderr, "Avg # of nonzeros per row    (-Z)= %ld\n", nz_per_row);
  fprintf(stderr, "models_mask (-M)                 = %ld or of 1,2,4,8,16,32 for gets,classic,exstack2,conveyor,ex2cyclic,ex2goto\n", models_mask);
  fprintf(stderr, "seed (-s)                        = %ld\n", seed);
  int64_t *rp = must(rand_permp(numrows, (int)seed), "rand_permp");
  int64_t *cp = must(rand_permp(numrows, (int)seed + 1), "rand_permp");
  sparsemat_t *inmat = gen_erdos_renyi_graph_dist((int)numrows, erdos_renyi_prob, 0, 3, seed + 2);
  if (!inmat) {
    printf("ERROR: inmat is null! (%s)\n", bale_hip_last_error());
    return -1;
  }
  sparsemat_t *want = must(permute_matrix_agi(inmat, rp, cp), "permute_matrix_agi");
  int64_t error = 0;
  for (int64_t use_model = 1L; use_model < 32; use_model *= 2) {
    double t1 = wall_seconds();
    sparsemat_t *outmat = NULL;
    switch (use_model & models_mask) {
      case AGI_Model:
        outmat = permute_matrix_agi(inmat, rp, cp);
        fprintf(stderr, "permute_matrix_AGI:           ");
        break;
      case EXSTACK_Model:
        outmat = permute_matrix_exstack(inmat, rp, cp, buf_cnt);
        fprintf(stderr, "permute_matrix_EXSTACK:       ");
        break;
      case EXSTACK2_Model:
        outmat = permute_matrix_exstack2(inmat, rp, cp, buf_cnt);
        fprintf(stderr, "permute_matrix_EXSTACK2:      ");
        break;
      case CONVEYOR_Model:
        outmat = permute_matrix_conveyor(inmat, rp, cp);
        fprintf(stderr, "permute_matrix_CONVEYOR:      ");
        break;
      case ALTERNATE_Model:
        fprintf(stderr, "There is no alternate model here!\n");
        continue;
      default:
        continue;
    }
    t1 = wall_seconds() - t1;
    fprintf(stderr, "%8.3lf\n", t1);
    if (!outmat || compare_matrix(outmat, want)) {
      fprintf(stderr, "ERROR: permuted matrix differs from permute_matrix_agi's %s\n",
              outmat ? "" : bale_hip_last_error());
      error++;
    }
    if (outmat) {
      clear_matrix(outmat);
      free(outmat);
    }
  }
  clear_matrix(want);
  free(want);
  clear_matrix(inmat);
  free(inmat);
  bale_all_free(rp);
  bale_all_free(cp);
  return (int)error;
}
