/*
 * transpose_matrix.c -- the reference's transpose_matrix driver
 * (apps/transpose_matrix_src/transpose_matrix.upc:78-200) on one MI355X through
 * libbale_hip.so: an Erdos-Renyi matrix (both triangles, seed + 2), each
 * model's seconds, and the check that the transpose of the transpose equals
 * the input.  The generator is the library's (splitmix64 draws), so a seed
 * gives another matrix of the same distribution than the reference's.
 */
#include <getopt.h>
#include <stdint.h>

#include "driver.h"

static void usage(void) {
  fprintf(stderr,
          "transpose_matrix: the transpose of an Erdos-Renyi matrix on one MI355X.\n"
          "options (defaults in brackets):\n"
          "  -n N     rows per PE [10000]\n"
          "  -Z N     mean nonzeros per row [10] (sets the edge probability)\n"
          "  -e P     edge probability (used when -Z is not given)\n"
          "  -M MASK  models to run, a sum of 1 AGI, 2 exstack, 4 exstack2, 8 conveyor [15]\n"
          "  -s SEED  seed of the matrix [101892]\n"
          "  -b N     buffer length of the buffered models (accepted; the GPU sizes its own)\n"
          "  -c N     cores per node (accepted, unused)\n"
          "  -C       skip the check that transposing twice gives the input back\n"
          "  -h       this text\n");
  exit(0);
}

int main(int argc, char *argv[]) {
  int64_t check = 1, models_mask = 0xF, nz_per_row = -1, buf_cnt = 1024, l_numrows = 10000;
  int64_t seed = 101892, cores = 1;
  double erdos_renyi_prob = 0.0;
  int printhelp = 0, opt;
  while ((opt = getopt(argc, argv, "hb:c:Ce:n:M:s:Z:")) != -1) {
    switch (opt) {
      case 'h': printhelp = 1; break;
      case 'b': sscanf(optarg, "%ld", &buf_cnt); break;
      case 'c': sscanf(optarg, "%ld", &cores); break;
      case 'C': check = 0; break;
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
  /* erdos_renyi_prob and nz_per_row made consistent (transpose_matrix.upc:131-139) */
  if (nz_per_row == -1 && erdos_renyi_prob == 0) nz_per_row = 10;
  else if (nz_per_row == -1) nz_per_row = (int64_t)(erdos_renyi_prob * numrows);
  erdos_renyi_prob = (2.0 * (nz_per_row - 1)) / numrows;
  if (erdos_renyi_prob > 1.0) erdos_renyi_prob = 1.0;
  fprintf(stderr, "Running transpose_matrix on %d threads\n", 1);
  fprintf(stderr, "buf_cnt (stack size)        (-b) = %ld\n", buf_cnt);
  fprintf(stderr, "Erdos-Renyi edge probability(-e) = %lf\n", erdos_renyi_prob);
  fprintf(stderr, "rows per PE (-n)                 = %ld\n", l_numrows);
  fprintf(stderr, "models_mask (-M)                 = %ld or of 1,2,4,8,16 for gets,classic,exstack2,conveyor,alternate\n", models_mask);
  fprintf(stderr, "seed (-s)                        = %ld\n", seed);
  fprintf(stderr, "Avg # of nonzeros per row   (-Z) = %ld\n", nz_per_row);
  sparsemat_t *inmat = gen_erdos_renyi_graph_dist((int)numrows, erdos_renyi_prob, 0, 3, seed + 2);
  if (!inmat) {
    printf("ERROR: inmat is null! (%s)\n", bale_hip_last_error());
    return -1;
  }
  int64_t error = 0;
  for (int64_t use_model = 1L; use_model < 32; use_model *= 2) {
    double t1 = wall_seconds();
    sparsemat_t *outmat = NULL;
    switch (use_model & models_mask) {
      case AGI_Model:
        outmat = transpose_matrix_agi(inmat);
        fprintf(stderr, "AGI:     ");
        break;
      case EXSTACK_Model:
        outmat = transpose_matrix_exstack(inmat, buf_cnt);
        fprintf(stderr, "Exstack: ");
        break;
      case EXSTACK2_Model:
        outmat = transpose_matrix_exstack2(inmat, buf_cnt);
        fprintf(stderr, "Exstack2:");
        break;
      case CONVEYOR_Model:
        outmat = transpose_matrix_conveyor(inmat);
        fprintf(stderr, "Conveyor:");
        break;
      case ALTERNATE_Model:
        fprintf(stderr, "There is no alternate model here!\n");
        continue;
      default:
        continue;
    }
    t1 = wall_seconds() - t1;
    fprintf(stderr, "%8.3lf\n", t1);
    if (!outmat) {
      fprintf(stderr, "ERROR: %s\n", bale_hip_last_error());
      error++;
      continue;
    }
    if (check) { /* transpose_matrix.upc:180-187 */
      sparsemat_t *outmatT = transpose_matrix(outmat);
      if (!outmatT || compare_matrix(outmatT, inmat)) {
        fprintf(stderr, "ERROR: transpose of transpose does not match!\n");
        error++;
      }
      if (outmatT) {
        clear_matrix(outmatT);
        free(outmatT);
      }
    }
    clear_matrix(outmat);
    free(outmat);
  }
  clear_matrix(inmat);
  free(inmat);
  return (int)error;
}
