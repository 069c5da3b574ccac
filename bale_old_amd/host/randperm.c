/*
 * randperm.c -- the reference's randperm driver (apps/randperm_src/randperm.upc:86-175)
 * on one MI355X through libbale_hip.so: the same options, one line per model
 * with its seconds, and is_perm on every result.  At one PE the four models
 * give the permutation rand_permp_agi gives, and the alternate model the one
 * rand_permp_agi_opt gives (a Fisher-Yates pass,
 * apps/randperm_src/alternates/randperm_agi_opt.upc).
 */
#include <getopt.h>
#include <stdint.h>

#include "driver.h"

static void usage(void) {
  fprintf(stderr,
          "randperm: a random permutation of 0..N-1 on one MI355X, by dart throwing\n"
          "(models 1-8) or by a Fisher-Yates pass (model 16).\n"
          "options (defaults in brackets):\n"
          "  -n N     entries per PE [1000000]\n"
          "  -M MASK  models to run, a sum of 1 AGI, 2 exstack, 4 exstack2, 8 conveyor, 16 alternate [15]\n"
          "  -s SEED  seed of the drand48 stream [101892]\n"
          "  -b N     buffer length of the buffered models (accepted; the GPU sizes its own)\n"
          "  -h       this text\n");
  exit(0);
}

int main(int argc, char *argv[]) {
  int64_t models_mask = 0xF, l_numrows = 1000000, buf_cnt = 1024, seed = 101892, cores = 1;
  int printhelp = 0, opt;
  while ((opt = getopt(argc, argv, "c:hn:M:s:b:")) != -1) {
    switch (opt) {
      case 'h': printhelp = 1; break;
      case 'b': sscanf(optarg, "%ld", &buf_cnt); break;
      case 'c': sscanf(optarg, "%ld", &cores); break;
      case 'n': sscanf(optarg, "%ld", &l_numrows); break;
      case 'M': sscanf(optarg, "%ld", &models_mask); break;
      case 's': sscanf(optarg, "%ld", &seed); break;
      default: break;
    }
  }
  if (printhelp) usage();
  driver_start(argc, argv);
  fprintf(stderr, "Running randperm on %d threads\n", 1);
  fprintf(stderr, "This is a demo program that runs various implementations of the randperm kernel.\n");
  fprintf(stderr, "Usage:\n");
  fprintf(stderr, "Permutation size per thread (-n) = %ld\n", l_numrows);
  fprintf(stderr, "models_mask (-M)                 = %ld or of 1,2,4,8 for atomics,classic,exstack2,conveyor\n", models_mask);
  fprintf(stderr, "seed (-s)                        = %ld\n", seed);
  const int64_t numrows = l_numrows;
  int64_t error = 0;
  for (int64_t use_model = 1L; use_model < 32; use_model *= 2) {
    double t1 = wall_seconds();
    int64_t *out = NULL;
    switch (use_model & models_mask) {
      case AGI_Model:
        out = rand_permp_agi(numrows, (int)seed);
        fprintf(stderr, "rand_permp_AGI:           ");
        break;
      case EXSTACK_Model:
        out = rand_permp_exstack(numrows, (int)seed, buf_cnt);
        fprintf(stderr, "rand_permp_EXSTACK:       ");
        break;
      case EXSTACK2_Model:
        out = rand_permp_exstack2(numrows, (int)seed, buf_cnt);
        fprintf(stderr, "rand_permp_EXSTACK2:      ");
        break;
      case CONVEYOR_Model:
        out = rand_permp_conveyor(numrows, (int)seed);
        fprintf(stderr, "rand_permp_CONVEYOR:      ");
        break;
      case ALTERNATE_Model:
        out = rand_permp_agi_opt(numrows, (int)seed);
        fprintf(stderr, "rand_permp_agi_opt :      ");
        break;
      default:
        continue;
    }
    t1 = wall_seconds() - t1;
    fprintf(stderr, "%8.3lf\n", t1);
    if (!out || !is_perm(out, numrows)) {
      error++;
      printf("\nERROR: rand_permp_%ld failed!%s%s\n\n", use_model & models_mask, out ? "" : " ",
             out ? "" : bale_hip_last_error());
    }
    bale_all_free(out);
  }
  if (error) fprintf(stderr, "BALE FAIL!!!!\n");
  return (int)error;
}
