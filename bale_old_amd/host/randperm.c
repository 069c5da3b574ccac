/*
 * randperm.c -- the reference's randperm driver (apps/randperm_src/randperm.upc:86-175)
 * on one MI355X through libbale_hip.so: the same options, one line per model
 * with its seconds, and is_perm on every result.  At one PE the four models
 * give the permutation rand_permp_agi gives; the alternate model
 * (rand_permp_agi_opt) is not provided.
 */
#include <getopt.h>
#include <stdint.h>

#include "driver.h"

static void usage(void) {
  fprintf(stderr,
          "Usage:\n"
          "randperm [-h][-n num][-M mask][-s seed]\n"
          " -h Print usage banner only\n"
          " -b count is the number of packages in an exstack(2) buffer\n"
          " -n num Set the permutation entries per PE to n (default = 1000).\n"
          " -M mask Set the models mask (1,2,4,8,16,32 for AGI, exstack, exstack2,conveyor,alternate)\n"
          " -s seed Set a seed for the random number generation.\n\n");
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
        fprintf(stderr, "There is no alternate model here!\n");
        continue;
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
