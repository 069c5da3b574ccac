/*
 * histo.c -- the reference's histogram driver, run on one MI355X through
 * libbale_hip.so.
 *
 * Same command line, output and self-check as apps/histo_src/histo.upc:93-235
 * at one PE (THREADS == 1): counts[] is allocated and zeroed, index[] and the
 * packed pckindx[] are generated with srand(120348); rand() % num_counts
 * (histo.upc:143-151), each model selected by -M adds one count per update
 * into the same counts[], and the run fails if subtracting num_models per
 * index does not leave every counter at zero (histo.upc:208-227).
 *
 * Extensions: -d keeps counts, index and pckindx in device memory;
 * -S seed draws the indices from splitmix64 stream `seed` on the GPU instead
 * of rand().  -M 16 (ALTERNATE_Model) runs the four alternates.
 */
#include <getopt.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bale_hip.h"

#define AGI_Model 1L
#define EXSTACK_Model 2L
#define EXSTACK2_Model 4L
#define CONVEYOR_Model 8L
#define ALTERNATE_Model 16L
#define ALL_Models 15L

static void usage(void) {
  fprintf(stderr,
          "histo: counts[index[i]] += 1 on one MI355X, for each selected model.\n"
          "options (defaults in brackets):\n"
          "  -n N     updates per PE [1000000]\n"
          "  -T N     counters per PE [1000]\n"
          "  -M MASK  models to run, a sum of 1 AGI, 2 exstack, 4 exstack2, 8 conveyor, 16 alternates [15]\n"
          "  -b N     buffer length of the buffered models (accepted; the GPU sizes its own)\n"
          "  -c N     cores per node: only scales the reported GB/s\n"
          "  -d       give the library device arrays instead of host arrays\n"
          "  -S SEED  indices from splitmix64 stream SEED instead of the reference's rand() sequence\n"
          "  -h       this text\n");
  exit(0);
}

static void *xalloc(int dev, int64_t n) {
  void *p = dev ? bale_all_alloc(n, sizeof(int64_t)) : calloc((size_t)(n ? n : 1), sizeof(int64_t));
  if (!p) {
    fprintf(stderr, "allocation of %ld int64 failed: %s\n", (long)n, bale_hip_last_error());
    exit(1);
  }
  return p;
}

int main(int argc, char *argv[]) {
  int64_t buf_cnt = 1024, models_mask = ALL_Models, l_num_ups = 1000000, lnum_counts = 1000;
  int64_t cores_per_node = 0, seed = -1;
  int printhelp = 0, on_dev = 0, opt;
  while ((opt = getopt(argc, argv, "hb:M:n:c:T:dS:")) != -1) {
    switch (opt) {
      case 'h': printhelp = 1; break;
      case 'b': sscanf(optarg, "%ld", &buf_cnt); break;
      case 'M': sscanf(optarg, "%ld", &models_mask); break;
      case 'n': sscanf(optarg, "%ld", &l_num_ups); break;
      case 'T': sscanf(optarg, "%ld", &lnum_counts); break;
      case 'c': sscanf(optarg, "%ld", &cores_per_node); break;
      case 'd': on_dev = 1; break;
      case 'S': sscanf(optarg, "%ld", &seed); break;
      default: break;
    }
  }
  if (printhelp) usage();
  if (bale_hip_init(0) != BALE_OK) {
    fprintf(stderr, "no GPU: %s\n", bale_hip_last_error());
    return 1;
  }
  time_t now = time(NULL);
  struct tm *date = localtime(&now);
  fprintf(stderr, "Bale Version 2.1 (on one MI355X, libbale_hip): %04d-%02d-%02d.%02d:%02d\n",
          date->tm_year + 1900, date->tm_mon + 1, date->tm_mday, date->tm_hour, date->tm_min);
  fprintf(stderr, "Command:");
  for (int i = 0; i < argc; i++) fprintf(stderr, " %s", argv[i]);
  fprintf(stderr, "\n");
  fprintf(stderr, "Running histo on %d threads\n", 1);
  fprintf(stderr, "buf_cnt (number of buffer pkgs)      (-b)= %ld\n", buf_cnt);
  fprintf(stderr, "Number updates / thread              (-n)= %ld\n", l_num_ups);
  fprintf(stderr, "Table size / thread                  (-T)= %ld\n", lnum_counts);
  fprintf(stderr, "models_mask                          (-M)= %ld\n", models_mask);
  fflush(stderr);

  int64_t num_counts = lnum_counts;
  /* host copies of the inputs (for generation and the final check) */
  int64_t *h_index = calloc((size_t)(l_num_ups ? l_num_ups : 1), sizeof(int64_t));
  int64_t *h_pck = calloc((size_t)(l_num_ups ? l_num_ups : 1), sizeof(int64_t));
  if (!h_index || !h_pck) return 1;
  if (seed < 0) {
    srand(0 + 120348);
    for (int64_t i = 0; i < l_num_ups; i++) {
      int64_t indx = rand() % num_counts;
      h_index[i] = indx;
      h_pck[i] = (indx << 16) | 0;
    }
  } else {
    int64_t *d = bale_all_alloc(l_num_ups, 8);
    if (!d || bale_gen_index(d, l_num_ups, num_counts, 0, 1.0, 0, (uint64_t)seed, NULL) != BALE_OK ||
        bale_memget(h_index, d, l_num_ups * 8) != BALE_OK) {
      fprintf(stderr, "index generation failed: %s\n", bale_hip_last_error());
      return 1;
    }
    bale_all_free(d);
    for (int64_t i = 0; i < l_num_ups; i++) h_pck[i] = h_index[i] << 16;
  }
  int64_t *counts = xalloc(on_dev, num_counts);
  int64_t *index = h_index, *pckindx = h_pck;
  if (on_dev) {
    index = xalloc(1, l_num_ups);
    pckindx = xalloc(1, l_num_ups);
    bale_memput(index, h_index, l_num_ups * 8);
    bale_memput(pckindx, h_pck, l_num_ups * 8);
  }
  double volume_per_node = (8 * l_num_ups * cores_per_node) * (1.0E-9);
  double laptime = 0.0;
  int64_t num_models = 0;
  int failed_call = 0;
  for (int64_t use_model = 1L; use_model < 32; use_model *= 2) {
    switch (use_model & models_mask) {
      case AGI_Model:
        fprintf(stderr, "      AGI: ");
        laptime = histo_agi(index, l_num_ups, counts);
        num_models++;
        break;
      case EXSTACK_Model:
        fprintf(stderr, "  Exstack: ");
        laptime = histo_exstack(pckindx, l_num_ups, counts, buf_cnt);
        num_models++;
        break;
      case EXSTACK2_Model:
        fprintf(stderr, " Exstack2: ");
        laptime = histo_exstack2(pckindx, l_num_ups, counts, buf_cnt);
        num_models++;
        break;
      case CONVEYOR_Model:
        fprintf(stderr, "Conveyors: ");
        laptime = histo_conveyor(pckindx, l_num_ups, counts);
        num_models++;
        break;
      case ALTERNATE_Model: {
        double (*alt[4])(int64_t *, int64_t, int64_t *, int64_t) = {
            histo_exstack2_cyclic, histo_exstack2_goto, histo_exstack_function,
            histo_exstack2_function};
        const char *names[4] = {"exstack2_cyclic", "exstack2_goto", "exstack_function",
                                "exstack2_function"};
        for (int a = 0; a < 4; a++) {
          fprintf(stderr, "Alternate %s: ", names[a]);
          laptime = alt[a](pckindx, l_num_ups, counts, buf_cnt);
          num_models++;
          if (laptime < 0) {
            fprintf(stderr, "ERROR: %s\n", bale_hip_last_error());
            failed_call = 1;
          } else {
            fprintf(stderr, "  %8.3lf seconds  %8.3lf GB/s injection bandwidth\n", laptime,
                    volume_per_node / laptime);
          }
        }
        continue;
      }
      default:
        continue;
    }
    if (laptime < 0) {
      fprintf(stderr, "ERROR: %s\n", bale_hip_last_error());
      failed_call = 1;
      continue;
    }
    fprintf(stderr, "  %8.3lf seconds  %8.3lf GB/s injection bandwidth\n", laptime,
            volume_per_node / laptime);
  }

  /* the reference's check: subtract num_models per index, expect all zeros */
  int64_t *lcounts = counts;
  if (on_dev) {
    lcounts = calloc((size_t)num_counts, sizeof(int64_t));
    bale_memget(lcounts, counts, num_counts * 8);
  }
  for (int64_t i = 0; i < l_num_ups; i++) lcounts[h_index[i]] -= num_models;
  int64_t num_errors = 0;
  for (int64_t i = 0; i < num_counts; i++) {
    if (lcounts[i] != 0L) {
      num_errors++;
      if (num_errors < 5)
        fprintf(stderr, "ERROR: Thread %d error at %ld (= %ld)\n", 0, i, lcounts[i]);
    }
  }
  if (num_errors) fprintf(stderr, "FAILED!!!! total errors = %ld\n", num_errors);
  if (on_dev) {
    free(lcounts);
    bale_all_free(counts);
    bale_all_free(index);
    bale_all_free(pckindx);
  } else {
    free(counts);
  }
  free(h_index);
  free(h_pck);
  return (int)(num_errors + failed_call);
}
