/*
 * ig.c -- the reference's index-gather driver, run on one MI355X through
 * libbale_hip.so.
 *
 * Same command line, output and check as apps/ig_src/ig.upc:116-239 at one PE:
 * table[j] = -(j+1) (ig.upc:159-161), index[] from srand(5); rand() % tab_siz
 * (ig.upc:168-175), every model selected by -M fills tgt[], which must equal
 * -(index[i]+1) and is then zeroed (ig_check_and_zero, ig.upc:96-114).
 * Unlike the reference (which always returns 0, ig.upc:238), the exit status
 * is the number of errors.
 *
 * Extensions: -d keeps table, index, pckindx and tgt in device memory;
 * -S seed draws the indices from splitmix64 stream `seed` on the GPU.
 * -M 16 (ALTERNATE_Model) runs the four alternates.
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
          "ig: tgt[i] = table[index[i]] on one MI355X, for each selected model.\n"
          "options (defaults in brackets):\n"
          "  -n N     requests per PE [1000000]\n"
          "  -T N     table entries per PE [100000]\n"
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

/* ig.upc:96-114 on a host copy of tgt; zeroes tgt afterwards */
static int64_t check_and_zero(int64_t use_model, int64_t *tgt, int on_dev, int64_t *h_tgt,
                              const int64_t *index, int64_t n) {
  if (on_dev) bale_memget(h_tgt, tgt, n * 8);
  else h_tgt = tgt;
  int64_t errors = 0;
  for (int64_t i = 0; i < n; i++) {
    if (h_tgt[i] != (-1) * (index[i] + 1)) {
      errors++;
      if (errors < 5)
        fprintf(stderr, "ERROR: model %ld: Thread %d: tgt[%ld] = %ld != %ld)\n", use_model, 0, i,
                h_tgt[i], (-1) * (index[i] + 1));
    }
    h_tgt[i] = 0;
  }
  if (on_dev) bale_memput(tgt, h_tgt, n * 8);
  if (errors > 0) fprintf(stderr, "ERROR: %ld: %ld total errors on thread %d\n", use_model, errors, 0);
  return errors;
}

int main(int argc, char *argv[]) {
  int64_t buf_cnt = 1024, models_mask = ALL_Models, ltab_siz = 100000, l_num_req = 1000000;
  int64_t cores_per_node = 0, seed = -1;
  int printhelp = 0, on_dev = 0, opt;
  while ((opt = getopt(argc, argv, "hb:M:n:c:T:dS:")) != -1) {
    switch (opt) {
      case 'h': printhelp = 1; break;
      case 'b': sscanf(optarg, "%ld", &buf_cnt); break;
      case 'M': sscanf(optarg, "%ld", &models_mask); break;
      case 'n': sscanf(optarg, "%ld", &l_num_req); break;
      case 'T': sscanf(optarg, "%ld", &ltab_siz); break;
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
  fprintf(stderr, "Running ig on %d threads\n", 1);
  fprintf(stderr, "buf_cnt (number of buffer pkgs)      (-b)= %ld\n", buf_cnt);
  fprintf(stderr, "Number of Request / thread           (-n)= %ld\n", l_num_req);
  fprintf(stderr, "Table size / thread                  (-T)= %ld\n", ltab_siz);
  fprintf(stderr, "models_mask                          (-M)= %ld\n", models_mask);
  fprintf(stderr, "models_mask is or of 1,2,4,8,16 for agi,exstack,exstack2,conveyor,alternate)\n");

  int64_t tab_siz = ltab_siz;
  int64_t *h_table = calloc((size_t)(tab_siz ? tab_siz : 1), sizeof(int64_t));
  int64_t *h_index = calloc((size_t)(l_num_req ? l_num_req : 1), sizeof(int64_t));
  int64_t *h_pck = calloc((size_t)(l_num_req ? l_num_req : 1), sizeof(int64_t));
  int64_t *h_tgt = calloc((size_t)(l_num_req ? l_num_req : 1), sizeof(int64_t));
  if (!h_table || !h_index || !h_pck || !h_tgt) return 1;
  for (int64_t i = 0; i < tab_siz; i++) h_table[i] = (-1) * (i + 1);
  if (seed < 0) {
    srand(0 + 5);
    for (int64_t i = 0; i < l_num_req; i++) {
      int64_t indx = rand() % tab_siz;
      h_index[i] = indx;
      h_pck[i] = (indx << 16) | 0;
    }
  } else {
    int64_t *d = bale_all_alloc(l_num_req, 8);
    if (!d || bale_gen_index(d, l_num_req, tab_siz, 0, 1.0, 0, (uint64_t)seed, NULL) != BALE_OK ||
        bale_memget(h_index, d, l_num_req * 8) != BALE_OK) {
      fprintf(stderr, "index generation failed: %s\n", bale_hip_last_error());
      return 1;
    }
    bale_all_free(d);
    for (int64_t i = 0; i < l_num_req; i++) h_pck[i] = h_index[i] << 16;
  }
  int64_t *table = h_table, *index = h_index, *pckindx = h_pck, *tgt = h_tgt;
  if (on_dev) {
    table = xalloc(1, tab_siz);
    index = xalloc(1, l_num_req);
    pckindx = xalloc(1, l_num_req);
    tgt = xalloc(1, l_num_req);
    bale_memput(table, h_table, tab_siz * 8);
    bale_memput(index, h_index, l_num_req * 8);
    bale_memput(pckindx, h_pck, l_num_req * 8);
  }
  double volume_per_node = (2 * 8 * l_num_req * cores_per_node) * (1.0E-9);
  double laptime = 0.0;
  int64_t num_errors = 0;
  int64_t *scratch = on_dev ? calloc((size_t)(l_num_req ? l_num_req : 1), sizeof(int64_t)) : NULL;
  for (int64_t use_model = 1L; use_model < 32; use_model *= 2) {
    switch (use_model & models_mask) {
      case AGI_Model:
        fprintf(stderr, "      AGI: ");
        laptime = ig_agi(tgt, index, l_num_req, table);
        break;
      case EXSTACK_Model:
        fprintf(stderr, "  Exstack: ");
        laptime = ig_exstack(tgt, pckindx, l_num_req, table, buf_cnt);
        break;
      case EXSTACK2_Model:
        fprintf(stderr, " Exstack2: ");
        laptime = ig_exstack2(tgt, pckindx, l_num_req, table, buf_cnt);
        break;
      case CONVEYOR_Model:
        fprintf(stderr, "Conveyors: ");
        laptime = ig_conveyor(tgt, pckindx, l_num_req, table);
        break;
      case ALTERNATE_Model: {
        double (*alt[4])(int64_t *, int64_t *, int64_t, int64_t *, int64_t) = {
            ig_exstack2_cyclic, ig_exstack2_goto, ig_exstack_function, ig_exstack_pkg};
        const char *names[4] = {"exstack2_cyclic", "exstack2_goto", "exstack_function",
                                "exstack_pkg"};
        for (int a = 0; a < 4; a++) {
          fprintf(stderr, "Alternate %s: ", names[a]);
          laptime = alt[a](tgt, pckindx, l_num_req, table, buf_cnt);
          if (laptime < 0) {
            fprintf(stderr, "ERROR: %s\n", bale_hip_last_error());
            num_errors++;
          } else {
            fprintf(stderr, "  %8.3lf seconds  %8.3lf GB/s/Node\n", laptime,
                    volume_per_node / laptime);
          }
          num_errors += check_and_zero(use_model, tgt, on_dev, scratch, h_index, l_num_req);
        }
        continue;
      }
      default:
        continue;
    }
    if (laptime < 0) {
      fprintf(stderr, "ERROR: %s\n", bale_hip_last_error());
      num_errors++;
    } else {
      fprintf(stderr, "  %8.3lf seconds  %8.3lf GB/s/Node\n", laptime, volume_per_node / laptime);
    }
    num_errors += check_and_zero(use_model, tgt, on_dev, scratch, h_index, l_num_req);
  }
  if (num_errors) fprintf(stderr, "YOU FAILED!!!!\n");
  if (on_dev) {
    bale_all_free(table);
    bale_all_free(index);
    bale_all_free(pckindx);
    bale_all_free(tgt);
    free(scratch);
  }
  free(h_table);
  free(h_index);
  free(h_pck);
  free(h_tgt);
  return (int)(num_errors > 255 ? 255 : num_errors);
}
