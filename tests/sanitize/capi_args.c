/*
 * capi_args.c -- every entry point of libbale_hip.so called with bad
 * arguments (each must return its error, never crash) and with good ones
 * (which must stop at BALE_ERR_NODEV or succeed without a device), built
 * with AddressSanitizer and UndefinedBehaviorSanitizer over a host-sanitized
 * build of the library, on a machine without a GPU.  The process exits 0
 * only if every expectation held and no sanitizer reported.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bale_hip.h"

static int failures = 0, checks = 0;

#define EXPECT(cond)                                                                      \
  do {                                                                                    \
    checks++;                                                                             \
    if (!(cond)) {                                                                        \
      fprintf(stderr, "FAIL %s:%d: %s (last error: %s)\n", __FILE__, __LINE__, #cond,     \
              bale_hip_last_error());                                                     \
      failures++;                                                                         \
    }                                                                                     \
  } while (0)
#define ERR(rc) ((rc) < 0)
#define ARG(rc) ((rc) == BALE_ERR_ARG)
#define NODEV(rc) ((rc) == BALE_ERR_NODEV)
#define ARG_OR_NODEV(rc) (ARG(rc) || NODEV(rc))

static bale_config cfg_of(int strategy) {
  bale_config c;
  memset(&c, 0, sizeof c);
  c.strategy = strategy;
  c.merge = -1;
  return c;
}

static void drop_ins(void) {
  int64_t idx[8] = {0, 1, 2, 3, 3, 2, 1, 0}, counts[4] = {0}, tgt[8], table[4] = {5, 6, 7, 8};
  double (*h[])(int64_t *, int64_t, int64_t *) = {histo_agi, histo_conveyor};
  double (*hb[])(int64_t *, int64_t, int64_t *, int64_t) = {
      histo_exstack, histo_exstack2, histo_exstack2_goto, histo_exstack2_cyclic,
      histo_exstack_function, histo_exstack2_function};
  for (int k = 0; k < 2; k++) {
    EXPECT(h[k](NULL, 8, counts) == -1.0);
    EXPECT(h[k](idx, -1, counts) == -1.0);
    EXPECT(h[k](idx, 8, counts) == -1.0); /* no device */
    EXPECT(h[k](idx, 0, counts) == -1.0 || h[k](idx, 0, counts) == 0.0);
  }
  for (int k = 0; k < 6; k++) {
    EXPECT(hb[k](idx, 8, NULL, 64) == -1.0);
    EXPECT(hb[k](idx, -3, counts, 64) == -1.0);
    EXPECT(hb[k](idx, 8, counts, 64) == -1.0);
  }
  EXPECT(counts[0] == 0 && counts[3] == 0); /* nothing counted without a device */
  double (*g[])(int64_t *, int64_t *, int64_t, int64_t *) = {ig_agi, ig_conveyor};
  double (*gb[])(int64_t *, int64_t *, int64_t, int64_t *, int64_t) = {
      ig_exstack, ig_exstack2, ig_exstack2_cyclic, ig_exstack2_goto, ig_exstack_function,
      ig_exstack_pkg};
  for (int k = 0; k < 2; k++) {
    EXPECT(g[k](NULL, idx, 8, table) == -1.0);
    EXPECT(g[k](tgt, idx, -1, table) == -1.0);
    EXPECT(g[k](tgt, idx, 8, NULL) == -1.0);
    EXPECT(g[k](tgt, idx, 8, table) == -1.0);
  }
  for (int k = 0; k < 6; k++) {
    EXPECT(gb[k](tgt, NULL, 8, table, 64) == -1.0);
    EXPECT(gb[k](tgt, idx, 8, table, 64) == -1.0);
  }
}

static void native_histo_gather(void) {
  int64_t idx[8] = {0, 1, 2, 3, 3, 2, 1, 0}, counts[4] = {0}, tgt[8], table[4] = {1, 2, 3, 4};
  bale_config c = cfg_of(BALE_STRAT_AUTO);
  EXPECT(ARG(bale_histo(idx, -1, counts, 4, BALE_I64, BALE_INDEX_PLAIN, NULL, NULL)));
  EXPECT(bale_histo(idx, 0, counts, 4, BALE_I64, BALE_INDEX_PLAIN, NULL, NULL) == BALE_OK);
  EXPECT(ARG(bale_histo(NULL, 8, counts, 4, BALE_I64, BALE_INDEX_PLAIN, NULL, NULL)));
  EXPECT(ARG(bale_histo(idx, 8, NULL, 4, BALE_I64, BALE_INDEX_PLAIN, NULL, NULL)));
  EXPECT(ARG(bale_histo(idx, 8, counts, 0, BALE_I64, BALE_INDEX_PLAIN, NULL, NULL)));
  EXPECT(ARG(bale_histo(idx, 8, counts, 4, 7, BALE_INDEX_PLAIN, NULL, NULL)));
  EXPECT(ARG(bale_histo(idx, 8, counts, 4, BALE_I64, 9, NULL, NULL)));
  EXPECT(NODEV(bale_histo(idx, 8, counts, 4, BALE_I64, BALE_INDEX_PLAIN, &c, NULL)));
  /* every field of a configuration out of its range */
  bale_config bad[9];
  for (int k = 0; k < 9; k++) bad[k] = cfg_of(BALE_STRAT_BUCKET);
  bad[0].strategy = 4;
  bad[1].strategy = -1;
  bad[2].block = 100;
  bad[3].unroll = 3;
  bad[4].tile_log2 = 31;
  bad[5].merge = 5;
  bad[6].groups_per_cu = 17;
  bad[7].task_log2 = 5;
  bad[8].strategy = BALE_STRAT_DIRECT;
  bad[8].hot = 2;
  for (int k = 0; k < 9; k++) {
    EXPECT(ARG(bale_histo(idx, 8, counts, 4, BALE_I64, BALE_INDEX_PLAIN, &bad[k], NULL)));
    EXPECT(ARG(bale_histo_noalloc(idx, 8, counts, 4, BALE_I32, BALE_INDEX_PLAIN, &bad[k], NULL, NULL)));
  }
  EXPECT(NODEV(bale_histo_noalloc(idx, 8, counts, 4, BALE_I64, BALE_INDEX_PACKED, NULL, NULL, NULL)));
  EXPECT(ARG(bale_gather(NULL, idx, 8, table, 4, BALE_INDEX_PLAIN, NULL, NULL)));
  EXPECT(ARG(bale_gather(tgt, idx, -2, table, 4, BALE_INDEX_PLAIN, NULL, NULL)));
  EXPECT(bale_gather(tgt, idx, 0, table, 4, BALE_INDEX_PLAIN, NULL, NULL) == BALE_OK);
  EXPECT(ARG(bale_gather(tgt, idx, 8, table, 0, BALE_INDEX_PLAIN, NULL, NULL)));
  EXPECT(ARG(bale_gather(tgt, idx, 8, table, 4, 3, NULL, NULL)));
  bale_config lds = cfg_of(BALE_STRAT_LDS);
  EXPECT(ARG(bale_gather(tgt, idx, 8, table, 4, BALE_INDEX_PLAIN, &lds, NULL)));
  for (int k = 0; k < 8; k++) EXPECT(ARG(bale_gather(tgt, idx, 8, table, 4, BALE_INDEX_PLAIN, &bad[k], NULL)));
  EXPECT(NODEV(bale_gather(tgt, idx, 8, table, 4, BALE_INDEX_PLAIN, NULL, NULL)));
  int64_t bad_count = 0;
  EXPECT(ERR(bale_sync_errors(NULL, &bad_count)));
  EXPECT(ERR(bale_hip_release_stream(NULL)));
  bale_config out;
  EXPECT(ERR(bale_histo_plan(-1, 4, BALE_I64, &out)));
  EXPECT(ERR(bale_histo_plan(8, 4, BALE_I64, NULL)));
  EXPECT(ERR(bale_gather_plan(8, 0, &out)));
  EXPECT(ERR(bale_gather_plan(8, 4, NULL)));
  (void)bale_hip_last_strategy();
}

static void perms_and_matrices(void) {
  int64_t perm[4] = {0, 1, 2, 3};
  EXPECT(ARG(bale_randperm(perm, -1, 1, NULL)));
  EXPECT(ARG(bale_randperm(NULL, 4, 1, NULL)));
  EXPECT(ARG(bale_randperm(perm, (int64_t)1 << 30, 1, NULL)));
  EXPECT(bale_randperm(perm, 0, 1, NULL) == BALE_OK);
  EXPECT(NODEV(bale_randperm(perm, 4, 1, NULL)));
  EXPECT(ARG(bale_randperm_shuffle(perm, -1, 1, NULL)));
  EXPECT(ARG(bale_randperm_shuffle(NULL, 4, 1, NULL)));
  EXPECT(ARG(bale_randperm_shuffle(perm, (int64_t)1 << 30, 1, NULL)));
  EXPECT(NODEV(bale_randperm_shuffle(perm, 4, 1, NULL)));
  EXPECT(ARG(bale_is_perm(NULL, 4, NULL)));
  EXPECT(ARG(bale_is_perm(perm, -4, NULL)));
  EXPECT(rand_permp_agi(-1, 1) == NULL);
  EXPECT(rand_permp_agi(10, 1) == NULL);
  EXPECT(rand_permp_exstack(10, 1, 64) == NULL);
  EXPECT(rand_permp_exstack2(10, 1, 64) == NULL);
  EXPECT(rand_permp_conveyor(10, 1) == NULL);
  EXPECT(rand_permp_agi_opt((int64_t)1 << 30, 1) == NULL);
  EXPECT(rand_permp(10, 1) == NULL);
  EXPECT(is_perm(NULL, 4) <= 0);
  EXPECT(init_matrix(-1, 4, 4) == NULL);
  EXPECT(init_matrix(4, 4, -1) == NULL);
  EXPECT(init_matrix(4, 4, 4) == NULL);
  clear_matrix(NULL);
  EXPECT(transpose_matrix(NULL) == NULL);
  EXPECT(transpose_matrix_agi(NULL) == NULL);
  EXPECT(transpose_matrix_exstack(NULL, 64) == NULL);
  EXPECT(transpose_matrix_exstack2(NULL, 64) == NULL);
  EXPECT(transpose_matrix_conveyor(NULL) == NULL);
  EXPECT(permute_matrix(NULL, perm, perm) == NULL);
  EXPECT(permute_matrix_agi(NULL, perm, perm) == NULL);
  EXPECT(permute_matrix_exstack(NULL, perm, perm, 64) == NULL);
  EXPECT(permute_matrix_exstack2(NULL, perm, perm, 64) == NULL);
  EXPECT(permute_matrix_conveyor(NULL, perm, perm) == NULL);
  EXPECT(gen_erdos_renyi_graph_dist(-5, 0.1, 1, 3, 1) == NULL);
  EXPECT(gen_erdos_renyi_graph_dist(10, 1.5, 1, 3, 1) == NULL);
  EXPECT(gen_erdos_renyi_graph_dist(10, 0.1, 1, 3, 1) == NULL);
  EXPECT(compare_matrix(NULL, NULL) != 0);
  /* 1 means triangular (spmat/spmat_utils.upc:446): a missing matrix is not */
  EXPECT(is_upper_triangular(NULL, 1) == 0);
  EXPECT(is_lower_triangular(NULL, 1) == 0);
  /* a 4x4 unit upper-triangular matrix in host memory */
  int64_t off[5] = {0, 2, 3, 5, 6}, nz[6] = {0, 3, 1, 2, 3, 3};
  sparsemat_t m = {0, 4, 4, 4, 6, 6, off, off, nz, nz};
  EXPECT(transpose_matrix(&m) == NULL);
  EXPECT(permute_matrix(&m, perm, perm) == NULL);
  EXPECT(permute_matrix(&m, NULL, perm) == NULL);
  EXPECT(toposort_matrix_agi(NULL, NULL, NULL, NULL) == -1.0);
  EXPECT(toposort_matrix_agi(perm, perm, &m, NULL) == -1.0);
  EXPECT(toposort_matrix_agi(perm, perm, &m, &m) == -1.0);
  EXPECT(toposort_matrix_exstack(perm, perm, &m, &m) == -1.0);
  EXPECT(toposort_matrix_exstack2(perm, perm, &m, &m) == -1.0);
  EXPECT(toposort_matrix_convey(perm, perm, &m, &m) == -1.0);
  EXPECT(toposort_matrix_oo(perm, perm, &m, &m) == -1.0);
  EXPECT(toposort_matrix_cooler(perm, perm, &m, &m) == -1.0);
  EXPECT(toposort_matrix_exstack_orig(perm, perm, &m, &m) == -1.0);
  int64_t t_off[5], t_nz[6], p_off[5], p_nz[6], nnz = 0, levels = 0, wrong = 0, nopiv = 0;
  EXPECT(ARG(bale_spmat_transpose(-1, 4, 6, off, nz, t_off, t_nz, NULL)));
  EXPECT(ARG(bale_spmat_transpose(4, 4, 6, NULL, nz, t_off, t_nz, NULL)));
  EXPECT(ARG_OR_NODEV(bale_spmat_transpose(4, 4, 6, off, nz, t_off, t_nz, NULL)));
  EXPECT(ARG(bale_spmat_permute(4, 4, 6, off, nz, NULL, perm, p_off, p_nz, NULL)));
  EXPECT(ARG(bale_spmat_permute(4, -4, 6, off, nz, perm, perm, p_off, p_nz, NULL)));
  EXPECT(ARG(bale_spmat_gen_er(-1, 0.1, 1, 3, 1, off, nz, 6, &nnz, NULL)));
  EXPECT(ARG(bale_spmat_gen_er(4, 0.1, 1, 3, 1, off, nz, 6, NULL, NULL)));
  EXPECT(ARG(bale_spmat_gen_er(4, -0.5, 1, 3, 1, off, nz, 6, &nnz, NULL)));
  EXPECT(ARG(bale_spmat_gen_er(4, 0.1, 1, 9, 1, off, nz, 6, &nnz, NULL)));
  EXPECT(ARG(bale_toposort(-1, off, nz, t_off, t_nz, p_off, p_nz, &levels, NULL)));
  EXPECT(ARG(bale_toposort(4, NULL, nz, t_off, t_nz, p_off, p_nz, &levels, NULL)));
  EXPECT(ARG(bale_spmat_triangular(-1, off, nz, 1, &wrong, &nopiv, NULL)));
  EXPECT(ARG(bale_spmat_triangular(4, NULL, nz, 1, &wrong, &nopiv, NULL)));
}

static void memory_and_generators(void) {
  int64_t buf[8] = {0};
  EXPECT(bale_all_alloc(-1, 8) == NULL);
  EXPECT(bale_all_alloc(4, 0) == NULL);
  EXPECT(bale_all_alloc(4, 8) == NULL);
  bale_all_free(NULL);
  EXPECT(ARG(bale_memput(NULL, buf, 8)));
  EXPECT(ARG(bale_memput(buf, buf, -8)));
  EXPECT(ARG(bale_memget(buf, NULL, 8)));
  EXPECT(bale_memput(buf, buf, 0) == BALE_OK || ERR(bale_memput(buf, buf, 0)));
  EXPECT(ARG(bale_memcpy_d2d(NULL, buf, 8, NULL)));
  EXPECT(ARG(bale_fill_splitmix(NULL, 8, NULL)));
  EXPECT(ARG(bale_fill_splitmix(buf, -1, NULL)));
  EXPECT(ARG(bale_fill_neg(NULL, 8, NULL)));
  EXPECT(ARG(bale_gen_index(NULL, 8, 4, 0, 1.0, 0, 1, NULL)));
  EXPECT(ARG(bale_gen_index(buf, 8, 0, 0, 1.0, 0, 1, NULL)));
  EXPECT(ARG(bale_gen_index(buf, 8, 4, 7, 1.0, 0, 1, NULL)));
  EXPECT(ARG(bale_gen_index(buf, -8, 4, 0, 1.0, 0, 1, NULL)));
  EXPECT(ARG(bale_pack(NULL, buf, 8, NULL)));
  EXPECT(ARG(bale_probe_copy(NULL, NULL, 16, NULL)));
  EXPECT(ARG(bale_probe_copy(buf, buf, 15, NULL)));
  EXPECT(ARG(bale_probe_random_read(NULL, 8, 8, buf, NULL)));
  EXPECT(ARG(bale_probe_random_read(buf, 0, 8, buf, NULL)));
  int cus = -1, lds = -1;
  EXPECT(ERR(bale_hip_device_info(&cus, &lds)));
}

static void batches_and_selector(void) {
  int64_t idx[8] = {0, 1, 2, 3, 3, 2, 1, 0}, counts[4] = {0};
  bale_instance inst[2];
  memset(inst, 0, sizeof inst);
  inst[0].kind = 0;
  inst[0].dtype = BALE_I64;
  inst[0].n = 8;
  inst[0].table_len = 4;
  inst[0].index = idx;
  inst[0].table = counts;
  inst[1] = inst[0];
  inst[1].kind = 5; /* bad kind */
  EXPECT(bale_batch_create(NULL, 2, NULL) == NULL);
  EXPECT(bale_batch_create(inst, -1, NULL) == NULL);
  EXPECT(bale_batch_create(inst, 2, NULL) == NULL);
  EXPECT(bale_batch_create(inst, 1, NULL) == NULL); /* no device */
  EXPECT(ARG(bale_batch_launch(NULL, NULL)));
  EXPECT(ARG(bale_batch_launch_direct(NULL, NULL)));
  int64_t st[6];
  EXPECT(ARG(bale_batch_stats(NULL, st, 6)));
  bale_batch_destroy(NULL);
  EXPECT(ARG(bale_batch_launch_plain(NULL, 2, NULL)));
  EXPECT(ERR(bale_batch_launch_plain(inst, 2, NULL)));
  EXPECT(ARG(bale_batch_launch_each(NULL, 2, NULL)));
  float costs[64];
  EXPECT(ARG(bale_tune(NULL, 2, 1 << 20, costs, 32, NULL)));
  EXPECT(ERR(bale_tune(inst, 2, 1 << 20, costs, 32, NULL)));
  /* the candidate lists are host arithmetic: check them whole */
  bale_config cand[64];
  for (int kind = 0; kind < 2; kind++)
    for (int64_t tl = 1; tl <= ((int64_t)1 << 30); tl *= 3) {
      int n = bale_tune_candidates(kind, tl, cand, 64);
      EXPECT(n >= 1 && n <= 64);
      for (int k = 0; k < n && k < 64; k++) {
        EXPECT(cand[k].strategy >= BALE_STRAT_DIRECT && cand[k].strategy <= BALE_STRAT_BUCKET);
        EXPECT(cand[k].task_log2 == 0 || (cand[k].task_log2 >= 8 && cand[k].task_log2 <= 30));
      }
      EXPECT(bale_tune_candidates(kind, tl, cand, 1) <= 64);
    }
  EXPECT(ARG(bale_tune_candidates(7, 100, cand, 64)));
  EXPECT(ARG(bale_tune_candidates(0, 0, cand, 64)));
  EXPECT(bale_sketch_feature_count() > 0);
  EXPECT(bale_mlp_param_count() > 0);
  float feats[64], logits[64], params[16], adam[32], loss[4];
  int32_t classes[8], labels[8];
  EXPECT(ARG(bale_sketch(NULL, 2, feats, NULL)));
  EXPECT(ARG(bale_sketch(inst, 1, NULL, NULL)));
  EXPECT(ARG(bale_mlp_infer(NULL, feats, 2, logits, classes, NULL)));
  EXPECT(ARG(bale_mlp_infer(params, feats, -2, logits, classes, NULL)));
  EXPECT(ARG(bale_mlp_train(NULL, adam, feats, labels, 2, 10, 2, 0.01f, 1, loss, NULL)));
  EXPECT(ARG(bale_mlp_train(params, adam, feats, labels, 2, -1, 2, 0.01f, 1, loss, NULL)));
  EXPECT(ARG(bale_mlp_train_cost(NULL, adam, feats, feats, 2, 2, 10, 2, 0.01f, 1, loss, feats, NULL)));
}

static void services(void) {
  EXPECT(bale_exstack_init(0, 10, 8) == NULL);
  EXPECT(bale_exstack_init(256, 10, 8) == NULL);
  EXPECT(bale_exstack_init(4, 0, 8) == NULL);
  EXPECT(bale_exstack_init(4, 10, 10) == NULL);
  EXPECT(bale_exstack_init(4, 10, 8) == NULL); /* no device */
  EXPECT(bale_exstack2_init(4, 10, 7) == NULL);
  EXPECT(bale_exstack2_init(4, 10, 8) == NULL);
  const void *items = NULL;
  const int64_t *from = NULL;
  int64_t pes[2] = {0, 1};
  EXPECT(ERR(bale_exstack_proceed(NULL, 0)));
  EXPECT(ERR(bale_exstack_push(NULL, pes, NULL, pes, 2, NULL)));
  EXPECT(ERR(bale_exstack_exchange(NULL, NULL)));
  EXPECT(ERR(bale_exstack_pop(NULL, 0, &items, &from)));
  EXPECT(ERR(bale_exstack_pop_thread(NULL, 0, 0, &items)));
  EXPECT(ERR(bale_exstack_reset(NULL)));
  bale_exstack_clear(NULL);
  EXPECT(ERR(bale_exstack2_push(NULL, pes, NULL, pes, 2, NULL)));
  EXPECT(ERR(bale_exstack2_proceed(NULL, 1, NULL)));
  EXPECT(ERR(bale_exstack2_pop(NULL, 0, &items, &from)));
  EXPECT(ERR(bale_exstack2_reset(NULL)));
  bale_exstack2_clear(NULL);
  EXPECT(bale_convey_new(0, 10, 8) == NULL);
  EXPECT(bale_convey_new(4, 0, 8) == NULL);
  EXPECT(bale_convey_new(4, 10, 0) == NULL);
  EXPECT(bale_convey_new(4, 10, 8) == NULL); /* no device */
  int64_t count = 0, fromp = 0;
  EXPECT(bale_convey_begin(NULL, 8) != BALE_CONVEY_OK);
  EXPECT(bale_convey_push(NULL, pes, pes, 2, NULL) != BALE_CONVEY_OK);
  EXPECT(bale_convey_advance(NULL, 1, NULL) < 0);
  EXPECT(bale_convey_pull(NULL, 0, &items, &count, &fromp) != BALE_CONVEY_OK);
  EXPECT(bale_convey_unpull(NULL, 0, 1) != BALE_CONVEY_OK);
  EXPECT(bale_convey_reset(NULL) != BALE_CONVEY_OK);
  EXPECT(bale_convey_free(NULL) == BALE_CONVEY_OK);
  bale_alc8r_t half = {NULL, NULL, NULL};
  EXPECT(ARG(bale_hip_set_alc8r(&half)));
  EXPECT(bale_hip_set_alc8r(NULL) == BALE_OK);
}

int main(void) {
  EXPECT(ERR(bale_hip_init(0)));
  EXPECT(bale_hip_last_error() != NULL);
  double g = -1, h = -1, d = -1;
  bale_hip_last_timing(&g, &h, &d);
  drop_ins();
  native_histo_gather();
  perms_and_matrices();
  memory_and_generators();
  batches_and_selector();
  services();
  fprintf(stderr, "capi_args: %d checks, %d failed\n", checks, failures);
  return failures ? 1 : 0;
}
