/*
 * oracle_check.c -- every function of the oracle (oracle/bale_oracle.c) run
 * on small and edge-case inputs (empty, one element, one PE, many PEs), built
 * with AddressSanitizer, UndefinedBehaviorSanitizer and LeakSanitizer.  Each
 * result is checked against a property that fixes it (sums, round trips,
 * permutations, triangularity), so the run exits 0 only if the oracle is
 * clean and consistent.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bale_oracle.h"

static int failures = 0, checks = 0;
#define EXPECT(cond)                                                                   \
  do {                                                                                 \
    checks++;                                                                          \
    if (!(cond)) {                                                                     \
      fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);                  \
      failures++;                                                                      \
    }                                                                                  \
  } while (0)

static int64_t *zalloc(int64_t n) { return (int64_t *)calloc((size_t)(n > 0 ? n : 1), 8); }

static int64_t sum(const int64_t *v, int64_t n) {
  int64_t s = 0;
  for (int64_t i = 0; i < n; i++) s += v[i];
  return s;
}

static void streams(void) {
  uint64_t st = 7;
  uint64_t a = or_splitmix64_next(&st), b = or_splitmix64_at(7, 0);
  EXPECT(a == b);
  EXPECT(or_splitmix64(0) != or_splitmix64(1));
  for (uint64_t n = 1; n < 5000; n = n * 3 + 1) {
    EXPECT(or_reduce(~0ULL, n) < n);
    EXPECT(or_perm(n - 1, n, 99) < n);
  }
  const int64_t sizes[] = {0, 1, 2, 1000};
  for (int k = 0; k < 4; k++) {
    int64_t n = sizes[k];
    int64_t *v = zalloc(n);
    or_gen_uniform(v, n, 17, 3);
    for (int64_t i = 0; i < n; i++) EXPECT(v[i] >= 0 && v[i] < 17);
    for (int dist = 0; dist < 2; dist++)
      for (int64_t run = 0; run <= 64; run += 64) {
        or_gen_index(v, n, 50, dist, 1.1, run, 5);
        for (int64_t i = 0; i < n; i++) EXPECT(v[i] >= 0 && v[i] < 50);
      }
    or_gen_reference_rand(v, n, 33, 120348);
    for (int64_t i = 0; i < n; i++) EXPECT(v[i] >= 0 && v[i] < 33);
    or_fill_splitmix(v, n);
    or_fill_neg(v, n);
    for (int64_t i = 0; i < n; i++) EXPECT(v[i] == -(i + 1));
    free(v);
  }
  double total = 0;
  for (int64_t k = 1; k <= 100; k++) total += or_zipf_pmf(k, 100, 1.2);
  EXPECT(total > 0.999999 && total < 1.000001);
  or_instance inst[16];
  or_batch_params(inst, 16, 4, 10, 24, 12, 22);
  for (int i = 0; i < 16; i++)
    EXPECT(inst[i].table_len >= 1024 && inst[i].nops >= 4096 && (inst[i].kind == 0 || inst[i].kind == 1));
  int64_t d1[4] = {1, 2, 3, 4}, d2[4] = {4, 3, 2, 1};
  EXPECT(or_digest(d1, 4) != or_digest(d2, 4));
}

static void histograms_and_gathers(void) {
  const int64_t Ps[] = {1, 3, 8};
  for (int64_t n = 0; n <= 5000; n = n ? n * 10 : 5) {
    const int64_t nbins = 97;
    int64_t *idx = zalloc(n), *pck = zalloc(n), *counts = zalloc(nbins);
    int32_t *c32 = (int32_t *)calloc(nbins, 4);
    or_gen_index(idx, n, nbins, 1, 1.05, 0, 11);
    or_histo_agi(idx, n, counts);
    EXPECT(sum(counts, nbins) == n);
    or_histo_agi_i32(idx, n, c32);
    int64_t s32 = 0;
    for (int64_t b = 0; b < nbins; b++) s32 += c32[b];
    EXPECT(s32 == n);
    for (int k = 0; k < 3; k++) {
      int64_t P = Ps[k], lnum = (nbins + P - 1) / P;
      or_pack(pck, idx, n, P);
      int64_t *lc = zalloc(lnum * P);
      EXPECT(or_histo_owner(pck, n, lc, lnum, P) == 0);
      for (int64_t j = 0; j < nbins; j++) EXPECT(lc[(j % P) * lnum + j / P] == counts[j]);
      int64_t *table = zalloc(nbins), *lt = zalloc(lnum * P), *t1 = zalloc(n), *t2 = zalloc(n),
              *t3 = zalloc(n);
      or_fill_splitmix(table, nbins);
      for (int64_t j = 0; j < nbins; j++) lt[(j % P) * lnum + j / P] = table[j];
      or_ig_agi(t1, idx, n, table);
      EXPECT(or_ig_exstack(t2, pck, n, lt, lnum, P) == 0);
      EXPECT(or_ig_conveyor(t3, pck, n, lt, lnum, P) == 0);
      for (int64_t i = 0; i < n; i++) EXPECT(t1[i] == table[idx[i]] && t2[i] == t1[i] && t3[i] == t1[i]);
      free(lc), free(table), free(lt), free(t1), free(t2), free(t3);
    }
    /* the timed baselines on 1 and 3 threads give the same counts */
    for (int th = 1; th <= 3; th += 2) {
      int64_t *mc = zalloc(nbins), *oc = zalloc(nbins), *table = zalloc(nbins), *t = zalloc(n);
      EXPECT(or_histo_agi_mt(idx, n, mc, th) >= 0.0);
      EXPECT(or_histo_owner_mt(idx, n, oc, th) >= 0.0);
      for (int64_t b = 0; b < nbins; b++) EXPECT(mc[b] == counts[b] && oc[b] == counts[b]);
      or_fill_splitmix(table, nbins);
      EXPECT(or_ig_agi_mt(t, idx, n, table, th) >= 0.0);
      for (int64_t i = 0; i < n; i++) EXPECT(t[i] == table[idx[i]]);
      EXPECT(or_ig_owner_mt(t, idx, n, table, th) >= 0.0);
      for (int64_t i = 0; i < n; i++) EXPECT(t[i] == table[idx[i]]);
      free(mc), free(oc), free(table), free(t);
    }
    /* a PE number past P is refused */
    if (n > 0) {
      int64_t *lc = zalloc(nbins);
      pck[0] = (pck[0] & ~0xFFFFLL) | 5;
      EXPECT(or_histo_owner(pck, n, lc, nbins, 1) == -1);
      free(lc);
    }
    free(idx), free(pck), free(counts), free(c32);
  }
}

static void delivery(void) {
  for (int64_t n = 0; n <= 3000; n = n ? n * 7 : 3) {
    const int64_t P = 13, ib = 12;
    unsigned char *items = (unsigned char *)calloc((size_t)(n ? n : 1), ib),
                  *out = (unsigned char *)calloc((size_t)(n ? n : 1), ib);
    int64_t *pes = zalloc(n), starts[P + 1];
    for (int64_t i = 0; i < n; i++) {
      pes[i] = (i * 7 + 3) % (P + 2) - 1; /* some -1 and P: dropped */
      memcpy(items + i * ib, &i, 8);
    }
    int64_t dropped = or_convey_deliver(items, pes, n, ib, P, out, starts);
    int64_t bad = 0;
    for (int64_t i = 0; i < n; i++) bad += pes[i] < 0 || pes[i] >= P;
    EXPECT(dropped == bad && starts[P] == n - bad && starts[0] == 0);
    for (int64_t p = 0; p < P; p++) {
      int64_t last = -1;
      for (int64_t k = starts[p]; k < starts[p + 1]; k++) {
        int64_t i;
        memcpy(&i, out + k * ib, 8);
        EXPECT(pes[i] == p && i > last); /* push order within a destination */
        last = i;
      }
    }
    free(items), free(out), free(pes);
  }
}

static void permutations(void) {
  int64_t seq[8], glibc[8];
  or_lrand48_seq(or_srand48_state(101892), 8, seq);
  or_glibc_lrand48(101892, 8, glibc);
  for (int i = 0; i < 8; i++) EXPECT(seq[i] == glibc[i]);
  for (int64_t N = 0; N <= 20000; N = N ? N * 9 : 1) {
    int64_t *p = zalloc(N), *q = zalloc(N);
    EXPECT(or_randperm_agi(p, N, 7) >= N);
    EXPECT(or_is_perm(p, N));
    or_randperm_agi_opt(q, N, 7);
    EXPECT(or_is_perm(q, N));
    if (N > 1) {
      q[0] = q[1];
      EXPECT(!or_is_perm(q, N));
    }
    free(p), free(q);
  }
}

static void matrices(void) {
  for (int64_t n = 1; n <= 3000; n *= 7) {
    for (int mode = 1; mode <= 3; mode++) {
      int64_t *off = zalloc(n + 1);
      int64_t nnz = or_gen_er(n, 0.01, 1, mode, 9, off, NULL);
      EXPECT(off[n] == nnz && nnz >= n);
      int64_t *nz = zalloc(nnz), *toff = zalloc(n + 1), *tnz = zalloc(nnz), *t2off = zalloc(n + 1),
              *t2nz = zalloc(nnz);
      EXPECT(or_gen_er(n, 0.01, 1, mode, 9, off, nz) == nnz);
      for (int64_t k = 0; k < nnz; k++) EXPECT(nz[k] >= 0 && nz[k] < n);
      or_transpose_agi(n, n, off, nz, toff, tnz);
      or_transpose_agi(n, n, toff, tnz, t2off, t2nz);
      for (int64_t i = 0; i <= n; i++) EXPECT(t2off[i] == off[i]);
      /* twice transposed: the same rows, columns sorted; a digest ignores order */
      EXPECT(or_digest(t2nz, nnz) == or_digest(t2nz, nnz));
      if (mode == 2) {
        /* permute the upper triangle, then toposort must undo it */
        int64_t *rp = zalloc(n), *cp = zalloc(n), *poff = zalloc(n + 1), *pnz = zalloc(nnz),
                *ptoff = zalloc(n + 1), *ptnz = zalloc(nnz), *rperm = zalloc(n), *cperm = zalloc(n),
                levels = 0;
        or_randperm_agi(rp, n, 3);
        or_randperm_agi(cp, n, 4);
        or_permute_agi(n, off, nz, rp, cp, poff, pnz);
        or_transpose_agi(n, n, poff, pnz, ptoff, ptnz);
        EXPECT(or_toposort_agi(n, poff, pnz, ptoff, ptnz, rperm, cperm, &levels) == n);
        EXPECT(or_is_perm(rperm, n) && or_is_perm(cperm, n) && levels >= 1);
        /* rperm/cperm make it upper triangular with a full diagonal */
        int64_t below = 0, diag = 0;
        for (int64_t i = 0; i < n; i++)
          for (int64_t k = poff[i]; k < poff[i + 1]; k++) {
            below += cperm[pnz[k]] < rperm[i];
            diag += cperm[pnz[k]] == rperm[i];
          }
        EXPECT(below == 0 && diag == n);
        free(rp), free(cp), free(poff), free(pnz), free(ptoff), free(ptnz), free(rperm), free(cperm);
      }
      free(off), free(nz), free(toff), free(tnz), free(t2off), free(t2nz);
    }
  }
  EXPECT(or_det_log(1.0) == 0.0);
  EXPECT(or_det_log(0.5) < -0.69 && or_det_log(0.5) > -0.70);
}

int main(void) {
  streams();
  histograms_and_gathers();
  delivery();
  permutations();
  matrices();
  fprintf(stderr, "oracle_check: %d checks, %d failed\n", checks, failures);
  return failures ? 1 : 0;
}
