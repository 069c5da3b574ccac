/*
 * bale_oracle.c -- CPU restatement of bale's histogram / index-gather hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Reference citations are file:line in the jdevinney/bale_old tree.
 */
#define _GNU_SOURCE
#include "bale_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#define GOLDEN 0x9E3779B97F4A7C15ULL

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t or_splitmix64_next(uint64_t *state) {
  *state += GOLDEN;
  return mix64(*state);
}

uint64_t or_splitmix64_at(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * GOLDEN);
}

uint64_t or_splitmix64(uint64_t j) { return mix64(j + GOLDEN); }

uint64_t or_reduce(uint64_t x, uint64_t n) {
  if (n <= 1) return 0;
  if ((n & (n - 1)) == 0) {
    int k = __builtin_ctzll(n);
    return x >> (64 - k);
  }
  return (uint64_t)(((unsigned __int128)x * n) >> 64);
}

/* ---- keyed permutation ------------------------------------------------ */
uint64_t or_perm(uint64_t x, uint64_t n, uint64_t key) {
  if (n <= 1) return 0;
  int bits = 64 - __builtin_clzll(n - 1); /* smallest b with 2^b >= n */
  int half = (bits + 1) / 2;
  if (half < 1) half = 1;
  uint64_t mask = (1ULL << half) - 1;
  do {
    uint64_t L = x >> half, R = x & mask;
    for (int r = 0; r < 4; r++) {
      uint64_t f = mix64(R ^ or_splitmix64_at(key, (uint64_t)r)) & mask;
      uint64_t t = R;
      R = L ^ f;
      L = t;
    }
    x = (L << half) | R;
  } while (x >= n);
  return x;
}

/* ---- Zipf by rejection-inversion -------------------------------------- */
/* h(x) = x^-s, H(x) = (x^(1-s) - 1)/(1-s) written with expm1/log1p so that
 * s -> 1 stays finite. */
typedef struct zipf_t {
  double s, hx1, hn, sc;
  int64_t n;
} zipf_t;

static double z_helper1(double t) { return fabs(t) > 1e-8 ? log1p(t) / t : 1.0 - t * (0.5 - t / 3.0); }
static double z_helper2(double t) { return fabs(t) > 1e-8 ? expm1(t) / t : 1.0 + t * 0.5 * (1.0 + t / 3.0); }
static double z_H(const zipf_t *z, double x) {
  double lx = log(x);
  return z_helper2((1.0 - z->s) * lx) * lx;
}
static double z_h(const zipf_t *z, double x) { return exp(-z->s * log(x)); }
static double z_Hinv(const zipf_t *z, double y) {
  double t = y * (1.0 - z->s);
  if (t < -1.0) t = -1.0;
  return exp(z_helper1(t) * y);
}
static void zipf_init(zipf_t *z, int64_t n, double s) {
  z->n = n;
  z->s = s;
  z->hx1 = z_H(z, 1.5) - 1.0;
  z->hn = z_H(z, (double)n + 0.5);
  z->sc = 2.0 - z_Hinv(z, z_H(z, 2.5) - z_h(z, 2.0));
}
static int64_t zipf_draw(const zipf_t *z, uint64_t subseed) {
  for (uint64_t k = 0;; k++) {
    uint64_t r = or_splitmix64_at(subseed, k);
    double u01 = (double)(r >> 11) * 0x1.0p-53;
    double u = z->hn + u01 * (z->hx1 - z->hn);
    double x = z_Hinv(z, u);
    int64_t kk = (int64_t)(x + 0.5);
    if (kk < 1) kk = 1;
    if (kk > z->n) kk = z->n;
    if ((double)kk - x <= z->sc || u >= z_H(z, (double)kk + 0.5) - z_h(z, (double)kk))
      return kk;
  }
}

double or_zipf_pmf(int64_t k, int64_t n, double s) {
  double norm = 0.0;
  for (int64_t j = n; j >= 1; j--) norm += pow((double)j, -s);
  return pow((double)k, -s) / norm;
}

#define RUN_CAP 4096
#define PERM_SALT 0x5045524D5045524DULL /* "PERMPERM" */
#define FLAG_SALT 0x52554E5352554E53ULL /* "RUNSRUNS" */

static int64_t draw_value(int64_t i, int64_t nbins, int dist, const zipf_t *z,
                          uint64_t seed, uint64_t pkey) {
  uint64_t r = or_splitmix64_at(seed, (uint64_t)i);
  if (dist == 0) return (int64_t)or_reduce(r, (uint64_t)nbins);
  int64_t rank = zipf_draw(z, r);
  return (int64_t)or_perm((uint64_t)(rank - 1), (uint64_t)nbins, pkey);
}

static int is_run_start(int64_t i, int64_t run_mean, uint64_t seed) {
  if (i % RUN_CAP == 0) return 1;
  uint64_t f = or_splitmix64_at(seed ^ FLAG_SALT, (uint64_t)i);
  return or_reduce(f, (uint64_t)run_mean) == 0;
}

void or_gen_index(int64_t *out, int64_t n, int64_t nbins, int dist, double s,
                  int64_t run_mean, uint64_t seed) {
  zipf_t z;
  if (dist == 1) zipf_init(&z, nbins, s);
  uint64_t pkey = mix64(seed ^ PERM_SALT);
  int64_t cur = 0;
  for (int64_t i = 0; i < n; i++) {
    if (run_mean > 1) {
      if (is_run_start(i, run_mean, seed)) cur = draw_value(i, nbins, dist, &z, seed, pkey);
      out[i] = cur;
    } else {
      out[i] = draw_value(i, nbins, dist, &z, seed, pkey);
    }
  }
}

void or_gen_uniform(int64_t *out, int64_t n, int64_t range, uint64_t seed) {
  for (int64_t i = 0; i < n; i++)
    out[i] = (int64_t)or_reduce(or_splitmix64_at(seed, (uint64_t)i), (uint64_t)range);
}

void or_fill_splitmix(int64_t *table, int64_t n) {
  for (int64_t j = 0; j < n; j++) table[j] = (int64_t)or_splitmix64((uint64_t)j);
}

void or_fill_neg(int64_t *table, int64_t n) {
  for (int64_t j = 0; j < n; j++) table[j] = -(j + 1);
}

void or_pack(int64_t *pck, const int64_t *idx, int64_t n, int64_t P) {
  for (int64_t i = 0; i < n; i++) {
    int64_t lindx = idx[i] / P, pe = idx[i] % P;
    pck[i] = (lindx << 16) | (pe & 0xffff);
  }
}

void or_gen_reference_rand(int64_t *out, int64_t n, int64_t range, unsigned seed0) {
  srand(seed0);
  for (int64_t i = 0; i < n; i++) out[i] = rand() % range;
}

/* ---- histogram --------------------------------------------------------- */
void or_histo_agi(const int64_t *index, int64_t n, int64_t *counts) {
  for (int64_t i = 0; i < n; i++) counts[index[i]] += 1;
}

void or_histo_agi_i32(const int64_t *index, int64_t n, int32_t *counts) {
  for (int64_t i = 0; i < n; i++) counts[index[i]] += 1;
}

int or_histo_owner(const int64_t *pckindx, int64_t n, int64_t *lcounts,
                   int64_t lnum, int64_t P) {
  /* push phase: route each decoded column to its owner's queue, in order */
  int64_t *cnt = calloc((size_t)P + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) {
    int64_t pe = pckindx[i] & 0xffff;
    if (pe >= P) { free(cnt); return -1; }
    cnt[pe + 1]++;
  }
  for (int64_t p = 0; p < P; p++) cnt[p + 1] += cnt[p];
  int64_t *q = malloc((size_t)(n ? n : 1) * sizeof(int64_t));
  int64_t *pos = malloc((size_t)P * sizeof(int64_t));
  memcpy(pos, cnt, (size_t)P * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) {
    int64_t pe = pckindx[i] & 0xffff;
    q[pos[pe]++] = pckindx[i] >> 16;
  }
  /* pull phase: every PE applies its queue to its local slice */
  for (int64_t p = 0; p < P; p++)
    for (int64_t k = cnt[p]; k < cnt[p + 1]; k++) lcounts[p * lnum + q[k]] += 1;
  free(q); free(pos); free(cnt);
  return 0;
}

/* ---- index gather ------------------------------------------------------ */
void or_ig_agi(int64_t *tgt, const int64_t *index, int64_t n, const int64_t *table) {
  for (int64_t i = 0; i < n; i++) tgt[i] = table[index[i]];
}

int or_ig_exstack(int64_t *tgt, const int64_t *pckindx, int64_t n,
                  const int64_t *ltables, int64_t lnum, int64_t P) {
  int64_t *cnt = calloc((size_t)P + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) {
    int64_t pe = pckindx[i] & 0xffff;
    if (pe >= P) { free(cnt); return -1; }
    cnt[pe + 1]++;
  }
  for (int64_t p = 0; p < P; p++) cnt[p + 1] += cnt[p];
  int64_t *q = malloc((size_t)(n ? n : 1) * sizeof(int64_t));
  int64_t *pos = malloc((size_t)P * sizeof(int64_t));
  memcpy(pos, cnt, (size_t)P * sizeof(int64_t));
  /* requests: raw local indices, FIFO per destination */
  for (int64_t i = 0; i < n; i++) {
    int64_t pe = pckindx[i] & 0xffff;
    q[pos[pe]++] = pckindx[i] >> 16;
  }
  /* owners replace each request in place by its table value (the reply) */
  for (int64_t p = 0; p < P; p++)
    for (int64_t k = cnt[p]; k < cnt[p + 1]; k++) q[k] = ltables[p * lnum + q[k]];
  /* retrace the requests in order, popping each source's FIFO */
  memcpy(pos, cnt, (size_t)P * sizeof(int64_t));
  for (int64_t j = 0; j < n; j++) {
    int64_t from = pckindx[j] & 0xffff;
    tgt[j] = q[pos[from]++];
  }
  free(q); free(pos); free(cnt);
  return 0;
}

int or_ig_conveyor(int64_t *tgt, const int64_t *pckindx, int64_t n,
                   const int64_t *ltables, int64_t lnum, int64_t P) {
  /* the reply carries its request slot, so delivery order is irrelevant:
   * serve the owners in reverse to make that visible */
  for (int64_t p = P - 1; p >= 0; p--)
    for (int64_t i = n - 1; i >= 0; i--) {
      int64_t pe = pckindx[i] & 0xffff;
      if (pe >= P) return -1;
      if (pe != p) continue;
      int64_t val = ltables[p * lnum + (pckindx[i] >> 16)];
      tgt[i] = val;
    }
  return 0;
}

/* ---- conveyor delivery (convey/narrative.h:135-137) ------------------- */
int64_t or_convey_deliver(const void *items, const int64_t *pes, int64_t n, int64_t item_bytes,
                          int64_t P, void *out, int64_t *starts) {
  /* one FIFO per destination, filled in push order and emptied in order */
  int64_t dropped = 0;
  for (int64_t p = 0; p <= P; p++) starts[p] = 0;
  for (int64_t i = 0; i < n; i++) {
    if (pes[i] < 0 || pes[i] >= P) dropped++;
    else starts[pes[i] + 1]++;
  }
  for (int64_t p = 0; p < P; p++) starts[p + 1] += starts[p];
  int64_t *tail = (int64_t *)malloc((size_t)(P > 0 ? P : 1) * sizeof(int64_t));
  if (!tail) return -1;
  for (int64_t p = 0; p < P; p++) tail[p] = starts[p];
  const unsigned char *src = (const unsigned char *)items;
  unsigned char *dst = (unsigned char *)out;
  for (int64_t i = 0; i < n; i++) {
    if (pes[i] < 0 || pes[i] >= P) continue;
    memcpy(dst + tail[pes[i]]++ * item_bytes, src + i * item_bytes, (size_t)item_bytes);
  }
  free(tail);
  return dropped;
}

/* ---- timed multi-thread CPU baseline ---------------------------------- */
typedef struct mt_arg {
  const int64_t *index;
  int64_t lo, hi;
  int64_t *counts;
  int64_t *tgt;
  const int64_t *table;
  pthread_barrier_t *bar;
  double secs;
  int gather;
} mt_arg;

static double wall_seconds(void) {
  struct timeval tp;
  gettimeofday(&tp, NULL);
  return (double)tp.tv_sec + (double)tp.tv_usec * 1.e-6;
}

static void *mt_body(void *p) {
  mt_arg *a = (mt_arg *)p;
  pthread_barrier_wait(a->bar);
  double t0 = wall_seconds();
  if (a->gather) {
    for (int64_t i = a->lo; i < a->hi; i++) a->tgt[i] = a->table[a->index[i]];
  } else {
    for (int64_t i = a->lo; i < a->hi; i++)
      __atomic_fetch_add(&a->counts[a->index[i]], 1, __ATOMIC_RELAXED);
  }
  pthread_barrier_wait(a->bar);
  a->secs = wall_seconds() - t0;
  return NULL;
}

static double run_mt(const int64_t *index, int64_t n, int64_t *counts, int64_t *tgt,
                     const int64_t *table, int nthreads, int gather) {
  if (nthreads < 1) nthreads = 1;
  pthread_t *th = malloc(sizeof(pthread_t) * (size_t)nthreads);
  mt_arg *args = malloc(sizeof(mt_arg) * (size_t)nthreads);
  pthread_barrier_t bar;
  pthread_barrier_init(&bar, NULL, (unsigned)nthreads);
  for (int t = 0; t < nthreads; t++) {
    args[t].index = index;
    args[t].lo = n * t / nthreads;
    args[t].hi = n * (t + 1) / nthreads;
    args[t].counts = counts;
    args[t].tgt = tgt;
    args[t].table = table;
    args[t].bar = &bar;
    args[t].gather = gather;
    pthread_create(&th[t], NULL, mt_body, &args[t]);
  }
  double sum = 0.0;
  for (int t = 0; t < nthreads; t++) {
    pthread_join(th[t], NULL);
    sum += args[t].secs;
  }
  pthread_barrier_destroy(&bar);
  free(th); free(args);
  return sum / nthreads;
}

double or_histo_agi_mt(const int64_t *index, int64_t n, int64_t *counts, int nthreads) {
  return run_mt(index, n, counts, NULL, NULL, nthreads, 0);
}

double or_ig_agi_mt(int64_t *tgt, const int64_t *index, int64_t n,
                    const int64_t *table, int nthreads) {
  return run_mt(index, n, NULL, tgt, table, nthreads, 1);
}

/* ---- timed multi-thread aggregate-by-owner baseline -------------------- */
/* P threads play P PEs; global index j belongs to PE j % P (the cyclic layout,
 * libgetput/libgetput.h:159,186-187).  Bulk-synchronous rounds as exstack runs
 * them (exstack/exstack.upc, apps/histo_src/histo_exstack.upc:64-76,
 * apps/ig_src/ig_exstack.upc:69-94): each PE pushes its items into a buffer of
 * at most kOwnBuf per destination until one fills or its input ends; after a
 * barrier every owner applies (histogram) or answers in place (gather) what was
 * sent to it; the gather adds a second barrier after which each requester
 * retraces its round's requests in order, popping each source's replies; a
 * closing barrier frees the buffers.  Rounds repeat until every PE's input is
 * done.  The shared arrays are the global counts / table / tgt: each owner
 * touches only its own entries, so no atomics are needed. */
enum { kOwnBuf = 1024 };
typedef struct own_arg {
  const int64_t *index;
  int64_t lo, hi;
  int64_t *counts;
  int64_t *tgt;
  const int64_t *table;
  int64_t *buf;    /* [P][P][kOwnBuf]: buf[src][dst] */
  int64_t *cnt;    /* [P][P] items in buf[src][dst] */
  int64_t *done;   /* [P] round in which the PE's input ended (or INT64_MAX) */
  pthread_barrier_t *bar;
  int me, P, gather;
  double secs;
} own_arg;

static void *own_body(void *p) {
  own_arg *a = (own_arg *)p;
  const int P = a->P, me = a->me;
  int64_t *mycnt = a->cnt + (int64_t)me * P;
  int64_t *pos = malloc(sizeof(int64_t) * (size_t)P);
  pthread_barrier_wait(a->bar);
  double t0 = wall_seconds();
  int64_t i = a->lo;
  for (int64_t round = 0;; round++) {
    /* push: until a destination buffer fills or the input ends */
    for (int d = 0; d < P; d++) mycnt[d] = 0;
    const int64_t i0 = i;
    for (; i < a->hi; i++) {
      const int64_t j = a->index[i];
      const int d = (int)(j % P);
      if (mycnt[d] == kOwnBuf) break;
      a->buf[((int64_t)me * P + d) * kOwnBuf + mycnt[d]++] = j / P;
    }
    if (i == a->hi && a->done[me] == INT64_MAX) a->done[me] = round;
    pthread_barrier_wait(a->bar);
    /* owners apply / answer what was sent to them */
    for (int src = 0; src < P; src++) {
      int64_t *b = a->buf + ((int64_t)src * P + me) * kOwnBuf;
      const int64_t c = a->cnt[(int64_t)src * P + me];
      if (a->gather)
        for (int64_t k = 0; k < c; k++) b[k] = a->table[b[k] * P + me];
      else
        for (int64_t k = 0; k < c; k++) a->counts[b[k] * P + me] += 1;
    }
    pthread_barrier_wait(a->bar);
    if (a->gather) {
      /* retrace this round's requests in order: each source's replies FIFO */
      for (int d = 0; d < P; d++) pos[d] = 0;
      for (int64_t k = i0; k < i; k++) {
        const int d = (int)(a->index[k] % P);
        a->tgt[k] = a->buf[((int64_t)me * P + d) * kOwnBuf + pos[d]++];
      }
      pthread_barrier_wait(a->bar);
    }
    int all = 1;
    for (int q = 0; q < P; q++)
      if (a->done[q] > round) all = 0;
    if (all) break;
  }
  pthread_barrier_wait(a->bar);
  a->secs = wall_seconds() - t0;
  free(pos);
  return NULL;
}

static double run_owner(const int64_t *index, int64_t n, int64_t *counts, int64_t *tgt,
                        const int64_t *table, int P, int gather) {
  if (P < 1) P = 1;
  int64_t *buf = malloc(sizeof(int64_t) * (size_t)P * (size_t)P * kOwnBuf);
  int64_t *cnt = calloc((size_t)P * (size_t)P, sizeof(int64_t));
  int64_t *done = malloc(sizeof(int64_t) * (size_t)P);
  pthread_t *th = malloc(sizeof(pthread_t) * (size_t)P);
  own_arg *args = malloc(sizeof(own_arg) * (size_t)P);
  if (!buf || !cnt || !done || !th || !args) {
    free(buf); free(cnt); free(done); free(th); free(args);
    return -1.0;
  }
  for (int q = 0; q < P; q++) done[q] = INT64_MAX;
  pthread_barrier_t bar;
  pthread_barrier_init(&bar, NULL, (unsigned)P);
  for (int t = 0; t < P; t++) {
    own_arg *g = &args[t];
    g->index = index;
    g->lo = n * t / P;
    g->hi = n * (t + 1) / P;
    g->counts = counts;
    g->tgt = tgt;
    g->table = table;
    g->buf = buf;
    g->cnt = cnt;
    g->done = done;
    g->bar = &bar;
    g->me = t;
    g->P = P;
    g->gather = gather;
    pthread_create(&th[t], NULL, own_body, g);
  }
  double sum = 0.0;
  for (int t = 0; t < P; t++) {
    pthread_join(th[t], NULL);
    sum += args[t].secs;
  }
  pthread_barrier_destroy(&bar);
  free(buf); free(cnt); free(done); free(th); free(args);
  return sum / P;
}

double or_histo_owner_mt(const int64_t *index, int64_t n, int64_t *counts, int nthreads) {
  return run_owner(index, n, counts, NULL, NULL, nthreads, 0);
}

double or_ig_owner_mt(int64_t *tgt, const int64_t *index, int64_t n, const int64_t *table,
                      int nthreads) {
  return run_owner(index, n, NULL, tgt, table, nthreads, 1);
}

/* ---- batched mix ------------------------------------------------------- */
void or_batch_params(or_instance *out, int64_t count, uint64_t seed,
                     int lo_tab, int hi_tab, int lo_ops, int hi_ops) {
  for (int64_t t = 0; t < count; t++) {
    uint64_t sub = or_splitmix64_at(seed, (uint64_t)t);
    uint64_t d0 = or_splitmix64_at(sub, 0), d1 = or_splitmix64_at(sub, 1);
    uint64_t d2 = or_splitmix64_at(sub, 2), d3 = or_splitmix64_at(sub, 3);
    uint64_t d4 = or_splitmix64_at(sub, 4), d5 = or_splitmix64_at(sub, 5);
    double u1 = (double)(d1 >> 11) * 0x1.0p-53, u2 = (double)(d2 >> 11) * 0x1.0p-53;
    double u4 = (double)(d4 >> 11) * 0x1.0p-53;
    or_instance *I = &out[t];
    I->kind = (int32_t)(d0 >> 63);
    I->table_len = (int64_t)floor(exp2((double)lo_tab + (double)(hi_tab - lo_tab) * u1));
    I->nops = (int64_t)floor(exp2((double)lo_ops + (double)(hi_ops - lo_ops) * u2));
    I->dist = (int32_t)or_reduce(d3, 3);
    I->s = 0.8 + 0.6 * u4;
    I->seed = d5;
  }
}

uint64_t or_digest(const int64_t *v, int64_t n) {
  uint64_t acc = 0;
  for (int64_t i = 0; i < n; i++) acc += mix64((uint64_t)v[i] ^ (uint64_t)i);
  return acc;
}

/* ---- random permutation by dart throwing ------------------------------ */
/* The drand48 family's generator (POSIX): X <- (0x5DEECE66D X + 0xB) mod 2^48,
 * lrand48 returns the top 31 bits of the new X, and srand48(s) sets
 * X = (s mod 2^32) << 16 | 0x330E.  Restated from the published definition so
 * a GPU sequence can be checked element by element; or_randperm_agi below calls
 * glibc's own lrand48, as the reference does. */
#define LCG48_A 0x5DEECE66DULL
#define LCG48_C 0xBULL
#define LCG48_MASK ((1ULL << 48) - 1)

uint64_t or_srand48_state(int64_t seed) {
  return ((uint64_t)(uint32_t)seed << 16) | 0x330EULL;
}

void or_lrand48_seq(uint64_t state, int64_t n, int64_t *out) {
  uint64_t x = state & LCG48_MASK;
  for (int64_t i = 0; i < n; i++) {
    x = (LCG48_A * x + LCG48_C) & LCG48_MASK;
    out[i] = (int64_t)(x >> 17);
  }
}

void or_glibc_lrand48(int64_t seed, int64_t n, int64_t *out) {
  srand48((long)seed);
  for (int64_t i = 0; i < n; i++) out[i] = lrand48();
}

/* rand_permp_agi at one PE (spmat/spmat_agi.upc:56-107): darts r = lrand48() % M,
 * M = 2N, thrown in turn; the i-th dart that finds target[r] empty places item
 * i there (:81-87, one PE: item i*THREADS + MYTHREAD = i); the permutation lists
 * the filled slots in order (:90-101).  srand48(seed) only when seed != 0 (:61),
 * on glibc's process-wide state, as in the reference.  Returns the number of
 * darts thrown, or -1 when memory runs out. */
int64_t or_randperm_agi(int64_t *perm, int64_t N, int64_t seed) {
  const int64_t M = 2 * N;
  int64_t *target = (int64_t *)malloc((size_t)(M ? M : 1) * sizeof(int64_t));
  if (!target) return -1;
  for (int64_t i = 0; i < M; i++) target[i] = -1;
  if (seed != 0) srand48((long)(int)seed);
  int64_t darts = 0;
  for (int64_t i = 0; i < N;) {
    const int64_t r = lrand48() % M;
    darts++;
    if (target[r] == -1) target[r] = i++;
  }
  int64_t pos = 0;
  for (int64_t s = 0; s < M; s++)
    if (target[s] != -1) perm[pos++] = target[s];
  free(target);
  return darts;
}

/* rand_permp_agi_opt at one PE (apps/randperm_src/alternates/randperm_agi_opt.upc):
 * every item draws its PE as lrand48() % THREADS (:38-42), N draws that leave
 * nothing to choose at one PE but advance the generator; the items reach PE 0
 * in order (:63-76), so tmp_perm = 0..N-1; then a Fisher-Yates pass
 * j = i + lrand48() % (N - i), swap tmp[i] and tmp[j] (:92-97), and the
 * result is tmp (:99-103).  srand48(seed) only when seed != 0 (:25). */
void or_randperm_agi_opt(int64_t *perm, int64_t N, int64_t seed) {
  if (seed != 0) srand48((long)(int)seed);
  for (int64_t i = 0; i < N; i++) {
    (void)(lrand48() % 1);
    perm[i] = i;
  }
  for (int64_t i = 0; i < N; i++) {
    const int64_t j = i + lrand48() % (N - i);
    const int64_t t = perm[j];
    perm[j] = perm[i];
    perm[i] = t;
  }
}

/* ---- sparse matrices ------------------------------------------------------ */

/* log(x), x in (0, 1]: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m =
 * 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172, by its series to s^25
 * (remainder below 2^-60 relative); plus e ln 2.  Only correctly rounded
 * operations (the Makefile turns contraction off), so any IEEE double
 * implementation gives the same bits. */
double or_det_log(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  int64_t e = (int64_t)((b >> 52) & 0x7FF) - 1022;
  b = (b & 0x000FFFFFFFFFFFFFULL) | (0x3FEULL << 52); /* m in [0.5, 1) */
  double m;
  memcpy(&m, &b, 8);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0), z = s * s;
  double t = 1.0 / 25.0;
  for (int k = 11; k >= 0; k--) t = t * z + 1.0 / (double)(2 * k + 1);
  return (double)e * 0.69314718055994530942 + 2.0 * s * t;
}

/* the generator's stream for (row, half) and its geometric gaps */
static uint64_t er_key(uint64_t seed, int64_t row, int half) {
  return or_splitmix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(2 * row + half + 1));
}
/* nonzeros of row `row` in columns [lo, hi), ascending; written to out when
 * not NULL; returns the count */
static int64_t er_half(uint64_t key, int64_t lo, int64_t hi, double p, double D, int64_t *out) {
  int64_t cnt = 0;
  if (lo >= hi || p <= 0.0) return 0;
  if (p >= 1.0) {
    for (int64_t c = lo; c < hi; c++) {
      if (out) out[cnt] = c;
      cnt++;
    }
    return cnt;
  }
  int64_t col = lo - 1;
  for (uint64_t i = 0;; i++) {
    const uint64_t x = or_splitmix64_at(key, i);
    const double u = (double)((x >> 11) + 1) * 0x1p-53; /* (0, 1] */
    const double g = floor(or_det_log(u) / D);
    if (g >= (double)(hi - col)) break;
    col += 1 + (int64_t)g;
    if (col >= hi) break;
    if (out) out[cnt] = col;
    cnt++;
  }
  return cnt;
}

int64_t or_gen_er(int64_t n, double p, int unit_diag, int mode, uint64_t seed, int64_t *offset,
                  int64_t *nonzero) {
  const double D = p > 0.0 && p < 1.0 ? or_det_log(1.0 - p) : -1.0;
  int64_t pos = 0;
  offset[0] = 0;
  for (int64_t r = 0; r < n; r++) {
    if (mode == 1 || mode == 3)
      pos += er_half(er_key(seed, r, 0), 0, r, p, D, nonzero ? nonzero + pos : NULL);
    if (unit_diag) {
      if (nonzero) nonzero[pos] = r;
      pos++;
    }
    if (mode == 2 || mode == 3)
      pos += er_half(er_key(seed, r, 1), r + 1, n, p, D, nonzero ? nonzero + pos : NULL);
    offset[r + 1] = pos;
  }
  return pos;
}

void or_transpose_agi(int64_t numrows, int64_t numcols, const int64_t *offset,
                      const int64_t *nonzero, int64_t *t_offset, int64_t *t_nonzero) {
  const int64_t nnz = offset[numrows];
  int64_t *cur = (int64_t *)calloc((size_t)(numcols ? numcols : 1), sizeof(int64_t));
  /* column counts (:196-201, the fetch_and_inc histogram) */
  for (int64_t j = 0; j < nnz; j++) cur[nonzero[j]]++;
  /* row starts of the transpose (:216-218) */
  t_offset[0] = 0;
  for (int64_t c = 0; c < numcols; c++) {
    t_offset[c + 1] = t_offset[c] + cur[c];
    cur[c] = t_offset[c];
  }
  /* rows in order, each nonzero to the next free slot of its column (:225-230) */
  for (int64_t r = 0; r < numrows; r++)
    for (int64_t j = offset[r]; j < offset[r + 1]; j++) t_nonzero[cur[nonzero[j]]++] = r;
  free(cur);
}

void or_permute_agi(int64_t numrows, const int64_t *offset, const int64_t *nonzero,
                    const int64_t *rperminv, const int64_t *cperminv, int64_t *p_offset,
                    int64_t *p_nonzero) {
  int64_t *rperm = (int64_t *)malloc((size_t)(numrows ? numrows : 1) * sizeof(int64_t));
  /* rperm from rperminv (:128-131) */
  for (int64_t i = 0; i < numrows; i++) rperm[rperminv[i]] = i;
  /* permuted rows in order (:145-154), then the columns mapped (:158-163) */
  int64_t pos = 0;
  p_offset[0] = 0;
  for (int64_t i = 0; i < numrows; i++) {
    const int64_t row = rperm[i];
    for (int64_t j = offset[row]; j < offset[row + 1]; j++) p_nonzero[pos++] = nonzero[j];
    p_offset[i + 1] = pos;
  }
  for (int64_t j = 0; j < pos; j++) p_nonzero[j] = cperminv[p_nonzero[j]];
  free(rperm);
}

int64_t or_toposort_agi(int64_t n, const int64_t *offset, const int64_t *nonzero,
                        const int64_t *t_offset, const int64_t *t_nonzero, int64_t *rperm,
                        int64_t *cperm, int64_t *levels) {
  const size_t sz = (size_t)(n ? n : 1) * sizeof(int64_t);
  int64_t *queue = (int64_t *)malloc(sz), *rowsum = (int64_t *)malloc(sz),
          *rowcnt = (int64_t *)malloc(sz);
  /* rowcnt, rowsum (the sum of a row's columns) and the degree-one rows (:79-87) */
  int64_t start = 0, end = 0, tail = 0;
  for (int64_t i = 0; i < n; i++) {
    rowcnt[i] = offset[i + 1] - offset[i];
    rowsum[i] = 0;
    for (int64_t j = offset[i]; j < offset[i + 1]; j++) rowsum[i] += nonzero[j];
    if (rowcnt[i] == 1) queue[end++] = i;
  }
  tail = end;
  /* levels (:110-138): the rows queued when a level starts are processed in
   * order; each claims the next position from the end, and the rows of its
   * column lose one nonzero; a row left with one joins the next level */
  int64_t pos = 0, nlev = 0;
  while (end > start) {
    nlev++;
    while (start < end) {
      const int64_t row = queue[start++];
      const int64_t col = rowsum[row];
      rperm[row] = n - 1 - pos;
      cperm[col] = n - 1 - pos;
      pos++;
      for (int64_t j = t_offset[col]; j < t_offset[col + 1]; j++) {
        const int64_t s = t_nonzero[j];
        const int64_t old = rowcnt[s]--;
        rowsum[s] -= col;
        if (old == 2) queue[tail++] = s;
      }
    }
    end = tail;
  }
  if (levels) *levels = nlev;
  free(queue);
  free(rowsum);
  free(rowcnt);
  return pos;
}

/* is_perm (spmat/spmat_utils.upc:508-530): every value 0..N-1 exactly once. */
int or_is_perm(const int64_t *perm, int64_t N) {
  unsigned char *seen = (unsigned char *)calloc((size_t)(N ? N : 1), 1);
  if (!seen) return 0;
  int ok = 1;
  for (int64_t i = 0; i < N && ok; i++) {
    if (perm[i] < 0 || perm[i] >= N || seen[perm[i]]) ok = 0;
    else seen[perm[i]] = 1;
  }
  free(seen);
  return ok;
}
