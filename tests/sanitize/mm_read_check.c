/*
 * mm_read_check.c -- the drivers' Matrix Market reader
 * (bale_old_amd/host/mm_read.h) on good, malformed and hostile files, built
 * with AddressSanitizer, UndefinedBehaviorSanitizer and LeakSanitizer.  The
 * files are written to the directory named by the first argument.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mm_read.h"

static int failures = 0, checks = 0;
#define EXPECT(cond)                                                                   \
  do {                                                                                 \
    checks++;                                                                          \
    if (!(cond)) {                                                                     \
      fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);                  \
      failures++;                                                                      \
    }                                                                                  \
  } while (0)

static char path[4096];

static const char *write_file(const char *dir, const char *name, const char *text) {
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "w");
  if (!f) {
    perror(path);
    exit(2);
  }
  fputs(text, f);
  fclose(f);
  return path;
}

static int reads(const char *dir, const char *name, const char *text, mm_csr *m) {
  return mm_read(write_file(dir, name, text), m);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const char *dir = argv[1];
  mm_csr m;
  /* a good file: columns kept in file order within each row */
  EXPECT(reads(dir, "good.mtx",
               "%%MatrixMarket matrix coordinate pattern general\n% comment\n%another\n"
               "3 4 5\n2 4\n1 1\n2 1\n3 3\n1 2\n", &m) == 0);
  EXPECT(m.nr == 3 && m.nc == 4 && m.nnz == 5);
  const int64_t off[4] = {0, 2, 4, 5}, nz[5] = {0, 1, 3, 0, 2};
  for (int i = 0; i < 4; i++) EXPECT(m.offset[i] == off[i]);
  for (int k = 0; k < 5; k++) EXPECT(m.nonzero[k] == nz[k]);
  mm_free(&m);
  EXPECT(reads(dir, "int.mtx", "%%MatrixMarket matrix coordinate integer general\n2 2 2\n1 1 7\n2 2 -3\n",
               &m) == 0);
  EXPECT(m.nnz == 2 && m.offset[2] == 2 && m.nonzero[1] == 1);
  mm_free(&m);
  EXPECT(reads(dir, "empty.mtx", "%%MatrixMarket matrix coordinate pattern general\n4 4 0\n", &m) == 0);
  EXPECT(m.nnz == 0 && m.offset[4] == 0);
  mm_free(&m);
  /* refusals */
  const char *bad[] = {
      "",
      "not a matrix market file\n",
      "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
      "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 0.5\n",
      "%%MatrixMarket matrix coordinate pattern general\n",
      "%%MatrixMarket matrix coordinate pattern general\n0 3 0\n",
      "%%MatrixMarket matrix coordinate pattern general\n3 3 -1\n",
      "%%MatrixMarket matrix coordinate pattern general\n3 3 2\n1 1\n",             /* too few */
      "%%MatrixMarket matrix coordinate pattern general\n3 3 1\n1 1\n2 2\n",        /* too many */
      "%%MatrixMarket matrix coordinate pattern general\n3 3 1\n4 1\n",             /* row range */
      "%%MatrixMarket matrix coordinate pattern general\n3 3 1\n1 0\n",             /* column range */
      "%%MatrixMarket matrix coordinate pattern general\n3 3 1\n1 x\n",             /* malformed */
      "%%MatrixMarket matrix coordinate integer general\n3 3 1\n1 1\n",             /* no value */
      /* sizes whose byte counts would wrap a size_t: refused before allocating */
      "%%MatrixMarket matrix coordinate pattern general\n5 5 2305843009213693953\n1 1\n2 2\n3 3\n",
      "%%MatrixMarket matrix coordinate pattern general\n9223372036854775807 5 1\n1 1\n",
      "%%MatrixMarket matrix coordinate pattern general\n5 5 99999999999999999999999\n1 1\n",
  };
  for (size_t k = 0; k < sizeof bad / sizeof bad[0]; k++) {
    char name[32];
    snprintf(name, sizeof name, "bad%zu.mtx", k);
    EXPECT(reads(dir, name, bad[k], &m) == -1);
    EXPECT(m.offset == NULL && m.nonzero == NULL);
  }
  snprintf(path, sizeof path, "%s/missing.mtx", dir);
  EXPECT(mm_read(path, &m) == -1);
  fprintf(stderr, "mm_read_check: %d checks, %d failed\n", checks, failures);
  return failures ? 1 : 0;
}
