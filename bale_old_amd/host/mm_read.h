/*
 * mm_read.h -- a Matrix Market coordinate file as host CSR arrays, read as
 * the reference's read_matrix_mm_to_dist reads it
 * (spmat/spmat_utils.upc:217-365): the banner "%%MatrixMarket matrix
 * coordinate pattern|integer ...", comment lines, the size line
 * "rows cols nnz", then nnz lines "row col [value]" (1-based; values are
 * ignored); each row keeps its columns in file order.  Host code only.
 *
 * Unlike the reference, a file whose entry count differs from its size line,
 * an entry out of range, or sizes past what the library takes (rows and
 * columns below 2^32, nonzeros below 2^31) is refused before anything is
 * allocated from them.
 */
#ifndef BALE_MM_READ_H
#define BALE_MM_READ_H
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct mm_csr {
  int64_t nr, nc, nnz;
  int64_t *offset;  /* [nr + 1] */
  int64_t *nonzero; /* [nnz] */
} mm_csr;

static void mm_free(mm_csr *m) {
  free(m->offset);
  free(m->nonzero);
  m->offset = m->nonzero = NULL;
}

/* 0 and *out filled, or -1 with an "ERROR: read_matrix_mm: ..." line on stderr. */
static int mm_read(const char *name, mm_csr *out) {
  memset(out, 0, sizeof *out);
  FILE *fp = fopen(name, "r");
  if (!fp) {
    fprintf(stderr, "ERROR: read_matrix_mm: cannot open %s\n", name);
    return -1;
  }
  char object[64] = "", format[64] = "", field[64] = "";
  if (fscanf(fp, "%%%%MatrixMarket %63s %63s %63s", object, format, field) != 3) {
    fprintf(stderr, "ERROR: read_matrix_mm: bad read of header!\n");
    fclose(fp);
    return -1;
  }
  if (strcmp(object, "matrix") || strcmp(format, "coordinate") ||
      (strcmp(field, "integer") && strcmp(field, "pattern"))) {
    fprintf(stderr, "ERROR: read_matrix_mm: unsupported matrix market format %s %s %s\n", object,
            format, field);
    fclose(fp);
    return -1;
  }
  /* the rest of the banner line, then comment lines */
  int c;
  while ((c = fgetc(fp)) != EOF && c != '\n') {}
  while ((c = fgetc(fp)) == '%')
    while ((c = fgetc(fp)) != EOF && c != '\n') {}
  if (c != EOF) ungetc(c, fp);
  long long nr = 0, nc = 0, nnz = 0;
  if (fscanf(fp, "%lld %lld %lld", &nr, &nc, &nnz) != 3 || nr <= 0 || nc <= 0 || nnz < 0 ||
      nr >= (1LL << 32) || nc >= (1LL << 32) || nnz >= (1LL << 31)) {
    fprintf(stderr, "ERROR: read_matrix_mm: bad size line (rows and columns in [1, 2^32), "
                    "nonzeros in [0, 2^31))\n");
    fclose(fp);
    return -1;
  }
  const size_t n1 = (size_t)(nnz ? nnz : 1);
  int64_t *rows = (int64_t *)malloc(8 * n1), *cols = (int64_t *)malloc(8 * n1);
  int64_t *off = (int64_t *)calloc((size_t)nr + 1, 8), *col = (int64_t *)malloc(8 * n1);
  int64_t *next = (int64_t *)malloc(8 * (size_t)nr);
  if (!rows || !cols || !off || !col || !next) {
    fprintf(stderr, "ERROR: read_matrix_mm: could not allocate arrays\n");
    fclose(fp);
    free(rows), free(cols), free(off), free(col), free(next);
    return -1;
  }
  const int pattern = !strcmp(field, "pattern");
  long long pos = 0, r, cl, val;
  int bad = 0;
  for (;;) {
    int got = pattern ? fscanf(fp, "%lld %lld", &r, &cl) : fscanf(fp, "%lld %lld %lld", &r, &cl, &val);
    if (got == EOF) break;
    if (got != (pattern ? 2 : 3) || r < 1 || r > nr || cl < 1 || cl > nc || pos == nnz) {
      bad = 1;
      break;
    }
    rows[pos] = r - 1;
    cols[pos++] = cl - 1;
    off[r]++;
  }
  fclose(fp);
  if (bad || pos != nnz) {
    fprintf(stderr, "ERROR: read_matrix_mm: %s entries (expected %lld)\n",
            bad ? "malformed, out-of-range or too many" : "too few", nnz);
    free(rows), free(cols), free(off), free(col), free(next);
    return -1;
  }
  for (long long i = 0; i < nr; i++) off[i + 1] += off[i];
  memcpy(next, off, 8 * (size_t)nr);
  for (long long k = 0; k < nnz; k++) col[next[rows[k]]++] = cols[k];
  free(rows), free(cols), free(next);
  out->nr = nr;
  out->nc = nc;
  out->nnz = nnz;
  out->offset = off;
  out->nonzero = col;
  return 0;
}
#endif
