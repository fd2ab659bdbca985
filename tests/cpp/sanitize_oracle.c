/* sanitize_oracle.c -- the oracle (oracle/tbsla_oracle.c, oracle/problems.c)
 * under AddressSanitizer and UBSan: every fill of every format over shapes,
 * block grids and rewiring probabilities, the conversions from COO, the
 * SpMV family, row and column operations, PageRank (three variants), CG and
 * PCG, the SPD problem generators and the Matrix Market reader (valid and
 * malformed files).  Any sanitizer report aborts the run; prints
 * "sanitize_oracle ok" otherwise.
 *
 *   sanitize_oracle <tmpdir> */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tbsla_oracle.h"

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

static void exercise(int fmt, void *m) {
  const orc_geom *g = (const orc_geom *)m;
  const int nin = fmt == ORC_COO ? g->n_col : g->ln_col;
  const int nout = orc_result_len(fmt, m);
  const int big = (g->n_row > g->n_col ? g->n_row : g->n_col) + 1;
  double *v = zalloc(sizeof(double) * (size_t)(nin > big ? nin : big));
  double *r = zalloc(sizeof(double) * (size_t)(nout > big ? nout : big));
  double *buf = zalloc(sizeof(double) * (size_t)big);
  for (int i = 0; i < nin; i++) v[i] = orc_unit_value(1, (uint64_t)i);
  orc_spmv(fmt, m, r, v, 0);
  orc_Ax(fmt, m, r, v, 0);
  if (g->ln_row == g->ln_col && fmt != ORC_COO) {
    orc_a_axpx_(fmt, m, r, v, 0);
    orc_AAxpAx(fmt, m, r, v, buf, 0);
    orc_AAxpAxpx(fmt, m, r, v, buf, 0);
  }
  double *s = zalloc(sizeof(double) * (size_t)big);
  orc_get_row_sums(fmt, m, s);
  orc_get_row_max_abs(fmt, m, s);
  orc_apply_exponential(fmt, m, s, -1);
  memset(s, 0, sizeof(double) * (size_t)big);
  orc_get_col_sums(fmt, m, s);
  orc_make_stochastic(fmt, m, s);
  if (g->NR == 1 && g->NC == 1 && g->n_row == g->n_col) {
    double err = 0;
    double *b = zalloc(sizeof(double) * (size_t)big);
    orc_page_rank(fmt, m, 0.85, 1e-10, 50, b, &err);
    orc_page_rank_max_norm(fmt, m, 0.85, 1e-10, 50, b, &err);
    orc_page_rank_source(fmt, m, 0.85, 1e-10, 50, 0, b, &err);
    free(b);
  }
  free(s);
  free(v);
  free(r);
  free(buf);
}

static void fills(void) {
  static const int shapes[][3] = {{30, 30, 6}, {20, 30, 8}, {30, 20, 12}, {10, 5, 12}, {64, 64, 1}};
  static const double qs[] = {0.0, 0.3, 1.0};
  static const int grids[][2] = {{1, 1}, {2, 2}, {3, 2}, {4, 4}};
  for (int si = 0; si < 5; si++)
    for (int qi = 0; qi < 3; qi++)
      for (int gi = 0; gi < 4; gi++)
        for (int pr = 0; pr < grids[gi][0]; pr++)
          for (int pc = 0; pc < grids[gi][1]; pc++) {
            const int nr = shapes[si][0], nc = shapes[si][1], c = shapes[si][2];
            const int NR = grids[gi][0], NC = grids[gi][1];
            const double q = qs[qi];
            orc_csr csr = {0};
            orc_ell ell = {0};
            orc_coo coo = {0}, scoo = {0};
            orc_dense dn = {0};
            orc_csr_fill_cqmat(&csr, nr, nc, c, q, 1, pr, pc, NR, NC, 7);
            orc_ell_fill_cqmat(&ell, nr, nc, c, q, 1, pr, pc, NR, NC, 7);
            orc_coo_fill_cqmat(&coo, nr, nc, c, q, 1, pr, pc, NR, NC, 7);
            orc_scoo_fill_cqmat(&scoo, nr, nc, c, q, 1, pr, pc, NR, NC, 7);
            orc_dense_fill_cqmat(&dn, nr, nc, c, q, 1, pr, pc, NR, NC, 7);
            exercise(ORC_CSR, &csr);
            exercise(ORC_ELL, &ell);
            exercise(ORC_COO, &coo);
            exercise(ORC_SCOO, &scoo);
            exercise(ORC_DENSE, &dn);
            orc_csr_free(&csr);
            orc_ell_free(&ell);
            orc_coo_free(&scoo);
            orc_dense_free(&dn);
            /* the conversions from a whole-matrix COO */
            orc_coo full = {0};
            orc_coo_fill_cqmat(&full, nr, nc, c, q, 0, 0, 0, 1, 1, -1);
            orc_csr a = {0};
            orc_ell b = {0};
            orc_coo d = {0};
            orc_dense e = {0};
            orc_csr_from_coo(&a, &full);
            orc_ell_from_coo(&b, &full);
            orc_scoo_from_coo(&d, &full);
            orc_dense_from_coo(&e, &full);
            exercise(ORC_CSR, &a);
            orc_csr_free(&a);
            orc_ell_free(&b);
            orc_coo_free(&d);
            orc_dense_free(&e);
            orc_coo_free(&full);
            orc_coo_free(&coo);
            /* cdiag */
            orc_csr cd = {0};
            orc_ell ed = {0};
            orc_coo qd = {0}, sd = {0};
            orc_dense dd = {0};
            orc_csr_fill_cdiag(&cd, nr, nc, c % 4, pr, pc, NR, NC);
            orc_ell_fill_cdiag(&ed, nr, nc, c % 4, pr, pc, NR, NC);
            orc_coo_fill_cdiag(&qd, nr, nc, c % 4, pr, pc, NR, NC);
            orc_scoo_fill_cdiag(&sd, nr, nc, c % 4, pr, pc, NR, NC);
            orc_dense_fill_cdiag(&dd, nr, nc, c % 4, pr, pc, NR, NC);
            exercise(ORC_CSR, &cd);
            exercise(ORC_DENSE, &dd);
            orc_csr_free(&cd);
            orc_ell_free(&ed);
            orc_coo_free(&qd);
            orc_coo_free(&sd);
            orc_dense_free(&dd);
          }
}

static void solvers(void) {
  const int nx = 9, ny = 7, nz = 5, n = nx * ny * nz;
  double *k = zalloc(sizeof(double) * n), *rhs = zalloc(sizeof(double) * n);
  double *x = zalloc(sizeof(double) * n);
  orc_lognormal_field(k, nx, ny, nz, 0.0, 1.0, 2.0, 3);
  orc_uniform_vector(rhs, n, 4, -1.0, 1.0);
  orc_csr p = {0};
  orc_poisson3d(&p, nx, ny, nz, k);
  double res = 0;
  orc_cg(ORC_CSR, &p, rhs, x, 500, 1e-10, 1, &res);
  orc_cg_clamp(ORC_CSR, &p, rhs, x, 500, 1e-10, 0, 1e-10, &res);
  orc_csr t = {0};
  orc_csr_transpose(&t, &p);
  for (int kind = 0; kind < 3; kind++)
    orc_pcg(&p, rhs, x, 500, 1e-10, 1, 1e-300, kind, kind == 2 ? &t : NULL, &res);
  double *B = zalloc(sizeof(double) * (size_t)n * 5), *C = zalloc(sizeof(double) * (size_t)n * 5);
  orc_csr_spmm(&p, B, C, 5);
  orc_csr bq = {0}, s = {0};
  orc_csr_fill_cqmat(&bq, 300, 300, 9, 0.2, 1, 0, 0, 1, 1, 2);
  orc_sym_cqmat(&s, &bq);
  orc_pcg(&s, rhs, x, 300, 1e-10, 1, 1e-300, 1, NULL, &res);
  orc_csr_free(&bq);
  orc_csr_free(&s);
  orc_csr_free(&t);
  orc_csr_free(&p);
  free(B);
  free(C);
  free(k);
  free(rhs);
  free(x);
}

static void matrix_market(const char *dir) {
  static const char *files[][2] = {
      {"g.mtx", "%%MatrixMarket matrix coordinate real general\n% c\n3 4 3\n1 1 1.5\n2 3 -2\n3 4 7\n"},
      {"s.mtx", "%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 4\n2 1 -1\n3 2 -1\n"},
      {"short.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1\n"},
      {"range.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n"},
      {"array.mtx", "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n"},
      {"empty.mtx", ""},
  };
  char path[4096];
  for (int i = 0; i < 6; i++) {
    snprintf(path, sizeof path, "%s/%s", dir, files[i][0]);
    FILE *f = fopen(path, "wb");
    if (!f) continue;
    fputs(files[i][1], f);
    fclose(f);
    orc_coo m = {0};
    if (orc_coo_read_mm(&m, path) == 0 && m.g.ln_row == m.g.ln_col) exercise(ORC_COO, &m);
    orc_coo_free(&m);
  }
  orc_coo m = {0};
  snprintf(path, sizeof path, "%s/does_not_exist.mtx", dir);
  orc_coo_read_mm(&m, path);
  orc_coo_free(&m);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    printf("usage: sanitize_oracle <tmpdir>\n");
    return 2;
  }
  orc_set_threads(2);
  fills();
  solvers();
  matrix_market(argv[1]);
  printf("sanitize_oracle ok\n");
  return 0;
}
