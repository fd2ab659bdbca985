/*
 * problems.c -- CPU restatement of the SPD problem generators and of the
 * preconditioned CG (test infrastructure, part of liboracle.so).
 *
 * These have no routine in the reference library: the reference's CG driver
 * (src/mpi/matrix/CG.cpp:32-73) is run here on SPD systems built from its
 * own cqmat generator (MatrixCSR.cpp:373-486) or from a 7-point stencil.
 * The loops fix the order of every floating-point operation; the GPU
 * generators follow the same order (libm's log/cos/exp in the noise field
 * aside).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "tbsla_oracle.h"

static void *zalloc2(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) abort();
  return p;
}

static void square_geom(orc_geom *g, int n, long nnz) {
  memset(g, 0, sizeof *g);
  g->n_row = g->n_col = g->ln_row = g->ln_col = n;
  g->NR = g->NC = 1;
  g->nnz = nnz;
}

static long cell3(int nx, int ny, int x, int y, int z) {
  return (long)x + (long)nx * ((long)y + (long)ny * z);
}

static double face_coef(const double *k, long a, long b) {
  if (!k) return 1.0;
  return 2.0 * k[a] * k[b] / (k[a] + k[b]);
}

int orc_poisson3d(orc_csr *m, int nx, int ny, int nz, const double *kappa) {
  const long n = (long)nx * ny * nz;
  memset(m, 0, sizeof *m);
  int *rp = (int *)zalloc2((size_t)n + 1, sizeof(int));
  for (long i = 0; i < n; i++) {
    int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
    rp[i + 1] = rp[i] + 1 + (x > 0) + (x < nx - 1) + (y > 0) + (y < ny - 1) + (z > 0) + (z < nz - 1);
  }
  long nnz = rp[n];
  square_geom(&m->g, (int)n, nnz);
  m->stored = nnz;
  m->rowptr = rp;
  m->colidx = (int *)zalloc2((size_t)nnz, sizeof(int));
  m->values = (double *)zalloc2((size_t)nnz, sizeof(double));
  for (long i = 0; i < n; i++) {
    int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
    double kc = kappa ? kappa[i] : 1.0;
    int in[6] = {z > 0, y > 0, x > 0, x < nx - 1, y < ny - 1, z < nz - 1};
    long nb[6] = {in[0] ? cell3(nx, ny, x, y, z - 1) : -1, in[1] ? cell3(nx, ny, x, y - 1, z) : -1,
                  in[2] ? cell3(nx, ny, x - 1, y, z) : -1, in[3] ? cell3(nx, ny, x + 1, y, z) : -1,
                  in[4] ? cell3(nx, ny, x, y + 1, z) : -1, in[5] ? cell3(nx, ny, x, y, z + 1) : -1};
    double f[6], d = 0.0;
    for (int t = 0; t < 6; t++) {
      f[t] = in[t] ? face_coef(kappa, i, nb[t]) : kc;
      d = d + f[t];
    }
    int pos = rp[i];
    for (int t = 0; t < 3; t++)
      if (in[t]) {
        m->colidx[pos] = (int)nb[t];
        m->values[pos++] = -f[t];
      }
    m->colidx[pos] = (int)i;
    m->values[pos++] = d;
    for (int t = 3; t < 6; t++)
      if (in[t]) {
        m->colidx[pos] = (int)nb[t];
        m->values[pos++] = -f[t];
      }
  }
  return 0;
}

void orc_uniform_vector(double *out, long n, uint64_t seed, double lo, double hi) {
  const double w = hi - lo;
  for (long i = 0; i < n; i++) out[i] = lo + w * orc_unit_value(seed, (uint64_t)i);
}

static int reflect_idx(long j, int n) {
  long p = 2L * n;
  long m = ((j % p) + p) % p;
  if (m >= n) m = p - 1 - m;
  return (int)m;
}

void orc_lognormal_field(double *out, int nx, int ny, int nz, double mu, double sigma,
                         double corr_len, uint64_t seed) {
  const long n = (long)nx * ny * nz;
  int R = 0;
  double w[129];
  double scale = 1.0;
  if (corr_len > 0) {
    R = (int)ceil(3.0 * corr_len);
    double sum = 0.0;
    for (int d = -R; d <= R; d++) {
      double u = (double)d / corr_len;
      w[d + R] = exp(-0.5 * (u * u));
      sum = sum + w[d + R];
    }
    double sq = 0.0;
    for (int d = 0; d <= 2 * R; d++) {
      w[d] = w[d] / sum;
      sq = sq + w[d] * w[d];
    }
    scale = 1.0 / sqrt(sq * sq * sq);
  } else {
    w[0] = 1.0;
  }
  double *a = (double *)zalloc2((size_t)n, sizeof(double));
  double *b = (double *)zalloc2((size_t)n, sizeof(double));
  for (long i = 0; i < n; i++) {
    double u1 = orc_unit_value(seed, 2 * (uint64_t)i);
    double u2 = orc_unit_value(seed, 2 * (uint64_t)i + 1);
    double r = sqrt(-2.0 * log(1.0 - u1));
    a[i] = r * cos(6.283185307179586 * u2);
  }
  double *src = a, *dst = b;
  int dims[3] = {nx, ny, nz};
  for (int axis = 0; axis < 3; axis++) {
    for (long i = 0; i < n; i++) {
      int c[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
      double acc = 0.0;
      for (int d = -R; d <= R; d++) {
        int cc[3] = {c[0], c[1], c[2]};
        cc[axis] = reflect_idx((long)c[axis] + d, dims[axis]);
        acc = acc + w[d + R] * src[cell3(nx, ny, cc[0], cc[1], cc[2])];
      }
      dst[i] = acc;
    }
    double *t = src;
    src = dst;
    dst = t;
  }
  for (long i = 0; i < n; i++) out[i] = exp(mu + sigma * (src[i] * scale));
  free(a);
  free(b);
}

/* (row, col, position) triplets for the stable sort of the symmetrisation */
typedef struct {
  long key;
  long pos;
} keypos;

static int cmp_keypos(const void *a, const void *b) {
  const keypos *x = (const keypos *)a, *y = (const keypos *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return (x->pos > y->pos) - (x->pos < y->pos);
}

int orc_sym_cqmat(orc_csr *A, const orc_csr *B) {
  const long n = B->g.ln_row, m = B->stored;
  keypos *e = (keypos *)zalloc2((size_t)(2 * m), sizeof(keypos));
  long k = 0;
  for (long r = 0; r < n; r++)
    for (int j = B->rowptr[r]; j < B->rowptr[r + 1]; j++) {
      long c = B->colidx[j];
      if (c == r) continue;
      e[k].key = r * n + c;
      e[k++].pos = j;
    }
  for (long r = 0; r < n; r++)
    for (int j = B->rowptr[r]; j < B->rowptr[r + 1]; j++) {
      long c = B->colidx[j];
      if (c == r) continue;
      e[k].key = c * n + r;
      e[k++].pos = m + j;
    }
  qsort(e, (size_t)k, sizeof(keypos), cmp_keypos);
  /* unique keys per row */
  int *rp = (int *)zalloc2((size_t)n + 1, sizeof(int));
  for (long t = 0; t < k; t++)
    if (t == 0 || e[t].key != e[t - 1].key) rp[e[t].key / n + 1]++;
  for (long r = 0; r < n; r++) rp[r + 1] += rp[r] + 1; /* + diagonal */
  long nnz = rp[n];
  memset(A, 0, sizeof *A);
  square_geom(&A->g, (int)n, nnz);
  A->stored = nnz;
  A->rowptr = rp;
  A->colidx = (int *)zalloc2((size_t)nnz, sizeof(int));
  A->values = (double *)zalloc2((size_t)nnz, sizeof(double));
  long t = 0;
  for (long r = 0; r < n; r++) {
    int pos = rp[r], diag_pos = -1;
    double abs_sum = 0.0;
    while (t < k && e[t].key / n == r) {
      long key = e[t].key;
      int c = (int)(key % n);
      double a = 0.0, b = 0.0;
      for (; t < k && e[t].key == key; t++) {
        if (e[t].pos < m)
          a = a + B->values[e[t].pos];
        else
          b = b + B->values[e[t].pos - m];
      }
      if (diag_pos < 0 && c > r) diag_pos = pos++;
      double v = (a + b) * 0.5;
      A->colidx[pos] = c;
      A->values[pos++] = v;
      abs_sum = abs_sum + fabs(v);
    }
    if (diag_pos < 0) diag_pos = pos;
    A->colidx[diag_pos] = (int)r;
    A->values[diag_pos] = abs_sum + 1.0;
  }
  free(e);
  return 0;
}

int orc_csr_transpose(orc_csr *T, const orc_csr *A) {
  const long nr = A->g.ln_row, nc = A->g.ln_col, m = A->stored;
  memset(T, 0, sizeof *T);
  T->g = A->g;
  T->g.n_row = A->g.n_col;
  T->g.n_col = A->g.n_row;
  T->g.ln_row = A->g.ln_col;
  T->g.ln_col = A->g.ln_row;
  T->g.f_row = A->g.f_col;
  T->g.f_col = A->g.f_row;
  T->g.pr = A->g.pc;
  T->g.pc = A->g.pr;
  T->g.NR = A->g.NC;
  T->g.NC = A->g.NR;
  T->g.nnz = m;
  T->stored = m;
  T->rowptr = (int *)zalloc2((size_t)nc + 1, sizeof(int));
  T->colidx = (int *)zalloc2((size_t)m, sizeof(int));
  T->values = (double *)zalloc2((size_t)m, sizeof(double));
  for (long j = 0; j < m; j++) T->rowptr[A->colidx[j] - A->g.f_col + 1]++;
  for (long c = 0; c < nc; c++) T->rowptr[c + 1] += T->rowptr[c];
  int *fill = (int *)zalloc2((size_t)nc, sizeof(int));
  for (long r = 0; r < nr; r++)
    for (int j = A->rowptr[r]; j < A->rowptr[r + 1]; j++) {
      long c = A->colidx[j] - A->g.f_col;
      long p = T->rowptr[c] + fill[c]++;
      T->colidx[p] = (int)(r + A->g.f_row);
      T->values[p] = A->values[j];
    }
  free(fill);
  return 0;
}

static double dotv(const double *a, const double *b, long n) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s = s + a[i] * b[i];
  return s;
}

static void csr_mv(const orc_csr *m, const double *x, double *y) {
  for (long i = 0; i < m->g.ln_row; i++) {
    double s = 0.0;
    for (int j = m->rowptr[i]; j < m->rowptr[i + 1]; j++)
      s = s + m->values[j] * x[m->colidx[j] - m->g.f_col];
    y[i] = s;
  }
}

/* z = M r: M = I, D^-1 (diagonal summed over duplicates; 1 where it is 0)
 * or G G^T (t = G^T r, then z = G t). */
static void apply_prec(int kind, const orc_csr *A, const orc_csr *G, const orc_csr *GT,
                       const double *dinv, const double *r, double *z, double *t, long n) {
  (void)A;
  if (kind == 0) {
    memcpy(z, r, sizeof(double) * (size_t)n);
  } else if (kind == 1) {
    for (long i = 0; i < n; i++) z[i] = r[i] * dinv[i];
  } else {
    csr_mv(GT, r, t);
    csr_mv(G, t, z);
  }
}

int orc_pcg(const orc_csr *A, const double *rhs, double *x, int max_iterations, double tolerance,
            int relative, double near_zero, int kind, const orc_csr *G, double *res_norm_out) {
  const long n = A->g.ln_row;
  double *R = (double *)zalloc2((size_t)n, sizeof(double));
  double *Z = (double *)zalloc2((size_t)n, sizeof(double));
  double *P = (double *)zalloc2((size_t)n, sizeof(double));
  double *AP = (double *)zalloc2((size_t)n, sizeof(double));
  double *T = (double *)zalloc2((size_t)n, sizeof(double));
  double *dinv = (double *)zalloc2((size_t)n, sizeof(double));
  orc_csr GT;
  memset(&GT, 0, sizeof GT);
  if (kind == 2) orc_csr_transpose(&GT, G);
  if (kind == 1)
    for (long i = 0; i < n; i++) {
      double d = 0.0;
      for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++)
        if (A->colidx[j] - A->g.f_col == i) d = d + A->values[j];
      dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
  memset(x, 0, sizeof(double) * (size_t)n);
  memcpy(R, rhs, sizeof(double) * (size_t)n);
  apply_prec(kind, A, G, &GT, dinv, R, Z, T, n);
  memcpy(P, Z, sizeof(double) * (size_t)n);
  double stop = tolerance;
  if (relative) stop = tolerance * sqrt(dotv(rhs, rhs, n));
  double rz = dotv(R, Z, n), rnorm = sqrt(dotv(R, R, n));
  int it = 0;
  while (it < max_iterations) {
    csr_mv(A, P, AP);
    double alpha = rz / fmax(dotv(P, AP, n), near_zero);
    for (long i = 0; i < n; i++) x[i] = x[i] + alpha * P[i];
    for (long i = 0; i < n; i++) R[i] = R[i] + (-alpha) * AP[i];
    apply_prec(kind, A, G, &GT, dinv, R, Z, T, n);
    double rr = dotv(R, R, n), rz_new = dotv(R, Z, n);
    rnorm = sqrt(rr);
    if (rnorm < stop) break;
    double beta = rz_new / fmax(rz, near_zero);
    for (long i = 0; i < n; i++) P[i] = Z[i] + beta * P[i];
    rz = rz_new;
    it++;
  }
  if (res_norm_out) *res_norm_out = rnorm;
  if (kind == 2) orc_csr_free(&GT);
  free(R);
  free(Z);
  free(P);
  free(AP);
  free(T);
  free(dinv);
  return it;
}
