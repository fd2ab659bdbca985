/*
 * tbsla_oracle.h -- CPU restatement of TBSLA's sparse hot path (test infrastructure).
 *
 * This header belongs to the oracle, not to the product: only tests/, the
 * smoke check and the CPU-baseline leg of bench.py load liboracle.so.  The
 * HIP library never links it.
 *
 * Every function restates one routine of the reference C++ library
 * (src/cpp/... in masudaaaaaaa/TBSLA); the citation sits next to the
 * declaration.  Types follow the reference: indices int32, values double,
 * nnz a long.
 */
#ifndef TBSLA_ORACLE_H
#define TBSLA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Block geometry, the protected fields of tbsla::cpp::Matrix
 * (src/cpp/include/tbsla/cpp/Matrix.hpp:54-55). */
typedef struct {
  int n_row, n_col, f_row, f_col, ln_row, ln_col, pr, pc, NR, NC;
  long nnz;
} orc_geom;

/* CSR block (src/cpp/include/tbsla/cpp/MatrixCSR.hpp).  `stored` is
 * rowptr[ln_row], the number of entries actually written; `nnz` in the
 * geometry is the reference's count-pass value, which can be larger. */
typedef struct {
  orc_geom g;
  long stored;
  int *rowptr;   /* ln_row + 1 */
  int *colidx;   /* stored */
  double *values;/* stored */
} orc_csr;

/* ELL block, row-major as in the reference (MatrixELL.hpp):
 * values[i * max_col + j]. */
typedef struct {
  orc_geom g;
  int max_col;
  int *columns;  /* ln_row * max_col */
  double *values;
} orc_ell;

/* COO and SCOO triplets (MatrixCOO.hpp / MatrixSCOO.hpp). */
typedef struct {
  orc_geom g;
  long len;      /* number of triplets held (= nnz after a fill) */
  int *row, *col;
  double *values;
} orc_coo;

/* DENSE block, row-major ln_row x ln_col (MatrixDENSE.hpp). */
typedef struct {
  orc_geom g;
  double *values;
} orc_dense;

enum { ORC_COO = 0, ORC_SCOO = 1, ORC_CSR = 2, ORC_ELL = 3, ORC_DENSE = 4 };

/* ---- generators and partition helpers ---------------------------------- */
int orc_rand_r(unsigned *seed);                                   /* glibc stdlib/rand_r.c */
void orc_cqmat_value(uint64_t k, uint64_t nr, uint64_t nc, uint64_t c, double q,
                     unsigned seed_mult, int64_t *row, int64_t *col);
void orc_cdiag_value(uint64_t i, uint64_t nv, uint64_t nr, uint64_t nc, uint64_t cdiag,
                     int64_t *row, int64_t *col);
long orc_lnv(long size, int pos, int npos);
long orc_pflv(long size, int pos, int npos);
double orc_unit_value(uint64_t seed, uint64_t k);                 /* project value generator */
long orc_cqmat_gnv(int n_row, int n_col, int c);
void orc_cqmat_prescan(int n_row, int n_col, int c, int f_row, int ln_row, long *incr, long *nv);

/* ---- fills (value_seed < 0: every value is 1.0, as the reference) -------- */
int orc_csr_fill_cqmat(orc_csr *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                       int pr, int pc, int NR, int NC, int64_t value_seed);
int orc_ell_fill_cqmat(orc_ell *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                       int pr, int pc, int NR, int NC, int64_t value_seed);
int orc_coo_fill_cqmat(orc_coo *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                       int pr, int pc, int NR, int NC, int64_t value_seed);
int orc_scoo_fill_cqmat(orc_coo *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                        int pr, int pc, int NR, int NC, int64_t value_seed);
int orc_dense_fill_cqmat(orc_dense *m, int n_row, int n_col, int c, double q, unsigned seed_mult,
                         int pr, int pc, int NR, int NC, int64_t value_seed);

int orc_csr_fill_cdiag(orc_csr *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC);
int orc_ell_fill_cdiag(orc_ell *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC);
int orc_coo_fill_cdiag(orc_coo *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC);
int orc_scoo_fill_cdiag(orc_coo *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC);
int orc_dense_fill_cdiag(orc_dense *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC);

/* ---- conversions from COO ------------------------------------------------ */
int orc_csr_from_coo(orc_csr *dst, const orc_coo *src);
int orc_ell_from_coo(orc_ell *dst, const orc_coo *src);
int orc_scoo_from_coo(orc_coo *dst, const orc_coo *src);
int orc_dense_from_coo(orc_dense *dst, const orc_coo *src);

/* MatrixCOO::readMM (MatrixCOO.cpp:206-278), symmetric overwrite fixed;
 * 0, or -1 (cannot open), -2 (header rejected), -3 (malformed) */
int orc_coo_read_mm(orc_coo *m, const char *path);

int orc_csr_alloc(orc_csr *m, int n_row, int n_col, long nnz);
int orc_coo_alloc(orc_coo *m, int n_row, int n_col, long nnz);
void orc_csr_free(orc_csr *m);
void orc_ell_free(orc_ell *m);
void orc_coo_free(orc_coo *m);
void orc_dense_free(orc_dense *m);

/* ---- SpMV and composites ------------------------------------------------- */
void orc_Ax(int fmt, const void *m, double *r, const double *v, int vect_incr);
int orc_result_len(int fmt, const void *m);
void orc_spmv(int fmt, const void *m, double *r, const double *v, int vect_incr);
void orc_a_axpx_(int fmt, const void *m, double *r, const double *v, int vect_incr);
void orc_AAxpAx(int fmt, const void *m, double *r, const double *v, double *buf, int vect_incr);
void orc_AAxpAxpx(int fmt, const void *m, double *r, const double *v, double *buf, int vect_incr);
void orc_set_threads(int n);
void orc_csr_Ax_omp(const orc_csr *m, double *r, const double *v);

/* ---- column normalisation, PageRank, CG, SpMM ---------------------------- */
void orc_get_col_sums(int fmt, const void *m, double *s);
void orc_normalize_cols(int fmt, void *m, const double *s);
void orc_make_stochastic(int fmt, void *m, double *s);
/* row operations (Matrix.hpp:46-51) */
void orc_get_row_sums(int fmt, const void *m, double *s);
void orc_normalize_rows(int fmt, void *m, const double *s);
void orc_get_row_max_abs(int fmt, const void *m, double *max_abs);
void orc_apply_exponential(int fmt, void *m, const double *max_abs, int base);
int orc_page_rank(int fmt, const void *m, double beta, double epsilon, int max_iterations,
                  double *b, double *err_out);
int orc_page_rank_max_norm(int fmt, const void *m, double beta, double epsilon,
                           int max_iterations, double *b, double *err_out);
int orc_page_rank_source(int fmt, const void *m, double beta, double epsilon, int max_iterations,
                         int source, double *b, double *err_out);
int orc_cg(int fmt, const void *m, const double *rhs, double *x, int max_iterations,
           double tolerance, int relative, double *res_norm_out);
int orc_cg_clamp(int fmt, const void *m, const double *rhs, double *x, int max_iterations,
                 double tolerance, int relative, double near_zero, double *res_norm_out);
/* problems.c: SPD generators and preconditioned CG (no reference routine;
 * the GPU's tbsla_hip_fill_poisson3d / _lognormal_field / _sym_cqmat /
 * _transpose / _uniform_vector and the PCG of tbsla_hip_cg). */
int orc_poisson3d(orc_csr *m, int nx, int ny, int nz, const double *kappa);
void orc_uniform_vector(double *out, long n, uint64_t seed, double lo, double hi);
void orc_lognormal_field(double *out, int nx, int ny, int nz, double mu, double sigma,
                         double corr_len, uint64_t seed);
int orc_sym_cqmat(orc_csr *A, const orc_csr *B);
int orc_csr_transpose(orc_csr *T, const orc_csr *A);
/* kind 0: none, 1: Jacobi, 2: M = G G^T; x starts at 0 */
int orc_pcg(const orc_csr *A, const double *rhs, double *x, int max_iterations, double tolerance,
            int relative, double near_zero, int kind, const orc_csr *G, double *res_norm_out);
void orc_csr_spmm(const orc_csr *m, const double *B, double *C, int k);

#ifdef __cplusplus
}
#endif
#endif
