// test_dropin.cpp -- the reference's cpp tests run through tbsla::hip classes.
//
// Mirrors src/cpp/tests/test_spmv_cdiag.cpp, test_Ax_cdiag.cpp,
// test_a_axpx__cdiag.cpp and test_spmv_cqmat.cpp: every format, generated
// directly and converted from COO, checked against the closed forms of
// src/cpp/utils/array.cpp:17-83 (exit codes 10/12/14 name the failing row
// band, 111 a non-square a_axpx_ case), and cross-format agreement on cqmat.
// It also checks the binary write/read round trip.  Exit 0 on success.
#include <tbsla/hip/Matrix.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <numeric>
#include <sstream>
#include <vector>

using namespace tbsla::hip;

static std::vector<double> spmv_cdiag_closed_form(int nr, int nc, int c, const double *v) {
  std::vector<double> r(nr, 0.0);
  int i;
  for (i = 0; i < std::min(c, nr); i++) r[i] = i < nc - c ? v[i + c] : 0;
  for (; i < std::min(nr, nc - c); i++) r[i] = c == 0 ? v[i] : v[i + c] + v[i - c];
  for (; i < nr; i++) r[i] = i < nc + c ? v[i - c] : 0;
  return r;
}

static int check_bands(int nr, int nc, int c, const double *r, const std::vector<double> &exp) {
  int i;
  for (i = 0; i < std::min(c, nr); i++)
    if (r[i] != exp[i]) return 10;
  for (; i < std::min(nr, nc - c); i++)
    if (r[i] != exp[i]) return 12;
  for (; i < nr; i++)
    if (r[i] != exp[i]) return 14;
  return 0;
}

static int test_matrix_spmv(Matrix &m, int c) {
  int nc = m.get_n_col(), nr = m.get_n_row();
  std::vector<double> v(nc);
  std::iota(v.begin(), v.end(), 0.0);
  std::unique_ptr<double[]> r(m.spmv(v.data()));
  int res = check_bands(nr, nc, c, r.get(), spmv_cdiag_closed_form(nr, nc, c, v.data()));
  if (res) return res;
  // Ax into a zeroed buffer (test_Ax_cdiag.cpp)
  std::vector<double> r2(nr, 0.0);
  m.Ax(r2.data(), v.data());
  return check_bands(nr, nc, c, r2.data(), spmv_cdiag_closed_form(nr, nc, c, v.data()));
}

static int test_matrix_a_axpx(Matrix &m, int c) {
  int nc = m.get_n_col(), nr = m.get_n_row();
  if (nr != nc) return 111;
  std::vector<double> v(nc);
  for (int i = 0; i < nc; i++) v[i] = 2 * i + 1;
  std::unique_ptr<double[]> r(m.a_axpx_(v.data()));
  std::vector<double> t = spmv_cdiag_closed_form(nr, nc, c, v.data());
  for (int i = 0; i < nr; i++) t[i] += v[i];
  return check_bands(nr, nc, c, r.get(), spmv_cdiag_closed_form(nr, nc, c, t.data()));
}

#define RUN(expr, what)                                                              \
  do {                                                                               \
    int rc__ = (expr);                                                               \
    if (rc__) {                                                                      \
      std::fprintf(stderr, "FAIL %s nr=%d nc=%d c=%d rc=%d\n", what, nr, nc, c, rc__); \
      return rc__;                                                                   \
    }                                                                                \
  } while (0)

static int test_cdiag(int nr, int nc, int c, bool axpx) {
  MatrixCOO mcoo;
  mcoo.fill_cdiag(nr, nc, c);
  MatrixSCOO mscoo(mcoo), mscoo2;
  mscoo2.fill_cdiag(nr, nc, c);
  MatrixCSR mcsr(mcoo), mcsr2;
  mcsr2.fill_cdiag(nr, nc, c);
  MatrixELL mell(mcoo), mell2;
  mell2.fill_cdiag(nr, nc, c);
  MatrixDENSE mdense(mcoo), mdense2;
  mdense2.fill_cdiag(nr, nc, c);
  Matrix *all[] = {&mcoo, &mscoo, &mscoo2, &mcsr, &mcsr2, &mell, &mell2, &mdense, &mdense2};
  for (Matrix *m : all) {
    RUN(test_matrix_spmv(*m, c), "spmv_cdiag");
    if (axpx) RUN(test_matrix_a_axpx(*m, c), "a_axpx_cdiag");
  }
  return 0;
}

// test_spmv_cqmat.cpp: every format agrees with COO on cqmat (exactly, since
// values are 1 and the input is iota).
static int test_cqmat(int nr, int nc, int c, double q, unsigned s) {
  std::vector<double> v(nc);
  std::iota(v.begin(), v.end(), 0.0);
  MatrixCOO mcoo;
  mcoo.fill_cqmat(nr, nc, c, q, s);
  std::unique_ptr<double[]> rcoo(mcoo.spmv(v.data()));
  MatrixSCOO mscoo(mcoo), mscoo2;
  mscoo2.fill_cqmat(nr, nc, c, q, s);
  MatrixCSR mcsr(mcoo), mcsr2;
  mcsr2.fill_cqmat(nr, nc, c, q, s);
  MatrixELL mell(mcoo), mell2;
  mell2.fill_cqmat(nr, nc, c, q, s);
  MatrixDENSE mdense(mcoo), mdense2;
  mdense2.fill_cqmat(nr, nc, c, q, s);
  Matrix *all[] = {&mscoo, &mscoo2, &mcsr, &mcsr2, &mell, &mell2, &mdense, &mdense2};
  for (Matrix *m : all) {
    std::unique_ptr<double[]> r(m->spmv(v.data()));
    for (int i = 0; i < nr; i++)
      if (r[i] != rcoo[i]) {
        std::fprintf(stderr, "FAIL cqmat nr=%d nc=%d c=%d q=%g s=%u row %d: %g vs %g\n", nr, nc, c,
                     q, s, i, r[i], rcoo[i]);
        return 1;
      }
  }
  return 0;
}

// write -> read -> identical SpMV and geometry, for every format.
static int test_roundtrip() {
  const int n = 500;
  std::vector<double> v(n);
  for (int i = 0; i < n; i++) v[i] = std::sin(i + 1.0);
  for (int fmt = TBSLA_HIP_COO; fmt <= TBSLA_HIP_DENSE; fmt++) {
    Matrix a(fmt), b(fmt);
    a.fill_cqmat_values(n, n, 9, 0.3, 2, 77, 0, 0, 1, 1);
    std::stringstream ss;
    a.write(ss);
    b.read(ss);
    tbsla_hip_geometry ga = a.geometry(), gb = b.geometry();
    if (ga.n_row != gb.n_row || ga.nnz != gb.nnz || ga.ln_col != gb.ln_col ||
        ga.max_col != gb.max_col) {
      std::fprintf(stderr, "FAIL roundtrip geometry fmt %d\n", fmt);
      return 2;
    }
    std::unique_ptr<double[]> ra(a.spmv(v.data())), rb(b.spmv(v.data()));
    for (int i = 0; i < n; i++)
      if (ra[i] != rb[i]) {
        std::fprintf(stderr, "FAIL roundtrip spmv fmt %d row %d\n", fmt, i);
        return 3;
      }
  }
  return 0;
}

// Row operations on cdiag blocks (values 1): row sums count the entries of
// each row (closed form); after normalize_rows every nonempty row sums to 1;
// CSR's max_abs is 1 on nonempty rows, exp(z - 1) maps every entry to 1 and
// pow(2, z) to 2.
static int test_row_ops(int nr, int nc, int c) {
  std::vector<double> cnt(nr, 0.0);
  for (int i = 0; i < nr; i++) {
    if (i - c >= 0 && i - c < nc) cnt[i] += 1;
    if (c > 0 && i + c < nc) cnt[i] += 1;
  }
  std::vector<double> ones(nc, 1.0);
  for (int fmt = TBSLA_HIP_COO; fmt <= TBSLA_HIP_DENSE; fmt++) {
    Matrix m(fmt);
    m.fill_cdiag(nr, nc, c);
    std::vector<double> s(nr, 0.0);
    m.get_row_sums(s.data());
    for (int i = 0; i < nr; i++)
      if (s[i] != (fmt == TBSLA_HIP_COO ? 0.0 : cnt[i])) return 20 + fmt;
    if (fmt == TBSLA_HIP_COO) continue;
    m.normalize_rows(s.data());
    if (fmt == TBSLA_HIP_CSR) {
      std::unique_ptr<double[]> r(m.spmv(ones.data()));
      for (int i = 0; i < nr; i++)
        if (r[i] != (cnt[i] ? 1.0 : 0.0)) return 30;
      m.fill_cdiag(nr, nc, c);
      std::vector<double> mx(nr, -1.0);
      m.get_row_max_abs(mx.data());
      for (int i = 0; i < nr; i++)
        if (mx[i] != (cnt[i] ? 1.0 : 0.0)) return 31;
      m.apply_exponential(mx.data(), -1);
      std::unique_ptr<double[]> r1(m.spmv(ones.data()));
      for (int i = 0; i < nr; i++)
        if (r1[i] != cnt[i]) return 32;
      m.apply_exponential(nullptr, 2);  // 2^1 = 2
      std::unique_ptr<double[]> r2(m.spmv(ones.data()));
      for (int i = 0; i < nr; i++)
        if (r2[i] != 2 * cnt[i]) return 33;
    }
  }
  return 0;
}

// print_as_dense, NUMAinit and the generator stubs.
static int test_misc_virtuals() {
  MatrixCSR m;
  m.fill_cdiag(3, 3, 1);
  std::ostringstream os;
  m.print_as_dense(os);
  if (os.str() != " 0.0   1.0   0.0  \n 1.0   0.0   1.0  \n 0.0   1.0   0.0  \n") {
    std::fprintf(stderr, "print_as_dense gave:\n%s", os.str().c_str());
    return 40;
  }
  MatrixCOO coo;
  coo.fill_cdiag(3, 3, 1);
  std::ostringstream os2;
  coo.print_as_dense(os2);
  if (os2.str() != os.str()) return 41;
  m.NUMAinit();
  std::vector<double> v = {1.0, 2.0, 3.0};
  std::unique_ptr<double[]> r(m.spmv(v.data()));
  if (r[0] != 2.0 || r[1] != 4.0 || r[2] != 2.0) return 42;
  try {
    m.fill_random(10, 10, 0.1);
    return 43;
  } catch (const std::runtime_error &) {
  }
  return 0;
}

// mpi::MatrixCOO overrides page_rank with the max-normalised iteration
// (src/mpi/matrix/MatrixCOO.cpp:108-175): a call through the base class on a
// COO object gives that iteration; another format keeps Matrix::page_rank.
int test_page_rank_dispatch() {
  MatrixCOO coo;
  coo.fill_cqmat(60, 60, 4, 0.5, 3);
  MatrixCSR csr(coo);
  Matrix *base = &coo;
  int it_coo = 0, it_max = 0, it_plain = 0;
  std::unique_ptr<double[]> a(base->page_rank(0.85, 1e-12, 500, it_coo));
  std::unique_ptr<double[]> b(csr.page_rank_max_norm(0.85, 1e-12, 500, it_max));
  std::unique_ptr<double[]> c(static_cast<Matrix &>(csr).page_rank(0.85, 1e-12, 500, it_plain));
  if (it_coo != it_max) return 50;
  double d_same = 0, d_other = 0;
  for (int i = 0; i < 60; i++) {
    d_same = std::max(d_same, std::fabs(a[i] - b[i]));
    d_other = std::max(d_other, std::fabs(a[i] - c[i]));
  }
  if (d_same > 1e-13) return 51;
  if (d_other < 1e-6) return 52;
  return 0;
}

int main() {
  int nr, nc, c;
  for (int i = 0; i <= 12; i++) {
    nr = 10, nc = 10, c = i;
    RUN(test_cdiag(nr, nc, c, true), "cdiag");
    nr = 5, nc = 10;
    RUN(test_cdiag(nr, nc, c, false), "cdiag");
    nr = 10, nc = 5;
    RUN(test_cdiag(nr, nc, c, false), "cdiag");
    nr = 30, nc = 30, c = 2 * i;
    RUN(test_cdiag(nr, nc, c, true), "cdiag");
    nr = 20, nc = 30;
    RUN(test_cdiag(nr, nc, c, false), "cdiag");
    nr = 30, nc = 20;
    RUN(test_cdiag(nr, nc, c, false), "cdiag");
  }
  for (int i = 0; i <= 12; i += 3)
    for (unsigned s = 0; s < 4; s++)
      for (int qi = 0; qi <= 10; qi += 5) {
        double q = qi / 10.0;
        nr = 10, nc = 10, c = i;
        RUN(test_cqmat(10, 10, i, q, s), "cqmat");
        RUN(test_cqmat(5, 10, i, q, s), "cqmat");
        RUN(test_cqmat(10, 5, i, q, s), "cqmat");
        RUN(test_cqmat(30, 30, 2 * i, q, s), "cqmat");
        RUN(test_cqmat(20, 30, 2 * i, q, s), "cqmat");
        RUN(test_cqmat(30, 20, 2 * i, q, s), "cqmat");
      }
  for (int i = 0; i <= 12; i += 4) {
    nr = 30, nc = 30, c = i;
    RUN(test_row_ops(nr, nc, c), "row_ops");
    nr = 20, nc = 30;
    RUN(test_row_ops(nr, nc, c), "row_ops");
    nr = 30, nc = 20;
    RUN(test_row_ops(nr, nc, c), "row_ops");
  }
  nr = nc = c = 0;
  RUN(test_misc_virtuals(), "misc_virtuals");
  RUN(test_page_rank_dispatch(), "page_rank_dispatch");
  RUN(test_roundtrip(), "roundtrip");
  std::printf("=== finished without error ===\n");
  return 0;
}
