// test_cpp_subclass.cpp -- tbsla::hip::CppMatrix as a tbsla::cpp::Matrix.
//
// TEST PROGRAM.  Built by oracle/Makefile (oracle/_ref/test_cpp_subclass)
// against the reference's headers and its compiled src/cpp library, where
// those sources are present; run by tests/test_gpu_cpp.py on a GPU.
//
// A GPU block and the reference's CPU block of the same format are driven
// through the same `tbsla::cpp::Matrix &`: the virtual fills, spmv and Ax,
// the reference's own non-virtual a_axpx_ / AAxpAx / AAxpAxpx /
// make_stochastic (src/cpp/matrix/Matrix.cpp:5-38, which call the
// virtuals), the base class's getters, write and read.  Values are 1.0 and
// x holds small integers, so every product and sum is exact and results
// must be equal -- until make_stochastic divides by column sums; from there
// on a SpMV is compared to 1e-12 relative (the GPU sums a row in another
// order).  Prints "test_cpp_subclass ok".
#include <tbsla/cpp/MatrixCOO.hpp>
#include <tbsla/cpp/MatrixCSR.hpp>
#include <tbsla/cpp/MatrixDENSE.hpp>
#include <tbsla/cpp/MatrixELL.hpp>
#include <tbsla/cpp/MatrixSCOO.hpp>
#include <tbsla/hip/CppMatrix.hpp>

#include <omp.h>

#include <cstdio>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace {

int failures = 0;
void expect(bool ok, const std::string &what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what.c_str());
    failures++;
  }
}

// tbsla::cpp::Matrix has no virtual destructor and the reference's classes
// inherit it virtually (MatrixCSR.hpp:10), so an object must be deleted
// through its own type, never through a tbsla::cpp::Matrix pointer.
void destroy(tbsla::cpp::Matrix *m) {
  if (auto *p = dynamic_cast<tbsla::cpp::MatrixCOO *>(m)) delete p;
  else if (auto *p = dynamic_cast<tbsla::cpp::MatrixSCOO *>(m)) delete p;
  else if (auto *p = dynamic_cast<tbsla::cpp::MatrixCSR *>(m)) delete p;
  else if (auto *p = dynamic_cast<tbsla::cpp::MatrixELL *>(m)) delete p;
  else if (auto *p = dynamic_cast<tbsla::cpp::MatrixDENSE *>(m)) delete p;
}
typedef std::unique_ptr<tbsla::cpp::Matrix, void (*)(tbsla::cpp::Matrix *)> RefPtr;

// |a[i] - b[i]| <= 1e-12 |b[i]| for every i (NaN fails)
bool close12(const double *a, const double *b, int n) {
  for (int i = 0; i < n; i++) {
    const double d = a[i] - b[i], t = 1e-12 * (b[i] < 0 ? -b[i] : b[i]) + 1e-300;
    if (!(d <= t && -d <= t)) return false;
  }
  return true;
}

RefPtr reference(int fmt) {
  switch (fmt) {
    case TBSLA_HIP_COO: return RefPtr(new tbsla::cpp::MatrixCOO(), destroy);
    case TBSLA_HIP_SCOO: return RefPtr(new tbsla::cpp::MatrixSCOO(), destroy);
    case TBSLA_HIP_CSR: return RefPtr(new tbsla::cpp::MatrixCSR(), destroy);
    case TBSLA_HIP_ELL: return RefPtr(new tbsla::cpp::MatrixELL(), destroy);
    default: return RefPtr(new tbsla::cpp::MatrixDENSE(), destroy);
  }
}

bool same(const double *a, const double *b, int n) {
  for (int i = 0; i < n; i++)
    if (a[i] != b[i]) return false;
  return true;
}

// the base-class interface only, on both objects
void drive(int fmt, tbsla::cpp::Matrix &G, tbsla::cpp::Matrix &R, const std::string &name) {
  expect(G.get_n_row() == R.get_n_row() && G.get_n_col() == R.get_n_col() &&
             G.get_f_row() == R.get_f_row() && G.get_f_col() == R.get_f_col() &&
             G.get_ln_row() == R.get_ln_row() && G.get_ln_col() == R.get_ln_col() &&
             G.get_pr() == R.get_pr() && G.get_pc() == R.get_pc() && G.get_NR() == R.get_NR() &&
             G.get_NC() == R.get_NC() && G.get_nnz() == R.get_nnz(),
         name + ": getters");
  const bool coo = fmt == TBSLA_HIP_COO;
  const int nin = coo ? R.get_n_col() : R.get_ln_col();
  const int nout = coo ? R.get_n_row() : R.get_ln_row();
  std::vector<double> x(nin > 0 ? nin : 1);
  for (int i = 0; i < nin; i++) x[i] = (double)(i % 7 + 1);
  {
    std::unique_ptr<double[]> a(G.spmv(x.data())), b(R.spmv(x.data()));
    expect(same(a.get(), b.get(), nout), name + ": spmv");
  }
  if (!coo && R.get_ln_row() == R.get_ln_col()) {
    std::unique_ptr<double[]> a(G.a_axpx_(x.data())), b(R.a_axpx_(x.data()));
    expect(same(a.get(), b.get(), nout), name + ": a_axpx_ (reference Matrix.cpp:5-14)");
    std::vector<double> r1(nout), r2(nout), b1(nout), b2(nout);
    G.AAxpAxpx(r1.data(), x.data(), b1.data());
    R.AAxpAxpx(r2.data(), x.data(), b2.data());
    expect(same(r1.data(), r2.data(), nout), name + ": AAxpAxpx (Matrix.cpp:26-31)");
  }
  // make_stochastic (Matrix.cpp:33-38): get_col_sums + normalize_cols, virtual
  std::vector<double> s1(R.get_ln_col() + 1, 0.0), s2(R.get_ln_col() + 1, 0.0);
  G.make_stochastic(s1.data());
  R.make_stochastic(s2.data());
  expect(same(s1.data(), s2.data(), R.get_ln_col()), name + ": make_stochastic column sums");
  {
    std::unique_ptr<double[]> a(G.spmv(x.data())), b(R.spmv(x.data()));
    expect(close12(a.get(), b.get(), nout), name + ": spmv after make_stochastic (1e-12)");
  }
  // the reference's image read by the GPU block gives the same SpMV
  std::ostringstream os;
  R.write(os);
  std::istringstream is(os.str());
  tbsla::hip::CppMatrix H(fmt);
  tbsla::cpp::Matrix &Hb = H;
  Hb.read(is);
  expect(Hb.get_nnz() == R.get_nnz() && Hb.get_ln_row() == R.get_ln_row(), name + ": read getters");
  {
    std::unique_ptr<double[]> a(Hb.spmv(x.data())), b(R.spmv(x.data()));
    expect(close12(a.get(), b.get(), nout), name + ": spmv of the reference's image read on the GPU (1e-12)");
  }
}

}  // namespace

int main() {
  omp_set_num_threads(1);  // the reference's CSR column sums race under OpenMP
  const int fmts[] = {TBSLA_HIP_CSR, TBSLA_HIP_ELL, TBSLA_HIP_COO, TBSLA_HIP_SCOO, TBSLA_HIP_DENSE};
  const char *names[] = {"COO", "SCOO", "CSR", "ELL", "DENSE"};
  struct Case {
    int n, c;
    double q;
    unsigned s;
    int pr, pc, NR, NC;
  };
  const Case cases[] = {{300, 8, 0.0, 0, 0, 0, 1, 1}, {300, 8, 0.2, 1, 0, 0, 1, 1},
                        {400, 12, 0.1, 2, 1, 1, 2, 2}, {400, 12, 0.1, 2, 0, 1, 2, 2},
                        {333, 5, 0.5, 3, 2, 2, 3, 3}};
  for (int f : fmts)
    for (const Case &c : cases) {
      const std::string name = std::string(names[f]) + " n=" + std::to_string(c.n) + " block (" +
                               std::to_string(c.pr) + "," + std::to_string(c.pc) + ") of " +
                               std::to_string(c.NR) + "x" + std::to_string(c.NC);
      tbsla::hip::CppMatrix gpu(f);
      RefPtr ref = reference(f);
      tbsla::cpp::Matrix &G = gpu;
      G.fill_cqmat(c.n, c.n, c.c, c.q, c.s, c.pr, c.pc, c.NR, c.NC);
      ref->fill_cqmat(c.n, c.n, c.c, c.q, c.s, c.pr, c.pc, c.NR, c.NC);
      drive(f, G, *ref, name);
      tbsla::hip::CppMatrix gd(f);
      RefPtr rd = reference(f);
      static_cast<tbsla::cpp::Matrix &>(gd).fill_cdiag(c.n, c.n, 3, c.pr, c.pc, c.NR, c.NC);
      rd->fill_cdiag(c.n, c.n, 3, c.pr, c.pc, c.NR, c.NC);
      drive(f, gd, *rd, name + " cdiag 3");
    }
  if (failures) {
    std::printf("%d checks failed\n", failures);
    return 1;
  }
  std::printf("test_cpp_subclass ok\n");
  return 0;
}
