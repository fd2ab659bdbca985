// ref_shim.cpp -- C entry points over the compiled reference library (oracle only).
//
// TEST INFRASTRUCTURE.  Built by oracle/Makefile together with the reference's
// own src/cpp sources (compiled where they lie, outputs in oracle/_ref/), and
// loaded only by tests/ and tests/golden/make_golden.py to pin the oracle
// against the real reference.  The product never loads it.
//
// The reference classes keep their arrays protected and have no getters
// (src/cpp/include/tbsla/cpp/MatrixCSR.hpp:41-44 etc.); `Peek` reads them
// through pointer-to-member access from a derived scope.
#include <tbsla/cpp/Matrix.hpp>
#include <tbsla/cpp/MatrixCOO.hpp>
#include <tbsla/cpp/MatrixCSR.hpp>
#include <tbsla/cpp/MatrixDENSE.hpp>
#include <tbsla/cpp/MatrixELL.hpp>
#include <tbsla/cpp/MatrixSCOO.hpp>
#include <tbsla/cpp/utils/values_generation.hpp>

#include <omp.h>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>

using namespace tbsla::cpp;

namespace {

enum { F_COO = 0, F_SCOO = 1, F_CSR = 2, F_ELL = 3, F_DENSE = 4 };

struct PeekBase : Matrix {
  static int *geom(Matrix &m, int i) {
    int Matrix::*f[10] = {&PeekBase::n_row, &PeekBase::n_col, &PeekBase::f_row,
                          &PeekBase::f_col, &PeekBase::ln_row, &PeekBase::ln_col,
                          &PeekBase::pr,    &PeekBase::pc,     &PeekBase::NR,
                          &PeekBase::NC};
    return &(m.*f[i]);
  }
  static long *nnz_of(Matrix &m) { return &(m.*(&PeekBase::nnz)); }
};
struct PeekCSR : MatrixCSR {
  static double *vals(MatrixCSR &m) { return m.*(&PeekCSR::values); }
  static int *rowptr_of(MatrixCSR &m) { return m.*(&PeekCSR::rowptr); }
  static int *colidx_of(MatrixCSR &m) { return m.*(&PeekCSR::colidx); }
};
struct PeekELL : MatrixELL {
  static double *vals(MatrixELL &m) { return m.*(&PeekELL::values); }
  static int *cols(MatrixELL &m) { return m.*(&PeekELL::columns); }
  static int maxc(MatrixELL &m) { return m.*(&PeekELL::max_col); }
};
struct PeekSCOO : MatrixSCOO {
  static double *vals(MatrixSCOO &m) { return m.*(&PeekSCOO::values); }
  static int *rows(MatrixSCOO &m) { return m.*(&PeekSCOO::row); }
  static int *cols(MatrixSCOO &m) { return m.*(&PeekSCOO::col); }
};
struct PeekDENSE : MatrixDENSE {
  static double *vals(MatrixDENSE &m) { return m.*(&PeekDENSE::values); }
};

Matrix *make(int fmt) {
  switch (fmt) {
    case F_COO: return new MatrixCOO();
    case F_SCOO: return new MatrixSCOO();
    case F_CSR: return new MatrixCSR();
    case F_ELL: return new MatrixELL();
    case F_DENSE: return new MatrixDENSE();
  }
  return nullptr;
}

// The reference prints progress on std::cout; keep the oracle's stdout clean.
struct Quiet {
  std::streambuf *old;
  std::ostringstream sink;
  Quiet() : old(std::cout.rdbuf(sink.rdbuf())) {}
  ~Quiet() { std::cout.rdbuf(old); }
};

}  // namespace

extern "C" {

void ref_set_threads(int n) { omp_set_num_threads(n); }

void *ref_new(int fmt) { return make(fmt); }

// Matrix has no virtual destructor (Matrix.hpp:11-57): delete through the
// concrete type.
void ref_delete(void *h) {
  Matrix *m = static_cast<Matrix *>(h);
  if (auto *x = dynamic_cast<MatrixCOO *>(m)) delete x;
  else if (auto *x = dynamic_cast<MatrixSCOO *>(m)) delete x;
  else if (auto *x = dynamic_cast<MatrixCSR *>(m)) delete x;
  else if (auto *x = dynamic_cast<MatrixELL *>(m)) delete x;
  else if (auto *x = dynamic_cast<MatrixDENSE *>(m)) delete x;
}

void ref_fill_cqmat(void *h, int n_row, int n_col, int c, double q, unsigned seed_mult, int pr,
                    int pc, int NR, int NC) {
  Quiet quiet;
  static_cast<Matrix *>(h)->fill_cqmat(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC);
}

void ref_fill_cdiag(void *h, int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC) {
  Quiet quiet;
  static_cast<Matrix *>(h)->fill_cdiag(n_row, n_col, cdiag, pr, pc, NR, NC);
}

// geometry: n_row n_col f_row f_col ln_row ln_col pr pc NR NC
void ref_geom(void *h, int *g, long *nnz) {
  Matrix &m = *static_cast<Matrix *>(h);
  for (int i = 0; i < 10; i++) g[i] = *PeekBase::geom(m, i);
  *nnz = *PeekBase::nnz_of(m);
}

// Array pointers of the object (owned by it).  COO/SCOO: a = row, b = col;
// CSR: a = rowptr, b = colidx; ELL: a = columns (row-major), *max_col set;
// DENSE: only values.
void ref_arrays(int fmt, void *h, double **values, int **a, int **b, int *max_col) {
  Matrix *m = static_cast<Matrix *>(h);
  *values = nullptr;
  *a = *b = nullptr;
  *max_col = 0;
  switch (fmt) {
    case F_COO: {
      MatrixCOO *x = dynamic_cast<MatrixCOO *>(m);
      *values = const_cast<double *>(x->get_values());
      *a = const_cast<int *>(x->get_row());
      *b = const_cast<int *>(x->get_col());
      break;
    }
    case F_SCOO: {
      MatrixSCOO *x = dynamic_cast<MatrixSCOO *>(m);
      *values = PeekSCOO::vals(*x);
      *a = PeekSCOO::rows(*x);
      *b = PeekSCOO::cols(*x);
      break;
    }
    case F_CSR: {
      MatrixCSR *x = dynamic_cast<MatrixCSR *>(m);
      *values = PeekCSR::vals(*x);
      *a = PeekCSR::rowptr_of(*x);
      *b = PeekCSR::colidx_of(*x);
      break;
    }
    case F_ELL: {
      MatrixELL *x = dynamic_cast<MatrixELL *>(m);
      *values = PeekELL::vals(*x);
      *a = PeekELL::cols(*x);
      *max_col = PeekELL::maxc(*x);
      break;
    }
    case F_DENSE: {
      *values = PeekDENSE::vals(*dynamic_cast<MatrixDENSE *>(m));
      break;
    }
  }
}

// <fmt>(const MatrixCOO&) conversion constructors.
void *ref_from_coo(int fmt, void *coo) {
  Quiet quiet;
  const MatrixCOO &c = *dynamic_cast<MatrixCOO *>(static_cast<Matrix *>(coo));
  // Handles are Matrix* (the virtual base), never the derived address.
  Matrix *m = nullptr;
  switch (fmt) {
    case F_SCOO: m = new MatrixSCOO(c); break;
    case F_CSR: m = new MatrixCSR(c); break;
    case F_ELL: m = new MatrixELL(c); break;
    case F_DENSE: m = new MatrixDENSE(c); break;
  }
  return m;
}

// Matrix::spmv into a caller buffer of `len` doubles.
void ref_spmv(void *h, const double *v, double *r, int len, int vect_incr) {
  Quiet quiet;
  double *out = static_cast<Matrix *>(h)->spmv(v, vect_incr);
  std::memcpy(r, out, sizeof(double) * (size_t)len);
  delete[] out;
}

void ref_Ax(void *h, double *r, const double *v, int vect_incr) {
  Quiet quiet;
  static_cast<Matrix *>(h)->Ax(r, v, vect_incr);
}

void ref_a_axpx(void *h, const double *v, double *r, int len, int vect_incr) {
  Quiet quiet;
  double *out = static_cast<Matrix *>(h)->a_axpx_(v, vect_incr);
  std::memcpy(r, out, sizeof(double) * (size_t)len);
  delete[] out;
}

void ref_AAxpAx(void *h, double *r, double *v, double *buf, int vect_incr) {
  Quiet quiet;
  static_cast<Matrix *>(h)->AAxpAx(r, v, buf, vect_incr);
}

void ref_AAxpAxpx(void *h, double *r, double *v, double *buf, int vect_incr) {
  Quiet quiet;
  static_cast<Matrix *>(h)->AAxpAxpx(r, v, buf, vect_incr);
}

// Matrix::make_stochastic; s holds ln_col zeros on entry.
void ref_make_stochastic(void *h, double *s) {
  Quiet quiet;
  static_cast<Matrix *>(h)->make_stochastic(s);
}

// Row operations (Matrix.hpp:46-51), through the virtual interface.
void ref_get_row_sums(void *h, double *s) {
  Quiet quiet;
  static_cast<Matrix *>(h)->get_row_sums(s);
}

void ref_normalize_rows(void *h, double *s) {
  Quiet quiet;
  static_cast<Matrix *>(h)->normalize_rows(s);
}

void ref_get_row_max_abs(void *h, double *m) {
  Quiet quiet;
  static_cast<Matrix *>(h)->get_row_max_abs(m);
}

void ref_apply_exponential(void *h, double *m, int base) {
  Quiet quiet;
  static_cast<Matrix *>(h)->apply_exponential(m, base);
}

// MatrixCOO::readMM into a COO handle: 0, or -2 when it throws
// MatrixFormatReadException.
int ref_read_mm(void *h, const char *path) {
  Quiet quiet;
  try {
    dynamic_cast<MatrixCOO *>(static_cast<Matrix *>(h))->readMM(std::string(path));
  } catch (const MatrixFormatReadException &) {
    return -2;
  }
  return 0;
}

// values_generation::cqmat_value, for single-index checks.
void ref_cqmat_value(unsigned long k, unsigned long nr, unsigned long nc, unsigned long c,
                     double q, unsigned seed_mult, long *row, long *col) {
  auto t = tbsla::utils::values_generation::cqmat_value(k, nr, nc, c, q, seed_mult);
  *row = (long)std::get<0>(t);
  *col = (long)std::get<1>(t);
}

void ref_cdiag_value(unsigned long i, unsigned long nv, unsigned long nr, unsigned long nc,
                     unsigned long cdiag, long *row, long *col) {
  auto t = tbsla::utils::values_generation::cdiag_value(i, nv, nr, nc, cdiag);
  *row = (long)std::get<0>(t);
  *col = (long)std::get<1>(t);
}

// The binary image written by Matrix::write (e.g. MatrixCSR.cpp:236-261),
// malloc'ed; the caller releases it with ref_free.
long ref_write(void *h, char **out) {
  std::ostringstream os;
  static_cast<Matrix *>(h)->write(os);
  std::string s = os.str();
  *out = static_cast<char *>(std::malloc(s.size() ? s.size() : 1));
  std::memcpy(*out, s.data(), s.size());
  return (long)s.size();
}

void ref_free(void *p) { std::free(p); }
}
