// tbsla/hip/Matrix.hpp -- drop-in C++ classes for TBSLA's matrix API on the GPU.
//
// Header-only layer over the C ABI of libtbsla_hip (include/tbsla_hip.h).
// tbsla::hip::Matrix keeps the interface of tbsla::cpp::Matrix
// (src/cpp/include/tbsla/cpp/Matrix.hpp:11-57): the same method names,
// argument meanings, host-pointer ownership rules (spmv and a_axpx_ return
// new[] host arrays the caller deletes; Ax writes a caller buffer,
// overwriting for CSR/ELL and accumulating for COO/SCOO/DENSE), getters and
// the binary write/read format; tbsla::mpi::Matrix's page_rank and CG
// (src/mpi/include/tbsla/mpi/Matrix.hpp:31-33) are available on one block.
// Device-resident overloads (*_device) take device pointers and a stream.
//
// Errors: a failing library call throws std::runtime_error with the
// library's message (the reference throws MatrixFormatReadException and
// std::runtime_error for unsupported operations).
#ifndef TBSLA_HIP_MATRIX_HPP
#define TBSLA_HIP_MATRIX_HPP

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../tbsla_hip.h"

namespace tbsla {
namespace hip {

// Matrix.hpp:59-63: thrown when a Matrix Market header is not supported.
struct MatrixFormatReadException : public std::exception {
  const char *what() const throw() { return "This Matrix format import is not implemented!"; }
};

inline void check(int status) {
  if (status != TBSLA_HIP_SUCCESS)
    throw std::runtime_error(std::string("tbsla_hip: ") + tbsla_hip_status_string(status) +
                             ": " + tbsla_hip_last_error());
}

// Device buffer owned by the caller's scope.
template <typename T>
class DeviceArray {
 public:
  explicit DeviceArray(size_t n = 0) : n_(n) {
    if (n) check(tbsla_hip_malloc(reinterpret_cast<void **>(&p_), n * sizeof(T)));
  }
  DeviceArray(const T *host, size_t n) : DeviceArray(n) {
    if (n) check(tbsla_hip_memcpy_htod(p_, host, n * sizeof(T), nullptr));
  }
  ~DeviceArray() {
    if (p_) tbsla_hip_free(p_);
  }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  void to_host(T *host, size_t n) const {
    if (n) check(tbsla_hip_memcpy_dtoh(host, p_, n * sizeof(T), nullptr));
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T *p_ = nullptr;
  size_t n_ = 0;
};

class Matrix {
 public:
  explicit Matrix(int format) { check(tbsla_hip_create(&h_, format)); }
  virtual ~Matrix() { tbsla_hip_destroy(h_); }
  Matrix(const Matrix &) = delete;
  Matrix &operator=(const Matrix &) = delete;

  friend std::ostream &operator<<(std::ostream &os, const Matrix &m) { return m.print(os); }

  // ---- SpMV (Matrix.hpp:14-15) -------------------------------------------
  double *spmv(const double *v, int vect_incr = 0) const {
    const tbsla_hip_geometry g = geometry();
    DeviceArray<double> dv(v, in_len(g)), dr(out_len(g));
    check(tbsla_hip_spmv(h_, dr.get(), dv.get(), vect_incr, nullptr));
    double *r = new double[out_len(g)]();
    dr.to_host(r, out_len(g));
    return r;
  }
  void Ax(double *r, const double *v, int vect_incr = 0) const {
    const tbsla_hip_geometry g = geometry();
    DeviceArray<double> dv(v, in_len(g)), dr(r, out_len(g));
    check(tbsla_hip_Ax(h_, dr.get(), dv.get(), vect_incr, nullptr));
    dr.to_host(r, out_len(g));
  }
  void Ax_device(double *d_r, const double *d_v, int vect_incr = 0, void *stream = nullptr) const {
    check(tbsla_hip_Ax(h_, d_r, d_v, vect_incr, stream));
  }
  void spmv_device(double *d_r, const double *d_v, int vect_incr = 0,
                   void *stream = nullptr) const {
    check(tbsla_hip_spmv(h_, d_r, d_v, vect_incr, stream));
  }

  // ---- composites (src/cpp/matrix/Matrix.cpp:5-38) -----------------------
  double *a_axpx_(const double *v, int vect_incr = 0) const {
    const tbsla_hip_geometry g = geometry();
    DeviceArray<double> dv(v, in_len(g)), dr(out_len(g)), dt(out_len(g));
    check(tbsla_hip_a_axpx(h_, dr.get(), dv.get(), dt.get(), vect_incr, nullptr));
    double *r = new double[out_len(g)]();
    dr.to_host(r, out_len(g));
    return r;
  }
  void AAxpAx(double *r, double *v, double *buffer, int vect_incr = 0) const {
    const tbsla_hip_geometry g = geometry();
    DeviceArray<double> dv(v, in_len(g)), dr(r, out_len(g)), db(buffer, out_len(g));
    check(tbsla_hip_AAxpAx(h_, dr.get(), dv.get(), db.get(), vect_incr, nullptr));
    dr.to_host(r, out_len(g));
    db.to_host(buffer, out_len(g));
  }
  void AAxpAxpx(double *r, double *v, double *buffer, int vect_incr = 0) const {
    const tbsla_hip_geometry g = geometry();
    DeviceArray<double> dv(v, in_len(g)), dr(r, out_len(g)), db(buffer, out_len(g));
    check(tbsla_hip_AAxpAxpx(h_, dr.get(), dv.get(), db.get(), vect_incr, nullptr));
    dr.to_host(r, out_len(g));
    db.to_host(buffer, out_len(g));
  }
  void make_stochastic(double *s) {
    const size_t n = (size_t)geometry().ln_col;
    DeviceArray<double> ds(s, n);
    check(tbsla_hip_make_stochastic(h_, ds.get(), nullptr));
    ds.to_host(s, n);
  }
  void get_col_sums(double *s) {
    const size_t n = (size_t)geometry().ln_col;
    DeviceArray<double> ds(s, n);
    check(tbsla_hip_get_col_sums(h_, ds.get(), nullptr));
    ds.to_host(s, n);
  }
  void normalize_cols(double *s) {
    const size_t n = (size_t)geometry().ln_col;
    DeviceArray<double> ds(s, n);
    check(tbsla_hip_normalize_cols(h_, ds.get(), nullptr));
  }

  // ---- row operations (Matrix.hpp:46-51) ---------------------------------
  // The buffer is indexed as the reference's format does: ln_row entries
  // (local rows) for CSR and DENSE, n_row entries (global rows) for ELL and
  // SCOO (MatrixELL.cpp:790, MatrixSCOO.cpp:703); COO leaves it alone.
  // SCOO accumulates into the buffer; the others overwrite their rows.
  void get_row_sums(double *s) {
    const tbsla_hip_geometry g = geometry();
    if (g.format == TBSLA_HIP_COO) return;
    const size_t n = row_sums_len(g);
    DeviceArray<double> ds(s, n);
    check(tbsla_hip_get_row_sums(h_, ds.get(), nullptr));
    ds.to_host(s, n);
  }
  void normalize_rows(double *s) {
    const tbsla_hip_geometry g = geometry();
    if (g.format == TBSLA_HIP_COO) return;
    DeviceArray<double> ds(s, row_sums_len(g));
    check(tbsla_hip_normalize_rows(h_, ds.get(), nullptr));
  }
  // CSR only (MatrixCSR.cpp:1285-1334); the other formats leave the
  // arguments and values untouched, as the reference's empty bodies do.
  void get_row_max_abs(double *max_abs) {
    const tbsla_hip_geometry g = geometry();
    if (g.format != TBSLA_HIP_CSR) return;
    const size_t n = (size_t)g.ln_row;
    DeviceArray<double> dm(n);
    check(tbsla_hip_get_row_max_abs(h_, dm.get(), nullptr));
    dm.to_host(max_abs, n);
  }
  void apply_exponential(double *max_abs, int base = -1) {
    const tbsla_hip_geometry g = geometry();
    if (g.format != TBSLA_HIP_CSR) return;
    if (!max_abs) {
      check(tbsla_hip_apply_exponential(h_, nullptr, base, nullptr));
      return;
    }
    DeviceArray<double> dm(max_abs, (size_t)g.ln_row);
    check(tbsla_hip_apply_exponential(h_, dm.get(), base, nullptr));
  }

  // ---- drivers (src/mpi/include/tbsla/mpi/Matrix.hpp:31-33), one block ---
  virtual double *page_rank(double beta, double epsilon, int max_iterations,
                            int &nb_iterations_done) {
    const size_t n = (size_t)geometry().ln_col;
    DeviceArray<double> db(n);
    double err = 0;
    check(tbsla_hip_page_rank(h_, beta, epsilon, max_iterations, db.get(), &nb_iterations_done,
                              &err, nullptr));
    double *b = new double[n];
    db.to_host(b, n);
    return b;
  }
  // the max-normalised iteration of tbsla::mpi::MatrixCOO::page_rank
  // (src/mpi/matrix/MatrixCOO.cpp:108-175); MatrixCOO::page_rank calls it
  double *page_rank_max_norm(double beta, double epsilon, int max_iterations,
                             int &nb_iterations_done) {
    const size_t n = (size_t)geometry().n_col;
    DeviceArray<double> db(n);
    double err = 0;
    check(tbsla_hip_page_rank_max_norm(h_, beta, epsilon, max_iterations, db.get(),
                                       &nb_iterations_done, &err, nullptr));
    double *b = new double[n];
    db.to_host(b, n);
    return b;
  }
  // CG(comm, v, r, ...) of the reference with v the right-hand side and r
  // the solution (CG.cpp:32); the Ax argument order is corrected.
  void CG(double *v, double *r, int max_iterations, int &nb_iterations_done,
          const tbsla_hip_cg_options *opt = nullptr) {
    const size_t n = (size_t)geometry().n_col;
    DeviceArray<double> dv(v, n), dr(r, n);
    tbsla_hip_cg_options o{};
    if (opt) o = *opt;
    else {
      o.tolerance = 1.0e-10;
      o.use_initial_guess = 0;
    }
    o.max_iterations = max_iterations;
    double rn = 0;
    check(tbsla_hip_cg(h_, dv.get(), dr.get(), &o, &nb_iterations_done, &rn, nullptr));
    dr.to_host(r, n);
  }

  // ---- generators (Matrix.hpp:41-42) -------------------------------------
  void fill_cdiag(int n_row, int n_col, int cdiag, int pr = 0, int pc = 0, int NR = 1,
                  int NC = 1) {
    check(tbsla_hip_fill_cdiag(h_, n_row, n_col, cdiag, pr, pc, NR, NC, nullptr));
  }
  void fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult = 1, int pr = 0,
                  int pc = 0, int NR = 1, int NC = 1) {
    check(tbsla_hip_fill_cqmat(h_, n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, -1, nullptr));
  }
  // fill_cqmat with values V(value_seed, k) uniform in [0,1) instead of 1.0
  void fill_cqmat_values(int n_row, int n_col, int c, double q, unsigned int seed_mult,
                         long long value_seed, int pr = 0, int pc = 0, int NR = 1, int NC = 1) {
    check(tbsla_hip_fill_cqmat(h_, n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, value_seed,
                               nullptr));
  }

  // Matrix.hpp:43-44.  The reference's generators for these draw their
  // entries per OpenMP thread, so their output depends on the thread count
  // (MatrixCSR.cpp:906-925) and has no fixed result to reproduce; COO and
  // DENSE implement them as empty stubs (MatrixCOO.cpp:356-379).  Here they
  // throw, as tbsla::mpi::Matrix does for operations it does not provide.
  void fill_random(int, int, double, unsigned int = 1, int = 0, int = 0, int = 1, int = 1) {
    throw std::runtime_error("tbsla_hip: fill_random is not provided (thread-count-dependent "
                             "generator); use fill_cqmat or read");
  }
  void fill_brain(int, int, int *, std::vector<std::vector<double> >,
                  std::vector<std::unordered_map<int, std::vector<int> > >, unsigned int,
                  int = 0, int = 0, int = 1, int = 1) {
    throw std::runtime_error("tbsla_hip: fill_brain is not provided (thread-count-dependent "
                             "generator); use fill_cqmat or read");
  }
  // Matrix.hpp:39.  First-touch placement of host arrays over NUMA nodes
  // (MatrixCSR.cpp:1234-1268) has no counterpart for arrays resident in
  // device memory: a no-op.
  void NUMAinit() {}

  // ---- getters (Matrix.hpp:21-32) ----------------------------------------
  int get_n_row() const { return geometry().n_row; }
  int get_n_col() const { return geometry().n_col; }
  int get_f_row() const { return geometry().f_row; }
  int get_f_col() const { return geometry().f_col; }
  int get_ln_row() const { return geometry().ln_row; }
  int get_ln_col() const { return geometry().ln_col; }
  int get_pr() const { return geometry().pr; }
  int get_pc() const { return geometry().pc; }
  int get_NR() const { return geometry().NR; }
  int get_NC() const { return geometry().NC; }
  long int get_nnz() const { return (long int)geometry().nnz; }
  std::string get_vectorization() const { return "HIP gfx950"; }

  // ---- binary format (MatrixCSR.cpp:236-295 and the per-format twins) ----
  // Header: n_row n_col ln_row ln_col f_row f_col (int32), nnz (int64),
  // pr pc NR NC (int32), ELL max_col (int32); then each array preceded by
  // its size_t length.
  std::ostream &write(std::ostream &os) const {
    size_t len = 0;
    check(tbsla_hip_write(h_, nullptr, 0, &len, nullptr));
    std::vector<char> img(len);
    check(tbsla_hip_write(h_, img.data(), img.size(), &len, nullptr));
    os.write(img.data(), (std::streamsize)len);
    return os;
  }

  std::istream &read(std::istream &is, std::size_t pos = 0, std::size_t n = 1) {
    (void)pos;
    (void)n;
    tbsla_hip_geometry g{};
    int hdr1[6], hdr2[4];
    long int nnz = 0;
    is.read(reinterpret_cast<char *>(hdr1), sizeof(hdr1));
    is.read(reinterpret_cast<char *>(&nnz), sizeof(nnz));
    is.read(reinterpret_cast<char *>(hdr2), sizeof(hdr2));
    g.n_row = hdr1[0];
    g.n_col = hdr1[1];
    g.ln_row = hdr1[2];
    g.ln_col = hdr1[3];
    g.f_row = hdr1[4];
    g.f_col = hdr1[5];
    g.nnz = nnz;
    g.pr = hdr2[0];
    g.pc = hdr2[1];
    g.NR = hdr2[2];
    g.NC = hdr2[3];
    const int fmt = geometry().format;
    g.format = fmt;
    if (fmt == TBSLA_HIP_ELL) is.read(reinterpret_cast<char *>(&g.max_col), sizeof(int));
    auto arr = [&is](auto &vec) {
      size_t len = 0;
      is.read(reinterpret_cast<char *>(&len), sizeof(len));
      vec.resize(len);
      if (len) is.read(reinterpret_cast<char *>(vec.data()), (std::streamsize)(len * sizeof(vec[0])));
    };
    std::vector<int> a, b;
    std::vector<double> v;
    switch (fmt) {
      case TBSLA_HIP_CSR: arr(v); arr(a); arr(b); break;
      case TBSLA_HIP_COO:
      case TBSLA_HIP_SCOO: arr(v); arr(a); arr(b); g.stored = (long long)v.size(); break;
      case TBSLA_HIP_ELL: arr(v); arr(a); break;
      case TBSLA_HIP_DENSE: arr(v); break;
    }
    if (!is) throw std::runtime_error("tbsla_hip: truncated matrix stream");
    check(tbsla_hip_import(h_, fmt, &g, a.empty() ? nullptr : a.data(),
                           b.empty() ? nullptr : b.data(), v.empty() ? nullptr : v.data(),
                           nullptr));
    return is;
  }

  std::ostream &print(std::ostream &os) const {
    const tbsla_hip_geometry g = geometry();
    static const char *names[] = {"COO", "SCOO", "CSR", "ELL", "DENSE"};
    os << "------ " << names[g.format] << " (HIP) ------\n"
       << "n_row " << g.n_row << " n_col " << g.n_col << " ln_row " << g.ln_row << " ln_col "
       << g.ln_col << " f_row " << g.f_row << " f_col " << g.f_col << " nnz " << g.nnz
       << " pr " << g.pr << " pc " << g.pc << " NR " << g.NR << " NC " << g.NC;
    if (g.format == TBSLA_HIP_ELL) os << " max_col " << g.max_col;
    return os << "\n";
  }

  // Matrix.hpp:37: the block as an n_row x n_col dense array, printed as
  // tbsla::utils::array::print_dense_matrix does (src/cpp/utils/array.cpp:
  // 85-104: fixed notation, precision = the widest integer part, width
  // precision + 3, two spaces after each value).  Entries sit at their
  // global row and column; DENSE prints its description instead, as
  // MatrixDENSE.cpp:59-61 does.
  std::ostream &print_as_dense(std::ostream &os) {
    const tbsla_hip_geometry g = geometry();
    if (g.format == TBSLA_HIP_DENSE) return print(os);
    std::vector<int> a, b;
    std::vector<double> v;
    exported(g, a, b, v);
    const size_t nr = (size_t)g.n_row, nc = (size_t)g.n_col;
    std::vector<double> d(nr * nc, 0.0);
    auto put = [&](long long r, long long c, double x) {
      if (r >= 0 && r < (long long)nr && c >= 0 && c < (long long)nc) d[r * nc + c] += x;
    };
    switch (g.format) {
      case TBSLA_HIP_CSR:
        for (int i = 0; i < g.ln_row; i++)
          for (int j = a[i]; j < a[i + 1]; j++) put((long long)g.f_row + i, b[j], v[j]);
        break;
      case TBSLA_HIP_ELL:
        for (int i = 0; i < g.ln_row; i++)
          for (int j = 0; j < g.max_col; j++) {
            const size_t k = (size_t)i * g.max_col + j;
            if (k < v.size() && v[k] != 0.0) put((long long)g.f_row + i, a[k], v[k]);
          }
        break;
      default:  // COO, SCOO: global triplets
        for (size_t k = 0; k < v.size(); k++) put(a[k], b[k], v[k]);
        break;
    }
    int max_p = 1;
    for (double x : d) {
      if (x == 0.0 || !std::isfinite(x)) continue;  // no integer part to widen for
      const int lr = (int)(std::log10(std::fabs(x)) + 1);
      if (lr > max_p) max_p = lr;
    }
    os << std::fixed;
    os.precision(max_p);
    for (size_t i = 0; i < nr; i++) {
      for (size_t j = 0; j < nc; j++) os << std::setw(max_p + 3) << d[i * nc + j] << "  ";
      os << std::endl;
    }
    return os;
  }

  tbsla_hip_geometry geometry() const {
    tbsla_hip_geometry g{};
    check(tbsla_hip_get_geometry(h_, &g));
    return g;
  }
  tbsla_hip_matrix_t handle() const { return h_; }

 protected:
  tbsla_hip_matrix_t h_ = nullptr;

  static size_t row_sums_len(const tbsla_hip_geometry &g) {
    return (size_t)(g.format == TBSLA_HIP_ELL || g.format == TBSLA_HIP_SCOO ? g.n_row : g.ln_row);
  }
  static size_t in_len(const tbsla_hip_geometry &g) { return (size_t)g.ln_col; }
  static size_t out_len(const tbsla_hip_geometry &g) {
    return (size_t)(g.format == TBSLA_HIP_COO ? g.n_row : g.ln_row);
  }
  void exported(const tbsla_hip_geometry &g, std::vector<int> &a, std::vector<int> &b,
                std::vector<double> &v) const {
    switch (g.format) {
      case TBSLA_HIP_CSR:
        a.resize((size_t)g.ln_row + 1);
        b.resize((size_t)g.stored);
        v.resize((size_t)g.stored);
        break;
      case TBSLA_HIP_COO:
      case TBSLA_HIP_SCOO:
        a.resize((size_t)g.stored);
        b.resize((size_t)g.stored);
        v.resize((size_t)g.stored);
        break;
      case TBSLA_HIP_ELL:
        a.resize((size_t)g.stored);
        v.resize((size_t)g.stored);
        break;
      case TBSLA_HIP_DENSE:
        v.resize((size_t)g.stored);
        break;
    }
    check(tbsla_hip_export(h_, a.empty() ? nullptr : a.data(), b.empty() ? nullptr : b.data(),
                           v.empty() ? nullptr : v.data(), nullptr));
  }
};

class MatrixCOO : public Matrix {
 public:
  MatrixCOO() : Matrix(TBSLA_HIP_COO) {}
  // MatrixCOO::readMM (MatrixCOO.cpp:206-278); an unsupported header throws
  // MatrixFormatReadException as the reference does; a symmetric file's
  // off-diagonal entries are kept with their mirrors.
  void readMM(std::string fname) {
    const int st = tbsla_hip_read_mm(h_, fname.c_str(), nullptr);
    if (st == TBSLA_HIP_ERR_UNSUPPORTED) throw MatrixFormatReadException();
    check(st);
  }
  // tbsla::mpi::MatrixCOO overrides page_rank with the max-normalised
  // iteration (src/mpi/matrix/MatrixCOO.cpp:108-175)
  double *page_rank(double beta, double epsilon, int max_iterations,
                    int &nb_iterations_done) override {
    return page_rank_max_norm(beta, epsilon, max_iterations, nb_iterations_done);
  }
};

// Conversion constructors <FMT>(const MatrixCOO&) (MatrixCSR.cpp:49 etc.).
class MatrixSCOO : public Matrix {
 public:
  MatrixSCOO() : Matrix(TBSLA_HIP_SCOO) {}
  explicit MatrixSCOO(const MatrixCOO &m) : Matrix(TBSLA_HIP_SCOO) {
    check(tbsla_hip_convert(h_, TBSLA_HIP_SCOO, m.handle(), nullptr));
  }
};
class MatrixCSR : public Matrix {
 public:
  MatrixCSR() : Matrix(TBSLA_HIP_CSR) {}
  explicit MatrixCSR(const MatrixCOO &m) : Matrix(TBSLA_HIP_CSR) {
    check(tbsla_hip_convert(h_, TBSLA_HIP_CSR, m.handle(), nullptr));
  }
  // MatrixCSR::dense_multiply (MatrixCSR.cpp:1270-1283, block form
  // src/mpi/matrix/MatrixCSR.cpp:11-30): C = A B with B ln_col x cols_B and
  // C ln_row x cols_B, row-major host arrays; C is overwritten.
  void dense_multiply(const double *B, double *C, int cols_B) const {
    const tbsla_hip_geometry g = geometry();
    const size_t nb = (size_t)g.ln_col * (size_t)cols_B, nc = (size_t)g.ln_row * (size_t)cols_B;
    DeviceArray<double> dB(B, nb), dC(nc);
    check(tbsla_hip_spmm(h_, dB.get(), dC.get(), cols_B, nullptr));
    dC.to_host(C, nc);
  }
  void dense_multiply_device(const double *d_B, double *d_C, int cols_B,
                             void *stream = nullptr) const {
    check(tbsla_hip_spmm(h_, d_B, d_C, cols_B, stream));
  }
  // The attention step's row softmax fused into one pass
  // (tbsla_attention_mpi.cpp:183-191): the result of get_row_max_abs,
  // apply_exponential(max_abs, base), get_row_sums and normalize_rows in
  // turn.  max_abs and s (ln_row doubles each, nullable) receive the row
  // maxima and sums.
  void softmax_rows(int base = -1, double *max_abs = nullptr, double *s = nullptr) {
    const size_t n = (size_t)geometry().ln_row;
    DeviceArray<double> dm(n), ds(n);
    check(tbsla_hip_softmax_rows(h_, base, dm.get(), ds.get(), nullptr));
    if (max_abs) dm.to_host(max_abs, n);
    if (s) ds.to_host(s, n);
  }
};
class MatrixELL : public Matrix {
 public:
  MatrixELL() : Matrix(TBSLA_HIP_ELL) {}
  explicit MatrixELL(const MatrixCOO &m) : Matrix(TBSLA_HIP_ELL) {
    check(tbsla_hip_convert(h_, TBSLA_HIP_ELL, m.handle(), nullptr));
  }
};
class MatrixDENSE : public Matrix {
 public:
  MatrixDENSE() : Matrix(TBSLA_HIP_DENSE) {}
  explicit MatrixDENSE(const MatrixCOO &m) : Matrix(TBSLA_HIP_DENSE) {
    check(tbsla_hip_convert(h_, TBSLA_HIP_DENSE, m.handle(), nullptr));
  }
};

}  // namespace hip
}  // namespace tbsla

#endif
