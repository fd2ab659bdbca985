// CppMatrix.hpp -- the GPU block as a subclass of the reference's abstract
// tbsla::cpp::Matrix (src/cpp/include/tbsla/cpp/Matrix.hpp:11-57).
//
// Code written against tbsla::cpp::Matrix (a pointer or reference to the
// base class, the reference's own non-virtual a_axpx_ / AAxpAx / AAxpAxpx /
// make_stochastic of src/cpp/matrix/Matrix.cpp:5-38, which call the
// virtuals) runs on the GPU unchanged with a tbsla::hip::CppMatrix in place
// of a tbsla::cpp::Matrix<FMT>.  Every pure virtual is overridden by the
// matching call of tbsla::hip::Matrix (include/tbsla/hip/Matrix.hpp), and
// the base class's protected geometry fields, which its non-virtual getters
// read, are refreshed after every call that changes the block.
//
// tbsla::cpp::Matrix has no virtual destructor: delete a CppMatrix through
// its own type (as the reference's own classes must be), never through a
// tbsla::cpp::Matrix pointer.
//
// It needs the reference's headers on the include path and its compiled
// src/cpp/matrix/Matrix.cpp at link time, so it is a header of its own;
// include/tbsla/hip/Matrix.hpp alone has no dependency on the reference.
#pragma once

#include <tbsla/cpp/Matrix.hpp>
#include <tbsla/hip/Matrix.hpp>

namespace tbsla {
namespace hip {

class CppMatrix : public tbsla::cpp::Matrix {
 public:
  explicit CppMatrix(int format) : g_(format) { sync(); }

  double *spmv(const double *v, int vect_incr = 0) const override { return g_.spmv(v, vect_incr); }
  void Ax(double *r, const double *v, int vect_incr = 0) const override { g_.Ax(r, v, vect_incr); }

  std::ostream &write(std::ostream &os) override { return g_.write(os); }
  std::istream &read(std::istream &is, std::size_t pos = 0, std::size_t n = 1) override {
    g_.read(is, pos, n);
    sync();
    return is;
  }
  std::ostream &print(std::ostream &os) const override { return g_.print(os); }
  std::ostream &print_as_dense(std::ostream &os) override { return g_.print_as_dense(os); }
  void NUMAinit() override { g_.NUMAinit(); }

  void fill_cdiag(int n_row, int n_col, int cdiag, int pr = 0, int pc = 0, int NR = 1,
                  int NC = 1) override {
    g_.fill_cdiag(n_row, n_col, cdiag, pr, pc, NR, NC);
    sync();
  }
  void fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult = 1, int pr = 0,
                  int pc = 0, int NR = 1, int NC = 1) override {
    g_.fill_cqmat(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC);
    sync();
  }
  void fill_random(int n_row, int n_col, double nnz_ratio, unsigned int seed_mult = 1, int pr = 0,
                   int pc = 0, int NR = 1, int NC = 1) override {
    g_.fill_random(n_row, n_col, nnz_ratio, seed_mult, pr, pc, NR, NC);
  }
  void fill_brain(int n_row, int n_col, int *neuron_type,
                  std::vector<std::vector<double> > proba_conn,
                  std::vector<std::unordered_map<int, std::vector<int> > > brain_struct,
                  unsigned int seed_mult, int pr = 0, int pc = 0, int NR = 1,
                  int NC = 1) override {
    g_.fill_brain(n_row, n_col, neuron_type, proba_conn, brain_struct, seed_mult, pr, pc, NR, NC);
  }

  void get_row_sums(double *buffer) override { g_.get_row_sums(buffer); }
  void normalize_rows(double *s) override { g_.normalize_rows(s); }
  void get_col_sums(double *buffer) override { g_.get_col_sums(buffer); }
  void normalize_cols(double *s) override { g_.normalize_cols(s); }
  void get_row_max_abs(double *max_abs) override { g_.get_row_max_abs(max_abs); }
  void apply_exponential(double *max_abs, int base) override { g_.apply_exponential(max_abs, base); }

  // the GPU object itself: device-vector calls, drivers, the C handle
  tbsla::hip::Matrix &gpu() { return g_; }
  const tbsla::hip::Matrix &gpu() const { return g_; }

 private:
  void sync() {
    const tbsla_hip_geometry g = g_.geometry();
    n_row = g.n_row;
    n_col = g.n_col;
    f_row = g.f_row;
    f_col = g.f_col;
    ln_row = g.ln_row;
    ln_col = g.ln_col;
    pr = g.pr;
    pc = g.pc;
    NR = g.NR;
    NC = g.NC;
    nnz = (long int)g.nnz;
  }
  tbsla::hip::Matrix g_;
};

}  // namespace hip
}  // namespace tbsla
