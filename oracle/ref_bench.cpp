// ref_bench.cpp -- times the reference's own CPU SpMV (oracle and CPU baseline only).
//
// TEST INFRASTRUCTURE.  Linked by oracle/Makefile against the reference's
// src/cpp sources compiled where they lie; bench.py runs the binary for its
// cpu_baseline leg.  It follows the reference harness tbsla_perf
// (src/cpp/apps/tbsla_perf.cpp:28-188): fill_cqmat on one block, optional
// NUMAinit first-touch copy (:78-80), then the mean time of `iters` calls of
// m.spmv(vec) (:94-100, allocation and zeroing included), repeated `reps`
// times; prints one JSON line with the median.
#include <tbsla/cpp/MatrixCSR.hpp>
#include <tbsla/cpp/MatrixELL.hpp>

#include <omp.h>
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv) {
  // args: format n c q seed_mult pr pc NR NC iters reps
  if (argc < 12) {
    std::fprintf(stderr, "usage: %s CSR|ELL n c q seed_mult pr pc NR NC iters reps\n", argv[0]);
    return 2;
  }
  std::string fmt = argv[1];
  int n = std::atoi(argv[2]), c = std::atoi(argv[3]);
  double q = std::atof(argv[4]);
  unsigned s = (unsigned)std::atoi(argv[5]);
  int pr = std::atoi(argv[6]), pc = std::atoi(argv[7]), NR = std::atoi(argv[8]),
      NC = std::atoi(argv[9]);
  int iters = std::atoi(argv[10]), reps = std::atoi(argv[11]);

  tbsla::cpp::Matrix *m = nullptr;
  if (fmt == "CSR") m = new tbsla::cpp::MatrixCSR();
  else if (fmt == "ELL") m = new tbsla::cpp::MatrixELL();
  else return 2;

  std::streambuf *old = std::cout.rdbuf();
  std::ostringstream sink;
  std::cout.rdbuf(sink.rdbuf());
  double t0 = now_s();
  m->fill_cqmat(n, n, c, q, s, pr, pc, NR, NC);
  double t_fill = now_s() - t0;
  m->NUMAinit();
  std::cout.rdbuf(old);

  std::vector<double> vec(m->get_n_col());
  for (size_t i = 0; i < vec.size(); i++) vec[i] = (double)((i * 2654435761u) % 1000) / 1000.0;
  const double *v = vec.data() + m->get_f_col();

  std::vector<double> per;
  double checksum = 0;
  for (int r = 0; r < reps; r++) {
    double t = 0;
    for (int i = 0; i < iters; i++) {
      double t1 = now_s();
      double *res = m->spmv(v);
      t += now_s() - t1;
      checksum += res[0];
      delete[] res;
    }
    per.push_back(t / iters);
  }
  std::sort(per.begin(), per.end());
  double med = per[per.size() / 2];
  std::fflush(stdout);
  std::printf(
      "{\"format\":\"%s\",\"n\":%d,\"c\":%d,\"q\":%g,\"NR\":%d,\"NC\":%d,\"nnz_field\":%ld,"
      "\"time_op\":%.9g,\"time_fill\":%.6g,\"iters\":%d,\"reps\":%d,\"omp_threads\":%d,"
      "\"checksum\":%.17g}\n",
      fmt.c_str(), n, c, q, NR, NC, m->get_nnz(), med, t_fill, iters, reps, omp_get_max_threads(),
      checksum);
  std::fflush(stdout);
  // tbsla::cpp::Matrix has no virtual destructor (Matrix.hpp:11-57)
  if (auto *x = dynamic_cast<tbsla::cpp::MatrixCSR *>(m)) delete x;
  else if (auto *x = dynamic_cast<tbsla::cpp::MatrixELL *>(m)) delete x;
  return 0;
}
