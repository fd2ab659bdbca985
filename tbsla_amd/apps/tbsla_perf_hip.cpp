// tbsla_perf_hip -- the reference perf harness on the GPU.
//
// Same flags and output as src/cpp/apps/tbsla_perf.cpp (dothings :28-188):
//   --format COO|SCOO|CSR|ELL|DENSE --matrix cqmat|cdiag --NR n_row --NC n_col
//   --C c --Q q --S seed_mult --op spmv|Ax|a_axpx|AAxpAx|AAxpAxpx
// plus the drivers of src/mpi/apps/tbsla_perf_page_rank_mpi.cpp and
// tbsla_perf_cg_mpi.cpp on one block: --op page_rank (--beta, --epsilon,
// --max-iterations) and --op cg.  Extensions: --GR/--GC/--pr/--pc pick one
// block of a grid, --values uniform stores V(0, k) instead of 1.0,
// --iterations sets the repeat count (default 100, as :93).
// Vectors stay on the GPU; timing is taken between stream synchronisations.
// Prints one JSON dict line with the reference's keys plus "gbs".
#include <tbsla/hip/Matrix.hpp>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

namespace {

struct Args {
  std::map<std::string, std::string> kv;
  Args(int argc, char **argv) {
    for (int i = 1; i < argc; i++) {
      std::string k = argv[i];
      if (k.rfind("--", 0) == 0) kv[k] = (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) ? argv[++i] : "";
    }
  }
  std::string get(const std::string &k, const std::string &d = "") const {
    auto it = kv.find(k);
    return it == kv.end() ? d : it->second;
  }
  bool has(const std::string &k) const { return kv.count(k) > 0; }
};

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int format_of(const std::string &f) {
  if (f == "COO" || f == "coo") return TBSLA_HIP_COO;
  if (f == "SCOO" || f == "scoo") return TBSLA_HIP_SCOO;
  if (f == "CSR" || f == "csr") return TBSLA_HIP_CSR;
  if (f == "ELL" || f == "ell") return TBSLA_HIP_ELL;
  if (f == "DENSE" || f == "dense") return TBSLA_HIP_DENSE;
  return -1;
}

double bytes_per_spmv(const tbsla_hip_geometry &g) {
  double x = 8.0 * g.ln_col, y = 8.0 * g.ln_row;
  switch (g.format) {
    case TBSLA_HIP_CSR: return 12.0 * g.stored + 4.0 * (g.ln_row + 1) + x + y;
    case TBSLA_HIP_ELL: return 12.0 * g.ln_row * g.max_col + x + y;
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO: return 16.0 * g.stored + x + y;
    default: return 8.0 * g.ln_row * g.ln_col + x + y;
  }
}

}  // namespace

int main(int argc, char **argv) {
  using namespace tbsla::hip;
  Args in(argc, argv);
  const std::string format = in.get("--format");
  const int fmt = format_of(format);
  if (fmt < 0) {
    std::fprintf(stderr, "A file format has to be given with the parameter --format format\n");
    return 1;
  }
  const std::string matrix = in.get("--matrix");
  const std::string op = in.get("--op");
  const int NR = std::stoi(in.get("--NR", "1024")), NC = std::stoi(in.get("--NC", "1024"));
  const int C = std::stoi(in.get("--C", "8"));
  const double Q = std::stod(in.get("--Q", "0.1"));
  const unsigned S = (unsigned)std::stoi(in.get("--S", "0"));
  const int GR = std::stoi(in.get("--GR", "1")), GC = std::stoi(in.get("--GC", "1"));
  const int pr = std::stoi(in.get("--pr", "0")), pc = std::stoi(in.get("--pc", "0"));
  const int iters = std::stoi(in.get("--iterations", "100"));
  if (matrix != "cqmat" && matrix != "cdiag") {
    std::fprintf(stderr, "No matrix has been given with the parameter --matrix matrix.\n");
    return 1;
  }
  const bool spmv_like = op == "spmv" || op == "Ax" || op == "a_axpx" || op == "AAxpAx" ||
                         op == "AAxpAxpx";
  if (!spmv_like && op != "page_rank" && op != "cg") {
    std::fprintf(stderr, "OP : %s unrecognized!\n", op.c_str());
    return 1;
  }
  Matrix m(fmt);
  const double t_app = now();
  if (matrix == "cdiag")
    m.fill_cdiag(NR, NC, C, pr, pc, GR, GC);
  else if (in.get("--values") == "uniform")
    m.fill_cqmat_values(NR, NC, C, Q, S, 0, pr, pc, GR, GC);
  else
    m.fill_cqmat(NR, NC, C, Q, S, pr, pc, GR, GC);
  const tbsla_hip_geometry g = m.geometry();
  const size_t nin = (size_t)g.ln_col;
  const size_t nout = (size_t)(fmt == TBSLA_HIP_COO ? g.n_row : g.ln_row);
  DeviceArray<double> dv(std::max(nin, nout)), dr(std::max(nin, nout)),
      dbuf(std::max(nin, nout));
  check(tbsla_hip_fill_unit_vector(dv.get(), (long long)nin, 1, nullptr));
  double t_op = 0, flops = 0;
  int nb_iter = 0;
  if (spmv_like) {
    auto once = [&]() {
      if (op == "spmv") check(tbsla_hip_spmv(m.handle(), dr.get(), dv.get(), 0, nullptr));
      else if (op == "Ax") check(tbsla_hip_Ax(m.handle(), dr.get(), dv.get(), 0, nullptr));
      else if (op == "a_axpx")
        check(tbsla_hip_a_axpx(m.handle(), dr.get(), dv.get(), dbuf.get(), 0, nullptr));
      else if (op == "AAxpAx")
        check(tbsla_hip_AAxpAx(m.handle(), dr.get(), dv.get(), dbuf.get(), 0, nullptr));
      else check(tbsla_hip_AAxpAxpx(m.handle(), dr.get(), dv.get(), dbuf.get(), 0, nullptr));
    };
    once();
    check(tbsla_hip_stream_synchronize(nullptr));
    const double t0 = now();
    for (int i = 0; i < iters; i++) once();
    check(tbsla_hip_stream_synchronize(nullptr));
    t_op = (now() - t0) / iters;
    // tbsla_perf.cpp:140-146
    if (op == "spmv" || op == "Ax") flops = 2.0 * g.nnz;
    else if (op == "AAxpAxpx") flops = 4.0 * g.nnz + g.n_col;
    else flops = 4.0 * g.nnz;
  } else if (op == "page_rank") {
    const double beta = std::stod(in.get("--beta", "0.85"));
    const double eps = std::stod(in.get("--epsilon", "1e-5"));
    const int max_it = std::stoi(in.get("--max-iterations", "10000"));
    check(tbsla_hip_make_stochastic(m.handle(), dbuf.get(), nullptr));
    double err = 0;
    const double t0 = now();
    check(tbsla_hip_page_rank(m.handle(), beta, eps, max_it, dr.get(), &nb_iter, &err, nullptr));
    t_op = now() - t0;
    // tbsla_perf_page_rank_mpi.cpp:145-152
    flops = (2.0 * g.nnz + 2.0 * g.n_col) * nb_iter + 2.0 * g.n_col;
  } else {
    tbsla_hip_cg_options o{};
    o.max_iterations = std::stoi(in.get("--max-iterations", "10000"));
    o.tolerance = std::stod(in.get("--tolerance", "1e-10"));
    o.relative = in.has("--relative") ? 1 : 0;
    double rn = 0;
    const double t0 = now();
    check(tbsla_hip_cg(m.handle(), dv.get(), dr.get(), &o, &nb_iter, &rn, nullptr));
    t_op = now() - t0;
    flops = (2.0 * g.nnz + 10.0 * g.n_col) * (nb_iter + 1);
  }
  const double t_app_end = now();
  std::printf(
      "{\"compiler\":\"hipcc gfx950\",\"gbs\":\"%f\",\"gflops\":\"%f\",\"lang\":\"HIP\","
      "\"matrix_format\":\"%s\",\"matrix_type\":\"%s\",\"n_col\":\"%d\",\"n_row\":\"%d\","
      "\"nnz\":\"%ld\",\"nb_iterations\":\"%d\",\"numa_init\":\"false\",\"test\":\"%s\","
      "\"time_app_in\":\"%f\",\"time_op\":\"%.9f\",\"vectorization\":\"HIP gfx950\","
      "\"%s_c\":\"%d\",\"cqmat_q\":\"%f\",\"cqmat_s\":\"%u\",\"GR\":\"%d\",\"GC\":\"%d\","
      "\"pr\":\"%d\",\"pc\":\"%d\"}\n",
      spmv_like && (op == "spmv" || op == "Ax") ? bytes_per_spmv(g) / t_op / 1e9 : 0.0,
      flops / t_op / 1e9, format.c_str(), matrix.c_str(), g.n_col, g.n_row, (long)g.nnz, nb_iter,
      op.c_str(), t_app_end - t_app, t_op, matrix.c_str(), C, Q, S, GR, GC, pr, pc);
  return 0;
}
