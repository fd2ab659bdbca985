// tbsla_perf_page_rank_hip -- src/mpi/apps/tbsla_perf_page_rank_mpi.cpp on one GPU.
//
// Flags as the reference (:204-272): --format (required), --matrix cdiag|
// cqmat|<name> (a binary image <--matrix_folder>/<name>.<format>),
// --matrix_dim (1024), --C, --Q, --S, --GR/--GC (1; a grid needs one
// process per block and the exchange step between them, so more than one
// block is refused as the reference refuses a process count that does not
// match, exit 99), --beta (0.85), --epsilon (1e-5), --max-iterations
// (10000), --numa-init (accepted; device memory has no host placement).
// Extension: --values uniform (cqmat values V(0, k) instead of 1.0).
//
// make_stochastic, then 10 timed runs of page_rank (the damped iteration of
// src/mpi/matrix/Matrix.cpp:540-602; the reference app calls
// page_rank_opticom, which drops the damping, :417 and Matrix.cpp:745-757);
// time_op and gflops are the medians of the 10 runs, gflops =
// ((2 nnz + 2 n) iterations + 2 n) / time (:145-152) with the nnz field, as
// the reference; gflops_stored uses the entries the block holds.
#include "app_common.hpp"

int main(int argc, char **argv) {
  using namespace tbsla_app;
  Args in(argc, argv);
  const std::string format = in.get("--format");
  const int fmt = format_of(format);
  if (fmt < 0) {
    std::fprintf(stderr, "A file format has to be given with the parameter --format format\n");
    return 1;
  }
  const int dim = std::stoi(in.get("--matrix_dim", "1024"));
  const std::string gr = in.get("--GR", "1"), gc = in.get("--GC", "1");
  const double beta = std::stod(in.get("--beta", "0.85"));
  const double eps = std::stod(in.get("--epsilon", "0.00001"));
  const int max_it = std::stoi(in.get("--max-iterations", "10000"));
  if (!grid_ok(std::stoi(gr), std::stoi(gc))) return 99;
  const std::string matrix = in.get("--matrix");
  std::map<std::string, std::string> out;
  tbsla::hip::Matrix m(fmt);
  const double t_app = now();
  try {
    if (!fill_source(m, in, matrix, format, dim, out)) return 1;
    const tbsla_hip_geometry g = m.geometry();
    tbsla::hip::DeviceArray<double> s((size_t)std::max(g.ln_col, 1)), b((size_t)std::max(g.n_col, 1));
    tbsla::hip::check(tbsla_hip_make_stochastic(m.handle(), s.get(), nullptr));
    tbsla::hip::check(tbsla_hip_stream_synchronize(nullptr));
    const double t_op_start = now();
    std::vector<double> runtimes, gflops, gflops_st;
    int total = 0;
    for (int run = 0; run < 10; run++) {
      int it = 0;
      double err = 0;
      const double t0 = now();
      tbsla::hip::check(tbsla_hip_page_rank(m.handle(), beta, eps, max_it, b.get(), &it, &err, nullptr));
      const double rt = now() - t0;
      runtimes.push_back(rt);
      const double n = g.n_col;
      gflops.push_back(((2.0 * g.nnz + 2.0 * n) * it + 2.0 * n) / (rt * 1e9));
      gflops_st.push_back(((2.0 * g.stored + 2.0 * n) * it + 2.0 * n) / (rt * 1e9));
      total += it;
    }
    common_keys(out, format, matrix, g, std::stoi(gr), std::stoi(gc), in.has("--numa-init"));
    out["test"] = "page_rank";
    out["nb_iterations"] = std::to_string(total);
    out["time_op"] = std::to_string(median(runtimes));
    out["gflops"] = std::to_string(median(gflops));
    out["gflops_stored"] = std::to_string(median(gflops_st));
    out["time_gen_mat"] = std::to_string(t_op_start - t_app);
    out["time_app_in"] = std::to_string(now() - t_app);
    out["beta"] = std::to_string(beta);
    out["epsilon"] = std::to_string(eps);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  print_json(out);
  return 0;
}
