// tbsla_perf_cg_hip -- src/mpi/apps/tbsla_perf_cg_mpi.cpp on one GPU.
//
// Flags as the reference (:26-112): --format (required; the solver runs on
// CSR blocks, other formats are refused), --matrix cdiag|cqmat|<name> (a
// binary image <--matrix_folder>/<name>.<format>), --matrix_dim (1024),
// --C, --Q, --S, --GR/--GC (1; more than one block is refused, exit 99, as
// the reference refuses a process count that does not match),
// --max-iterations (10000), --numa-init (accepted).  The right-hand side is
// all ones and x starts at zero (:117-122); the test is the reference's
// absolute |r| < 1e-10 (CG.cpp:34) with its Ax call corrected (:61).
// Extensions: --matrix poisson (7-point Poisson on a --side^3 grid) and
// sym_cqmat ((B + B^T)/2 of a cqmat, --C --Q --S, diagonal row-abs-sum + 1),
// SPD problems on which CG converges; --precision fp64|mixed,
// --preconditioner none|jacobi, --tolerance, --relative, --replace-every.
// Output: the reference's keys (:134-176) plus status and res_norm.
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
  if (fmt != TBSLA_HIP_CSR) {
    std::fprintf(stderr, "%s: the CG solver runs on CSR blocks\n", format.c_str());
    return 1;
  }
  int dim = std::stoi(in.get("--matrix_dim", "1024"));
  const std::string gr = in.get("--GR", "1"), gc = in.get("--GC", "1");
  const int max_it = std::stoi(in.get("--max-iterations", "10000"));
  if (!grid_ok(std::stoi(gr), std::stoi(gc))) return 99;
  const std::string matrix = in.get("--matrix");
  std::map<std::string, std::string> out;
  tbsla::hip::Matrix m(fmt);
  const double t_app = now();
  try {
    if (matrix == "poisson") {
      const int side = std::stoi(in.get("--side", "64"));
      tbsla::hip::check(tbsla_hip_fill_poisson3d(m.handle(), side, side, side, nullptr, nullptr));
      out["poisson_side"] = std::to_string(side);
    } else if (matrix == "sym_cqmat") {
      const int C = std::stoi(in.get("--C", "8"));
      const double Q = std::stod(in.get("--Q", "0.1"));
      const int S = std::stoi(in.get("--S", "1"));
      tbsla::hip::check(tbsla_hip_fill_sym_cqmat(m.handle(), dim, C, Q, (unsigned)S, 0, nullptr));
      out["cqmat_c"] = std::to_string(C);
      out["cqmat_q"] = std::to_string(Q);
      out["cqmat_s"] = std::to_string(S);
    } else if (!fill_source(m, in, matrix, format, dim, out)) {
      return 1;
    }
    const tbsla_hip_geometry g = m.geometry();
    if (g.ln_row != g.ln_col) {
      std::fprintf(stderr, "CG needs a square block\n");
      return 1;
    }
    const size_t n = (size_t)g.ln_row;
    std::vector<double> ones(n, 1.0), zeros(n, 0.0);
    tbsla::hip::DeviceArray<double> v(ones.data(), n), r(zeros.data(), n);
    tbsla_hip_cg_options o{};
    o.max_iterations = max_it;
    o.tolerance = std::stod(in.get("--tolerance", "1e-10"));
    o.relative = in.has("--relative") ? 1 : 0;
    o.precision = in.get("--precision", "fp64") == "mixed" ? TBSLA_HIP_CG_MIXED : TBSLA_HIP_CG_FP64;
    o.preconditioner =
        in.get("--preconditioner", "none") == "jacobi" ? TBSLA_HIP_PRECOND_JACOBI : TBSLA_HIP_PRECOND_NONE;
    o.replace_every = std::stoi(in.get("--replace-every", o.precision == TBSLA_HIP_CG_MIXED ? "50" : "0"));
    tbsla::hip::check(tbsla_hip_stream_synchronize(nullptr));
    int it = 0;
    double rn = 0;
    const double t_op_start = now();
    const int st = tbsla_hip_cg(m.handle(), v.get(), r.get(), &o, &it, &rn, nullptr);
    const double t_op_end = now();
    if (st != TBSLA_HIP_SUCCESS && st != TBSLA_HIP_ERR_BREAKDOWN) tbsla::hip::check(st);
    common_keys(out, format, matrix, g, std::stoi(gr), std::stoi(gc), in.has("--numa-init"));
    out["test"] = "conjugate_gradient";
    out["nb_iterations"] = std::to_string(it);
    out["time_op"] = std::to_string(t_op_end - t_op_start);
    out["time_gen_mat"] = std::to_string(t_op_start - t_app);
    out["time_app_in"] = std::to_string(now() - t_app);
    out["status"] = st == TBSLA_HIP_SUCCESS ? "ok" : "breakdown";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.6e", rn);
    out["res_norm"] = buf;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  print_json(out);
  return 0;
}
