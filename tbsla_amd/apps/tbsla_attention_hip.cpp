// tbsla_attention_hip -- the reference's sparse attention benchmark on one GPU.
//
// Same flags and JSON output as src/mpi/apps/tbsla_attention_mpi.cpp
// (main :126-277) with one process (world_size 1, so the row reductions
// reduce_row_max_abs / reduce_row_sums and row_sum_reduction_for_dense_multiply
// are identities):
//   --matrix_dim n (1024) --NNZ entries per row (10) --cols_B k (512)
//   --base b (-1: exp) --nb_multiplication m (1) --skip_softmax
//   --skip_multiplication
// A = an n x n CSR block; softmax over its rows (row max of |z|, exponential,
// row sums, normalisation: :183-191); then m times C = A B with B n x k
// filled as fill_matrix_by_blocks (:95-112) does for one block (every entry
// 1.0).  The reference fills A with fill_random, whose entries depend on the
// OpenMP thread count (MatrixCSR.cpp:906-925); here A is fill_cqmat with NNZ
// entries per row and column-randomisation probability --Q (default 1.0:
// every column drawn at random), seed --S, values V(--values_seed, k)
// uniform in [0, 1) so that the softmax is not uniform.  --separate runs the
// four row operations as four calls instead of the fused kernel.
// Extra keys: "check" = max |C_ij - 1| over rows with entries (softmax rows
// sum to 1, and B is all ones), and GB/s of each phase.
#include <tbsla/hip/Matrix.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace {

struct Args {
  std::map<std::string, std::string> kv;
  Args(int argc, char **argv) {
    for (int i = 1; i < argc; i++) {
      std::string k = argv[i];
      if (k.rfind("--", 0) == 0)
        kv[k] = (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) ? argv[++i] : "";
    }
  }
  std::string get(const std::string &k, const std::string &d) const {
    auto it = kv.find(k);
    return it == kv.end() ? d : it->second;
  }
  bool has(const std::string &k) const { return kv.count(k) > 0; }
};

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void sync() { tbsla::hip::check(tbsla_hip_stream_synchronize(nullptr)); }

}  // namespace

int main(int argc, char **argv) {
  using namespace tbsla::hip;
  Args in(argc, argv);
  const int matrix_dim = std::stoi(in.get("--matrix_dim", "1024"));
  const int nnz_per_row = std::stoi(in.get("--NNZ", "10"));
  const int cols_B = std::stoi(in.get("--cols_B", "512"));
  const int base = std::stoi(in.get("--base", "-1"));
  const int nb_multiplication = std::stoi(in.get("--nb_multiplication", "1"));
  const bool skip_softmax = in.has("--skip_softmax");
  const bool skip_multiplication = in.has("--skip_multiplication");
  const double Q = std::stod(in.get("--Q", "1.0"));
  const unsigned S = (unsigned)std::stoi(in.get("--S", "0"));
  const long long value_seed = std::stoll(in.get("--values_seed", "1"));
  const bool separate = in.has("--separate");
  const int world = 1;

  MatrixCSR m;
  double t0 = now();
  m.fill_cqmat_values(matrix_dim, matrix_dim, nnz_per_row, Q, S, value_seed, 0, 0, 1, 1);
  sync();
  const double t_init = now() - t0;
  const tbsla_hip_geometry g = m.geometry();
  const size_t n = (size_t)g.ln_row;

  double t_softmax = 0;
  if (!skip_softmax) {
    DeviceArray<double> dmax(n), ds(n);
    t0 = now();
    if (separate) {
      check(tbsla_hip_get_row_max_abs(m.handle(), dmax.get(), nullptr));
      check(tbsla_hip_apply_exponential(m.handle(), dmax.get(), base, nullptr));
      check(tbsla_hip_get_row_sums(m.handle(), ds.get(), nullptr));
      check(tbsla_hip_normalize_rows(m.handle(), ds.get(), nullptr));
    } else {
      check(tbsla_hip_softmax_rows(m.handle(), base, dmax.get(), ds.get(), nullptr));
    }
    sync();
    t_softmax = now() - t0;
  }

  // B: fill_matrix_by_blocks with one block -> every entry 1.0
  const size_t nbv = (size_t)g.ln_col * (size_t)cols_B, ncv = n * (size_t)cols_B;
  std::vector<double> B(nbv, 1.0);
  t0 = now();
  DeviceArray<double> dB(B.data(), nbv), dC(ncv);
  sync();
  const double t_distribute = now() - t0;

  double t_multiply = 0;
  double check_err = 0;
  if (!skip_multiplication) {
    t0 = now();
    for (int i = 0; i < nb_multiplication; i++) m.dense_multiply_device(dB.get(), dC.get(), cols_B);
    sync();
    t_multiply = now() - t0;
    std::vector<double> C(ncv);
    dC.to_host(C.data(), ncv);
    std::vector<int> rp(n + 1), ci((size_t)g.stored);
    std::vector<double> va((size_t)g.stored);
    check(tbsla_hip_export(m.handle(), rp.data(), ci.data(), va.data(), nullptr));
    for (size_t i = 0; i < n; i++) {
      if (rp[i + 1] == rp[i]) continue;
      double rs = 0;
      for (int j = rp[i]; j < rp[i + 1]; j++) rs += va[j];
      if (rs == 0) continue;  // a row the softmax zeroed (or --skip_softmax)
      for (int k = 0; k < cols_B; k++)
        check_err = std::max(check_err, std::fabs(C[i * (size_t)cols_B + k] - (skip_softmax ? rs : 1.0)));
    }
  }

  const double nnz = (double)g.stored;
  const double gflops_mult =
      t_multiply > 0 ? 2.0 * nnz_per_row * (double)matrix_dim * cols_B * nb_multiplication / (t_multiply * 1e9) : 0;
  const double gflops_soft = t_softmax > 0 ? 6.0 * nnz_per_row * (double)matrix_dim / (t_softmax * 1e9) : 0;
  // bytes: softmax reads rowptr and the values and writes the values; the
  // product reads A once, one B row per entry and writes C
  const double soft_bytes = 4.0 * (n + 1) + 16.0 * nnz;
  const double mult_bytes = (12.0 * nnz + 4.0 * (n + 1) + 8.0 * nnz * cols_B + 8.0 * ncv) * nb_multiplication;
  std::printf(
      "{\"parameters\": {\"matrix_dim\": %d,\"nnz_per_row\": %d,\"cols_B\": %d,\"base\": %d,"
      "\"world_size\": %d,\"nb_multiplication\": %d},"
      "\"timings\": {\"initialization\": %f,\"softmax_operations\": %f,"
      "\"initial_dense_matrix_distribution\": %f,\"multiplication\": %f,\"finalization\": %f},"
      "\"gflops\": {\"softmax_operations\": %f,\"multiplication\": %f,\"total\": %f},"
      "\"nnz\": %.0f,\"fused\": %s,\"check\": %.3e,"
      "\"gbs\": {\"softmax_operations\": %f,\"multiplication\": %f},\"lang\": \"HIP\"}\n",
      matrix_dim, nnz_per_row, cols_B, base, world, nb_multiplication, t_init, t_softmax,
      t_distribute, t_multiply, 0.0, gflops_soft, gflops_mult,
      skip_softmax ? 0.0 : gflops_soft + gflops_mult, nnz, separate ? "false" : "true", check_err,
      t_softmax > 0 ? soft_bytes / t_softmax / 1e9 : 0.0,
      t_multiply > 0 ? mult_bytes / t_multiply / 1e9 : 0.0);
  return 0;
}
