// app_common.hpp -- what the driver apps share: the reference's flag
// parsing (--key value), steady-clock time, the median of repeated runs
// (tbsla_perf_page_rank_mpi.cpp compute_median), the matrix sources of the
// MPI apps (cdiag, cqmat, or a binary file <folder>/<matrix>.<format>) and
// their output: one JSON object of string values, keys in std::map order.
#pragma once

#include <tbsla/hip/Matrix.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace tbsla_app {

struct Args {
  std::map<std::string, std::string> kv;
  Args(int argc, char **argv) {
    for (int i = 1; i < argc; i++) {
      std::string k = argv[i];
      if (k.rfind("--", 0) == 0)
        kv[k] = (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) ? argv[++i] : "";
    }
  }
  std::string get(const std::string &k, const std::string &d = "") const {
    auto it = kv.find(k);
    return it == kv.end() ? d : it->second;
  }
  bool has(const std::string &k) const { return kv.count(k) > 0; }
};

inline double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline int format_of(const std::string &f) {
  if (f == "COO" || f == "coo") return TBSLA_HIP_COO;
  if (f == "SCOO" || f == "scoo") return TBSLA_HIP_SCOO;
  if (f == "CSR" || f == "csr") return TBSLA_HIP_CSR;
  if (f == "ELL" || f == "ell") return TBSLA_HIP_ELL;
  if (f == "DENSE" || f == "dense") return TBSLA_HIP_DENSE;
  return -1;
}

// the middle value, or the mean of the two middle values
inline double median(std::vector<double> v) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const size_t n = v.size();
  return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

inline void print_json(const std::map<std::string, std::string> &out) {
  std::string s = "{";
  for (auto it = out.begin(); it != out.end(); ++it) {
    if (it != out.begin()) s += ",";
    s += "\"" + it->first + "\":\"" + it->second + "\"";
  }
  std::printf("%s}\n", s.c_str());
}

// The MPI apps' process-grid check (one process here): a GR x GC grid
// needs GR * GC processes and the exchange step between their blocks.
inline bool grid_ok(int GR, int GC) {
  if (GR * GC != 1) {
    std::printf("The number of processes (%d) does not match the grid dimensions (%d x %d = %d).\n",
                1, GR, GC, GR * GC);
    return false;
  }
  return true;
}

// --matrix cdiag (--C), cqmat (--C --Q --S; --values uniform stores V(0, k)
// instead of 1.0), or a binary image <--matrix_folder>/<matrix>.<format> as
// Matrix::write writes it; returns false (message printed) when none fits.
inline bool fill_source(tbsla::hip::Matrix &m, const Args &in, const std::string &matrix,
                        const std::string &format, int dim, std::map<std::string, std::string> &out) {
  if (matrix == "cdiag") {
    const int C = std::stoi(in.get("--C", "8"));
    m.fill_cdiag(dim, dim, C);
    out["cdiag_c"] = std::to_string(C);
  } else if (matrix == "cqmat") {
    const int C = std::stoi(in.get("--C", "8"));
    const double Q = std::stod(in.get("--Q", "0.1"));
    const int S = std::stoi(in.get("--S", "0"));
    if (in.get("--values") == "uniform") m.fill_cqmat_values(dim, dim, C, Q, (unsigned)S, 0);
    else m.fill_cqmat(dim, dim, C, Q, (unsigned)S);
    out["cqmat_c"] = std::to_string(C);
    out["cqmat_q"] = std::to_string(Q);
    out["cqmat_s"] = std::to_string(S);
  } else if (!matrix.empty()) {
    const std::string path = in.get("--matrix_folder", ".") + "/" + matrix + "." + format;
    const int st = tbsla_hip_read_file(m.handle(), path.c_str(), 0, nullptr);
    if (st != TBSLA_HIP_SUCCESS) {
      std::fprintf(stderr, "%s cannot be open! (%s)\n", path.c_str(), tbsla_hip_last_error());
      return false;
    }
  } else {
    std::fprintf(stderr, "No matrix has been given with the parameter --matrix matrix.\n");
    return false;
  }
  return true;
}

// keys every MPI app prints (tbsla_perf_page_rank_mpi.cpp:448-494), for one
// process on one GPU
inline void common_keys(std::map<std::string, std::string> &out, const std::string &format,
                        const std::string &matrix, const tbsla_hip_geometry &g, int GR, int GC,
                        bool numa) {
  out["matrix_format"] = format;
  out["matrix_dim"] = std::to_string(g.n_col);
  out["g_row"] = std::to_string(GR);
  out["g_col"] = std::to_string(GC);
  out["nnz"] = std::to_string((long)g.nnz);
  out["nnz_min"] = std::to_string((long)g.nnz);
  out["nnz_max"] = std::to_string((long)g.nnz);
  out["nnz_stored"] = std::to_string((long long)g.stored);
  out["gnnz"] = std::to_string((long)g.nnz);
  out["processes"] = "1";
  out["lang"] = "HIP";
  out["matrix_type"] = matrix;
  out["compiler"] = "hipcc gfx950";
  out["compile_options"] = "-O3 --offload-arch=gfx950";
  out["vectorization"] = "HIP gfx950";
  if (numa) out["numa_init"] = "true";
}

}  // namespace tbsla_app
