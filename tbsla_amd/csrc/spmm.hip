// spmm.hip -- C = A B (MatrixCSR::dense_multiply) and personalised PageRank
// over k teleport vectors as one SpMM per iteration.
//
// One wavefront per row, lane t owning columns t, t+64, ... of the dense
// operands (row-major n x k).  A row's (column, value) pairs are
// wave-uniform and come through the scalar cache; each nonzero reads one
// contiguous k-double row of B.  Products are added in the row's storage
// order with separate multiply and add (no FMA), so C is bit-identical to
// the reference's loop (src/mpi/matrix/MatrixCSR.cpp:11-30).
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

constexpr int kMaxTiles = 4;  // k <= 256

template <int KT>
__global__ __launch_bounds__(kBlock) void k_csr_spmm(long long n_rows, int k, int f_col, int ln_col,
                                                     const int *__restrict__ rowptr,
                                                     const int *__restrict__ colidx,
                                                     const double *__restrict__ vals,
                                                     const double *__restrict__ B,
                                                     double *__restrict__ C) {
  TBSLA_NO_FMA
  const long long row = __builtin_amdgcn_readfirstlane(
      (int)(((long long)blockIdx.x * kBlock + threadIdx.x) / kWave));
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= n_rows) return;
  double acc[KT];
#pragma unroll
  for (int t = 0; t < KT; t++) acc[t] = 0.0;
  const int beg = rowptr[row], end = rowptr[row + 1];
  for (int j = beg; j < end; j++) {
    const int c = colidx[j] - f_col;
    if (c < 0 || c >= ln_col) continue;  // block filter (:19)
    const double v = vals[j];
    const double *b = B + (size_t)c * k;
#pragma unroll
    for (int t = 0; t < KT; t++) {
      const int col = lane + t * kWave;
      if (col < k) acc[t] = acc[t] + v * b[col];
    }
  }
  double *out = C + (size_t)row * k;
#pragma unroll
  for (int t = 0; t < KT; t++) {
    const int col = lane + t * kWave;
    if (col < k) out[col] = acc[t];
  }
}

struct PprState {
  int iters;
  int done;
  double max_err;
};

// One personalised-PageRank step for k = 64 columns fused with the SpMM:
// Y[i,:] = A[i,:] X; X'[i,k] = (beta Y[i,k] + S_k (1-beta) [i == src_k]) / S_k;
// per-block partials of |X' - X| and of X' per column.  The page_rank update
// (src/mpi/matrix/Matrix.cpp:123-141) with a one-hot teleport per column.
__global__ __launch_bounds__(kBlock) void k_ppr_step(long long n_rows, int f_col, int ln_col,
                                                     const int *__restrict__ rowptr,
                                                     const int *__restrict__ colidx,
                                                     const double *__restrict__ vals,
                                                     const double *__restrict__ X,
                                                     double *__restrict__ Xn,
                                                     const int *__restrict__ src,
                                                     const double *__restrict__ S, double beta,
                                                     const PprState *__restrict__ st,
                                                     double *__restrict__ p_err,
                                                     double *__restrict__ p_sum) {
  TBSLA_NO_FMA
  if (st->done) return;
  __shared__ double lds_e[kBlock / kWave][kWave];
  __shared__ double lds_s[kBlock / kWave][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const double Sk = S[lane];
  const int sk = src[lane];
  const double tel = Sk * (1 - beta);
  double e = 0.0, sm = 0.0;
  const long long waves = (long long)gridDim.x * (kBlock / kWave);
  for (long long row = (long long)blockIdx.x * (kBlock / kWave) + wave; row < n_rows;
       row += waves) {
    const int r = __builtin_amdgcn_readfirstlane((int)row);
    double acc = 0.0;
    const int beg = rowptr[r], end = rowptr[r + 1];
    for (int j = beg; j < end; j++) {
      const int c = colidx[j] - f_col;
      if (c < 0 || c >= ln_col) continue;
      acc = acc + vals[j] * X[(size_t)c * kWave + lane];
    }
    double v = beta * acc + ((r + f_col == sk) ? tel : 0.0);
    v = v / Sk;
    const size_t o = (size_t)r * kWave + lane;
    e += fabs(v - X[o]);
    sm += v;
    Xn[o] = v;
  }
  lds_e[wave][lane] = e;
  lds_s[wave][lane] = sm;
  __syncthreads();
  if (wave == 0) {
    double te = 0.0, ts = 0.0;
    for (int w = 0; w < kBlock / kWave; w++) {
      te += lds_e[w][lane];
      ts += lds_s[w][lane];
    }
    p_err[(size_t)blockIdx.x * kWave + lane] = te;
    p_sum[(size_t)blockIdx.x * kWave + lane] = ts;
  }
}

// Column totals in block order; converged when every column's change is
// below epsilon.
__global__ __launch_bounds__(kWave) void k_ppr_finalize(PprState *__restrict__ st,
                                                        const double *__restrict__ p_err,
                                                        const double *__restrict__ p_sum, int nb,
                                                        double *__restrict__ S, double epsilon) {
  if (st->done) return;
  const int lane = threadIdx.x;
  double e = 0.0, sm = 0.0;
  for (int b = 0; b < nb; b++) {
    e += p_err[(size_t)b * kWave + lane];
    sm += p_sum[(size_t)b * kWave + lane];
  }
  S[lane] = sm;
  double m = e;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
  if (lane == 0) {
    st->max_err = m;
    st->iters += 1;
    if (m < epsilon) st->done = 1;
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_d(double *__restrict__ a, long long n, double v) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    a[i] = v;
}

// per-column sums of the final iterate, then X[:, k] /= sum_k
// (pagerank_norma_end, Matrix.cpp:149-171)
__global__ __launch_bounds__(kBlock) void k_col_partial(const double *__restrict__ X, long long n_rows,
                                                        double *__restrict__ part) {
  __shared__ double lds[kBlock / kWave][kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  double sm = 0.0;
  for (long long r = (long long)blockIdx.x * (kBlock / kWave) + wave; r < n_rows;
       r += (long long)gridDim.x * (kBlock / kWave))
    sm += X[(size_t)r * kWave + lane];
  lds[wave][lane] = sm;
  __syncthreads();
  if (wave == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; w++) t += lds[w][lane];
    part[(size_t)blockIdx.x * kWave + lane] = t;
  }
}

__global__ __launch_bounds__(kWave) void k_col_total(const double *__restrict__ part, int nb,
                                                     double *__restrict__ tot) {
  double t = 0.0;
  for (int b = 0; b < nb; b++) t += part[(size_t)b * kWave + threadIdx.x];
  tot[threadIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_col_divide(double *__restrict__ X, long long n,
                                                       const double *__restrict__ tot) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    X[i] /= tot[i % kWave];
}

unsigned stride_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 65536));
}

}  // namespace

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_spmm(tbsla_hip_matrix_t m, const double *d_B, double *d_C, int k,
                              void *stream) {
  if (!m || !d_B || !d_C) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_CSR) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "SpMM needs a CSR block");
  if (k < 1 || k > kMaxTiles * kWave) return fail(TBSLA_HIP_ERR_INVALID, "k must be in 1..256");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long rows = m->ln_row;
  if (rows == 0) return TBSLA_HIP_SUCCESS;
  const unsigned grid = grid_for(rows * kWave);
  const int tiles = (k + kWave - 1) / kWave;
  switch (tiles) {
    case 1: k_csr_spmm<1><<<grid, kBlock, 0, s>>>(rows, k, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values, d_B, d_C); break;
    case 2: k_csr_spmm<2><<<grid, kBlock, 0, s>>>(rows, k, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values, d_B, d_C); break;
    case 3: k_csr_spmm<3><<<grid, kBlock, 0, s>>>(rows, k, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values, d_B, d_C); break;
    default: k_csr_spmm<4><<<grid, kBlock, 0, s>>>(rows, k, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values, d_B, d_C); break;
  }
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_personalized_page_rank(tbsla_hip_matrix_t m, double beta, double epsilon,
                                                int max_iterations, const int *h_sources,
                                                double *d_X, int *nb_iterations,
                                                double *max_error, void *stream) {
  if (!m || !h_sources || !d_X) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_CSR)
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "personalised PageRank needs a CSR block");
  if (m->ln_row != m->ln_col || m->f_row != m->f_col)
    return fail(TBSLA_HIP_ERR_INVALID, "personalised PageRank needs a square diagonal block");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = m->ln_row;
  const unsigned nb = (unsigned)std::max<long long>(1, std::min<long long>(2048, (n + 3) / 4));
  double *buf = nullptr, *part = nullptr, *S = nullptr;
  int *src = nullptr;
  PprState *st = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(s);
    hipFree(buf);
    hipFree(part);
    hipFree(S);
    hipFree(src);
    hipFree(st);
  };
  int rc;
  if ((rc = device_alloc((void **)&buf, (size_t)n * kWave * 8)) ||
      (rc = device_alloc((void **)&part, (size_t)2 * nb * kWave * 8)) ||
      (rc = device_alloc((void **)&S, kWave * 8)) || (rc = device_alloc((void **)&src, kWave * 4)) ||
      (rc = device_alloc((void **)&st, sizeof(PprState)))) {
    cleanup();
    return rc;
  }
  std::vector<double> hS(kWave, (double)n);  // X starts at 1: column sums n
  PprState hst{0, 0, 0.0};
  hipMemcpyAsync(S, hS.data(), kWave * 8, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(src, h_sources, kWave * 4, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(st, &hst, sizeof(hst), hipMemcpyHostToDevice, s);
  k_fill_d<<<stride_grid(n * kWave), kBlock, 0, s>>>(d_X, n * kWave, 1.0);
  double *Xb[2] = {d_X, buf};
  int launched = 0;
  while (launched < max_iterations) {
    const int batch = std::min(16, max_iterations - launched);
    for (int i = 0; i < batch; i++, launched++) {
      k_ppr_step<<<nb, kBlock, 0, s>>>(n, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values,
                                       Xb[launched % 2], Xb[(launched + 1) % 2], src, S, beta, st,
                                       part, part + (size_t)nb * kWave);
      k_ppr_finalize<<<1, kWave, 0, s>>>(st, part, part + (size_t)nb * kWave, (int)nb, S, epsilon);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      cleanup();
      return fail(TBSLA_HIP_ERR_HIP, std::string("personalised PageRank: ") + hipGetErrorString(e));
    }
    hipMemcpyAsync(&hst, st, sizeof(hst), hipMemcpyDeviceToHost, s);
    hipStreamSynchronize(s);
    if (hst.done) break;
  }
  double *X = Xb[hst.iters % 2];
  k_col_partial<<<nb, kBlock, 0, s>>>(X, n, part);
  k_col_total<<<1, kWave, 0, s>>>(part, (int)nb, S);
  k_col_divide<<<stride_grid(n * kWave), kBlock, 0, s>>>(X, n * kWave, S);
  if (X != d_X) hipMemcpyAsync(d_X, X, (size_t)n * kWave * 8, hipMemcpyDeviceToDevice, s);
  hipError_t e = hipGetLastError();
  cleanup();
  if (e != hipSuccess) return fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  if (nb_iterations) *nb_iterations = hst.iters;
  if (max_error) *max_error = hst.max_err;
  return TBSLA_HIP_SUCCESS;
}
