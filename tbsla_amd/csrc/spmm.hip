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
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

constexpr int kMaxTiles = 4;  // columns per launch row: 256; wider B in chunks (grid y)

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
  const int k0 = blockIdx.y * (kMaxTiles * kWave);  // this launch row's column chunk
  double acc[KT];
#pragma unroll
  for (int t = 0; t < KT; t++) acc[t] = 0.0;
  const int beg = rowptr[row], end = rowptr[row + 1];
  int j = beg;
  // E entries per step: their B rows' loads are in flight together; the
  // products are still added in entry order
  constexpr int E = 4;
  for (; j + E <= end; j += E) {
    bool in[E];
    double v[E];
    const double *bp[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int c = colidx[j + e] - f_col;
      in[e] = c >= 0 && c < ln_col;  // block filter (:19)
      v[e] = vals[j + e];
      bp[e] = B + (size_t)(in[e] ? c : 0) * k;
    }
    double xv[E][KT];
#pragma unroll
    for (int e = 0; e < E; e++)
#pragma unroll
      for (int t = 0; t < KT; t++) {
        const int col = k0 + lane + t * kWave;
        xv[e][t] = col < k ? bp[e][col] : 0.0;
      }
#pragma unroll
    for (int e = 0; e < E; e++)
#pragma unroll
      for (int t = 0; t < KT; t++)
        if (in[e]) acc[t] = acc[t] + v[e] * xv[e][t];
  }
  for (; j < end; j++) {
    const int c = colidx[j] - f_col;
    if (c >= 0 && c < ln_col) {
      const double v = vals[j];
      const double *b = B + (size_t)c * k;
#pragma unroll
      for (int t = 0; t < KT; t++) {
        const int col = k0 + lane + t * kWave;
        if (col < k) acc[t] = acc[t] + v * b[col];
      }
    }
  }
  double *out = C + (size_t)row * k;
#pragma unroll
  for (int t = 0; t < KT; t++) {
    const int col = k0 + lane + t * kWave;
    if (col < k) out[col] = acc[t];
  }
}

struct PprState {
  int iters;
  int done;
  double max_err;
};

// Row slots: a workgroup holds kPprSlots rows at a time (kPprLanes lanes
// per row, each lane owning 4 of the 64 columns).  Workgroup b runs on XCD
// b mod 8 (workgroups are dispatched round-robin over the eight XCDs): XCD x
// takes the x-th contiguous eighth of the rows and its slots take those rows
// in turn (slot v of the XCD's W slots: rows v, v + W, ...), so at any
// moment an XCD works on a window of about W consecutive rows, whose
// near-diagonal X rows stay in its L2.  Other grid sizes: every slot of the
// grid in turn.
constexpr int kPprLanes = 16;                  // lanes per row
constexpr int kPprSlots = kBlock / kPprLanes;  // rows in flight per workgroup
constexpr int kPprDepth = 8;                   // entries whose X rows load together
constexpr int kPprGrid = 2048;                 // workgroups of a step

struct RowWalk {
  long long first, stride, end;
};
__device__ inline RowWalk xcd_walk(long long n_rows, int nb, int slot) {
  const int b = blockIdx.x;
  if (nb % 8 != 0) return RowWalk{(long long)b * kPprSlots + slot, (long long)nb * kPprSlots, n_rows};
  const long long per = (n_rows + 7) / 8;
  const long long x0 = min(n_rows, (long long)(b % 8) * per);
  return RowWalk{x0 + (long long)(b / 8) * kPprSlots + slot, (long long)(nb / 8) * kPprSlots,
                 min(n_rows, x0 + per)};
}

constexpr int kPprCols = kWave / kPprLanes;  // columns per lane

// kPprCols consecutive doubles (aligned to their size) in one load / store
template <int N>
struct DVec {
  typedef double type __attribute__((ext_vector_type(N)));
};
template <>
struct DVec<1> {
  typedef double type;
};
typedef DVec<kPprCols>::type pprvec_t;
template <typename VT>
__device__ inline double vget(const VT &v, int q) {
  if constexpr (std::is_same<VT, double>::value) return v;
  else return v[q];
}
template <typename VT>
__device__ inline void vset(VT &v, int q, double x) {
  if constexpr (std::is_same<VT, double>::value) v = x;
  else v[q] = x;
}

// One personalised-PageRank step for k = 64 columns fused with the SpMM:
// Y[i,:] = A[i,:] X; X'[i,k] = (beta Y[i,k] + S_k (1-beta) [i == src_k]) / S_k;
// per-block partials of |X' - X| and of X' per column.  The page_rank update
// (src/mpi/matrix/Matrix.cpp:123-141) with a one-hot teleport per column.
// Each row's entries are read kPprDepth at a time and their X rows loaded
// before any is added; products are added in storage order with separate
// multiply and add (the reference's arithmetic).
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
  __shared__ double lds_e[kPprSlots][kWave];
  __shared__ double lds_s[kPprSlots][kWave];
  const int lane = threadIdx.x & (kPprLanes - 1);
  const int slot = threadIdx.x / kPprLanes;
  const int c0 = kPprCols * lane;  // this lane's columns c0..c0+kPprCols-1
  double e[kPprCols], sm[kPprCols];
#pragma unroll
  for (int q = 0; q < kPprCols; q++) e[q] = sm[q] = 0.0;
  const RowWalk wk = xcd_walk(n_rows, gridDim.x, slot);
  for (long long row = wk.first; row < wk.end; row += wk.stride) {
    const int r = (int)row;
    double acc[kPprCols];
#pragma unroll
    for (int q = 0; q < kPprCols; q++) acc[q] = 0.0;
    const int beg = rowptr[r], end = rowptr[r + 1];
    for (int j = beg; j < end; j += kPprDepth) {
      int c[kPprDepth];
      double v[kPprDepth];
      pprvec_t xb[kPprDepth];
#pragma unroll
      for (int u = 0; u < kPprDepth; u++) {
        const int jj = min(j + u, end - 1);
        c[u] = colidx[jj] - f_col;
        v[u] = vals[jj];
      }
#pragma unroll
      for (int u = 0; u < kPprDepth; u++) {
        const bool ok = c[u] >= 0 && c[u] < ln_col;  // block filter (MatrixCSR.cpp:19)
        xb[u] = *reinterpret_cast<const pprvec_t *>(X + (size_t)(ok ? c[u] : 0) * kWave + c0);
      }
#pragma unroll
      for (int u = 0; u < kPprDepth; u++)
        if (j + u < end && c[u] >= 0 && c[u] < ln_col)
#pragma unroll
          for (int q = 0; q < kPprCols; q++) acc[q] = acc[q] + v[u] * vget(xb[u], q);
    }
    const size_t o = (size_t)r * kWave + c0;
    const pprvec_t old = *reinterpret_cast<const pprvec_t *>(X + o);
    pprvec_t nv;
#pragma unroll
    for (int q = 0; q < kPprCols; q++) {
      const double Sk = S[c0 + q];
      double t = beta * acc[q] + ((r + f_col == src[c0 + q]) ? Sk * (1 - beta) : 0.0);
      t = t / Sk;
      e[q] += fabs(t - vget(old, q));
      sm[q] += t;
      vset(nv, q, t);
    }
    *reinterpret_cast<pprvec_t *>(Xn + o) = nv;
  }
#pragma unroll
  for (int q = 0; q < kPprCols; q++) {
    lds_e[slot][c0 + q] = e[q];
    lds_s[slot][c0 + q] = sm[q];
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    const int col = threadIdx.x;
    double te = 0.0, ts = 0.0;
    for (int w = 0; w < kPprSlots; w++) {
      te += lds_e[w][col];
      ts += lds_s[w][col];
    }
    p_err[(size_t)blockIdx.x * kWave + col] = te;
    p_sum[(size_t)blockIdx.x * kWave + col] = ts;
  }
}

// Sum over nb per-block rows of 64 partials, in block order within each of
// kBlock / kWave quarters, the quarters then in order (a fixed order);
// result for column threadIdx.x (< 64) in thread threadIdx.x.
__device__ inline double block_order_sum(const double *__restrict__ part, int nb, double *lds) {
  const int col = threadIdx.x & (kWave - 1), qt = threadIdx.x / kWave;
  constexpr int kQ = kBlock / kWave;
  const int per = (nb + kQ - 1) / kQ;
  const int b0 = min(nb, qt * per), b1 = min(nb, b0 + per);
  double t = 0.0;
#pragma unroll 8
  for (int b = b0; b < b1; b++) t += part[(size_t)b * kWave + col];
  lds[qt * kWave + col] = t;
  __syncthreads();
  double tot = 0.0;
  if (qt == 0)
    for (int k = 0; k < kQ; k++) tot += lds[k * kWave + col];
  return tot;
}

// Column totals in a fixed order; converged when every column's change is
// below epsilon.
__global__ __launch_bounds__(kBlock) void k_ppr_finalize(PprState *__restrict__ st,
                                                         const double *__restrict__ p_err,
                                                         const double *__restrict__ p_sum, int nb,
                                                         double *__restrict__ S, double epsilon) {
  if (st->done) return;
  __shared__ double lds[2][kBlock];
  const double e = block_order_sum(p_err, nb, lds[0]);
  const double sm = block_order_sum(p_sum, nb, lds[1]);
  if (threadIdx.x >= kWave) return;
  S[threadIdx.x] = sm;
  double m = e;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
  if (threadIdx.x == 0) {
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
  __shared__ double lds[kPprSlots][kWave];
  const int lane = threadIdx.x & (kPprLanes - 1), slot = threadIdx.x / kPprLanes;
  const int c0 = kPprCols * lane;
  double sm[kPprCols];
#pragma unroll
  for (int q = 0; q < kPprCols; q++) sm[q] = 0.0;
  const RowWalk wk = xcd_walk(n_rows, gridDim.x, slot);
  for (long long r = wk.first; r < wk.end; r += wk.stride) {
    const pprvec_t v = *reinterpret_cast<const pprvec_t *>(X + (size_t)r * kWave + c0);
#pragma unroll
    for (int q = 0; q < kPprCols; q++) sm[q] += vget(v, q);
  }
#pragma unroll
  for (int q = 0; q < kPprCols; q++) lds[slot][c0 + q] = sm[q];
  __syncthreads();
  if (threadIdx.x < kWave) {
    double t = 0.0;
    for (int w = 0; w < kPprSlots; w++) t += lds[w][threadIdx.x];
    part[(size_t)blockIdx.x * kWave + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_col_total(const double *__restrict__ part, int nb,
                                                      double *__restrict__ tot) {
  __shared__ double lds[kBlock];
  const double t = block_order_sum(part, nb, lds);
  if (threadIdx.x < kWave) tot[threadIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_col_divide(double *__restrict__ X, long long n,
                                                       const double *__restrict__ tot) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    X[i] /= tot[i & (kWave - 1)];
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
  if (k < 1) return fail(TBSLA_HIP_ERR_INVALID, "k must be at least 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long rows = m->ln_row;
  if (rows == 0) return TBSLA_HIP_SUCCESS;
  if (m->ln_col == 0 || !m->has_arrays || m->stored == 0)  // no products: C = 0
    return launch_fill(d_C, rows * (long long)k, 0.0, s);
  const int chunk = kMaxTiles * kWave;
  const dim3 grid(grid_for(rows * kWave), (unsigned)((k + chunk - 1) / chunk));
  const int tiles = k > chunk ? kMaxTiles : (k + kWave - 1) / kWave;
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
  // kPprGrid workgroups (a multiple of 8: one contiguous eighth of the rows
  // per XCD) unless the block has fewer rows than slots
  const unsigned nb = n >= (long long)kPprSlots * kPprGrid
                          ? (unsigned)kPprGrid
                          : (unsigned)std::max<long long>(1, (n + kPprSlots - 1) / kPprSlots);
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
      k_ppr_finalize<<<1, kBlock, 0, s>>>(st, part, part + (size_t)nb * kWave, (int)nb, S, epsilon);
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
  k_col_total<<<1, kBlock, 0, s>>>(part, (int)nb, S);
  k_col_divide<<<stride_grid(n * kWave), kBlock, 0, s>>>(X, n * kWave, S);
  if (X != d_X) hipMemcpyAsync(d_X, X, (size_t)n * kWave * 8, hipMemcpyDeviceToDevice, s);
  hipError_t e = hipGetLastError();
  cleanup();
  if (e != hipSuccess) return fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  if (nb_iterations) *nb_iterations = hst.iters;
  if (max_error) *max_error = hst.max_err;
  return TBSLA_HIP_SUCCESS;
}
