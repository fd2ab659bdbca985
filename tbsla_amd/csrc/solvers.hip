// solvers.hip -- column normalisation and PageRank on one block.
//
// Deterministic by construction: every reduction writes one partial per
// workgroup of a fixed-size grid and a one-workgroup finaliser combines them
// in a fixed order, so repeated runs give identical bits.  Iterations are
// queued in batches without host round-trips: finalisers keep the solver
// scalars (sums, iteration count, convergence flag) in device memory, and
// every kernel of an iteration returns at once when the flag says the solve
// has finished.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

constexpr int kRedBlocks = 512;  // reduction grid (fixed: the partial layout is part of the order)

struct SolverState {
  double S;      // sum of the current iterate
  double err;    // last L1 change
  int iters;     // iterations done (reference counting)
  int done;      // set when the solve has finished
};

__device__ inline double block_sum(double v) {
  __shared__ double lds[kBlock / kWave];
  v = group_sum<kWave>(v);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / kWave; w++) t += lds[w];
  __syncthreads();
  return t;  // valid in thread 0
}

// Fixed-order sum of nb partials by one workgroup.
__device__ inline double sum_partials(const double *__restrict__ p, int nb) {
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) v += p[i];
  return block_sum(v);
}

unsigned red_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, kRedBlocks));
}

// ---- column sums --------------------------------------------------------

// key = local column, payload = the entry's position in the reference's loop
// order (CSR/SCOO: storage order; ELL: row-major slot i*max_col+j).  Entries
// the reference skips -- ELL slots whose value is not positive
// (MatrixELL.cpp:812), which include the padding (column 0, possibly left of
// the block) -- and any column outside [0, ln_col) get the key ln_col, a
// bucket after the last column that no sum reads.
__global__ __launch_bounds__(kBlock) void k_colkeys(long long n, int fmt, long long ln_row,
                                                    long long ln_col, int max_col,
                                                    const int *__restrict__ cols,
                                                    const double *__restrict__ vals, int f_col,
                                                    unsigned long long *__restrict__ keys,
                                                    int *__restrict__ pos) {
  long long p = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  long long src = p;
  if (fmt == TBSLA_HIP_ELL) src = (p % max_col) * ln_row + p / max_col;
  long long c = (long long)cols[src] - f_col;
  bool keep = c >= 0 && c < ln_col && (fmt != TBSLA_HIP_ELL || vals[src] > 0);
  keys[p] = keep ? (unsigned long long)c : (unsigned long long)ln_col;
  pos[p] = (int)p;
}

// counts[key]++ for keys <= ln_col (the last bucket collects skipped entries)
__global__ __launch_bounds__(kBlock) void k_count_cols(const unsigned long long *__restrict__ keys,
                                                       long long n, long long ln_col,
                                                       int *__restrict__ counts) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    unsigned long long k = keys[i];
    if (k <= (unsigned long long)ln_col) atomicAdd(counts + k, 1);
  }
}

// s[c] += the column's values in reference order (one thread per column):
// the serial loop of MatrixCSR::get_col_sums (MatrixCSR.cpp:1200-1206) and
// MatrixSCOO::get_col_sums (MatrixSCOO.cpp:719-723); ELL adds positive
// values only (MatrixELL.cpp:809-815).
__global__ __launch_bounds__(kBlock) void k_col_sums(long long ln_col, int fmt, long long ln_row,
                                                     int max_col, const int *__restrict__ colptr,
                                                     const int *__restrict__ pos,
                                                     const double *__restrict__ vals,
                                                     double *__restrict__ s) {
  long long c = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= ln_col) return;
  double acc = s[c];
  for (int j = colptr[c]; j < colptr[c + 1]; j++) {
    long long p = pos[j];
    if (fmt == TBSLA_HIP_ELL) {
      double v = vals[(p % max_col) * ln_row + p / max_col];
      if (v > 0) acc += v;
    } else {
      acc += vals[p];
    }
  }
  s[c] = acc;
}

// normalize_cols: CSR (MatrixCSR.cpp:1221-1231) and SCOO
// (MatrixSCOO.cpp:731-736) divide where s > 0; ELL also needs value > 0
// (MatrixELL.cpp:824-832).
// The value is tested before s is read, so padding slots (column 0, which
// lies left of a block with f_col > 0) never index s.
__global__ __launch_bounds__(kBlock) void k_normalize_cols(long long n, int ell, long long ln_col,
                                                           const int *__restrict__ cols, int f_col,
                                                           double *__restrict__ vals,
                                                           const double *__restrict__ s) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    double v = vals[i];
    if (ell && !(v > 0)) continue;
    long long c = (long long)cols[i] - f_col;
    if (c < 0 || c >= ln_col) continue;
    double sv = s[c];
    if (sv > 0) vals[i] = v / sv;
  }
}

// ---- PageRank -----------------------------------------------------------

// b = (beta * y|window + S (1 - beta) / n_col) / S; partials of |b - b_t|
// and of sum(b) (Matrix.cpp:123-141 with the copy of :289-291).
__global__ __launch_bounds__(kBlock) void k_pr_update(long long ln_col, const double *__restrict__ y,
                                                      const double *__restrict__ bt,
                                                      double *__restrict__ b, long long r_start,
                                                      long long s_start, long long length,
                                                      double beta, double n_col,
                                                      const SolverState *__restrict__ st,
                                                      double *__restrict__ p_err,
                                                      double *__restrict__ p_sum) {
  TBSLA_NO_FMA
  if (st->done) return;
  const double S = st->S;
  const double t_sum = S * ((1 - beta) / n_col);  // t_sum *= (1-beta)/n_col
  double e = 0.0, sm = 0.0;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < ln_col;
       k += (long long)gridDim.x * kBlock) {
    double v = (k >= r_start && k < r_start + length) ? y[s_start + k - r_start] : 0.0;
    v = beta * v + t_sum;
    v = v / S;
    b[k] = v;
    e += fabs(v - bt[k]);
    sm += v;
  }
  e = block_sum(e);
  sm = block_sum(sm);
  if (threadIdx.x == 0) {
    p_err[blockIdx.x] = e;
    p_sum[blockIdx.x] = sm;
  }
}

__global__ __launch_bounds__(kBlock) void k_pr_finalize(SolverState *__restrict__ st,
                                                        const double *__restrict__ p_err,
                                                        const double *__restrict__ p_sum, int nb,
                                                        int NC, double epsilon) {
  if (st->done) return;
  double e = sum_partials(p_err, nb);
  double sm = sum_partials(p_sum, nb);
  if (threadIdx.x == 0) {
    e /= NC;
    st->err = e;
    st->S = sm;
    st->iters += 1;
    if (e < epsilon) st->done = 1;
  }
}

__global__ __launch_bounds__(kBlock) void k_partial_sum(const double *__restrict__ a, long long n,
                                                        double *__restrict__ part) {
  double v = 0.0;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < n;
       k += (long long)gridDim.x * kBlock)
    v += a[k];
  v = block_sum(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_total(const double *__restrict__ part, int nb,
                                                  double *__restrict__ out) {
  double v = sum_partials(part, nb);
  if (threadIdx.x == 0) *out = v;
}

__global__ __launch_bounds__(kBlock) void k_divide(double *__restrict__ a, long long n,
                                                   const double *__restrict__ by) {
  const double d = *by;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < n;
       k += (long long)gridDim.x * kBlock)
    a[k] /= d;
}

// ---- max-normalised PageRank ---------------------------------------------
//
// tbsla::mpi::MatrixCOO::page_rank (src/mpi/matrix/MatrixCOO.cpp:108-175)
// on one block covering a square matrix: b = 1; per iteration b_t = b,
// y = A b_t, T = (T + sum(b_t)) (1 - beta) / n (the reduction adds into the
// previous iteration's T, :131-135; T starts at 0 here, the reference leaves
// it uninitialised), b = beta y + T, b /= max(b), err = sum |b - b_t|;
// finally b /= sum(b).  max(b) is taken from the extremes of y:
// fl(beta y + T) is monotone in y (nondecreasing for beta >= 0,
// nonincreasing otherwise), so max_k fl(beta y_k + T) is fl(beta y_max + T)
// (or y_min), the same double the reference's loop finds.

struct PrmState {
  double S;     // sum of the current iterate b_t
  double T;     // teleportation term of the current iteration
  double maxv;  // max of beta y + T
  double err;   // last L1 change
  int iters;
  int done;
};

// partial max and min of y
__global__ __launch_bounds__(kBlock) void k_prm_extremes(const double *__restrict__ y, long long n,
                                                         const PrmState *__restrict__ st,
                                                         double *__restrict__ p_max,
                                                         double *__restrict__ p_min) {
  __shared__ double lmax[kBlock / kWave], lmin[kBlock / kWave];
  if (st->done) return;
  double hi = -INFINITY, lo = INFINITY;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < n;
       k += (long long)gridDim.x * kBlock) {
    const double v = y[k];
    hi = fmax(hi, v);
    lo = fmin(lo, v);
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    hi = fmax(hi, __shfl_xor(hi, o, kWave));
    lo = fmin(lo, __shfl_xor(lo, o, kWave));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    lmax[threadIdx.x / kWave] = hi;
    lmin[threadIdx.x / kWave] = lo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; w++) {
      hi = fmax(hi, lmax[w]);
      lo = fmin(lo, lmin[w]);
    }
    p_max[blockIdx.x] = hi;
    p_min[blockIdx.x] = lo;
  }
}

__global__ __launch_bounds__(kBlock) void k_prm_max(PrmState *__restrict__ st,
                                                    const double *__restrict__ p_max,
                                                    const double *__restrict__ p_min, int nb,
                                                    double beta) {
  TBSLA_NO_FMA
  if (st->done || threadIdx.x != 0) return;
  double hi = -INFINITY, lo = INFINITY;
  for (int i = 0; i < nb; i++) {
    hi = fmax(hi, p_max[i]);
    lo = fmin(lo, p_min[i]);
  }
  const double e = beta >= 0 ? hi : lo;
  st->maxv = beta * e + st->T;
}

// b = (beta y + T) / maxv; partials of |b - b_t| and of sum(b)
__global__ __launch_bounds__(kBlock) void k_prm_update(long long n, const double *__restrict__ y,
                                                       const double *__restrict__ bt,
                                                       double *__restrict__ b, double beta,
                                                       const PrmState *__restrict__ st,
                                                       double *__restrict__ p_err,
                                                       double *__restrict__ p_sum) {
  TBSLA_NO_FMA
  if (st->done) return;
  const double T = st->T, mx = st->maxv;
  double e = 0.0, sm = 0.0;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < n;
       k += (long long)gridDim.x * kBlock) {
    double v = beta * y[k] + T;
    v = v / mx;
    b[k] = v;
    e += fabs(v - bt[k]);
    sm += v;
  }
  e = block_sum(e);
  sm = block_sum(sm);
  if (threadIdx.x == 0) {
    p_err[blockIdx.x] = e;
    p_sum[blockIdx.x] = sm;
  }
}

__global__ __launch_bounds__(kBlock) void k_prm_finalize(PrmState *__restrict__ st,
                                                         const double *__restrict__ p_err,
                                                         const double *__restrict__ p_sum,
                                                         int nb, double scale, double epsilon) {
  TBSLA_NO_FMA
  if (st->done) return;
  const double e = sum_partials(p_err, nb);
  const double sm = sum_partials(p_sum, nb);
  if (threadIdx.x == 0) {
    st->err = e;
    st->S = sm;
    st->T = (st->T + sm) * scale;  // the next iteration's T (:131-135)
    st->iters += 1;
    if (e < epsilon) st->done = 1;
  }
}

// ---- batched PageRank ----------------------------------------------------
//
// Instances are the diagonal blocks of one square CSR block.  Instance k
// owns rows [row0, row0 + n) and the partial slots [part0, part0 + nb) with
// nb = red_grid(n): its update workgroups walk its rows exactly as the
// single-block update walks a block of n rows, and its finaliser sums its
// partials in the same order, so each instance follows the single-block
// solve of its own block bit for bit (given the same SpMV row kernel).

struct PrInstance {
  double S;     // sum of the current iterate
  double err;   // last L1 change
  int iters;    // iterations done
  int done;     // converged or at the iteration cap
  long long row0, n;
  int part0, nb;
  long long t_done;  // device wall clock at the test that finished the instance
};

struct PrCommon {
  int ndone;     // instances finished
  int all_done;  // every instance finished: the SpMV row kernel returns at once
  long long t0;  // device wall clock before the first iteration
};

__global__ void k_prb_start(PrCommon *__restrict__ com) { com->t0 = wall_clock64(); }

__global__ __launch_bounds__(kBlock) void k_prb_update(PrInstance *__restrict__ inst,
                                                       const int *__restrict__ block_inst,
                                                       const double *__restrict__ y,
                                                       const double *__restrict__ bt,
                                                       double *__restrict__ bn, double beta,
                                                       double *__restrict__ p_err,
                                                       double *__restrict__ p_sum) {
  TBSLA_NO_FMA
  const int gb = blockIdx.x;
  const PrInstance &I = inst[block_inst[gb]];
  if (I.done) return;
  const int b = gb - I.part0;
  const double S = I.S;
  const double t_sum = S * ((1 - beta) / (double)I.n);
  double e = 0.0, sm = 0.0;
  for (long long k = (long long)b * kBlock + threadIdx.x; k < I.n; k += (long long)I.nb * kBlock) {
    const long long i = I.row0 + k;
    double v = y[i];
    v = beta * v + t_sum;
    v = v / S;
    bn[i] = v;
    e += fabs(v - bt[i]);
    sm += v;
  }
  e = block_sum(e);
  sm = block_sum(sm);
  if (threadIdx.x == 0) {
    p_err[gb] = e;
    p_sum[gb] = sm;
  }
}

// one workgroup per instance
__global__ __launch_bounds__(kBlock) void k_prb_finalize(PrInstance *__restrict__ inst, int ninst,
                                                         const double *__restrict__ p_err,
                                                         const double *__restrict__ p_sum,
                                                         double epsilon, int max_iterations,
                                                         PrCommon *__restrict__ com,
                                                         int *__restrict__ done_flags) {
  PrInstance &I = inst[blockIdx.x];
  if (I.done) return;
  double e = sum_partials(p_err + I.part0, I.nb);
  double sm = sum_partials(p_sum + I.part0, I.nb);
  if (threadIdx.x == 0) {
    e /= 1;  // NC = 1: one block per instance
    I.err = e;
    I.S = sm;
    I.iters += 1;
    if (e < epsilon || I.iters >= max_iterations) {
      I.done = 1;
      I.t_done = wall_clock64();
      done_flags[blockIdx.x] = 1;  // the SpMV skips the instance's rows from now on
      if (atomicAdd(&com->ndone, 1) + 1 == ninst) com->all_done = 1;
    }
  }
}

// sum of each instance's final iterate (buffer iters % 2), per-workgroup partials
__global__ __launch_bounds__(kBlock) void k_prb_norm_partial(const PrInstance *__restrict__ inst,
                                                             const int *__restrict__ block_inst,
                                                             const double *__restrict__ b0,
                                                             const double *__restrict__ b1,
                                                             double *__restrict__ part) {
  const int gb = blockIdx.x;
  const PrInstance &I = inst[block_inst[gb]];
  const double *src = (I.iters % 2) ? b1 : b0;
  const int b = gb - I.part0;
  double v = 0.0;
  for (long long k = (long long)b * kBlock + threadIdx.x; k < I.n; k += (long long)I.nb * kBlock)
    v += src[I.row0 + k];
  v = block_sum(v);
  if (threadIdx.x == 0) part[gb] = v;
}

__global__ __launch_bounds__(kBlock) void k_prb_norm_total(const PrInstance *__restrict__ inst,
                                                           const double *__restrict__ part,
                                                           double *__restrict__ total) {
  const PrInstance &I = inst[blockIdx.x];
  const double v = sum_partials(part + I.part0, I.nb);
  if (threadIdx.x == 0) total[blockIdx.x] = v;
}

// out = final iterate / its sum (pagerank_norma_end, Matrix.cpp:149-171)
__global__ __launch_bounds__(kBlock) void k_prb_divide(const PrInstance *__restrict__ inst,
                                                       const int *__restrict__ block_inst,
                                                       const double *__restrict__ b0,
                                                       const double *__restrict__ b1,
                                                       const double *__restrict__ total,
                                                       double *__restrict__ out) {
  const int gb = blockIdx.x;
  const int k = block_inst[gb];
  const PrInstance &I = inst[k];
  const double *src = (I.iters % 2) ? b1 : b0;
  const double d = total[k];
  const int b = gb - I.part0;
  for (long long j = (long long)b * kBlock + threadIdx.x; j < I.n; j += (long long)I.nb * kBlock)
    out[I.row0 + j] = src[I.row0 + j] / d;
}

// ---- CG -----------------------------------------------------------------

unsigned stride_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  int alloc(long long n) { return device_alloc((void **)&p, (size_t)std::max(n, 1LL) * sizeof(T)); }
};

}  // namespace

}  // namespace tbsla_hip

namespace tbsla_hip {

int values_changed(tbsla_hip_matrix_s *m, hipStream_t s) {
  int win = 0, sl = 0, lay = 0;  // the split layouts copy values: rebuild
  if (split_params(m, &win, &sl, &lay)) TBSLA_TRY(split_build(m, win, sl, s, lay));
  m->values_version++;  // copies made from the old values are stale
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_values_changed(tbsla_hip_matrix_t m, void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null handle");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  // the COO/SCOO row pointer is derived from the row array: rebuilt (and the
  // triplets' order checked again) at the next SpMV
  if (m->coo_rowptr) {
    (void)hipFree(m->coo_rowptr);
    m->coo_rowptr = nullptr;
  }
  if (m->format == TBSLA_HIP_COO || m->format == TBSLA_HIP_SCOO) {
    m->row_sorted = 1;
    // a derived layout over the triplets' rows is rebuilt from the new row
    // pointer, or dropped when the triplets are no longer sorted by row
    if (m->split) {
      const long long n = (m->format == TBSLA_HIP_SCOO && m->nnz == 0) ? 0 : m->stored;
      bool ok = false;
      if (m->has_arrays && n > 0) TBSLA_TRY(coo_rowptr(m, n, &ok, s));
      if (!ok) split_release(m);
    }
  }
  return values_changed(m, s);
}

extern "C" int tbsla_hip_get_col_sums(tbsla_hip_matrix_t m, double *d_s, void *stream) {
  if (!m || !d_s) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int fmt = m->format;
  if (fmt == TBSLA_HIP_COO || fmt == TBSLA_HIP_DENSE) return TBSLA_HIP_SUCCESS;  // reference no-ops
  if (m->nnz == 0 || !m->has_arrays) return TBSLA_HIP_SUCCESS;
  const long long n = fmt == TBSLA_HIP_ELL ? (long long)m->ln_row * m->max_col : m->stored;
  const long long ln_col = m->ln_col;
  if (n == 0) return TBSLA_HIP_SUCCESS;
  DevBuf<unsigned long long> keys;
  DevBuf<int> pos, colptr;
  TBSLA_TRY(keys.alloc(n));
  TBSLA_TRY(pos.alloc(n));
  TBSLA_TRY(colptr.alloc(ln_col + 2));  // ln_col columns + the skipped-entry bucket
  k_colkeys<<<grid_for(n), kBlock, 0, s>>>(n, fmt, m->ln_row, ln_col, m->max_col, m->colidx,
                                           m->values, m->f_col, keys.p, pos.p);
  TBSLA_HIP_LAUNCH_CHECK();
  int bits = 1;
  while (bits < 62 && (1LL << bits) <= ln_col) bits++;  // keys <= ln_col < 2^bits
  TBSLA_TRY(radix_sort_pairs(keys.p, pos.p, n, bits, s));
  TBSLA_HIP_CHECK(hipMemsetAsync(colptr.p, 0, (size_t)(ln_col + 2) * 4, s));
  k_count_cols<<<stride_grid(n), kBlock, 0, s>>>(keys.p, n, ln_col, colptr.p + 1);
  TBSLA_HIP_LAUNCH_CHECK();
  DevBuf<char> ws;
  TBSLA_TRY(ws.alloc((long long)scan_workspace_bytes(ln_col + 1) + 64));
  TBSLA_TRY(scan<int>(colptr.p + 1, colptr.p + 1, ln_col + 1, true,
                      reinterpret_cast<long long *>(ws.p + scan_workspace_bytes(ln_col + 1)), ws.p,
                      s));
  if (ln_col > 0) {
    k_col_sums<<<grid_for(ln_col), kBlock, 0, s>>>(ln_col, fmt, m->ln_row, m->max_col, colptr.p,
                                                   pos.p, m->values, d_s);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_normalize_cols(tbsla_hip_matrix_t m, const double *d_s, void *stream) {
  if (!m || !d_s) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int fmt = m->format;
  if (fmt == TBSLA_HIP_COO || fmt == TBSLA_HIP_DENSE) return TBSLA_HIP_SUCCESS;
  if (m->nnz == 0 || !m->has_arrays) return TBSLA_HIP_SUCCESS;
  const long long n = fmt == TBSLA_HIP_ELL ? (long long)m->ln_row * m->max_col : m->stored;
  if (n > 0) {
    k_normalize_cols<<<stride_grid(n), kBlock, 0, s>>>(n, fmt == TBSLA_HIP_ELL, m->ln_col,
                                                       m->colidx, m->f_col, m->values, d_s);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  return values_changed(m, s);
}

extern "C" int tbsla_hip_make_stochastic(tbsla_hip_matrix_t m, double *d_s, void *stream) {
  TBSLA_TRY(tbsla_hip_get_col_sums(m, d_s, stream));
  return tbsla_hip_normalize_cols(m, d_s, stream);
}

extern "C" int tbsla_hip_page_rank(tbsla_hip_matrix_t m, double beta, double epsilon,
                                   int max_iterations, double *d_b, int *nb_iterations,
                                   double *final_error, void *stream) {
  if (!m || !d_b) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long ln_col = m->ln_col, ln_row = m->ln_row;
  const long long ylen = m->format == TBSLA_HIP_COO ? m->n_row : ln_row;
  // redistribute_vector's copy window on one block (Matrix.cpp:266-292)
  long long r_start = 0, s_start = 0, length = 0;
  {
    long long rs = m->f_col, re = (long long)m->f_col + ln_col;
    long long bs = m->f_row, be = (long long)m->f_row + ln_row;
    if (rs <= be && re >= bs) {
      if (bs <= rs) {
        r_start = 0;
        s_start = rs - bs;
      } else {
        r_start = bs - rs;
        s_start = 0;
      }
      length = std::min(be, re) - (rs + r_start);
    }
  }
  // a square block on the diagonal in the diagonal layout's LDS-slice form:
  // SpMV and update in one pass (split_launch_pr), one partial per row block
  const int fused = (r_start == 0 && s_start == 0 && length == ln_col && ln_row == ln_col)
                        ? split_pr_blocks(m)
                        : 0;
  const long long np = std::max<long long>(kRedBlocks, fused);  // partials of each kind
  DevBuf<double> bufA, bufB, y, part;
  DevBuf<SolverState> st;
  TBSLA_TRY(bufA.alloc(ln_col));
  TBSLA_TRY(bufB.alloc(ln_col));
  if (!fused) TBSLA_TRY(y.alloc(ylen));
  TBSLA_TRY(part.alloc(2 * np + 8));
  TBSLA_TRY(st.alloc(1));
  TBSLA_TRY(launch_fill(bufA.p, ln_col, 1.0, s));  // b = 1 (Matrix.cpp:549-552)
  SolverState h{};
  h.S = (double)ln_col;
  TBSLA_HIP_CHECK(hipMemcpyAsync(st.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
  const unsigned nb = red_grid(ln_col);
  double *buf[2] = {bufA.p, bufB.p};
  int launched = 0;
  while (launched < max_iterations) {
    const int batch = std::min(16, max_iterations - launched);
    for (int i = 0; i < batch; i++, launched++) {
      double *bt = buf[launched % 2], *bn = buf[(launched + 1) % 2];
      if (fused) {
        const PrFuse pf{bn, &st.p->S, beta, (double)m->n_col, part.p, part.p + np};
        TBSLA_TRY(split_launch_pr(m, bt, pf, &st.p->done, s));
      } else {
        TBSLA_TRY(spmv_launch(m, y.p, bt, 0, true, s, &st.p->done));
        k_pr_update<<<nb, kBlock, 0, s>>>(ln_col, y.p, bt, bn, r_start, s_start, length, beta,
                                          (double)m->n_col, st.p, part.p, part.p + np);
      }
      k_pr_finalize<<<1, kBlock, 0, s>>>(st.p, part.p, part.p + np, fused ? fused : (int)nb,
                                         m->NC, epsilon);
    }
    TBSLA_HIP_LAUNCH_CHECK();
    TBSLA_HIP_CHECK(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    if (h.done) break;
  }
  double *b = buf[h.iters % 2];
  // pagerank_norma_end (Matrix.cpp:149-171): b /= sum(b)
  k_partial_sum<<<nb, kBlock, 0, s>>>(b, ln_col, part.p);
  k_total<<<1, kBlock, 0, s>>>(part.p, (int)nb, part.p + 2 * np);
  k_divide<<<stride_grid(ln_col), kBlock, 0, s>>>(b, ln_col, part.p + 2 * np);
  TBSLA_HIP_LAUNCH_CHECK();
  TBSLA_HIP_CHECK(hipMemcpyAsync(d_b, b, (size_t)ln_col * 8, hipMemcpyDeviceToDevice, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  if (nb_iterations) *nb_iterations = h.iters;
  if (final_error) *final_error = h.err;
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_page_rank_max_norm(tbsla_hip_matrix_t m, double beta, double epsilon,
                                            int max_iterations, double *d_b, int *nb_iterations,
                                            double *final_error, void *stream) {
  if (!m || !d_b) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  const long long n = m->n_col;
  if (m->n_row != m->n_col ||
      (m->format != TBSLA_HIP_COO && (m->ln_row != m->n_row || m->ln_col != m->n_col)))
    return fail(TBSLA_HIP_ERR_UNSUPPORTED,
                "max-normalised PageRank needs one block covering a square matrix");
  if (n == 0) {  // nothing to iterate (the reference reads b[0])
    if (nb_iterations) *nb_iterations = 0;
    if (final_error) *final_error = 0.0;
    return TBSLA_HIP_SUCCESS;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevBuf<double> bufA, bufB, y, part;
  DevBuf<PrmState> st;
  TBSLA_TRY(bufA.alloc(n));
  TBSLA_TRY(bufB.alloc(n));
  TBSLA_TRY(y.alloc(n));
  TBSLA_TRY(part.alloc(4 * kRedBlocks + 8));
  TBSLA_TRY(st.alloc(1));
  TBSLA_TRY(launch_fill(bufA.p, n, 1.0, s));  // b = 1 (:115-118)
  const double scale = (1 - beta) / (double)n;
  PrmState h{};
  h.S = (double)n;
  h.T = h.S * scale;  // (0 + sum(b_t)) (1 - beta) / n
  TBSLA_HIP_CHECK(hipMemcpyAsync(st.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
  const unsigned nb = red_grid(n);
  double *p_a = part.p, *p_b = part.p + kRedBlocks;
  double *buf[2] = {bufA.p, bufB.p};
  int launched = 0;
  while (launched < max_iterations) {
    const int batch = std::min(16, max_iterations - launched);
    for (int i = 0; i < batch; i++, launched++) {
      double *bt = buf[launched % 2], *bn = buf[(launched + 1) % 2];
      TBSLA_TRY(spmv_launch(m, y.p, bt, 0, true, s, &st.p->done));
      k_prm_extremes<<<nb, kBlock, 0, s>>>(y.p, n, st.p, p_a, p_b);
      k_prm_max<<<1, kBlock, 0, s>>>(st.p, p_a, p_b, (int)nb, beta);
      k_prm_update<<<nb, kBlock, 0, s>>>(n, y.p, bt, bn, beta, st.p, p_a, p_b);
      k_prm_finalize<<<1, kBlock, 0, s>>>(st.p, p_a, p_b, (int)nb, scale, epsilon);
    }
    TBSLA_HIP_LAUNCH_CHECK();
    TBSLA_HIP_CHECK(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    if (h.done) break;
  }
  double *b = buf[h.iters % 2];
  // final b /= sum(b) (:159-169)
  k_partial_sum<<<nb, kBlock, 0, s>>>(b, n, part.p);
  k_total<<<1, kBlock, 0, s>>>(part.p, (int)nb, part.p + 2 * kRedBlocks);
  k_divide<<<stride_grid(n), kBlock, 0, s>>>(b, n, part.p + 2 * kRedBlocks);
  TBSLA_HIP_LAUNCH_CHECK();
  TBSLA_HIP_CHECK(hipMemcpyAsync(d_b, b, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  if (nb_iterations) *nb_iterations = h.iters;
  if (final_error) *final_error = h.err;
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_page_rank_batched(tbsla_hip_matrix_t m, int ninst, const int *h_row_offsets,
                                           double beta, double epsilon, int max_iterations,
                                           double *d_b, int *h_iterations, double *h_errors,
                                           double *h_time_us, void *stream) {
  if (!m || !h_row_offsets || !d_b) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_CSR)
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "batched PageRank needs a CSR block");
  if (ninst < 1 || max_iterations < 1) return fail(TBSLA_HIP_ERR_INVALID, "bad counts");
  if (m->ln_row != m->ln_col || m->f_row != m->f_col || m->f_row != 0)
    return fail(TBSLA_HIP_ERR_INVALID, "batched PageRank needs a square block at (0, 0)");
  if (h_row_offsets[0] != 0 || h_row_offsets[ninst] != m->ln_row)
    return fail(TBSLA_HIP_ERR_INVALID, "row offsets must run from 0 to the block's rows");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // instance table and the workgroup -> instance map of the update kernels
  std::vector<PrInstance> hi(ninst);
  std::vector<int> hb;
  for (int k = 0; k < ninst; k++) {
    const long long n = (long long)h_row_offsets[k + 1] - h_row_offsets[k];
    if (n < 1) return fail(TBSLA_HIP_ERR_INVALID, "every instance needs at least one row");
    PrInstance &I = hi[k];
    I.S = (double)n;  // b = 1 (Matrix.cpp:549-552)
    I.err = 0.0;
    I.iters = 0;
    I.done = 0;
    I.row0 = h_row_offsets[k];
    I.n = n;
    I.part0 = (int)hb.size();
    I.nb = (int)red_grid(n);
    I.t_done = 0;
    for (int b = 0; b < I.nb; b++) hb.push_back(k);
  }
  const int nblocks = (int)hb.size();
  const long long rows = m->ln_row;
  DevBuf<double> bufA, bufB, y, p_err, p_sum, total;
  DevBuf<PrInstance> inst;
  DevBuf<int> binst, d_offs, done_flags;
  DevBuf<PrCommon> com;
  TBSLA_TRY(bufA.alloc(rows));
  TBSLA_TRY(bufB.alloc(rows));
  TBSLA_TRY(y.alloc(rows));
  TBSLA_TRY(p_err.alloc(nblocks));
  TBSLA_TRY(p_sum.alloc(nblocks));
  TBSLA_TRY(total.alloc(ninst));
  TBSLA_TRY(inst.alloc(ninst));
  TBSLA_TRY(binst.alloc(nblocks));
  TBSLA_TRY(d_offs.alloc(ninst + 1));
  TBSLA_TRY(done_flags.alloc(ninst));
  TBSLA_TRY(com.alloc(1));
  const PrCommon hc{0, 0, 0};
  TBSLA_HIP_CHECK(hipMemcpyAsync(inst.p, hi.data(), sizeof(PrInstance) * ninst, hipMemcpyHostToDevice, s));
  TBSLA_HIP_CHECK(hipMemcpyAsync(binst.p, hb.data(), sizeof(int) * nblocks, hipMemcpyHostToDevice, s));
  TBSLA_HIP_CHECK(hipMemcpyAsync(com.p, &hc, sizeof(hc), hipMemcpyHostToDevice, s));
  TBSLA_HIP_CHECK(hipMemcpyAsync(d_offs.p, h_row_offsets, sizeof(int) * (ninst + 1),
                                 hipMemcpyHostToDevice, s));
  {
    const std::vector<int> zeros(ninst, 0);
    TBSLA_HIP_CHECK(hipMemcpyAsync(done_flags.p, zeros.data(), sizeof(int) * ninst,
                                   hipMemcpyHostToDevice, s));
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));  // zeros leaves scope
  }
  TBSLA_TRY(launch_fill(bufA.p, rows, 1.0, s));
  double *buf[2] = {bufA.p, bufB.p};
  // one batch of 16 iterations (an even count: every batch starts on bufA)
  constexpr int kBatch = 16;
  auto enqueue_batch = [&](hipStream_t q) -> int {
    for (int i = 0; i < kBatch; i++) {
      double *bt = buf[i % 2], *bn = buf[(i + 1) % 2];
      TBSLA_TRY(spmv_launch(m, y.p, bt, 0, true, q, &com.p->all_done, d_offs.p, ninst,
                            done_flags.p));
      k_prb_update<<<nblocks, kBlock, 0, q>>>(inst.p, binst.p, y.p, bt, bn, beta, p_err.p,
                                              p_sum.p);
      k_prb_finalize<<<ninst, kBlock, 0, q>>>(inst.p, ninst, p_err.p, p_sum.p, epsilon,
                                              max_iterations, com.p, done_flags.p);
    }
    return TBSLA_HIP_SUCCESS;
  };
  // the batch is captured once into a hipGraph and replayed (on a stream of
  // our own when the caller passes the null stream, which cannot capture);
  // the split layout's launch is enqueued directly
  hipStream_t q = s;
  hipStream_t own = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  auto cleanup = [&]() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (own) {
      (void)hipStreamSynchronize(own);
      (void)hipStreamDestroy(own);
    }
  };
  bool graphs = m->split == nullptr;
  if (graphs && !s) {
    if (hipStreamCreateWithFlags(&own, hipStreamNonBlocking) != hipSuccess) graphs = false;
    else {
      TBSLA_HIP_CHECK(hipStreamSynchronize(s));
      q = own;
    }
  }
  if (graphs) {
    int st = TBSLA_HIP_SUCCESS;
    if (hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      cleanup();
      return fail(TBSLA_HIP_ERR_HIP, "stream capture failed to begin");
    }
    st = enqueue_batch(q);
    hipError_t e = hipStreamEndCapture(q, &graph);
    if (st == TBSLA_HIP_SUCCESS && e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (st != TBSLA_HIP_SUCCESS || e != hipSuccess) {
      cleanup();
      return st != TBSLA_HIP_SUCCESS ? st : fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
    }
  }
  PrCommon h{};
  k_prb_start<<<1, 1, 0, q>>>(com.p);
  int launched = 0;
  while (launched < max_iterations) {
    hipError_t e = hipSuccess;
    if (graphs) e = hipGraphLaunch(exec, q);
    else {
      const int st = enqueue_batch(q);
      if (st) {
        cleanup();
        return st;
      }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, com.p, sizeof(h), hipMemcpyDeviceToHost, q);
    if (e == hipSuccess) e = hipStreamSynchronize(q);
    if (e != hipSuccess) {
      cleanup();
      return fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
    }
    launched += kBatch;
    if (h.all_done) break;
  }
  // every instance: b /= sum(b) from its final buffer
  k_prb_norm_partial<<<nblocks, kBlock, 0, q>>>(inst.p, binst.p, bufA.p, bufB.p, p_sum.p);
  k_prb_norm_total<<<ninst, kBlock, 0, q>>>(inst.p, p_sum.p, total.p);
  k_prb_divide<<<nblocks, kBlock, 0, q>>>(inst.p, binst.p, bufA.p, bufB.p, total.p, d_b);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(hi.data(), inst.p, sizeof(PrInstance) * ninst, hipMemcpyDeviceToHost, q);
  if (e == hipSuccess) e = hipStreamSynchronize(q);
  cleanup();
  if (e != hipSuccess) return fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  int dev = 0, khz = 0;
  if (h_time_us &&
      (hipGetDevice(&dev) != hipSuccess ||
       hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess))
    khz = 0;
  for (int k = 0; k < ninst; k++) {
    if (h_iterations) h_iterations[k] = hi[k].iters;
    if (h_errors) h_errors[k] = hi[k].err;
    if (h_time_us)
      h_time_us[k] = (hi[k].t_done > 0 && khz > 0) ? (double)(hi[k].t_done - h.t0) * 1000.0 / khz
                                                   : -1.0;
  }
  return TBSLA_HIP_SUCCESS;
}
