// spmv.hip -- y = A x for the five formats (Matrix::Ax, Matrix.hpp:15).
//
// All kernels are HBM-bound streams over the matrix arrays with gathers of x
// (which stays in L2 / the 256 MB MALL):
//   CSR    L lanes per row, R rows per lane group in flight, V entries per
//          lane load, optional non-temporal loads (csr_rows.hpp); shuffle
//          reduction.
//   ELL    one lane per row over column-major slots: every slot load is a
//          contiguous 64-row stripe.
//   COO/SCOO  row-sorted triplets: SpMV (which zeroes y first) derives a
//          row pointer over the triplets' row range once, kept with the
//          block, and runs the CSR row kernel on the column and value arrays
//          (the row array is not read again).  Accumulating Ax uses a
//          segmented scan: each wavefront reduces a 512-entry chunk; rows
//          complete inside a chunk are written directly, rows crossing chunk
//          borders are summed in chunk order by a fix-up kernel --
//          deterministic.  Unsorted triplets use FP64 atomics (unordered, as
//          the reference's OpenMP array reduction).
//   DENSE  one wavefront per row (GEMV).
// Floating point: FMA accumulation; the per-row order differs from the
// reference's sequential loop only in the CSR/COO/DENSE lane split.
#include <algorithm>
#include <climits>

#include <mutex>
#include <vector>

#include "common.hpp"
#include "csr_rows.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

// MatrixELL::Ax (MatrixELL.cpp:181-205): a negative local column (padding
// of a block with f_col > 0) reads x[0], multiplied by the zero padding.
__global__ __launch_bounds__(kBlock) void k_ell_spmv(long long n_rows, int max_col,
                                                     const int *__restrict__ cols,
                                                     const double *__restrict__ vals,
                                                     const double *__restrict__ x, int f_col,
                                                     double *__restrict__ y,
                                                     const int *__restrict__ done) {
  if (done && *done) return;
  const long long row = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (row >= n_rows) return;
  double sum = 0.0;
#pragma unroll 4
  for (int j = 0; j < max_col; j++) {
    const size_t s = (size_t)j * n_rows + row;
    int idx = ldg_stream(cols + s) - f_col;
    idx = idx < 0 ? 0 : idx;
    sum = fma(ldg_stream(vals + s), x[idx], sum);
  }
  y[row] = sum;
}

constexpr int kSegSteps = 8;
constexpr int kSegChunk = kWave * kSegSteps;

// Segmented reduction of row-sorted triplets; r[row + row_off] +=
// val * x[col + col_off].  One wavefront per chunk of kSegChunk entries.
// Each chunk reports its first segment (head) and, when it has more than one,
// its last (tail); the fix-up kernel completes those.
__global__ __launch_bounds__(kBlock) void k_coo_seg(long long nnz, const int *__restrict__ rows,
                                                    const int *__restrict__ cols,
                                                    const double *__restrict__ vals,
                                                    const double *__restrict__ x, int row_off,
                                                    int col_off, double *__restrict__ y,
                                                    int *__restrict__ head_key,
                                                    double *__restrict__ head_val,
                                                    int *__restrict__ tail_key,
                                                    double *__restrict__ tail_val) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long chunk = ((long long)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const long long base = chunk * kSegChunk;
  if (base >= nnz) return;
  const long long end = min(nnz, base + kSegChunk);
  const int first_key = rows[base];
  int carry_key = -1;
  double carry_val = 0.0;
  bool head_written = false;
  for (int t = 0; t < kSegSteps; t++) {
    const long long e = base + (long long)t * kWave + lane;
    const bool valid = e < end;
    if (!__any(valid)) break;
    const int key = valid ? rows[e] : INT_MAX;
    double p = valid ? ldg_stream(vals + e) * x[ldg_stream(cols + e) + col_off] : 0.0;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const double o = __shfl_up(p, off, kWave);
      const int ok = __shfl_up(key, off, kWave);
      if (lane >= off && ok == key) p += o;
    }
    // the carried segment either continues into this step or is complete
    const int key0 = __shfl(key, 0, kWave);
    if (carry_key >= 0) {
      if (key0 == carry_key) {
        if (key == carry_key) p += carry_val;
      } else if (lane == 0) {
        if (carry_key == first_key) {
          head_key[chunk] = carry_key;
          head_val[chunk] = carry_val;
        } else {
          y[carry_key + row_off] += carry_val;
        }
      }
      if (key0 != carry_key && carry_key == first_key) head_written = true;
    }
    const int next_key = __shfl_down(key, 1, kWave);
    const bool seg_end = valid && lane != kWave - 1 && next_key != key;
    if (seg_end) {
      if (key == first_key) {
        head_key[chunk] = key;
        head_val[chunk] = p;
      } else {
        y[key + row_off] += p;
      }
    }
    if (__any(seg_end && key == first_key)) head_written = true;
    carry_key = __shfl(key, kWave - 1, kWave);
    carry_val = __shfl(p, kWave - 1, kWave);
    if (carry_key == INT_MAX) carry_key = -1;  // partial last step: nothing carried
  }
  if (lane == 0) {
    int tk = -1;
    double tv = 0.0;
    if (carry_key >= 0) {
      if (carry_key == first_key && !head_written) {
        head_key[chunk] = carry_key;
        head_val[chunk] = carry_val;
      } else {
        tk = carry_key;
        tv = carry_val;
      }
    }
    tail_key[chunk] = tk;
    tail_val[chunk] = tv;
  }
}

// Completes the rows that cross chunk borders, each summed by one thread in
// chunk order.
__global__ __launch_bounds__(kBlock) void k_coo_fixup(long long nchunks, const int *__restrict__ head_key,
                                                      const double *__restrict__ head_val,
                                                      const int *__restrict__ tail_key,
                                                      const double *__restrict__ tail_val,
                                                      int row_off, double *__restrict__ y) {
  const long long c = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (c >= nchunks) return;
  const int hk = head_key[c];
  const int prev_last = c > 0 ? (tail_key[c - 1] >= 0 ? tail_key[c - 1] : head_key[c - 1]) : -1;
  if (hk != prev_last) {
    double sum = head_val[c];
    long long cc = c;
    while (tail_key[cc] < 0 && cc + 1 < nchunks && head_key[cc + 1] == hk) {
      cc++;
      sum += head_val[cc];
    }
    y[hk + row_off] += sum;
  }
  const int tk = tail_key[c];
  if (tk >= 0) {
    double sum = tail_val[c];
    long long cc = c;
    while (cc + 1 < nchunks && head_key[cc + 1] == tk) {
      cc++;
      sum += head_val[cc];
      if (tail_key[cc] >= 0) break;
    }
    y[tk + row_off] += sum;
  }
}

__global__ __launch_bounds__(kBlock) void k_coo_atomic(long long nnz, const int *__restrict__ rows,
                                                       const int *__restrict__ cols,
                                                       const double *__restrict__ vals,
                                                       const double *__restrict__ x, int row_off,
                                                       int col_off, double *__restrict__ y) {
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < nnz;
       e += (long long)gridDim.x * kBlock)
    atomicAdd(y + rows[e] + row_off, vals[e] * x[cols[e] + col_off]);
}

// MatrixDENSE::Ax (MatrixDENSE.cpp:77-86): r[i] += A[i,:] . x.
// Row pointer of row-sorted triplets: rp[i] = the first k with
// rows[k] >= row0 + i, for i in [0, nrows]; entry k writes the pointers of
// the rows after its predecessor's, up to its own.  A decrease sets *unsorted.
__global__ __launch_bounds__(kBlock) void k_coo_rowptr(long long n, const int *__restrict__ rows,
                                                       int row0, int nrows, int *__restrict__ rp,
                                                       int *__restrict__ unsorted) {
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k <= n;
       k += (long long)gridDim.x * kBlock) {
    const long long lo = k == 0 ? -1 : (long long)rows[k - 1] - row0;
    const long long hi = k == n ? nrows : (long long)rows[k] - row0;
    if (hi < lo) {
      *unsorted = 1;
      continue;
    }
    for (long long r = lo + 1 < 0 ? 0 : lo + 1; r <= hi && r <= nrows; r++) rp[r] = (int)k;
  }
}

// One wavefront per row.  With vec (an even row length, so 16-byte aligned
// rows, and a 16-byte aligned x) each lane reads two values per load and
// keeps four loads in flight, with four partial sums added in a fixed
// order; otherwise one value per load.
__global__ __launch_bounds__(kBlock) void k_dense_gemv(long long n_rows, long long n_cols,
                                                       const double *__restrict__ A,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ y, int vec) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  const long long row = ((long long)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= n_rows) return;
  const double *a = A + (size_t)row * n_cols;
  double sum = 0.0;
  if (vec) {
    constexpr int U = 4, STEP = 2 * kWave;  // doubles per wave load
    double s[U] = {0.0, 0.0, 0.0, 0.0};
    long long j = 2 * lane;
    for (; j + (U - 1) * STEP + 1 < n_cols; j += U * STEP) {
      v2 av[U], xv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        av[u] = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(a + j + u * STEP));
        xv[u] = *reinterpret_cast<const v2 *>(x + j + u * STEP);
      }
#pragma unroll
      for (int u = 0; u < U; u++) s[u] = fma(av[u].y, xv[u].y, fma(av[u].x, xv[u].x, s[u]));
    }
    for (; j + 1 < n_cols; j += STEP) {
      const v2 av = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(a + j));
      const v2 xv = *reinterpret_cast<const v2 *>(x + j);
      s[0] = fma(av.y, xv.y, fma(av.x, xv.x, s[0]));
    }
    sum = (s[0] + s[1]) + (s[2] + s[3]);
  } else {
    for (long long j = lane; j < n_cols; j += kWave) sum = fma(ldg_stream(a + j), x[j], sum);
  }
  sum = group_sum<kWave>(sum);
  if (lane == 0) y[row] += sum;
}

__global__ __launch_bounds__(kBlock) void k_fill(double *__restrict__ d, long long n, double v) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    d[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_add_inplace(double *__restrict__ y,
                                                        const double *__restrict__ x, long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    y[i] += x[i];
}

unsigned stride_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

int launch_fill(double *d, long long n, double v, hipStream_t s) {
  if (n <= 0) return TBSLA_HIP_SUCCESS;
  k_fill<<<stride_grid(n), kBlock, 0, s>>>(d, n, v);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

int launch_axpy_inplace(double *y, const double *x, long long n, hipStream_t s) {
  if (n <= 0) return TBSLA_HIP_SUCCESS;
  k_add_inplace<<<stride_grid(n), kBlock, 0, s>>>(y, x, n);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

// Lanes per row from the mean row length (two entries per lane per row,
// clamped to 1..64) unless set; rows per group from the lanes.
// The row pointer of a row-sorted COO/SCOO block (built once; false when
// the triplets turn out not to be sorted, which also clears row_sorted).
int coo_rowptr(tbsla_hip_matrix_s *m, long long n, bool *ok, hipStream_t s) {
  *ok = false;
  if (m->coo_rowptr) {
    *ok = true;
    return TBSLA_HIP_SUCCESS;
  }
  int ends[2] = {0, 0};
  TBSLA_HIP_CHECK(hipMemcpyAsync(&ends[0], m->rowidx, 4, hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipMemcpyAsync(&ends[1], m->rowidx + n - 1, 4, hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  if (ends[1] < ends[0]) {
    m->row_sorted = 0;
    return TBSLA_HIP_SUCCESS;
  }
  const int nrows = ends[1] - ends[0] + 1;
  int *rp = nullptr, *flag = nullptr;
  TBSLA_TRY(device_alloc((void **)&rp, (size_t)(nrows + 1) * 4));
  int st = device_alloc((void **)&flag, 4);
  if (st) {
    (void)hipFree(rp);
    return st;
  }
  hipError_t e = hipMemsetAsync(flag, 0, 4, s);
  unsigned g = (unsigned)std::min<long long>((n + kBlock) / kBlock, 65536);
  if (e == hipSuccess) {
    k_coo_rowptr<<<g, kBlock, 0, s>>>(n, m->rowidx, ends[0], nrows, rp, flag);
    e = hipGetLastError();
  }
  int unsorted = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&unsorted, flag, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(flag);
  if (e != hipSuccess) {
    (void)hipFree(rp);
    return fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  }
  if (unsorted) {
    (void)hipFree(rp);
    m->row_sorted = 0;
    return TBSLA_HIP_SUCCESS;
  }
  m->coo_rowptr = rp;
  m->coo_row0 = ends[0];
  m->coo_nrows = nrows;
  *ok = true;
  return TBSLA_HIP_SUCCESS;
}

int pick_csr_lanes(const tbsla_hip_matrix_s *m) {
  if (m->lanes_per_row > 0) return m->lanes_per_row;
  double mean = m->ln_row > 0 ? (double)m->stored / m->ln_row : 1.0;
  int L = 1;
  while (L < 64 && L < mean / 2) L <<= 1;
  return L;
}

int pick_csr_rows(const tbsla_hip_matrix_s *m, int L) {
  if (m->rows_per_group > 0) return m->rows_per_group;
  return L >= 64 ? 1 : (L >= 32 ? 2 : 4);
}

// entries per lane load (1 unless set) and the non-temporal load hint (on
// unless set)
int pick_csr_vec(const tbsla_hip_matrix_s *m) { return m->csr_vec > 0 ? m->csr_vec : 1; }
int pick_csr_nt(const tbsla_hip_matrix_s *m) { return m->csr_nt >= 0 ? m->csr_nt : 1; }

// Scratch of the segmented COO/SCOO accumulation of n row-sorted triplets:
// each chunk's head and tail keys (int) and values (double).
size_t coo_seg_scratch_bytes(long long n) {
  const long long nchunks = (n + kSegChunk - 1) / kSegChunk;
  return (size_t)nchunks * 24 + 256;
}

// k_coo_seg + k_coo_fixup with their per-chunk heads and tails in sc (every
// chunk writes both, so sc needs no clearing)
int coo_seg_launch(const tbsla_hip_matrix_s *m, long long n, int row_off, int col_off,
                   const double *d_v, double *d_r, char *sc, hipStream_t s) {
  const long long nchunks = (n + kSegChunk - 1) / kSegChunk;
  int *head_key = reinterpret_cast<int *>(sc);
  int *tail_key = head_key + nchunks;
  const size_t off = ((size_t)nchunks * 8 + 15) & ~(size_t)15;
  double *head_val = reinterpret_cast<double *>(sc + off);
  double *tail_val = head_val + nchunks;
  k_coo_seg<<<grid_for(nchunks * kWave), kBlock, 0, s>>>(n, m->rowidx, m->colidx, m->values, d_v,
                                                         row_off, col_off, d_r, head_key, head_val,
                                                         tail_key, tail_val);
  TBSLA_HIP_LAUNCH_CHECK();
  k_coo_fixup<<<grid_for(nchunks), kBlock, 0, s>>>(nchunks, head_key, head_val, tail_key, tail_val,
                                                   row_off, d_r);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

int spmv_launch(tbsla_hip_matrix_s *m, double *d_r, const double *d_v, int vect_incr,
                bool zero_first, hipStream_t s, const int *d_done, const int *seg_offs, int nseg,
                const int *seg_done) {
  switch (m->format) {
    case TBSLA_HIP_CSR: {
      // MatrixCSR::Ax returns when nnz == 0 (MatrixCSR.cpp:210-211)
      if (m->nnz == 0 || m->ln_row == 0) {
        if (zero_first) TBSLA_TRY(launch_fill(d_r, m->ln_row, 0.0, s));
        return TBSLA_HIP_SUCCESS;
      }
      if (m->split) return split_launch(m, d_r, d_v, d_done, s);
      CsrRowsArgs a{m->ln_row, m->rowptr, m->colidx, m->values, d_v, m->f_col, d_r, d_done,
                    nullptr, seg_offs, nseg, seg_done};
      const int L = pick_csr_lanes(m);
      csr_rows_launch<false>(a, pick_csr_vec(m), L, pick_csr_rows(m, L), pick_csr_nt(m), s);
      TBSLA_HIP_LAUNCH_CHECK();
      return TBSLA_HIP_SUCCESS;
    }
    case TBSLA_HIP_ELL: {
      // MatrixELL.cpp:182-183
      if (m->nnz == 0 || m->max_col == 0 || !m->has_arrays || m->ln_row == 0) {
        if (zero_first) TBSLA_TRY(launch_fill(d_r, m->ln_row, 0.0, s));
        return TBSLA_HIP_SUCCESS;
      }
      if (m->split) return split_launch(m, d_r, d_v, d_done, s);
      k_ell_spmv<<<grid_for(m->ln_row), kBlock, 0, s>>>(m->ln_row, m->max_col, m->colidx,
                                                         m->values, d_v, m->f_col, d_r, d_done);
      TBSLA_HIP_LAUNCH_CHECK();
      return TBSLA_HIP_SUCCESS;
    }
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO: {
      const bool coo = m->format == TBSLA_HIP_COO;
      const long long len = coo ? m->n_row : m->ln_row;
      if (zero_first) TBSLA_TRY(launch_fill(d_r, len, 0.0, s));
      // SCOO returns when nnz == 0 (MatrixSCOO.cpp:155-157); COO loops over nnz
      // the triplets held (an import may carry a larger nnz field than arrays)
      const long long n = (!coo && m->nnz == 0) ? 0 : m->stored;
      if (n <= 0 || !m->has_arrays) return TBSLA_HIP_SUCCESS;
      const int row_off = coo ? vect_incr : vect_incr - m->f_row;
      const int col_off = coo ? 0 : -m->f_col;
      bool rows_ok = false;
      if (zero_first && m->row_sorted) TBSLA_TRY(coo_rowptr(m, n, &rows_ok, s));
      if (rows_ok && m->split)  // a derived layout over the triplets' rows
        return split_launch(m, d_r + ((long long)m->coo_row0 + row_off), d_v, nullptr, s);
      if (rows_ok) {
        // y is zero: the row kernel's stores over the triplets' row range
        // equal the accumulation.  V from the mean row length, L the
        // fewest lanes whose V L entries cover a mean row in one trip.
        const double mean = (double)n / m->coo_nrows;
        int V = mean >= 16 ? 4 : (mean >= 6 ? 2 : 1);
        int L = 1;
        while (L < 64 && L * V < mean) L <<= 1;
        if (V == 2) L = std::min(std::max(L, 4), 32);
        if (V == 4) L = std::min(std::max(L, 4), 16);
        const int R = V > 1 ? 1 : (L >= 64 ? 1 : (L >= 32 ? 2 : 4));
        CsrRowsArgs a{m->coo_nrows, m->coo_rowptr, m->colidx, m->values, d_v, -col_off,
                      d_r + ((long long)m->coo_row0 + row_off), nullptr, nullptr};
        csr_rows_launch<false>(a, V, L, R, 1, s);
        TBSLA_HIP_LAUNCH_CHECK();
        return TBSLA_HIP_SUCCESS;
      }
      if (!m->row_sorted) {
        k_coo_atomic<<<stride_grid(n), kBlock, 0, s>>>(n, m->rowidx, m->colidx, m->values, d_v,
                                                        row_off, col_off, d_r);
        TBSLA_HIP_LAUNCH_CHECK();
        return TBSLA_HIP_SUCCESS;
      }
      TBSLA_TRY(ensure_scratch(m, coo_seg_scratch_bytes(n)));
      return coo_seg_launch(m, n, row_off, col_off, d_v, d_r, static_cast<char *>(m->scratch), s);
    }
    case TBSLA_HIP_DENSE: {
      if (zero_first) TBSLA_TRY(launch_fill(d_r, m->ln_row, 0.0, s));
      // MatrixDENSE.cpp:78-79
      if (m->nnz == 0 || !m->has_arrays || m->ln_row == 0) return TBSLA_HIP_SUCCESS;
      const int vec = (m->ln_col % 2 == 0) && (reinterpret_cast<uintptr_t>(d_v) % 16 == 0);
      k_dense_gemv<<<grid_for((long long)m->ln_row * kWave), kBlock, 0, s>>>(
          m->ln_row, m->ln_col, m->values, d_v, d_r, vec);
      TBSLA_HIP_LAUNCH_CHECK();
      return TBSLA_HIP_SUCCESS;
    }
  }
  return fail(TBSLA_HIP_ERR_INVALID, "unknown format");
}

// ---- Ax without allocation -----------------------------------------------
//
// One scratch array per GPU for the segmented COO/SCOO accumulation:
// kNoAllocScratch bytes, i.e. up to (kNoAllocScratch - 256) / 24 chunks of
// 512 triplets (about 358 M).  Each GPU the library runs on holds it until
// the process ends.  Host code uses the device address resolved once per
// device by hipGetSymbolAddress, never the array's name.
constexpr size_t kNoAllocScratch = 16u << 20;
__device__ __align__(256) char scratch[kNoAllocScratch];

namespace {

std::mutex &scratch_mutex() {
  static std::mutex *mu = new std::mutex;
  return *mu;
}

// the current device's copy of `scratch`, resolved on the first call on it
int scratch_address(char **out) {
  int dev = 0;
  TBSLA_HIP_CHECK(hipGetDevice(&dev));
  static std::vector<char *> *ptrs = new std::vector<char *>;
  std::lock_guard<std::mutex> g(scratch_mutex());
  if ((int)ptrs->size() <= dev) ptrs->resize((size_t)dev + 1, nullptr);
  if (!(*ptrs)[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(TBSLA_HIP_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    (*ptrs)[dev] = static_cast<char *>(p);
  }
  *out = (*ptrs)[dev];
  return TBSLA_HIP_SUCCESS;
}

}  // namespace

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_Ax_noalloc(tbsla_hip_matrix_t m, double *d_r, const double *d_v,
                                    int vect_incr, void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool coo = m->format == TBSLA_HIP_COO, scoo = m->format == TBSLA_HIP_SCOO;
  // CSR, ELL and DENSE, and unsorted triplets (atomics), allocate nothing
  // in tbsla_hip_Ax: the same launch
  if (!(coo || scoo) || !m->row_sorted) return spmv_launch(m, d_r, d_v, vect_incr, false, s);
  const long long n = (!coo && m->nnz == 0) ? 0 : m->stored;
  if (n <= 0 || !m->has_arrays) return TBSLA_HIP_SUCCESS;
  const size_t need = coo_seg_scratch_bytes(n);
  if (need > kNoAllocScratch)
    return fail(TBSLA_HIP_ERR_RANGE, "tbsla_hip_Ax_noalloc: " + std::to_string(n) +
                                         " triplets need " + std::to_string(need) +
                                         " bytes of scratch, more than its " +
                                         std::to_string(kNoAllocScratch));
  char *sc = nullptr;
  TBSLA_TRY(scratch_address(&sc));
  const int row_off = coo ? vect_incr : vect_incr - m->f_row;
  const int col_off = coo ? 0 : -m->f_col;
  return coo_seg_launch(m, n, row_off, col_off, d_v, d_r, sc, s);
}
