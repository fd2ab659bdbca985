// problems.hip -- SPD problem generators and CSR assembly on the device.
//
// * 3-D 7-point Poisson -div(k grad u) on an nx x ny x nz cell grid with
//   Dirichlet boundaries (k = 1 gives the textbook 6/-1 stencil).
// * A smoothed log-normal coefficient field k = exp(mu + sigma z), z white
//   noise under a separable Gaussian filter, scaled to unit variance.
// * The symmetrised cqmat A = (B + B^T)/2 with its diagonal replaced by the
//   row's off-diagonal absolute sum + 1 (strictly diagonally dominant, SPD),
//   B = fill_cqmat's CSR block (MatrixCSR.cpp:373-486) with project values.
// * Uniform vectors, block-diagonal concatenation, the CSR transpose.
//
// Every value is produced by a fixed sequence of IEEE operations without
// contraction, so the oracle's C loops reproduce them bit for bit (the log,
// cos and exp of the noise field aside, which are libm-accurate).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

struct Grid3 {
  int nx, ny, nz;
};

__device__ inline long long cell(const Grid3 &g, int x, int y, int z) {
  return (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z);
}

// Face coefficient between cells a and b: the harmonic mean of k.
__device__ inline double face_coef(const double *k, long long a, long long b) {
  TBSLA_NO_FMA
  if (!k) return 1.0;
  const double ka = k[a], kb = k[b];
  return 2.0 * ka * kb / (ka + kb);
}

__global__ __launch_bounds__(kBlock) void k_poisson_count(Grid3 g, long long n, int *__restrict__ cnt) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((long long)g.nx * g.ny));
    cnt[i] = 1 + (x > 0) + (x < g.nx - 1) + (y > 0) + (y < g.ny - 1) + (z > 0) + (z < g.nz - 1);
  }
}

// Row i in ascending column order: -z, -y, -x, diagonal, +x, +y, +z.
// Off-diagonal -k_face; a boundary face contributes k_cell to the diagonal.
// Diagonal = faces summed in the order -z, -y, -x, +x, +y, +z.
__global__ __launch_bounds__(kBlock) void k_poisson_fill(Grid3 g, long long n,
                                                         const double *__restrict__ k,
                                                         const int *__restrict__ rowptr,
                                                         int *__restrict__ colidx,
                                                         double *__restrict__ vals) {
  TBSLA_NO_FMA
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / ((long long)g.nx * g.ny));
    const double kc = k ? k[i] : 1.0;
    long long nb[6];
    bool in[6];
    in[0] = z > 0;        nb[0] = in[0] ? cell(g, x, y, z - 1) : -1;
    in[1] = y > 0;        nb[1] = in[1] ? cell(g, x, y - 1, z) : -1;
    in[2] = x > 0;        nb[2] = in[2] ? cell(g, x - 1, y, z) : -1;
    in[3] = x < g.nx - 1; nb[3] = in[3] ? cell(g, x + 1, y, z) : -1;
    in[4] = y < g.ny - 1; nb[4] = in[4] ? cell(g, x, y + 1, z) : -1;
    in[5] = z < g.nz - 1; nb[5] = in[5] ? cell(g, x, y, z + 1) : -1;
    double f[6];
    double d = 0.0;
#pragma unroll
    for (int t = 0; t < 6; t++) {
      f[t] = in[t] ? face_coef(k, i, nb[t]) : kc;
      d = d + f[t];
    }
    int pos = rowptr[i];
#pragma unroll
    for (int t = 0; t < 3; t++)
      if (in[t]) {
        colidx[pos] = (int)nb[t];
        vals[pos] = -f[t];
        pos++;
      }
    colidx[pos] = (int)i;
    vals[pos] = d;
    pos++;
#pragma unroll
    for (int t = 3; t < 6; t++)
      if (in[t]) {
        colidx[pos] = (int)nb[t];
        vals[pos] = -f[t];
        pos++;
      }
  }
}

// Standard normal noise by Box-Muller from the project value generator:
// u1 = V(seed, 2i), u2 = V(seed, 2i+1); z = sqrt(-2 log(1-u1)) cos(2 pi u2).
__global__ __launch_bounds__(kBlock) void k_normal_noise(long long n, uint64_t seed_hash,
                                                         double *__restrict__ z) {
  TBSLA_NO_FMA
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const double u1 = unit_value(seed_hash, 2 * (uint64_t)i);
    const double u2 = unit_value(seed_hash, 2 * (uint64_t)i + 1);
    const double r = sqrt(-2.0 * log(1.0 - u1));
    z[i] = r * cos(6.283185307179586 * u2);
  }
}

// Reflect an index into [0, n) ("half-sample symmetric": -1 -> 0, n -> n-1).
__device__ inline int reflect(long long j, int n) {
  const long long p = 2LL * n;
  long long m = ((j % p) + p) % p;
  if (m >= n) m = p - 1 - m;
  return (int)m;
}

constexpr int kMaxTaps = 129;

struct Taps {
  int radius;
  double w[kMaxTaps];
};

// One axis of the separable filter: out = sum_{d=-R..R} w[d] in[reflect(.+d)],
// summed in increasing d.
__global__ __launch_bounds__(kBlock) void k_filter_axis(Grid3 g, int axis, Taps t,
                                                        const double *__restrict__ in,
                                                        double *__restrict__ out) {
  TBSLA_NO_FMA
  const long long n = (long long)g.nx * g.ny * g.nz;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    int c[3] = {(int)(i % g.nx), (int)((i / g.nx) % g.ny), (int)(i / ((long long)g.nx * g.ny))};
    const int len = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
    double acc = 0.0;
    for (int d = -t.radius; d <= t.radius; d++) {
      int cc[3] = {c[0], c[1], c[2]};
      cc[axis] = reflect((long long)c[axis] + d, len);
      acc = acc + t.w[d + t.radius] * in[cell(g, cc[0], cc[1], cc[2])];
    }
    out[i] = acc;
  }
}

__global__ __launch_bounds__(kBlock) void k_lognormal(long long n, double mu, double sigma,
                                                      double scale, double *__restrict__ z) {
  TBSLA_NO_FMA
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    z[i] = exp(mu + sigma * (z[i] * scale));
}

__global__ __launch_bounds__(kBlock) void k_uniform(double *__restrict__ d, long long n,
                                                    uint64_t seed_hash, double lo, double hi) {
  TBSLA_NO_FMA
  const double w = hi - lo;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    d[i] = lo + w * unit_value(seed_hash, (uint64_t)i);
}

// ---- symmetrised cqmat -----------------------------------------------------

// Entry j of B (row r, column c) gives key (r,c) at position j and key (c,r)
// at position m + j; diagonal entries get the sentinel key (sorted last).
__global__ __launch_bounds__(kBlock) void k_sym_keys(long long n, long long m,
                                                     const int *__restrict__ rowptr,
                                                     const int *__restrict__ colidx, int cb,
                                                     unsigned long long sentinel,
                                                     unsigned long long *__restrict__ keys,
                                                     int *__restrict__ payload,
                                                     int *__restrict__ cnt) {
  for (long long r = (long long)blockIdx.x * kBlock + threadIdx.x; r < n;
       r += (long long)gridDim.x * kBlock) {
    for (int j = rowptr[r]; j < rowptr[r + 1]; j++) {
      const long long c = colidx[j];
      payload[j] = j;
      payload[m + j] = (int)(m + j);
      if (c == r) {
        keys[j] = sentinel;
        keys[m + j] = sentinel;
      } else {
        keys[j] = ((unsigned long long)r << cb) | (unsigned long long)c;
        keys[m + j] = ((unsigned long long)c << cb) | (unsigned long long)r;
        atomicAdd(cnt + r, 1);
        atomicAdd(cnt + c, 1);
      }
    }
  }
}

// distinct off-diagonal columns of each row, + 1 for the diagonal
__global__ __launch_bounds__(kBlock) void k_sym_count(long long n, const int *__restrict__ start,
                                                      const unsigned long long *__restrict__ keys,
                                                      int *__restrict__ ucnt) {
  for (long long r = (long long)blockIdx.x * kBlock + threadIdx.x; r < n;
       r += (long long)gridDim.x * kBlock) {
    int u = 0;
    for (int j = start[r]; j < start[r + 1]; j++) u += (j == start[r] || keys[j] != keys[j - 1]);
    ucnt[r] = u + 1;
  }
}

// One thread per row: each run of equal keys gives A_rc = (a + b) * 0.5 with
// a the sum of the direct entries (B_rc, generation order) and b of the
// transposed ones (B_cr); the diagonal is the off-diagonal |A_rc| summed in
// column order, + 1.
__global__ __launch_bounds__(kBlock) void k_sym_fill(long long n, long long m, int cb,
                                                     const int *__restrict__ start,
                                                     const unsigned long long *__restrict__ keys,
                                                     const int *__restrict__ payload,
                                                     const double *__restrict__ bvals,
                                                     const int *__restrict__ rowptr,
                                                     int *__restrict__ colidx,
                                                     double *__restrict__ vals) {
  TBSLA_NO_FMA
  const unsigned long long cmask = (1ull << cb) - 1;
  for (long long r = (long long)blockIdx.x * kBlock + threadIdx.x; r < n;
       r += (long long)gridDim.x * kBlock) {
    int pos = rowptr[r];
    int diag_pos = -1;
    double abs_sum = 0.0;
    const int e = start[r + 1];
    int j = start[r];
    while (j < e) {
      const unsigned long long key = keys[j];
      const int c = (int)(key & cmask);
      double a = 0.0, b = 0.0;
      for (; j < e && keys[j] == key; j++) {
        const long long p = payload[j];
        if (p < m)
          a = a + bvals[p];
        else
          b = b + bvals[p - m];
      }
      if (diag_pos < 0 && c > r) diag_pos = pos++;
      const double v = (a + b) * 0.5;
      colidx[pos] = c;
      vals[pos] = v;
      pos++;
      abs_sum = abs_sum + fabs(v);
    }
    if (diag_pos < 0) diag_pos = pos;
    colidx[diag_pos] = (int)r;
    vals[diag_pos] = abs_sum + 1.0;
  }
}

// ---- transpose ---------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_transpose_keys(long long n_rows,
                                                           const int *__restrict__ rowptr,
                                                           const int *__restrict__ colidx,
                                                           int f_col, int rb,
                                                           unsigned long long *__restrict__ keys,
                                                           int *__restrict__ payload,
                                                           int *__restrict__ cnt) {
  for (long long r = (long long)blockIdx.x * kBlock + threadIdx.x; r < n_rows;
       r += (long long)gridDim.x * kBlock) {
    for (int j = rowptr[r]; j < rowptr[r + 1]; j++) {
      const long long c = colidx[j] - f_col;
      keys[j] = ((unsigned long long)c << rb) | (unsigned long long)r;
      payload[j] = j;
      atomicAdd(cnt + c, 1);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_transpose_fill(long long m,
                                                           const unsigned long long *__restrict__ keys,
                                                           const int *__restrict__ payload,
                                                           const double *__restrict__ v, int rb,
                                                           int f_row, int *__restrict__ colidx,
                                                           double *__restrict__ vals) {
  const unsigned long long rmask = (1ull << rb) - 1;
  for (long long j = (long long)blockIdx.x * kBlock + threadIdx.x; j < m;
       j += (long long)gridDim.x * kBlock) {
    colidx[j] = (int)(keys[j] & rmask) + f_row;
    vals[j] = v[payload[j]];
  }
}

// ---- block diagonal --------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_bd_rows(long long n, const int *__restrict__ rp,
                                                    int base, int *__restrict__ out) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    out[i] = rp[i] + base;
}

__global__ __launch_bounds__(kBlock) void k_bd_entries(long long m, const int *__restrict__ ci,
                                                       const double *__restrict__ v, int shift,
                                                       int *__restrict__ co,
                                                       double *__restrict__ vo) {
  for (long long j = (long long)blockIdx.x * kBlock + threadIdx.x; j < m;
       j += (long long)gridDim.x * kBlock) {
    co[j] = ci[j] + shift;
    vo[j] = v[j];
  }
}

unsigned sgrid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 65536));
}

int bits_of(long long v) {
  int b = 1;
  while ((1LL << b) <= v) b++;
  return b;
}

void set_square(tbsla_hip_matrix_s *m, long long n) {
  m->n_row = m->n_col = m->ln_row = m->ln_col = (int)n;
  m->f_row = m->f_col = 0;
  m->pr = m->pc = 0;
  m->NR = m->NC = 1;
  m->max_col = 0;
  m->row_sorted = 1;
}

}  // namespace

// exclusive scan of n counts into rowptr[0..n], total into *total (host)
int counts_to_rowptr(const int *cnt, int *rowptr, long long n, long long *total, hipStream_t s) {
  void *ws = nullptr;
  TBSLA_TRY(device_alloc(&ws, scan_workspace_bytes(n + 1) + 64));
  long long *d_total = reinterpret_cast<long long *>(static_cast<char *>(ws) + scan_workspace_bytes(n + 1));
  int st = scan<int>(cnt, rowptr, n, false, d_total, ws, s);
  long long t = 0;
  if (st == TBSLA_HIP_SUCCESS) {
    hipError_t e = hipMemcpyAsync(&t, d_total, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(ws);
  if (st) return st;
  if (t > INT32_MAX) return fail(TBSLA_HIP_ERR_RANGE, "more than 2^31-1 entries");
  int ti = (int)t;
  TBSLA_HIP_CHECK(hipMemcpyAsync(rowptr + n, &ti, 4, hipMemcpyHostToDevice, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  *total = t;
  return TBSLA_HIP_SUCCESS;
}

// CSR transpose into freshly allocated arrays (rows of the result = local
// columns of src; entries of a result row in increasing source row, ties in
// source order).  Used for G^T of the factored preconditioner.
int csr_transpose_arrays(const tbsla_hip_matrix_s *src, int **rowptr_t, int **colidx_t,
                         double **values_t, hipStream_t s) {
  const long long nr = src->ln_row, nc = src->ln_col, m = src->stored;
  int *cnt = nullptr, *rp = nullptr, *ci = nullptr, *pay = nullptr;
  double *v = nullptr;
  unsigned long long *keys = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    (void)hipFree(cnt);
    (void)hipFree(pay);
    (void)hipFree(keys);
  };
  int st;
  if ((st = device_alloc((void **)&cnt, (size_t)(nc + 1) * 4)) ||
      (st = device_alloc((void **)&rp, (size_t)(nc + 1) * 4)) ||
      (st = device_alloc((void **)&ci, (size_t)std::max(m, 1LL) * 4)) ||
      (st = device_alloc((void **)&v, (size_t)std::max(m, 1LL) * 8)) ||
      (st = device_alloc((void **)&pay, (size_t)std::max(m, 1LL) * 4)) ||
      (st = device_alloc((void **)&keys, (size_t)std::max(m, 1LL) * 8))) {
    cleanup();
    (void)hipFree(rp);
    (void)hipFree(ci);
    (void)hipFree(v);
    return st;
  }
  (void)hipMemsetAsync(cnt, 0, (size_t)(nc + 1) * 4, s);
  const int rb = bits_of(std::max(nr, 1LL));
  const int cbits = bits_of(std::max(nc, 1LL));
  if (m > 0) {
    k_transpose_keys<<<sgrid(nr), kBlock, 0, s>>>(nr, src->rowptr, src->colidx, src->f_col, rb,
                                                  keys, pay, cnt);
    st = radix_sort_pairs(keys, pay, m, rb + cbits, s);
  }
  long long total = 0;
  if (!st) st = counts_to_rowptr(cnt, rp, nc, &total, s);
  if (!st && m > 0) {
    k_transpose_fill<<<sgrid(m), kBlock, 0, s>>>(m, keys, pay, src->values, rb, src->f_row, ci, v);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) st = fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  }
  cleanup();
  if (st) {
    (void)hipFree(rp);
    (void)hipFree(ci);
    (void)hipFree(v);
    return st;
  }
  *rowptr_t = rp;
  *colidx_t = ci;
  *values_t = v;
  return TBSLA_HIP_SUCCESS;
}

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_fill_poisson3d(tbsla_hip_matrix_t m, int nx, int ny, int nz,
                                        const double *d_kappa, void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null handle");
  if (m->format != TBSLA_HIP_CSR) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "Poisson fills a CSR block");
  if (nx < 1 || ny < 1 || nz < 1) return fail(TBSLA_HIP_ERR_INVALID, "grid sizes must be >= 1");
  const long long n = (long long)nx * ny * nz;
  if (7 * n > INT32_MAX) return fail(TBSLA_HIP_ERR_RANGE, "grid too large for int32 indices");
  hipStream_t s = static_cast<hipStream_t>(stream);
  release_arrays(m);
  set_square(m, n);
  int *cnt = nullptr;
  TBSLA_TRY(device_alloc((void **)&cnt, (size_t)(n + 1) * 4));
  TBSLA_TRY(device_alloc((void **)&m->rowptr, (size_t)(n + 1) * 4));
  Grid3 g{nx, ny, nz};
  k_poisson_count<<<sgrid(n), kBlock, 0, s>>>(g, n, cnt);
  long long total = 0;
  int st = counts_to_rowptr(cnt, m->rowptr, n, &total, s);
  (void)hipFree(cnt);
  if (st) return st;
  TBSLA_TRY(device_alloc((void **)&m->colidx, (size_t)total * 4));
  TBSLA_TRY(device_alloc((void **)&m->values, (size_t)total * 8));
  k_poisson_fill<<<sgrid(n), kBlock, 0, s>>>(g, n, d_kappa, m->rowptr, m->colidx, m->values);
  TBSLA_HIP_LAUNCH_CHECK();
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  m->nnz = m->stored = total;
  m->has_arrays = 1;
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_fill_lognormal_field(double *d_kappa, int nx, int ny, int nz, double mu,
                                              double sigma, double corr_len,
                                              unsigned long long seed, void *stream) {
  if (!d_kappa) return fail(TBSLA_HIP_ERR_INVALID, "null field");
  if (nx < 1 || ny < 1 || nz < 1) return fail(TBSLA_HIP_ERR_INVALID, "grid sizes must be >= 1");
  if (!(corr_len >= 0) || corr_len > (kMaxTaps - 1) / 6.0)
    return fail(TBSLA_HIP_ERR_INVALID, "correlation length out of range");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = (long long)nx * ny * nz;
  Taps t{};
  double scale = 1.0;
  if (corr_len > 0) {
    t.radius = (int)std::ceil(3.0 * corr_len);
    double sum = 0.0;
    for (int d = -t.radius; d <= t.radius; d++) {
      const double u = (double)d / corr_len;
      t.w[d + t.radius] = std::exp(-0.5 * (u * u));
      sum = sum + t.w[d + t.radius];
    }
    double sq = 0.0;
    for (int d = 0; d <= 2 * t.radius; d++) {
      t.w[d] = t.w[d] / sum;
      sq = sq + t.w[d] * t.w[d];
    }
    scale = 1.0 / std::sqrt(sq * sq * sq);
  } else {
    t.radius = 0;
    t.w[0] = 1.0;
  }
  double *tmp = nullptr;
  TBSLA_TRY(device_alloc((void **)&tmp, (size_t)n * 8));
  Grid3 g{nx, ny, nz};
  k_normal_noise<<<sgrid(n), kBlock, 0, s>>>(n, mix64(seed), d_kappa);
  k_filter_axis<<<sgrid(n), kBlock, 0, s>>>(g, 0, t, d_kappa, tmp);
  k_filter_axis<<<sgrid(n), kBlock, 0, s>>>(g, 1, t, tmp, d_kappa);
  k_filter_axis<<<sgrid(n), kBlock, 0, s>>>(g, 2, t, d_kappa, tmp);
  k_lognormal<<<sgrid(n), kBlock, 0, s>>>(n, mu, sigma, scale, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(d_kappa, tmp, (size_t)n * 8, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_fill_uniform_vector(double *d_v, long long n, unsigned long long seed,
                                             double lo, double hi, void *stream) {
  if (n < 0 || (!d_v && n > 0)) return fail(TBSLA_HIP_ERR_INVALID, "bad vector");
  if (n == 0) return TBSLA_HIP_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_uniform<<<sgrid(n), kBlock, 0, s>>>(d_v, n, mix64(seed), lo, hi);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_fill_sym_cqmat(tbsla_hip_matrix_t m, int n, int c, double q,
                                        unsigned seed_mult, long long value_seed, void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null handle");
  if (m->format != TBSLA_HIP_CSR) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "fills a CSR block");
  if (n < 1 || c < 0) return fail(TBSLA_HIP_ERR_INVALID, "bad size");
  hipStream_t s = static_cast<hipStream_t>(stream);
  tbsla_hip_matrix_t B = nullptr;
  TBSLA_TRY(tbsla_hip_create(&B, TBSLA_HIP_CSR));
  int st = tbsla_hip_fill_cqmat(B, n, n, c, q, seed_mult, 0, 0, 1, 1, value_seed, stream);
  if (st) {
    tbsla_hip_destroy(B);
    return st;
  }
  const long long mb = B->stored;
  if (2 * mb + n > INT32_MAX) {
    tbsla_hip_destroy(B);
    return fail(TBSLA_HIP_ERR_RANGE, "symmetrised matrix too large for int32 positions");
  }
  const int cb = bits_of(n);
  const unsigned long long sentinel = (cb * 2 >= 64) ? ~0ull : ((1ull << (2 * cb)) - 1);
  unsigned long long *keys = nullptr;
  int *pay = nullptr, *cnt = nullptr, *start = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    (void)hipFree(keys);
    (void)hipFree(pay);
    (void)hipFree(cnt);
    (void)hipFree(start);
    tbsla_hip_destroy(B);
  };
  const long long m2 = std::max(2 * mb, 1LL);
  if ((st = device_alloc((void **)&keys, (size_t)m2 * 8)) ||
      (st = device_alloc((void **)&pay, (size_t)m2 * 4)) ||
      (st = device_alloc((void **)&cnt, (size_t)(n + 1) * 4)) ||
      (st = device_alloc((void **)&start, (size_t)(n + 1) * 4))) {
    cleanup();
    return st;
  }
  (void)hipMemsetAsync(cnt, 0, (size_t)(n + 1) * 4, s);
  k_sym_keys<<<sgrid(n), kBlock, 0, s>>>(n, mb, B->rowptr, B->colidx, cb, sentinel, keys, pay, cnt);
  if (2 * mb > 1) st = radix_sort_pairs(keys, pay, 2 * mb, 2 * cb, s);
  long long total = 0;
  if (!st) st = counts_to_rowptr(cnt, start, n, &total, s);
  release_arrays(m);
  set_square(m, n);
  if (!st) {
    k_sym_count<<<sgrid(n), kBlock, 0, s>>>(n, start, keys, cnt);
    st = device_alloc((void **)&m->rowptr, (size_t)(n + 1) * 4);
  }
  long long nnz = 0;
  if (!st) st = counts_to_rowptr(cnt, m->rowptr, n, &nnz, s);
  if (!st) st = device_alloc((void **)&m->colidx, (size_t)nnz * 4);
  if (!st) st = device_alloc((void **)&m->values, (size_t)nnz * 8);
  if (!st) {
    k_sym_fill<<<sgrid(n), kBlock, 0, s>>>(n, mb, cb, start, keys, pay, B->values, m->rowptr,
                                           m->colidx, m->values);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  }
  cleanup();
  if (st) return st;
  m->nnz = m->stored = nnz;
  m->has_arrays = 1;
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_transpose(tbsla_hip_matrix_t dst, tbsla_hip_matrix_t src, void *stream) {
  if (!dst || !src || dst == src) return fail(TBSLA_HIP_ERR_INVALID, "bad handles");
  if (src->format != TBSLA_HIP_CSR || dst->format != TBSLA_HIP_CSR)
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "transpose works on CSR blocks");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int *rp = nullptr, *ci = nullptr;
  double *v = nullptr;
  if (src->stored > 0 || src->ln_col > 0)
    TBSLA_TRY(csr_transpose_arrays(src, &rp, &ci, &v, s));
  release_arrays(dst);
  dst->n_row = src->n_col;
  dst->n_col = src->n_row;
  dst->ln_row = src->ln_col;
  dst->ln_col = src->ln_row;
  dst->f_row = src->f_col;
  dst->f_col = src->f_row;
  dst->pr = src->pc;
  dst->pc = src->pr;
  dst->NR = src->NC;
  dst->NC = src->NR;
  dst->max_col = 0;
  dst->row_sorted = 1;
  dst->rowptr = rp;
  dst->colidx = ci;
  dst->values = v;
  dst->nnz = dst->stored = src->stored;
  dst->has_arrays = 1;
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_block_diag(tbsla_hip_matrix_t dst, const tbsla_hip_matrix_t *parts,
                                    int k, void *stream) {
  if (!dst || !parts || k < 1) return fail(TBSLA_HIP_ERR_INVALID, "bad arguments");
  if (dst->format != TBSLA_HIP_CSR) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "builds a CSR block");
  long long rows = 0, cols = 0, nnz = 0;
  for (int i = 0; i < k; i++) {
    if (!parts[i] || parts[i] == dst || parts[i]->format != TBSLA_HIP_CSR)
      return fail(TBSLA_HIP_ERR_INVALID, "parts must be other CSR blocks");
    rows += parts[i]->ln_row;
    cols += parts[i]->ln_col;
    nnz += parts[i]->stored;
  }
  if (rows > INT32_MAX || cols > INT32_MAX || nnz > INT32_MAX)
    return fail(TBSLA_HIP_ERR_RANGE, "concatenation beyond int32");
  hipStream_t s = static_cast<hipStream_t>(stream);
  release_arrays(dst);
  dst->n_row = dst->ln_row = (int)rows;
  dst->n_col = dst->ln_col = (int)cols;
  dst->f_row = dst->f_col = 0;
  dst->pr = dst->pc = 0;
  dst->NR = dst->NC = 1;
  dst->max_col = 0;
  dst->row_sorted = 1;
  TBSLA_TRY(device_alloc((void **)&dst->rowptr, (size_t)(rows + 1) * 4));
  TBSLA_TRY(device_alloc((void **)&dst->colidx, (size_t)std::max(nnz, 1LL) * 4));
  TBSLA_TRY(device_alloc((void **)&dst->values, (size_t)std::max(nnz, 1LL) * 8));
  long long r0 = 0, c0 = 0, e0 = 0;
  for (int i = 0; i < k; i++) {
    const tbsla_hip_matrix_s *p = parts[i];
    if (p->ln_row > 0)
      k_bd_rows<<<sgrid(p->ln_row), kBlock, 0, s>>>(p->ln_row, p->rowptr, (int)e0, dst->rowptr + r0);
    if (p->stored > 0)
      k_bd_entries<<<sgrid(p->stored), kBlock, 0, s>>>(p->stored, p->colidx, p->values,
                                                       (int)(c0 - p->f_col), dst->colidx + e0,
                                                       dst->values + e0);
    r0 += p->ln_row;
    c0 += p->ln_col;
    e0 += p->stored;
  }
  const int last = (int)nnz;
  TBSLA_HIP_CHECK(hipMemcpyAsync(dst->rowptr + rows, &last, 4, hipMemcpyHostToDevice, s));
  TBSLA_HIP_LAUNCH_CHECK();
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  dst->nnz = dst->stored = nnz;
  dst->has_arrays = 1;
  return TBSLA_HIP_SUCCESS;
}
