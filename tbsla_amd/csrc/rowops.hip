// rowops.hip -- the row operations of tbsla::cpp::Matrix on one block:
// get_row_sums, normalize_rows, get_row_max_abs and apply_exponential
// (src/cpp/include/tbsla/cpp/Matrix.hpp:46-51).
//
// Each format keeps the reference's own semantics, including where they
// differ between formats:
//   get_row_sums     CSR, DENSE: s[i] = the row's sum (local row index);
//                    ELL: s[f_row + i] = the sum over all max_col slots;
//                    SCOO: s[row[k]] += values[k] (global row, not zeroed);
//                    COO: nothing.
//   normalize_rows   CSR: divide by s[i], or set 0 where s[i] == 0;
//                    ELL: divide every slot by s[f_row + i];
//                    DENSE: divide by s[i];  SCOO: divide by s[row[k]];
//                    COO: nothing.
//   get_row_max_abs  CSR only: the largest |z| over the row's nonzero
//                    values (0 for an empty row); other formats nothing.
//   apply_exponential CSR only: every nonzero z becomes exp(z - max_abs[i])
//                    (or pow(base, ...) when base > 0; no shift when
//                    max_abs is null); other formats nothing.
//
// Sums are formed in the reference's loop order (one thread per row walking
// it, or for SCOO the entries sorted stably by row), so they are
// bit-identical to a serial run of the reference.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

unsigned rgrid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 65536));
}

// CSR: s[i] = sum of row i in storage order (MatrixCSR.cpp:1147-1163)
__global__ __launch_bounds__(kBlock) void k_csr_row_sums(long long n, const int *__restrict__ rp,
                                                         const double *__restrict__ v,
                                                         double *__restrict__ s) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    double sum = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++) sum += v[j];
    s[i] = sum;
  }
}

// ELL (slot j of row i at j * ln_row + i): s[i] = sum over j = 0..max_col-1
// (MatrixELL.cpp:783-792); the caller offsets s by f_row.
__global__ __launch_bounds__(kBlock) void k_ell_row_sums(long long n, int max_col,
                                                         const double *__restrict__ v,
                                                         double *__restrict__ s) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    double sum = 0.0;
    for (int j = 0; j < max_col; j++) sum += v[(long long)j * n + i];
    s[i] = sum;
  }
}

// DENSE (row-major): s[i] = sum of row i (MatrixDENSE.cpp:292-301).  Each
// workgroup takes 64 rows and walks their columns in tiles of 64: the tile
// is read coalesced into LDS and each of 64 threads adds its row's 64
// values in column order, so the sum keeps the serial order.
constexpr int kDT = 64;
__global__ __launch_bounds__(kBlock) void k_dense_row_sums(long long n, long long ln_col,
                                                           const double *__restrict__ v,
                                                           double *__restrict__ s) {
  __shared__ double tile[kDT][kDT + 1];
  const long long r0 = (long long)blockIdx.x * kDT;
  const int t = threadIdx.x;
  double sum = 0.0;
  for (long long c0 = 0; c0 < ln_col; c0 += kDT) {
    for (int e = t; e < kDT * kDT; e += kBlock) {
      const int rr = e / kDT, cc = e % kDT;
      const long long row = r0 + rr, col = c0 + cc;
      tile[rr][cc] = (row < n && col < ln_col) ? v[row * ln_col + col] : 0.0;
    }
    __syncthreads();
    if (t < kDT) {
      const int m = ln_col - c0 < kDT ? (int)(ln_col - c0) : kDT;
      for (int cc = 0; cc < m; cc++) sum += tile[t][cc];
    }
    __syncthreads();
  }
  if (t < kDT && r0 + t < n) s[r0 + t] = sum;
}

// CSR normalize_rows (MatrixCSR.cpp:1177-1190): L lanes per row.
__global__ __launch_bounds__(kBlock) void k_csr_normalize_rows(long long n, int L,
                                                               const int *__restrict__ rp,
                                                               double *__restrict__ v,
                                                               const double *__restrict__ s) {
  const long long gid = ((long long)blockIdx.x * kBlock + threadIdx.x);
  const int lane = (int)(gid % L);
  for (long long i = gid / L; i < n; i += (long long)gridDim.x * kBlock / L) {
    const double d = s[i];
    for (int j = rp[i] + lane; j < rp[i + 1]; j += L) v[j] = d != 0 ? v[j] / d : 0.0;
  }
}

// ELL normalize_rows (MatrixELL.cpp:794-801): every slot, padding included.
__global__ __launch_bounds__(kBlock) void k_ell_normalize_rows(long long n, int max_col,
                                                               double *__restrict__ v,
                                                               const double *__restrict__ s) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const double d = s[i];
    for (int j = 0; j < max_col; j++) v[(long long)j * n + i] /= d;
  }
}

// DENSE normalize_rows (MatrixDENSE.cpp:303-310): a workgroup per row.
__global__ __launch_bounds__(kBlock) void k_dense_normalize_rows(long long n, long long ln_col,
                                                                 double *__restrict__ v,
                                                                 const double *__restrict__ s) {
  for (long long i = blockIdx.x; i < n; i += gridDim.x) {
    const double d = s[i];
    for (long long j = threadIdx.x; j < ln_col; j += kBlock) v[i * ln_col + j] /= d;
  }
}

// SCOO normalize_rows (MatrixSCOO.cpp:707-711): values[k] /= s[row[k]];
// rows outside [0, n_row) are left alone.
__global__ __launch_bounds__(kBlock) void k_scoo_normalize_rows(long long n, long long n_row,
                                                                const int *__restrict__ row,
                                                                double *__restrict__ v,
                                                                const double *__restrict__ s) {
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < n;
       k += (long long)gridDim.x * kBlock) {
    const long long r = row[k];
    if (r >= 0 && r < n_row) v[k] /= s[r];
  }
}

// SCOO row sums: key = row[k] - f_row (entries of other rows get the key
// ln_row, a bucket no sum reads), payload = k.
__global__ __launch_bounds__(kBlock) void k_scoo_rowkeys(long long n, const int *__restrict__ row,
                                                         int f_row, long long ln_row,
                                                         unsigned long long *__restrict__ keys,
                                                         int *__restrict__ pos) {
  long long k = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const long long r = (long long)row[k] - f_row;
  keys[k] = (r >= 0 && r < ln_row) ? (unsigned long long)r : (unsigned long long)ln_row;
  pos[k] = (int)k;
}

__global__ __launch_bounds__(kBlock) void k_count_keys(const unsigned long long *__restrict__ keys,
                                                       long long n, long long nb,
                                                       int *__restrict__ counts) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const unsigned long long k = keys[i];
    if (k <= (unsigned long long)nb) atomicAdd(counts + k, 1);
  }
}

// s[b] += the bucket's values in entry order (MatrixSCOO.cpp:701-705)
__global__ __launch_bounds__(kBlock) void k_bucket_sums(long long nb, const int *__restrict__ ptr,
                                                        const int *__restrict__ pos,
                                                        const double *__restrict__ v,
                                                        double *__restrict__ s) {
  const long long b = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (b >= nb) return;
  if (ptr[b] == ptr[b + 1]) return;
  double acc = s[b];
  for (int j = ptr[b]; j < ptr[b + 1]; j++) acc += v[pos[j]];
  s[b] = acc;
}

// CSR get_row_max_abs (MatrixCSR.cpp:1285-1315).  std::max(m, |z|) keeps m
// when |z| is NaN; the comparison below does the same.
__global__ __launch_bounds__(kBlock) void k_csr_row_max_abs(long long n, const int *__restrict__ rp,
                                                            const double *__restrict__ v,
                                                            double *__restrict__ out) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    double m = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++) {
      const double z = v[j];
      if (z != 0) {
        const double a = fabs(z);
        m = (m < a) ? a : m;
      }
    }
    out[i] = m;
  }
}

// CSR apply_exponential (MatrixCSR.cpp:1318-1334): L lanes per row.
__global__ __launch_bounds__(kBlock) void k_csr_exponential(long long n, int L,
                                                            const int *__restrict__ rp,
                                                            double *__restrict__ v,
                                                            const double *__restrict__ mx,
                                                            int base) {
  const long long gid = ((long long)blockIdx.x * kBlock + threadIdx.x);
  const int lane = (int)(gid % L);
  const double b = (double)base;
  for (long long i = gid / L; i < n; i += (long long)gridDim.x * kBlock / L) {
    const double shift = mx ? mx[i] : 0.0;
    for (int j = rp[i] + lane; j < rp[i + 1]; j += L) {
      const double z = v[j];
      if (z == 0) continue;
      const double e = mx ? z - shift : z;
      v[j] = base <= 0 ? exp(e) : pow(b, e);
    }
  }
}

// ---- fused row softmax ----------------------------------------------------
//
// The attention step's four row operations on one CSR block
// (tbsla_attention_mpi.cpp:183-191; with one block the row reductions across
// the process row are identities): get_row_max_abs, apply_exponential(
// max_abs, base), get_row_sums, normalize_rows -- in one pass over the
// values instead of six, bit-identical to the four calls.
//
// A group of L lanes takes a row; each lane holds V consecutive entries per
// trip (vector loads from the row's first entry rounded down to a multiple
// of V; entries outside the row are masked).  The row maximum of |z| is a
// group reduction (the maximum does not depend on order; NaN magnitudes are
// skipped as std::max skips them).  w = exp(z - m), or pow(base, z - m), is
// elementwise.  The row sum must follow the serial order of the reference
// loop, so it is a chain through the group: lane k adds its V values, in
// order, to the running sum it receives from lane k - 1, trip after trip.
// Then every entry becomes w / s, or 0 where s == 0.  Rows that fit in one
// trip (V L entries) keep their values in registers: one read and one write
// of each value.  Longer rows take three passes (the later two from cache),
// recomputing w, which is the same function of the same operands.
__device__ inline double softmax_w(double z, double m, int base, double b) {
  return z != 0 ? (base <= 0 ? exp(z - m) : pow(b, z - m)) : z;
}

template <int V, int L>
struct SmLoad {
  double z[V];
  bool in[V];
  __device__ void load(const double *__restrict__ v, int j, int beg, int end) {
    typedef double vd __attribute__((ext_vector_type(V)));
    if constexpr (V == 1) {
      in[0] = j >= beg && j < end;
      z[0] = in[0] ? v[j] : 0.0;
    } else if (j + V <= end && j + V > beg) {
      const vd x = *reinterpret_cast<const vd *>(v + j);
#pragma unroll
      for (int k = 0; k < V; k++) {
        z[k] = x[k];
        in[k] = j + k >= beg;
      }
    } else {
#pragma unroll
      for (int k = 0; k < V; k++) {
        in[k] = j + k >= beg && j + k < end;
        z[k] = in[k] ? v[j + k] : 0.0;
      }
    }
  }
  __device__ void store(double *__restrict__ v, int j, int beg, int end, const double (&o)[V]) {
    typedef double vd __attribute__((ext_vector_type(V)));
    if constexpr (V == 1) {
      if (in[0]) v[j] = o[0];
    } else if (j >= beg && j + V <= end) {
      vd x;
#pragma unroll
      for (int k = 0; k < V; k++) x[k] = o[k];
      *reinterpret_cast<vd *>(v + j) = x;
    } else {
#pragma unroll
      for (int k = 0; k < V; k++)
        if (in[k]) v[j + k] = o[k];
    }
  }
};

// x of the next lane down (lane i receives lane i - 1's value; lane 0
// receives 0): a DPP wavefront shift, two 32-bit moves, no LDS traffic
__device__ inline double lane_shr1(double x) {
  const long long bits = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)bits, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
  const int hi =
      __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The running sum through the group's lanes, in entry order: at step k lane
// k takes the sum of lane k - 1 (a one-lane shift of every lane's partial)
// and adds its V values; s, the sum so far (uniform in the group), starts
// lane 0.  Returns the group's sum to every lane of the group.
template <int V, int L>
__device__ inline double chain_sum(double s, const double (&w)[V], const bool (&in)[V], int lane,
                                   int gbase) {
  double acc = s;
#pragma unroll
  for (int k = 0; k < L; k++) {
    const double prev = k == 0 ? s : lane_shr1(acc);
    if (lane == k) {
      acc = prev;
#pragma unroll
      for (int q = 0; q < V; q++)
        if (in[q]) acc += w[q];
    }
  }
  return L == 1 ? acc : __shfl(acc, gbase + L - 1, kWave);
}

template <int L>
__device__ inline double group_max(double m) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) {
    const double o = __shfl_xor(m, off, kWave);
    m = (m < o) ? o : m;
  }
  return m;
}

template <int V, int L>
__global__ __launch_bounds__(kBlock) void k_csr_softmax(long long n, const int *__restrict__ rp,
                                                        double *__restrict__ v, int base,
                                                        double *__restrict__ mx_out,
                                                        double *__restrict__ s_out) {
  constexpr int G = kWave / L;
  const int wl = threadIdx.x & (kWave - 1);
  const int lane = wl & (L - 1), gbase = wl & ~(L - 1);
  const long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) / kWave;
  if (wave * G >= n) return;  // wave-uniform
  const long long row = wave * G + wl / L;
  const bool live = row < n;
  const int beg = live ? rp[row] : 0, end = live ? rp[row + 1] : 0;
  const int a0 = beg & ~(V - 1);
  int len = end - a0;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off, kWave));
  const int trips = (len + V * L - 1) / (V * L);  // wave-uniform: every shuffle sees all lanes
  const double b = (double)base;
  double m = 0.0, s = 0.0;
  if (trips <= 1) {
    SmLoad<V, L> e;
    e.load(v, a0 + V * lane, beg, end);
#pragma unroll
    for (int k = 0; k < V; k++)
      if (e.in[k] && e.z[k] != 0) {
        const double ab = fabs(e.z[k]);
        m = (m < ab) ? ab : m;
      }
    m = group_max<L>(m);
    double w[V];
#pragma unroll
    for (int k = 0; k < V; k++) w[k] = softmax_w(e.z[k], m, base, b);
    s = chain_sum<V, L>(0.0, w, e.in, lane, gbase);
    double o[V];
#pragma unroll
    for (int k = 0; k < V; k++) o[k] = s != 0 ? w[k] / s : 0.0;
    if (trips == 1) e.store(v, a0 + V * lane, beg, end, o);
  } else {
    for (int t = 0; t < trips; t++) {
      SmLoad<V, L> e;
      e.load(v, a0 + V * (t * L + lane), beg, end);
#pragma unroll
      for (int k = 0; k < V; k++)
        if (e.in[k] && e.z[k] != 0) {
          const double ab = fabs(e.z[k]);
          m = (m < ab) ? ab : m;
        }
    }
    m = group_max<L>(m);
    for (int t = 0; t < trips; t++) {
      SmLoad<V, L> e;
      e.load(v, a0 + V * (t * L + lane), beg, end);
      double w[V];
#pragma unroll
      for (int k = 0; k < V; k++) w[k] = softmax_w(e.z[k], m, base, b);
      s = chain_sum<V, L>(s, w, e.in, lane, gbase);
    }
    for (int t = 0; t < trips; t++) {
      SmLoad<V, L> e;
      const int j = a0 + V * (t * L + lane);
      e.load(v, j, beg, end);
      double o[V];
#pragma unroll
      for (int k = 0; k < V; k++) {
        const double w = softmax_w(e.z[k], m, base, b);
        o[k] = s != 0 ? w / s : 0.0;
      }
      e.store(v, j, beg, end, o);
    }
  }
  if (live && lane == 0) {
    if (mx_out) mx_out[row] = m;
    if (s_out) s_out[row] = s;
  }
}

template <int V, int L>
void softmax_go(long long n, const int *rp, double *v, int base, double *mx, double *sm,
                hipStream_t s) {
  const long long waves = (n + kWave / L - 1) / (kWave / L);
  k_csr_softmax<V, L><<<grid_for(waves * kWave), kBlock, 0, s>>>(n, rp, v, base, mx, sm);
}

template <int V>
void softmax_v(int L, long long n, const int *rp, double *v, int base, double *mx, double *sm,
               hipStream_t s) {
  switch (L) {
    case 1: softmax_go<V, 1>(n, rp, v, base, mx, sm, s); break;
    case 2: softmax_go<V, 2>(n, rp, v, base, mx, sm, s); break;
    case 4: softmax_go<V, 4>(n, rp, v, base, mx, sm, s); break;
    case 8: softmax_go<V, 8>(n, rp, v, base, mx, sm, s); break;
    case 16: softmax_go<V, 16>(n, rp, v, base, mx, sm, s); break;
    case 32: softmax_go<V, 32>(n, rp, v, base, mx, sm, s); break;
    default: softmax_go<V, 64>(n, rp, v, base, mx, sm, s); break;
  }
}

// V and L from the mean row length: one trip (V L entries) covers the mean
// row plus the alignment slack V - 1 (longer rows take several trips), and
// the pair minimises the wave instructions per entry of a simple model: the
// sum chain costs about V + 6 per lane step, the exponential and division
// about 52 per value held, 30 for the rest, over the entries a wave holds.
void softmax_launch(long long n, double mean, const int *rp, double *v, int base, double *mx,
                    double *sm, hipStream_t s) {
  int bestV = 1, bestL = 64;
  double best = 1e300;
  for (int V = 1; V <= 4; V *= 2) {
    int L = 1;
    while (L < 64 && V * L < mean + V - 1) L *= 2;
    const double held = (double)(kWave / L) * std::max(1.0, std::min(mean, (double)(V * L)));
    const double cost = (L * (V + 6.0) + 52.0 * V + 30.0) / held;
    if (cost < best) {
      best = cost;
      bestV = V;
      bestL = L;
    }
  }
  if (bestV == 4) softmax_v<4>(bestL, n, rp, v, base, mx, sm, s);
  else if (bestV == 2) softmax_v<2>(bestL, n, rp, v, base, mx, sm, s);
  else softmax_v<1>(bestL, n, rp, v, base, mx, sm, s);
}

// lanes per row for the entry-wise CSR kernels: the power of two nearest the
// mean row length, 1..64
int row_lanes(const tbsla_hip_matrix_s *m) {
  const double mean = m->ln_row > 0 ? (double)m->stored / m->ln_row : 1.0;
  int L = 1;
  while (L < 64 && 2 * L <= mean) L *= 2;
  return L;
}

unsigned lane_grid(long long n, int L) {
  long long g = (n * L + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 65536));
}

template <typename T>
struct Buf {
  T *p = nullptr;
  ~Buf() { if (p) (void)hipFree(p); }
  int alloc(long long n) { return device_alloc((void **)&p, (size_t)std::max(n, 1LL) * sizeof(T)); }
};

int scoo_row_sums(tbsla_hip_matrix_s *m, double *d_s, hipStream_t s) {
  const long long n = m->stored, nb = m->ln_row;
  Buf<unsigned long long> keys;
  Buf<int> pos, ptr;
  Buf<char> ws;
  TBSLA_TRY(keys.alloc(n));
  TBSLA_TRY(pos.alloc(n));
  TBSLA_TRY(ptr.alloc(nb + 2));  // ln_row buckets + the bucket of other rows
  k_scoo_rowkeys<<<grid_for(n), kBlock, 0, s>>>(n, m->rowidx, m->f_row, nb, keys.p, pos.p);
  TBSLA_HIP_LAUNCH_CHECK();
  int bits = 1;
  while (bits < 62 && (1LL << bits) <= nb) bits++;  // keys <= ln_row < 2^bits
  TBSLA_TRY(radix_sort_pairs(keys.p, pos.p, n, bits, s));
  TBSLA_HIP_CHECK(hipMemsetAsync(ptr.p, 0, (size_t)(nb + 2) * 4, s));
  k_count_keys<<<rgrid(n), kBlock, 0, s>>>(keys.p, n, nb, ptr.p + 1);
  TBSLA_HIP_LAUNCH_CHECK();
  const size_t wsb = scan_workspace_bytes(nb + 1);
  TBSLA_TRY(ws.alloc((long long)wsb + 64));
  TBSLA_TRY(scan<int>(ptr.p + 1, ptr.p + 1, nb + 1, true,
                      reinterpret_cast<long long *>(ws.p + wsb), ws.p, s));
  if (nb > 0) {
    k_bucket_sums<<<grid_for(nb), kBlock, 0, s>>>(nb, ptr.p, pos.p, m->values, d_s + m->f_row);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

}  // namespace

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_get_row_sums(tbsla_hip_matrix_t m, double *d_s, void *stream) {
  if (!m || !d_s) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = m->ln_row;
  switch (m->format) {
    case TBSLA_HIP_COO:
      return TBSLA_HIP_SUCCESS;  // MatrixCOO.cpp:382-384
    case TBSLA_HIP_CSR:
      if (!m->has_arrays || !m->rowptr) {
        TBSLA_TRY(launch_fill(d_s, n, 0.0, s));
      } else if (n > 0) {
        k_csr_row_sums<<<rgrid(n), kBlock, 0, s>>>(n, m->rowptr, m->values, d_s);
        TBSLA_HIP_LAUNCH_CHECK();
      }
      break;
    case TBSLA_HIP_ELL:
      if (!m->has_arrays || m->max_col == 0) {
        TBSLA_TRY(launch_fill(d_s + m->f_row, n, 0.0, s));
      } else if (n > 0) {
        k_ell_row_sums<<<rgrid(n), kBlock, 0, s>>>(n, m->max_col, m->values, d_s + m->f_row);
        TBSLA_HIP_LAUNCH_CHECK();
      }
      break;
    case TBSLA_HIP_DENSE:
      if (!m->has_arrays || !m->values) {
        TBSLA_TRY(launch_fill(d_s, n, 0.0, s));
      } else if (n > 0) {
        k_dense_row_sums<<<grid_for(n, kDT), kBlock, 0, s>>>(n, m->ln_col, m->values, d_s);
        TBSLA_HIP_LAUNCH_CHECK();
      }
      break;
    case TBSLA_HIP_SCOO:
      if (m->nnz == 0 || m->stored == 0 || !m->has_arrays) return TBSLA_HIP_SUCCESS;
      return scoo_row_sums(m, d_s, s);
    default:
      return fail(TBSLA_HIP_ERR_INVALID, "unknown format");
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_normalize_rows(tbsla_hip_matrix_t m, const double *d_s, void *stream) {
  if (!m || !d_s) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = m->ln_row;
  if (m->format == TBSLA_HIP_COO || !m->has_arrays) return TBSLA_HIP_SUCCESS;
  switch (m->format) {
    case TBSLA_HIP_CSR:
      if (n == 0 || !m->rowptr) return TBSLA_HIP_SUCCESS;
      {
        const int L = row_lanes(m);
        k_csr_normalize_rows<<<lane_grid(n, L), kBlock, 0, s>>>(n, L, m->rowptr, m->values, d_s);
      }
      break;
    case TBSLA_HIP_ELL:
      if (n == 0 || m->max_col == 0) return TBSLA_HIP_SUCCESS;
      k_ell_normalize_rows<<<rgrid(n), kBlock, 0, s>>>(n, m->max_col, m->values, d_s + m->f_row);
      break;
    case TBSLA_HIP_DENSE:
      if (n == 0 || m->ln_col == 0 || !m->values) return TBSLA_HIP_SUCCESS;
      k_dense_normalize_rows<<<(unsigned)std::min<long long>(n, 65536), kBlock, 0, s>>>(
          n, m->ln_col, m->values, d_s);
      break;
    case TBSLA_HIP_SCOO:
      if (m->nnz == 0 || m->stored == 0) return TBSLA_HIP_SUCCESS;
      k_scoo_normalize_rows<<<rgrid(m->stored), kBlock, 0, s>>>(m->stored, m->n_row, m->rowidx,
                                                                m->values, d_s);
      break;
    default:
      return fail(TBSLA_HIP_ERR_INVALID, "unknown format");
  }
  TBSLA_HIP_LAUNCH_CHECK();
  return values_changed(m, s);
}

extern "C" int tbsla_hip_get_row_max_abs(tbsla_hip_matrix_t m, double *d_max_abs, void *stream) {
  if (!m || !d_max_abs) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_CSR) return TBSLA_HIP_SUCCESS;  // no-ops in the other formats
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = m->ln_row;
  if (!m->has_arrays || !m->rowptr) {
    TBSLA_TRY(launch_fill(d_max_abs, n, 0.0, s));
  } else if (n > 0) {
    k_csr_row_max_abs<<<rgrid(n), kBlock, 0, s>>>(n, m->rowptr, m->values, d_max_abs);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_apply_exponential(tbsla_hip_matrix_t m, const double *d_max_abs, int base,
                                           void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_CSR) return TBSLA_HIP_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = m->ln_row;
  if (!m->has_arrays || !m->rowptr || n == 0) return TBSLA_HIP_SUCCESS;
  const int L = row_lanes(m);
  k_csr_exponential<<<lane_grid(n, L), kBlock, 0, s>>>(n, L, m->rowptr, m->values, d_max_abs,
                                                       base);
  TBSLA_HIP_LAUNCH_CHECK();
  return values_changed(m, s);
}

extern "C" int tbsla_hip_softmax_rows(tbsla_hip_matrix_t m, int base, double *d_max_abs,
                                      double *d_s, void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_CSR)
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "the fused row softmax needs a CSR block");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = m->ln_row;
  if (n == 0) return TBSLA_HIP_SUCCESS;
  if (!m->has_arrays || !m->rowptr) {  // an empty block: every row max and sum is 0
    if (d_max_abs) TBSLA_TRY(launch_fill(d_max_abs, n, 0.0, s));
    if (d_s) TBSLA_TRY(launch_fill(d_s, n, 0.0, s));
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    return TBSLA_HIP_SUCCESS;
  }
  softmax_launch(n, (double)m->stored / (double)n, m->rowptr, m->values, base, d_max_abs, d_s, s);
  TBSLA_HIP_LAUNCH_CHECK();
  return values_changed(m, s);
}
