// gen.hip -- fill_cqmat / fill_cdiag generators for the five formats.
//
// Bit-exact restatement of the reference fills on the GPU.  Every generator
// index k is independent (its rand_r seed is seed_mult * k,
// values_generation.cpp:29-32), so each row decodes its own contiguous
// index range; the per-row in-block counts are prefix-summed into rowptr /
// triplet offsets, and a second pass writes the entries in generation order.
// The structure is therefore identical to the reference's sequential loops,
// whatever the launch shape.
#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

// Block and generator-index layout of one fill_cqmat block
// (MatrixCSR.cpp:388-417 and the loop bounds at :423, :448, :465).
struct CqmatBlock {
  CqmatParams p;
  long long f_row, ln_row, f_col, ln_col;
  long long cc;        // min(c, n_col)
  long long min_;      // first tail row: min(n_col - cc + 1, n_row)
  long long full_end;  // rows [f_row, full_end) hold cc generator entries
  long long row_end;   // rows beyond hold none: min(n_row, n_col, f_row + ln_row)
  long long incr;      // generator index of row f_row's first entry
  long long nv;        // generator entries of the block's rows
  long long base_tail; // generator index of the first tail row of the block
  long long tail_a;    // its tail position (row - min_)
};

__host__ __device__ inline long long tail_prefix(long long cc, long long b) {
  // sum_{u<b} max(0, cc - 1 - u)
  if (cc <= 0) return 0;
  if (b <= cc - 1) return b * (cc - 1) - b * (b - 1) / 2;
  return (cc - 1) * cc / 2;
}

// Generator range [k0, k0 + cnt) of global row i.
__device__ inline void row_range(const CqmatBlock &b, long long i, long long &k0, long long &cnt) {
  if (i < b.full_end) {
    k0 = b.incr + (i - b.f_row) * b.cc;
    cnt = b.cc;
  } else if (i < b.row_end) {
    long long u = i - b.min_;
    cnt = b.cc - 1 - u;
    if (cnt < 0) cnt = 0;
    k0 = b.base_tail + tail_prefix(b.cc, u) - tail_prefix(b.cc, b.tail_a);
  } else {
    k0 = 0;
    cnt = 0;
  }
}

__device__ inline bool in_block(const CqmatBlock &b, int64_t ii, int64_t jj) {
  return ii >= b.f_row && ii < b.f_row + b.ln_row && jj >= b.f_col && jj < b.f_col + b.ln_col;
}

__device__ inline double entry_value(bool unit, uint64_t seed_hash, long long k) {
  return unit ? unit_value(seed_hash, (uint64_t)k) : 1.0;
}

// counts[r] = in-block entries of local row r; optionally the ELL width
// including the tail-row slot quirk (MatrixELL.cpp:408-412).
__global__ __launch_bounds__(kBlock) void k_cqmat_count_rows(CqmatBlock b, int *__restrict__ counts,
                                                             int *__restrict__ ell_width) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= b.ln_row) return;
  long long i = b.f_row + r, k0, cnt;
  row_range(b, i, k0, cnt);
  int n = 0;
  for (long long j = 0; j < cnt; j++) {
    int64_t ii, jj;
    cqmat_decode(b.p, (uint64_t)(k0 + j), ii, jj);
    n += in_block(b, ii, jj) ? 1 : 0;
  }
  counts[r] = n;
  if (ell_width) {
    int nbc = n;
    if (i >= b.full_end && i < b.row_end)
      for (long long j = cnt; j < b.cc; j++)
        if (j >= b.f_col && j < b.f_col + b.ln_col) nbc++;
    if (nbc > 0) atomicMax(ell_width, nbc);
  }
}

// Number of generator indices in [k_begin, k_end) whose entry lies in the
// block (the count passes at MatrixCSR.cpp:423-434, MatrixSCOO.cpp:338-350).
__global__ __launch_bounds__(kBlock) void k_cqmat_count_range(CqmatBlock b, long long k_begin,
                                                              long long k_end,
                                                              unsigned long long *__restrict__ out) {
  __shared__ unsigned long long lds[kBlock / kWave];
  unsigned long long n = 0;
  for (long long k = k_begin + (long long)blockIdx.x * kBlock + threadIdx.x; k < k_end;
       k += (long long)gridDim.x * kBlock) {
    int64_t ii, jj;
    cqmat_decode(b.p, (uint64_t)k, ii, jj);
    n += in_block(b, ii, jj) ? 1 : 0;
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) n += __shfl_xor(n, off, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / kWave; w++) t += lds[w];
    if (t) atomicAdd(out, t);
  }
}

// CSR (MatrixCSR.cpp:446-484) and SCOO (MatrixSCOO.cpp:367-401): the row's
// in-block entries, in generation order, from offs[r].
__global__ __launch_bounds__(kBlock) void k_cqmat_fill_rows(CqmatBlock b, const int *__restrict__ offs,
                                                            int *__restrict__ rows,
                                                            int *__restrict__ cols,
                                                            double *__restrict__ vals, int unit,
                                                            uint64_t seed_hash) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= b.ln_row) return;
  long long i = b.f_row + r, k0, cnt;
  row_range(b, i, k0, cnt);
  long long pos = offs[r];
  for (long long j = 0; j < cnt; j++) {
    int64_t ii, jj;
    cqmat_decode(b.p, (uint64_t)(k0 + j), ii, jj);
    if (in_block(b, ii, jj)) {
      if (rows) rows[pos] = (int)ii;
      cols[pos] = (int)jj;
      vals[pos] = entry_value(unit, seed_hash, k0 + j);
      pos++;
    }
  }
}

// ELL (MatrixELL.cpp:426-465), column-major slots; padding (col 0, value 0).
__global__ __launch_bounds__(kBlock) void k_cqmat_fill_ell(CqmatBlock b, int max_col,
                                                           int *__restrict__ cols,
                                                           double *__restrict__ vals, int unit,
                                                           uint64_t seed_hash) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= b.ln_row) return;
  long long i = b.f_row + r, k0, cnt;
  row_range(b, i, k0, cnt);
  int slot = 0;
  for (long long j = 0; j < cnt; j++) {
    int64_t ii, jj;
    cqmat_decode(b.p, (uint64_t)(k0 + j), ii, jj);
    if (in_block(b, ii, jj)) {
      size_t s = (size_t)slot * b.ln_row + r;
      cols[s] = (int)jj;
      vals[s] = entry_value(unit, seed_hash, k0 + j);
      slot++;
    }
  }
  for (; slot < max_col; slot++) {
    size_t s = (size_t)slot * b.ln_row + r;
    cols[s] = 0;
    vals[s] = 0.0;
  }
}

// DENSE (MatrixDENSE.cpp:238-269): duplicates accumulate in generation
// order, one thread per row, so the order of the additions is the reference's.
__global__ __launch_bounds__(kBlock) void k_cqmat_fill_dense(CqmatBlock b, double *__restrict__ dense,
                                                             int unit, uint64_t seed_hash) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= b.ln_row) return;
  long long i = b.f_row + r, k0, cnt;
  row_range(b, i, k0, cnt);
  double *row = dense + (size_t)r * b.ln_col;
  for (long long j = 0; j < cnt; j++) {
    int64_t ii, jj;
    cqmat_decode(b.p, (uint64_t)(k0 + j), ii, jj);
    if (in_block(b, ii, jj)) row[jj - b.f_col] += entry_value(unit, seed_hash, k0 + j);
  }
}

// COO (MatrixCOO.cpp:348-353): entries s .. s+n-1 of the generation order.
__global__ __launch_bounds__(kBlock) void k_cqmat_fill_coo(CqmatParams p, long long s, long long n,
                                                           int *__restrict__ rows,
                                                           int *__restrict__ cols,
                                                           double *__restrict__ vals, int unit,
                                                           uint64_t seed_hash) {
  long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  int64_t ii, jj;
  cqmat_decode(p, (uint64_t)(s + e), ii, jj);
  rows[e] = (int)ii;
  cols[e] = (int)jj;
  vals[e] = entry_value(unit, seed_hash, s + e);
}

// ---- cdiag ---------------------------------------------------------------

struct CdiagBlock {
  long long f_row, ln_row, f_col, ln_col, cd;
};

__device__ inline bool cd_in(const CdiagBlock &b, long long ii, long long jj) {
  return ii >= b.f_row && ii < b.f_row + b.ln_row && jj >= b.f_col && jj < b.f_col + b.ln_col;
}

// counts per row; rows/cols/vals written when offs != nullptr
// (MatrixCSR.cpp:323-368, MatrixSCOO.cpp:241-292).
__global__ __launch_bounds__(kBlock) void k_cdiag_rows(CdiagBlock b, int *__restrict__ counts,
                                                       const int *__restrict__ offs,
                                                       int *__restrict__ rows,
                                                       int *__restrict__ cols,
                                                       double *__restrict__ vals) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= b.ln_row) return;
  long long i = b.f_row + r;
  bool lo = cd_in(b, i, i - b.cd), hi = b.cd != 0 && cd_in(b, i, i + b.cd);
  if (counts) counts[r] = (lo ? 1 : 0) + (hi ? 1 : 0);
  if (offs) {
    long long pos = offs[r];
    if (lo) {
      if (rows) rows[pos] = (int)i;
      cols[pos] = (int)(i - b.cd);
      vals[pos++] = 1.0;
    }
    if (hi) {
      if (rows) rows[pos] = (int)i;
      cols[pos] = (int)(i + b.cd);
      vals[pos] = 1.0;
    }
  }
}

// ELL (MatrixELL.cpp:307-332): slot 0 the i-cdiag entry, slot 1 the i+cdiag
// entry, each (0, 0) when outside the block.
__global__ __launch_bounds__(kBlock) void k_cdiag_fill_ell(CdiagBlock b, int max_col,
                                                           int *__restrict__ cols,
                                                           double *__restrict__ vals) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= b.ln_row) return;
  long long i = b.f_row + r;
  bool lo = cd_in(b, i, i - b.cd);
  cols[r] = lo ? (int)(i - b.cd) : 0;
  vals[r] = lo ? 1.0 : 0.0;
  if (max_col > 1) {
    bool hi = cd_in(b, i, i + b.cd);
    cols[b.ln_row + r] = hi ? (int)(i + b.cd) : 0;
    vals[b.ln_row + r] = hi ? 1.0 : 0.0;
  }
}

// DENSE (MatrixDENSE.cpp:170-184): entries set to 1 (assignment).
__global__ __launch_bounds__(kBlock) void k_cdiag_fill_dense(CdiagBlock b, double *__restrict__ dense) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= b.ln_row) return;
  long long i = b.f_row + r;
  double *row = dense + (size_t)r * b.ln_col;
  if (cd_in(b, i, i - b.cd)) row[i - b.cd - b.f_col] = 1.0;
  if (b.cd != 0 && cd_in(b, i, i + b.cd)) row[i + b.cd - b.f_col] = 1.0;
}

// COO (MatrixCOO.cpp:312-317).
__global__ __launch_bounds__(kBlock) void k_cdiag_fill_coo(long long s, long long n, long long nc,
                                                           long long cd, int *__restrict__ rows,
                                                           int *__restrict__ cols,
                                                           double *__restrict__ vals) {
  long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  int64_t ii, jj;
  cdiag_decode((uint64_t)(s + e), nc, cd, ii, jj);
  rows[e] = (int)ii;
  cols[e] = (int)jj;
  vals[e] = 1.0;
}

__global__ __launch_bounds__(kBlock) void k_zero(double *__restrict__ d, long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    d[i] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_unit_vector(double *__restrict__ x, long long n,
                                                        uint64_t seed_hash) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    x[i] = unit_value(seed_hash, (uint64_t)i);
}

unsigned stride_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ---- host helpers --------------------------------------------------------

void set_geometry(tbsla_hip_matrix_s *m, int n_row, int n_col, int pr, int pc, int NR, int NC) {
  m->n_row = n_row;
  m->n_col = n_col;
  m->pr = pr;
  m->pc = pc;
  m->NR = NR;
  m->NC = NC;
  m->ln_row = (int)lnv(n_row, pr, NR);
  m->f_row = (int)pflv(n_row, pr, NR);
  m->ln_col = (int)lnv(n_col, pc, NC);
  m->f_col = (int)pflv(n_col, pc, NC);
}

// MatrixCSR.cpp:393-417 (identical loops in every format).
void prescan(long long n_row, long long n_col, long long cc, long long f_row, long long ln_row,
             long long &incr, long long &nv) {
  long long min_ = std::min(n_col - cc + 1, n_row);
  incr = 0;
  nv = 0;
  for (long long i = 0; i < min_; i++) {
    if (i < f_row) incr += cc;
    if (i >= f_row && i < f_row + ln_row) nv += cc;
    if (i >= f_row + ln_row) break;
  }
  for (long long i = 0; i < std::min(n_row, n_col) - min_; i++) {
    if (i + min_ < f_row) incr += cc - i - 1;
    if (i + min_ >= f_row && i + min_ < f_row + ln_row) nv += cc - i - 1;
    if (i + min_ >= f_row + ln_row) break;
  }
}

CqmatBlock make_block(const tbsla_hip_matrix_s *m, int c, double q, unsigned seed_mult) {
  CqmatBlock b;
  long long cc = std::min(c, m->n_col);
  b.p.nr = (uint64_t)m->n_row;
  b.p.nc = (uint64_t)m->n_col;
  b.p.c = (uint64_t)cc;
  b.p.nfull = (uint64_t)std::min<long long>((long long)m->n_col - cc + 1, m->n_row);
  b.p.q = q;
  b.p.seed_mult = seed_mult;
  b.f_row = m->f_row;
  b.ln_row = m->ln_row;
  b.f_col = m->f_col;
  b.ln_col = m->ln_col;
  b.cc = cc;
  b.min_ = std::min<long long>((long long)m->n_col - cc + 1, m->n_row);
  b.full_end = std::min<long long>(b.min_, (long long)m->f_row + m->ln_row);
  b.row_end = std::min<long long>(std::min(m->n_row, m->n_col), (long long)m->f_row + m->ln_row);
  prescan(m->n_row, m->n_col, cc, m->f_row, m->ln_row, b.incr, b.nv);
  long long nfull_block = std::max<long long>(0, b.full_end - b.f_row);
  b.base_tail = b.incr + nfull_block * cc;
  b.tail_a = std::max<long long>(b.f_row, b.min_) - b.min_;
  return b;
}

int count_range(const CqmatBlock &b, long long k_begin, long long k_end, long long &out,
                unsigned long long *d_counter, hipStream_t s) {
  out = 0;
  if (k_end <= k_begin) return TBSLA_HIP_SUCCESS;
  TBSLA_HIP_CHECK(hipMemsetAsync(d_counter, 0, sizeof(unsigned long long), s));
  k_cqmat_count_range<<<stride_grid(k_end - k_begin), kBlock, 0, s>>>(b, k_begin, k_end,
                                                                        d_counter);
  TBSLA_HIP_LAUNCH_CHECK();
  unsigned long long h = 0;
  TBSLA_HIP_CHECK(hipMemcpyAsync(&h, d_counter, sizeof(h), hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  out = (long long)h;
  return TBSLA_HIP_SUCCESS;
}

// Scratch layout of a fill: [0,64) a 64-bit counter, [64,128) the ELL
// width, [128,192) the scan total, [256, ...) the scan workspace.
constexpr size_t kScratchHead = 256;

int fill_scratch(tbsla_hip_matrix_s *m, long long rows) {
  return ensure_scratch(m, kScratchHead + scan_workspace_bytes(rows));
}
unsigned long long *scratch_counter(tbsla_hip_matrix_s *m) {
  return static_cast<unsigned long long *>(m->scratch);
}
int *scratch_width(tbsla_hip_matrix_s *m) {
  return reinterpret_cast<int *>(static_cast<char *>(m->scratch) + 64);
}

// counts (ln_row ints, at offs + 1) -> inclusive scan in place -> offs;
// returns the total.  The scratch must come from fill_scratch(m, n).
int counts_to_offsets(tbsla_hip_matrix_s *m, int *offs, long long n, long long &total,
                      hipStream_t s) {
  long long *d_total = reinterpret_cast<long long *>(static_cast<char *>(m->scratch) + 128);
  void *ws = static_cast<char *>(m->scratch) + kScratchHead;
  TBSLA_HIP_CHECK(hipMemsetAsync(offs, 0, sizeof(int), s));
  TBSLA_TRY(scan<int>(offs + 1, offs + 1, n, true, d_total, ws, s));
  TBSLA_HIP_CHECK(hipMemcpyAsync(&total, d_total, sizeof(long long), hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

int check_block_args(tbsla_hip_matrix_s *m, int n_row, int n_col, int pr, int pc, int NR,
                     int NC) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  if (n_row < 0 || n_col < 0 || NR < 1 || NC < 1 || pr < 0 || pc < 0 || pr >= NR || pc >= NC)
    return fail(TBSLA_HIP_ERR_INVALID, "bad block geometry");
  return TBSLA_HIP_SUCCESS;
}

}  // namespace

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_fill_cqmat(tbsla_hip_matrix_t m, int n_row, int n_col, int c, double q,
                                    unsigned seed_mult, int pr, int pc, int NR, int NC,
                                    long long value_seed, void *stream) {
  TBSLA_TRY(check_block_args(m, n_row, n_col, pr, pc, NR, NC));
  if (c < 0) return fail(TBSLA_HIP_ERR_INVALID, "c must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int fmt = m->format;
  release_arrays(m);
  const int unit = value_seed >= 0 ? 1 : 0;
  const uint64_t seed_hash = unit ? mix64((uint64_t)value_seed) : 0;

  if (fmt == TBSLA_HIP_COO) {
    // MatrixCOO::fill_cqmat (MatrixCOO.cpp:320-354): a slice of the global
    // generation order; MatrixCOO::init resets the grid fields (:14-34).
    long long cc = std::min(c, n_col);
    long long full = std::min<long long>((long long)n_col - cc + 1, n_row);
    long long gnv = full * cc;
    for (long long i = 0; i < std::min<long long>(n_row, n_col) - full; i++) gnv += cc - i - 1;
    m->n_row = m->ln_row = n_row;
    m->n_col = m->ln_col = n_col;
    m->f_row = m->f_col = 0;
    m->pr = pr;
    m->pc = pc;
    m->NR = NR;
    m->NC = NC;
    m->nnz = 0;
    m->stored = 0;
    m->row_sorted = 1;
    if (gnv == 0) return TBSLA_HIP_SUCCESS;
    long long st = pflv(gnv, pr * NC + pc, NR * NC), n = lnv(gnv, pr * NC + pc, NR * NC);
    if (n > INT32_MAX) return fail(TBSLA_HIP_ERR_RANGE, "COO slice beyond int32");
    m->pr = m->pc = 0;
    m->NR = m->NC = 1;
    m->nnz = n;
    m->stored = n;
    m->has_arrays = 1;
    TBSLA_TRY(device_alloc((void **)&m->rowidx, (size_t)std::max(n, 1LL) * sizeof(int)));
    TBSLA_TRY(device_alloc((void **)&m->colidx, (size_t)std::max(n, 1LL) * sizeof(int)));
    TBSLA_TRY(device_alloc((void **)&m->values, (size_t)std::max(n, 1LL) * sizeof(double)));
    CqmatParams p;
    p.nr = (uint64_t)n_row;
    p.nc = (uint64_t)n_col;
    p.c = (uint64_t)cc;
    p.nfull = (uint64_t)full;
    p.q = q;
    p.seed_mult = seed_mult;
    k_cqmat_fill_coo<<<grid_for(n), kBlock, 0, s>>>(p, st, n, m->rowidx, m->colidx, m->values,
                                                     unit, seed_hash);
    TBSLA_HIP_LAUNCH_CHECK();
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    return TBSLA_HIP_SUCCESS;
  }

  set_geometry(m, n_row, n_col, pr, pc, NR, NC);
  CqmatBlock b = make_block(m, c, q, seed_mult);
  m->nnz = 0;
  m->stored = 0;
  m->max_col = 0;
  m->row_sorted = 1;
  const long long ln_row = m->ln_row;
  const long long walk_end = b.incr + std::max<long long>(0, b.row_end - b.f_row) * b.cc;
  unsigned long long *d_counter = nullptr;

  // per-row in-block counts (every format needs them)
  int *offs = nullptr;
  TBSLA_TRY(device_alloc((void **)&offs, (size_t)(ln_row + 1) * sizeof(int)));
  int *d_width = nullptr;
  TBSLA_TRY(fill_scratch(m, ln_row));
  d_counter = scratch_counter(m);
  if (fmt == TBSLA_HIP_ELL) {
    d_width = scratch_width(m);
    TBSLA_HIP_CHECK(hipMemsetAsync(d_width, 0, sizeof(int), s));
  }
  if (ln_row > 0) {
    k_cqmat_count_rows<<<grid_for(ln_row), kBlock, 0, s>>>(b, offs + 1, d_width);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  int width = 0;
  if (d_width) {
    TBSLA_HIP_CHECK(hipMemcpyAsync(&width, d_width, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  long long total = 0;
  int st = counts_to_offsets(m, offs, ln_row, total, s);
  if (st != TBSLA_HIP_SUCCESS) {
    (void)hipFree(offs);
    return st;
  }
  if (total > INT32_MAX) {
    (void)hipFree(offs);
    return fail(TBSLA_HIP_ERR_RANGE, "block holds more than 2^31-1 entries");
  }

  if (fmt == TBSLA_HIP_CSR) {
    // nnz field: the count pass walks cc indices for every row up to row_end
    // (:423-434): the stored entries plus those of the extra indices.
    long long extra = 0;
    TBSLA_TRY(count_range(b, b.incr + b.nv, walk_end, extra, d_counter, s));
    m->nnz = total + extra;
    m->rowptr = offs;
    m->stored = total;
    m->has_arrays = b.nv > 0;
    TBSLA_TRY(device_alloc((void **)&m->colidx, (size_t)std::max(total, 1LL) * sizeof(int)));
    TBSLA_TRY(device_alloc((void **)&m->values, (size_t)std::max(total, 1LL) * sizeof(double)));
    if (b.nv > 0 && ln_row > 0) {
      k_cqmat_fill_rows<<<grid_for(ln_row), kBlock, 0, s>>>(b, offs, nullptr, m->colidx,
                                                             m->values, unit, seed_hash);
      TBSLA_HIP_LAUNCH_CHECK();
    }
  } else if (fmt == TBSLA_HIP_SCOO) {
    // nnz: the stored count after the fill (:402); with nothing generated
    // for the block's rows the count pass value stays (:338-355).
    if (b.nv == 0) {
      long long walk = 0;
      TBSLA_TRY(count_range(b, b.incr, walk_end, walk, d_counter, s));
      m->nnz = walk;
      m->stored = 0;
      m->has_arrays = 0;
    } else {
      m->nnz = total;
      m->stored = total;
      m->has_arrays = 1;
      TBSLA_TRY(device_alloc((void **)&m->rowidx, (size_t)std::max(total, 1LL) * sizeof(int)));
      TBSLA_TRY(device_alloc((void **)&m->colidx, (size_t)std::max(total, 1LL) * sizeof(int)));
      TBSLA_TRY(device_alloc((void **)&m->values, (size_t)std::max(total, 1LL) * sizeof(double)));
      if (ln_row > 0) {
        k_cqmat_fill_rows<<<grid_for(ln_row), kBlock, 0, s>>>(b, offs, m->rowidx, m->colidx,
                                                               m->values, unit, seed_hash);
        TBSLA_HIP_LAUNCH_CHECK();
      }
    }
    (void)hipFree(offs);
  } else if (fmt == TBSLA_HIP_ELL) {
    m->nnz = total;
    m->max_col = width;
    m->has_arrays = b.nv > 0;
    if (b.nv > 0) {
      size_t slots = (size_t)ln_row * (size_t)width;
      TBSLA_TRY(device_alloc((void **)&m->colidx, std::max<size_t>(slots, 1) * sizeof(int)));
      TBSLA_TRY(device_alloc((void **)&m->values, std::max<size_t>(slots, 1) * sizeof(double)));
      m->stored = (long long)slots;
      if (ln_row > 0) {
        k_cqmat_fill_ell<<<grid_for(ln_row), kBlock, 0, s>>>(b, width, m->colidx, m->values,
                                                              unit, seed_hash);
        TBSLA_HIP_LAUNCH_CHECK();
      }
    }
    (void)hipFree(offs);
  } else if (fmt == TBSLA_HIP_DENSE) {
    if (b.nv > 0) {
      m->nnz = total;
      m->has_arrays = 1;
      size_t cells = (size_t)ln_row * (size_t)m->ln_col;
      m->stored = (long long)cells;
      TBSLA_TRY(device_alloc((void **)&m->values, std::max<size_t>(cells, 1) * sizeof(double)));
      k_zero<<<stride_grid((long long)cells), kBlock, 0, s>>>(m->values, (long long)cells);
      TBSLA_HIP_LAUNCH_CHECK();
      if (ln_row > 0) {
        k_cqmat_fill_dense<<<grid_for(ln_row), kBlock, 0, s>>>(b, m->values, unit, seed_hash);
        TBSLA_HIP_LAUNCH_CHECK();
      }
    }
    (void)hipFree(offs);
  } else {
    (void)hipFree(offs);
    return fail(TBSLA_HIP_ERR_INVALID, "unknown format");
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_fill_cdiag(tbsla_hip_matrix_t m, int n_row, int n_col, int cdiag, int pr,
                                    int pc, int NR, int NC, void *stream) {
  TBSLA_TRY(check_block_args(m, n_row, n_col, pr, pc, NR, NC));
  if (cdiag < 0) return fail(TBSLA_HIP_ERR_INVALID, "cdiag must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int fmt = m->format;
  release_arrays(m);
  m->row_sorted = 1;
  m->stored = 0;
  m->max_col = 0;

  if (fmt == TBSLA_HIP_COO) {
    // MatrixCOO::fill_cdiag (MatrixCOO.cpp:280-318); grid fields restored
    // after init (:303-306).
    long long gnv = std::max<long long>(std::min<long long>(n_row, (long long)n_col - cdiag), 0) +
                    std::max<long long>(std::min<long long>((long long)n_row - cdiag, n_col), 0);
    if (cdiag == 0) gnv /= 2;
    m->n_row = m->ln_row = n_row;
    m->n_col = m->ln_col = n_col;
    m->f_row = m->f_col = 0;
    m->pr = pr;
    m->pc = pc;
    m->NR = NR;
    m->NC = NC;
    m->nnz = 0;
    if (gnv == 0) return TBSLA_HIP_SUCCESS;
    long long st = pflv(gnv, pr * NC + pc, NR * NC), n = lnv(gnv, pr * NC + pc, NR * NC);
    m->nnz = n;
    m->stored = n;
    m->has_arrays = 1;
    TBSLA_TRY(device_alloc((void **)&m->rowidx, (size_t)std::max(n, 1LL) * sizeof(int)));
    TBSLA_TRY(device_alloc((void **)&m->colidx, (size_t)std::max(n, 1LL) * sizeof(int)));
    TBSLA_TRY(device_alloc((void **)&m->values, (size_t)std::max(n, 1LL) * sizeof(double)));
    k_cdiag_fill_coo<<<grid_for(n), kBlock, 0, s>>>(st, n, n_col, cdiag, m->rowidx, m->colidx,
                                                     m->values);
    TBSLA_HIP_LAUNCH_CHECK();
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    return TBSLA_HIP_SUCCESS;
  }

  set_geometry(m, n_row, n_col, pr, pc, NR, NC);
  CdiagBlock b{m->f_row, m->ln_row, m->f_col, m->ln_col, cdiag};
  const long long ln_row = m->ln_row;
  int *offs = nullptr;
  TBSLA_TRY(fill_scratch(m, ln_row));
  TBSLA_TRY(device_alloc((void **)&offs, (size_t)(ln_row + 1) * sizeof(int)));
  if (ln_row > 0) {
    k_cdiag_rows<<<grid_for(ln_row), kBlock, 0, s>>>(b, offs + 1, nullptr, nullptr, nullptr,
                                                      nullptr);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  long long total = 0;
  int st = counts_to_offsets(m, offs, ln_row, total, s);
  if (st != TBSLA_HIP_SUCCESS) {
    (void)hipFree(offs);
    return st;
  }
  m->nnz = total;
  if (fmt == TBSLA_HIP_CSR || fmt == TBSLA_HIP_SCOO) {
    m->stored = total;
    m->has_arrays = total > 0;
    if (fmt == TBSLA_HIP_SCOO)
      TBSLA_TRY(device_alloc((void **)&m->rowidx, (size_t)std::max(total, 1LL) * sizeof(int)));
    TBSLA_TRY(device_alloc((void **)&m->colidx, (size_t)std::max(total, 1LL) * sizeof(int)));
    TBSLA_TRY(device_alloc((void **)&m->values, (size_t)std::max(total, 1LL) * sizeof(double)));
    if (total > 0) {
      k_cdiag_rows<<<grid_for(ln_row), kBlock, 0, s>>>(b, nullptr, offs, m->rowidx, m->colidx,
                                                        m->values);
      TBSLA_HIP_LAUNCH_CHECK();
    }
    if (fmt == TBSLA_HIP_CSR)
      m->rowptr = offs;
    else
      (void)hipFree(offs);
  } else if (fmt == TBSLA_HIP_ELL) {
    (void)hipFree(offs);
    m->max_col = cdiag == 0 ? 1 : 2;
    if (total > 0) {
      size_t slots = (size_t)ln_row * m->max_col;
      m->stored = (long long)slots;
      m->has_arrays = 1;
      TBSLA_TRY(device_alloc((void **)&m->colidx, slots * sizeof(int)));
      TBSLA_TRY(device_alloc((void **)&m->values, slots * sizeof(double)));
      k_cdiag_fill_ell<<<grid_for(ln_row), kBlock, 0, s>>>(b, m->max_col, m->colidx, m->values);
      TBSLA_HIP_LAUNCH_CHECK();
    }
  } else if (fmt == TBSLA_HIP_DENSE) {
    (void)hipFree(offs);
    if (total > 0) {
      size_t cells = (size_t)ln_row * (size_t)m->ln_col;
      m->stored = (long long)cells;
      m->has_arrays = 1;
      TBSLA_TRY(device_alloc((void **)&m->values, std::max<size_t>(cells, 1) * sizeof(double)));
      k_zero<<<stride_grid((long long)cells), kBlock, 0, s>>>(m->values, (long long)cells);
      TBSLA_HIP_LAUNCH_CHECK();
      k_cdiag_fill_dense<<<grid_for(ln_row), kBlock, 0, s>>>(b, m->values);
      TBSLA_HIP_LAUNCH_CHECK();
    }
  } else {
    (void)hipFree(offs);
    return fail(TBSLA_HIP_ERR_INVALID, "unknown format");
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_fill_unit_vector(double *d_x, long long n, long long seed, void *stream) {
  if (n < 0 || (n > 0 && !d_x)) return fail(TBSLA_HIP_ERR_INVALID, "bad vector");
  if (n == 0) return TBSLA_HIP_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_unit_vector<<<stride_grid(n), kBlock, 0, s>>>(d_x, n, mix64((uint64_t)seed));
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}
