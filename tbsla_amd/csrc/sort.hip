// sort.hip -- stable LSD radix sort and the COO conversion constructors.
//
// Radix sort of 64-bit keys with a 32-bit payload, 8-bit digits.  Per pass:
// a per-tile digit histogram, one exclusive scan over the digit-major
// histogram, and a stable scatter.  Stability inside a tile comes from
// processing elements in index order: each wavefront owns a contiguous
// 1024-element sub-tile, walks it in 64-element rounds, finds the lanes that
// share its digit with eight ballots and ranks itself among them; per-wave
// running bases live in LDS.  Integer work only: the permutation is unique.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

constexpr int kRadix = 256;
constexpr int kRounds = 16;                       // rounds of 64 per wave
constexpr int kSubTile = kWave * kRounds;         // 1024 elements per wave
constexpr int kTileElems = kSubTile * (kBlock / kWave);  // 4096 per workgroup

__global__ __launch_bounds__(kBlock) void k_digit_hist(const unsigned long long *__restrict__ keys,
                                                       long long n, int shift, long long ntiles,
                                                       int *__restrict__ hist) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kTileElems;
  for (int i = threadIdx.x; i < kTileElems; i += kBlock) {
    long long e = base + i;
    if (e < n) atomicAdd(&h[(keys[e] >> shift) & 0xFF], 1);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_digit_scatter(
    const unsigned long long *__restrict__ keys_in, const int *__restrict__ pay_in,
    unsigned long long *__restrict__ keys_out, int *__restrict__ pay_out, long long n, int shift,
    long long ntiles, const int *__restrict__ offsets) {
  __shared__ int wbase[kBlock / kWave][kRadix];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  for (int i = threadIdx.x; i < (kBlock / kWave) * kRadix; i += kBlock) (&wbase[0][0])[i] = 0;
  __syncthreads();
  const long long sub = (long long)blockIdx.x * kTileElems + (long long)wave * kSubTile;
  // phase A: digit counts of each wave's sub-tile
  for (int r = 0; r < kRounds; r++) {
    long long e = sub + r * kWave + lane;
    if (e < n) atomicAdd(&wbase[wave][(keys_in[e] >> shift) & 0xFF], 1);
  }
  __syncthreads();
  // per-wave starting position of each digit
  {
    const int d = threadIdx.x;
    int run = offsets[(long long)d * ntiles + blockIdx.x];
    for (int w = 0; w < kBlock / kWave; w++) {
      int c = wbase[w][d];
      wbase[w][d] = run;
      run += c;
    }
  }
  __syncthreads();
  // phase B: stable scatter, rounds in index order
  const unsigned long long lanes_below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  for (int r = 0; r < kRounds; r++) {
    long long e = sub + r * kWave + lane;
    const bool valid = e < n;
    unsigned long long key = valid ? keys_in[e] : 0ull;
    int d = (int)((key >> shift) & 0xFF);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __popcll(peers & lanes_below);
    const int base = valid ? wbase[wave][d] : 0;
    if (valid) {
      keys_out[base + rank] = key;
      pay_out[base + rank] = pay_in[e];
    }
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) wbase[wave][d] = base + __popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ __launch_bounds__(kBlock) void k_iota(int *__restrict__ p, long long n) {
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = (int)i;
}

__global__ __launch_bounds__(kBlock) void k_coo_keys(const int *__restrict__ rows,
                                                     const int *__restrict__ cols, long long n,
                                                     int col_bits,
                                                     unsigned long long *__restrict__ keys) {
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) keys[i] = ((unsigned long long)(unsigned)rows[i] << col_bits) | (unsigned)cols[i];
}

__global__ __launch_bounds__(kBlock) void k_gather_sorted(const int *__restrict__ perm,
                                                          const int *__restrict__ cols,
                                                          const double *__restrict__ vals,
                                                          long long n, int *__restrict__ cols_out,
                                                          double *__restrict__ vals_out) {
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  int p = perm[i];
  cols_out[i] = cols[p];
  vals_out[i] = vals[p];
}

__global__ __launch_bounds__(kBlock) void k_count_keys(const int *__restrict__ idx, long long n,
                                                       int *__restrict__ counts) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    atomicAdd(counts + idx[i], 1);
}

__global__ __launch_bounds__(kBlock) void k_max_int(const int *__restrict__ a, long long n,
                                                    int *__restrict__ out) {
  int m = 0;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    m = max(m, a[i]);
  if (m > 0) atomicMax(out, m);
}

// ELL slots from row-sorted (col, val) with rowptr: column-major, padded.
__global__ __launch_bounds__(kBlock) void k_rows_to_ell(long long n_rows, int max_col,
                                                        const int *__restrict__ rowptr,
                                                        const int *__restrict__ cols,
                                                        const double *__restrict__ vals,
                                                        int *__restrict__ ecols,
                                                        double *__restrict__ evals) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= n_rows) return;
  int b = rowptr[r], e = rowptr[r + 1];
  for (int j = 0; j < max_col; j++) {
    size_t s = (size_t)j * n_rows + r;
    bool in = b + j < e;
    ecols[s] = in ? cols[b + j] : 0;
    evals[s] = in ? vals[b + j] : 0.0;
  }
}

// DENSE(COO) (MatrixDENSE.cpp:34-36): per row, the row's triplets in COO
// order (stable row sort), so repeated cells add in the reference's order.
__global__ __launch_bounds__(kBlock) void k_rows_to_dense(long long n_rows, long long ln_col,
                                                          const int *__restrict__ rowptr,
                                                          const int *__restrict__ perm,
                                                          const int *__restrict__ cols,
                                                          const double *__restrict__ vals,
                                                          double *__restrict__ dense) {
  long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= n_rows) return;
  double *row = dense + (size_t)r * ln_col;
  for (int j = rowptr[r]; j < rowptr[r + 1]; j++) {
    int p = perm[j];
    row[cols[p]] += vals[p];
  }
}

__global__ __launch_bounds__(kBlock) void k_zero_d(double *__restrict__ d, long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    d[i] = 0.0;
}

unsigned stride_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

int bits_for(long long v) {
  int b = 0;
  while (b < 62 && (1LL << b) <= v) b++;
  return std::max(b, 1);
}

}  // namespace

int radix_sort_pairs(unsigned long long *keys, int *payload, long long n, int key_bits,
                     hipStream_t s) {
  if (n <= 1) return TBSLA_HIP_SUCCESS;
  const long long ntiles = (n + kTileElems - 1) / kTileElems;
  const long long hist_n = (long long)kRadix * ntiles;
  unsigned long long *keys_alt = nullptr;
  int *pay_alt = nullptr, *hist = nullptr;
  void *ws = nullptr;
  long long *d_total = nullptr;
  int st = TBSLA_HIP_SUCCESS;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    (void)hipFree(keys_alt);
    (void)hipFree(pay_alt);
    (void)hipFree(hist);
    (void)hipFree(ws);
  };
  if ((st = device_alloc((void **)&keys_alt, (size_t)n * 8)) ||
      (st = device_alloc((void **)&pay_alt, (size_t)n * 4)) ||
      (st = device_alloc((void **)&hist, (size_t)hist_n * 4)) ||
      (st = device_alloc(&ws, scan_workspace_bytes(hist_n) + 64))) {
    cleanup();
    return st;
  }
  d_total = reinterpret_cast<long long *>(static_cast<char *>(ws) + scan_workspace_bytes(hist_n));
  unsigned long long *kin = keys, *kout = keys_alt;
  int *pin = payload, *pout = pay_alt;
  const int passes = (key_bits + 7) / 8;
  for (int p = 0; p < passes && st == TBSLA_HIP_SUCCESS; p++) {
    const int shift = 8 * p;
    k_digit_hist<<<(unsigned)ntiles, kBlock, 0, s>>>(kin, n, shift, ntiles, hist);
    st = scan<int>(hist, hist, hist_n, false, d_total, ws, s);
    if (st) break;
    k_digit_scatter<<<(unsigned)ntiles, kBlock, 0, s>>>(kin, pin, kout, pout, n, shift, ntiles,
                                                        hist);
    std::swap(kin, kout);
    std::swap(pin, pout);
  }
  if (st == TBSLA_HIP_SUCCESS && kin != keys) {
    (void)hipMemcpyAsync(keys, kin, (size_t)n * 8, hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(payload, pin, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
  }
  hipError_t e = hipGetLastError();
  cleanup();
  if (st) return st;
  if (e != hipSuccess) return fail(TBSLA_HIP_ERR_HIP, std::string("radix sort: ") + hipGetErrorString(e));
  return TBSLA_HIP_SUCCESS;
}

// Stable permutation of COO triplets by (row, col) (csr::compare_row,
// src/cpp/utils/csr.cpp:3-8).  perm receives n ints.
static int coo_order(const tbsla_hip_matrix_s *coo, int *perm, hipStream_t s) {
  const long long n = std::min(coo->nnz, coo->stored);
  unsigned long long *keys = nullptr;
  TBSLA_TRY(device_alloc((void **)&keys, (size_t)std::max(n, 1LL) * 8));
  const int col_bits = bits_for(std::max(coo->n_col, 1));
  const int row_bits = bits_for(std::max(coo->n_row, 1));
  k_coo_keys<<<grid_for(n), kBlock, 0, s>>>(coo->rowidx, coo->colidx, n, col_bits, keys);
  k_iota<<<grid_for(n), kBlock, 0, s>>>(perm, n);
  int st = radix_sort_pairs(keys, perm, n, col_bits + row_bits, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(keys);
  return st;
}

}  // namespace tbsla_hip

using namespace tbsla_hip;

// <FMT>(const MatrixCOO&): MatrixCSR.cpp:49-96, MatrixELL.cpp:25-79,
// MatrixSCOO.cpp:75-101, MatrixDENSE.cpp:15-37.  The geometry is the COO's,
// with ln_row = n_row and ln_col = n_col.
extern "C" int tbsla_hip_convert(tbsla_hip_matrix_t dst, int format, tbsla_hip_matrix_t coo,
                                 void *stream) {
  if (!dst || !coo || dst == coo) return fail(TBSLA_HIP_ERR_INVALID, "bad handles");
  if (coo->format != TBSLA_HIP_COO) return fail(TBSLA_HIP_ERR_INVALID, "source must be COO");
  if (format < TBSLA_HIP_SCOO || format > TBSLA_HIP_DENSE)
    return fail(TBSLA_HIP_ERR_INVALID, "target must be SCOO, CSR, ELL or DENSE");
  hipStream_t s = static_cast<hipStream_t>(stream);
  release_arrays(dst);
  dst->format = format;
  dst->n_row = coo->n_row;
  dst->n_col = coo->n_col;
  dst->ln_row = coo->n_row;
  dst->ln_col = coo->n_col;
  dst->f_row = coo->f_row;
  dst->f_col = coo->f_col;
  dst->pr = coo->pr;
  dst->pc = coo->pc;
  dst->NR = coo->NR;
  dst->NC = coo->NC;
  dst->nnz = coo->nnz;
  dst->max_col = 0;
  dst->row_sorted = coo->row_sorted;
  // the reference converts nnz triplets; never more than are held
  const long long n = coo->has_arrays ? std::min(coo->nnz, coo->stored) : 0;
  const long long rows = dst->n_row;

  if (format == TBSLA_HIP_SCOO) {
    dst->stored = n;
    dst->has_arrays = n > 0;
    TBSLA_TRY(device_alloc((void **)&dst->rowidx, (size_t)std::max(n, 1LL) * 4));
    TBSLA_TRY(device_alloc((void **)&dst->colidx, (size_t)std::max(n, 1LL) * 4));
    TBSLA_TRY(device_alloc((void **)&dst->values, (size_t)std::max(n, 1LL) * 8));
    if (n > 0) {
      TBSLA_HIP_CHECK(hipMemcpyAsync(dst->rowidx, coo->rowidx, n * 4, hipMemcpyDeviceToDevice, s));
      TBSLA_HIP_CHECK(hipMemcpyAsync(dst->colidx, coo->colidx, n * 4, hipMemcpyDeviceToDevice, s));
      TBSLA_HIP_CHECK(hipMemcpyAsync(dst->values, coo->values, n * 8, hipMemcpyDeviceToDevice, s));
    }
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    return TBSLA_HIP_SUCCESS;
  }
  if (format == TBSLA_HIP_ELL && n == 0) {
    dst->nnz = 0;  // MatrixELL.cpp:39-41
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    return TBSLA_HIP_SUCCESS;
  }

  // row pointers from the row counts; stable (row, col) order
  int *rowptr = nullptr, *perm = nullptr;
  TBSLA_TRY(device_alloc((void **)&rowptr, (size_t)(rows + 1) * 4));
  TBSLA_TRY(device_alloc((void **)&perm, (size_t)std::max(n, 1LL) * 4));
  TBSLA_HIP_CHECK(hipMemsetAsync(rowptr, 0, (size_t)(rows + 1) * 4, s));
  if (n > 0) {
    k_count_keys<<<stride_grid(n), kBlock, 0, s>>>(coo->rowidx, n, rowptr + 1);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  TBSLA_TRY(ensure_scratch(dst, scan_workspace_bytes(rows) + 256));
  TBSLA_TRY(scan<int>(rowptr + 1, rowptr + 1, rows, true,
                      reinterpret_cast<long long *>(dst->scratch),
                      static_cast<char *>(dst->scratch) + 256, s));
  if (format == TBSLA_HIP_DENSE) {
    // DENSE needs only a stable row order (cells add in COO order)
    if (n > 0) {
      unsigned long long *keys = nullptr;
      TBSLA_TRY(device_alloc((void **)&keys, (size_t)n * 8));
      k_coo_keys<<<grid_for(n), kBlock, 0, s>>>(coo->rowidx, coo->rowidx, n, 0, keys);
      k_iota<<<grid_for(n), kBlock, 0, s>>>(perm, n);
      // key = row only (col_bits 0 and row OR row == row)
      int st = radix_sort_pairs(keys, perm, n, bits_for(std::max(coo->n_row, 1)), s);
      (void)hipStreamSynchronize(s);
      (void)hipFree(keys);
      if (st) return st;
    }
    const size_t cells = (size_t)rows * (size_t)dst->ln_col;
    TBSLA_TRY(device_alloc((void **)&dst->values, std::max<size_t>(cells, 1) * 8));
    k_zero_d<<<stride_grid((long long)cells), kBlock, 0, s>>>(dst->values, (long long)cells);
    if (rows > 0)
      k_rows_to_dense<<<grid_for(rows), kBlock, 0, s>>>(rows, dst->ln_col, rowptr, perm,
                                                        coo->colidx, coo->values, dst->values);
    TBSLA_HIP_LAUNCH_CHECK();
    dst->stored = (long long)cells;
    dst->has_arrays = 1;
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    (void)hipFree(perm);
    (void)hipFree(rowptr);
    return TBSLA_HIP_SUCCESS;
  }

  int *scol = nullptr;
  double *sval = nullptr;
  TBSLA_TRY(device_alloc((void **)&scol, (size_t)std::max(n, 1LL) * 4));
  TBSLA_TRY(device_alloc((void **)&sval, (size_t)std::max(n, 1LL) * 8));
  if (n > 0) {
    TBSLA_TRY(coo_order(coo, perm, s));
    k_gather_sorted<<<grid_for(n), kBlock, 0, s>>>(perm, coo->colidx, coo->values, n, scol, sval);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(perm);
  if (format == TBSLA_HIP_CSR) {
    dst->rowptr = rowptr;
    dst->colidx = scol;
    dst->values = sval;
    dst->stored = n;
    dst->has_arrays = 1;
    return TBSLA_HIP_SUCCESS;
  }
  // ELL: width = longest row
  int *d_w = reinterpret_cast<int *>(dst->scratch);
  TBSLA_HIP_CHECK(hipMemsetAsync(d_w, 0, 4, s));
  // row lengths = rowptr differences: recount into a temporary
  int *lens = nullptr;
  TBSLA_TRY(device_alloc((void **)&lens, (size_t)std::max(rows, 1LL) * 4));
  TBSLA_HIP_CHECK(hipMemsetAsync(lens, 0, (size_t)std::max(rows, 1LL) * 4, s));
  k_count_keys<<<stride_grid(n), kBlock, 0, s>>>(coo->rowidx, n, lens);
  k_max_int<<<stride_grid(rows), kBlock, 0, s>>>(lens, rows, d_w);
  int width = 0;
  TBSLA_HIP_CHECK(hipMemcpyAsync(&width, d_w, 4, hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  (void)hipFree(lens);
  dst->max_col = width;
  const size_t slots = (size_t)rows * (size_t)width;
  TBSLA_TRY(device_alloc((void **)&dst->colidx, std::max<size_t>(slots, 1) * 4));
  TBSLA_TRY(device_alloc((void **)&dst->values, std::max<size_t>(slots, 1) * 8));
  if (rows > 0 && width > 0)
    k_rows_to_ell<<<grid_for(rows), kBlock, 0, s>>>(rows, width, rowptr, scol, sval, dst->colidx,
                                                    dst->values);
  TBSLA_HIP_LAUNCH_CHECK();
  dst->stored = (long long)slots;
  dst->has_arrays = 1;
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  (void)hipFree(rowptr);
  (void)hipFree(scol);
  (void)hipFree(sval);
  return TBSLA_HIP_SUCCESS;
}
