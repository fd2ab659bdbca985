// capi.hip -- handle lifecycle, memory helpers, import/export and the SpMV
// entry points of the C ABI (include/tbsla_hip.h).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include <atomic>

#include "common.hpp"
#include "csr_rows.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int status, const std::string &msg) {
  set_error(msg);
  return status;
}

int device_alloc(void **p, size_t bytes) {
  *p = nullptr;
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(TBSLA_HIP_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(bytes) +
                                         "): " + hipGetErrorString(e));
  }
  return TBSLA_HIP_SUCCESS;
}

int ensure_scratch(tbsla_hip_matrix_s *m, size_t bytes) {
  if (m->scratch_bytes >= bytes) return TBSLA_HIP_SUCCESS;
  if (m->scratch) {
    // the previous scratch may still be read by queued work
    TBSLA_HIP_CHECK(hipDeviceSynchronize());
    (void)hipFree(m->scratch);
    m->scratch = nullptr;
    m->scratch_bytes = 0;
  }
  size_t want = std::max<size_t>(bytes, 4096);
  TBSLA_TRY(device_alloc(&m->scratch, want));
  m->scratch_bytes = want;
  return TBSLA_HIP_SUCCESS;
}

void release_arrays(tbsla_hip_matrix_s *m) {
  cg_cache_release(m);  // holds graphs that name these arrays
  split_release(m);     // a copy of the entries in another order
  if (m->rowptr || m->colidx || m->rowidx || m->values || m->coo_rowptr) (void)hipDeviceSynchronize();
  (void)hipFree(m->rowptr);
  (void)hipFree(m->colidx);
  (void)hipFree(m->rowidx);
  (void)hipFree(m->values);
  (void)hipFree(m->coo_rowptr);
  m->coo_rowptr = nullptr;
  m->coo_row0 = m->coo_nrows = 0;
  m->rowptr = m->colidx = m->rowidx = nullptr;
  m->values = nullptr;
  m->values_version++;
  m->has_arrays = 0;
  m->stored = 0;
}

}  // namespace tbsla_hip

using namespace tbsla_hip;

static hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

typedef double v2d_t __attribute__((ext_vector_type(2)));

// 16-byte pieces, four per thread in flight: piece i + u * stride for u < 4
__global__ __launch_bounds__(kBlock) void k_copy16(const v2d_t *__restrict__ a,
                                                   v2d_t *__restrict__ b, long long n16) {
  const long long stride = (long long)gridDim.x * kBlock;
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    v2d_t t[4];
#pragma unroll
    for (int u = 0; u < 4; u++) t[u] = __builtin_nontemporal_load(a + i + u * stride);
#pragma unroll
    for (int u = 0; u < 4; u++) b[i + u * stride] = t[u];
  }
  for (; i < n16; i += stride) b[i] = __builtin_nontemporal_load(a + i);
}

// One pass over n16 16-byte pieces, four per thread in flight; each
// workgroup writes one value (the xor of what it read) so the loads are kept.
__global__ __launch_bounds__(kBlock) void k_read16(const v2d_t *__restrict__ a, long long n16,
                                                   unsigned long long *__restrict__ sink) {
  __shared__ unsigned long long lds[kBlock / kWave];
  const long long stride = (long long)gridDim.x * kBlock;
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long h = 0;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    v2d_t t[4];
#pragma unroll
    for (int u = 0; u < 4; u++) t[u] = __builtin_nontemporal_load(a + i + u * stride);
#pragma unroll
    for (int u = 0; u < 4; u++)
      h ^= (unsigned long long)__double_as_longlong(t[u][0]) ^
           (unsigned long long)__double_as_longlong(t[u][1]);
  }
  for (; i < n16; i += stride) {
    const v2d_t t = __builtin_nontemporal_load(a + i);
    h ^= (unsigned long long)__double_as_longlong(t[0]) ^ (unsigned long long)__double_as_longlong(t[1]);
  }
  for (int off = kWave / 2; off > 0; off >>= 1) h ^= __shfl_xor(h, off, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / kWave; w++) t ^= lds[w];
    sink[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_copy1(const char *__restrict__ a, char *__restrict__ b,
                                                  long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    b[i] = a[i];
}

extern "C" {

const char *tbsla_hip_status_string(int status) {
  switch (status) {
    case TBSLA_HIP_SUCCESS: return "success";
    case TBSLA_HIP_ERR_INVALID: return "invalid argument";
    case TBSLA_HIP_ERR_HIP: return "HIP runtime error";
    case TBSLA_HIP_ERR_ALLOC: return "device allocation failed";
    case TBSLA_HIP_ERR_UNSUPPORTED: return "unsupported for this format";
    case TBSLA_HIP_ERR_RANGE: return "size beyond int32 indexing";
    case TBSLA_HIP_ERR_BREAKDOWN: return "CG breakdown";
  }
  return "unknown status";
}

const char *tbsla_hip_last_error(void) { return g_last_error.c_str(); }

int tbsla_hip_version(void) { return 1; }

int tbsla_hip_malloc(void **d_ptr, size_t bytes) {
  if (!d_ptr) return fail(TBSLA_HIP_ERR_INVALID, "null output pointer");
  return device_alloc(d_ptr, bytes);
}

int tbsla_hip_free(void *d_ptr) {
  TBSLA_HIP_CHECK(hipFree(d_ptr));
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_memcpy_htod(void *d_dst, const void *h_src, size_t bytes, void *stream) {
  if (bytes == 0) return TBSLA_HIP_SUCCESS;
  TBSLA_HIP_CHECK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
  TBSLA_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes, void *stream) {
  if (bytes == 0) return TBSLA_HIP_SUCCESS;
  TBSLA_HIP_CHECK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
  TBSLA_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream) {
  if (bytes == 0) return TBSLA_HIP_SUCCESS;
  TBSLA_HIP_CHECK(
      hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_copy_kernel(void *d_dst, const void *d_src, size_t bytes, void *stream) {
  if (bytes == 0) return TBSLA_HIP_SUCCESS;
  if (!d_dst || !d_src) return fail(TBSLA_HIP_ERR_INVALID, "null pointer");
  hipStream_t s = as_stream(stream);
  const bool aligned = ((uintptr_t)d_dst % 16 == 0) && ((uintptr_t)d_src % 16 == 0);
  const long long n16 = aligned ? (long long)(bytes / 16) : 0;
  if (n16 > 0) {
    k_copy16<<<8192, kBlock, 0, s>>>(static_cast<const v2d_t *>(d_src), static_cast<v2d_t *>(d_dst),
                                     n16);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  const long long rest = (long long)bytes - 16 * n16;
  if (rest > 0) {
    k_copy1<<<grid_for(rest), kBlock, 0, s>>>(static_cast<const char *>(d_src) + 16 * n16,
                                              static_cast<char *>(d_dst) + 16 * n16, rest);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_read_kernel(const void *d_src, size_t bytes, void *d_sink, void *stream) {
  if (bytes == 0) return TBSLA_HIP_SUCCESS;
  if (!d_src || !d_sink) return fail(TBSLA_HIP_ERR_INVALID, "null pointer");
  if ((uintptr_t)d_src % 16 != 0 || (uintptr_t)d_sink % 8 != 0)
    return fail(TBSLA_HIP_ERR_INVALID, "d_src must be 16-byte and d_sink 8-byte aligned");
  hipStream_t s = as_stream(stream);
  k_read16<<<8192, kBlock, 0, s>>>(static_cast<const v2d_t *>(d_src), (long long)(bytes / 16),
                                   static_cast<unsigned long long *>(d_sink));
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_stream_synchronize(void *stream) {
  TBSLA_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_device_count(int *count) {
  if (!count) return fail(TBSLA_HIP_ERR_INVALID, "null output pointer");
  *count = 0;
  TBSLA_HIP_CHECK(hipGetDeviceCount(count));
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_create(tbsla_hip_matrix_t *out, int format) {
  if (!out) return fail(TBSLA_HIP_ERR_INVALID, "null output pointer");
  if (format < TBSLA_HIP_COO || format > TBSLA_HIP_DENSE)
    return fail(TBSLA_HIP_ERR_INVALID, "unknown format");
  static std::atomic<unsigned long long> next_uid{1};
  *out = new tbsla_hip_matrix_s();
  (*out)->format = format;
  (*out)->uid = next_uid++;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_destroy(tbsla_hip_matrix_t m) {
  if (!m) return TBSLA_HIP_SUCCESS;
  release_arrays(m);
  (void)hipFree(m->scratch);
  delete m;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_get_geometry(tbsla_hip_matrix_t m, tbsla_hip_geometry *g) {
  if (!m || !g) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  g->n_row = m->n_row;
  g->n_col = m->n_col;
  g->f_row = m->f_row;
  g->f_col = m->f_col;
  g->ln_row = m->ln_row;
  g->ln_col = m->ln_col;
  g->pr = m->pr;
  g->pc = m->pc;
  g->NR = m->NR;
  g->NC = m->NC;
  g->nnz = m->nnz;
  g->stored = m->stored;
  g->max_col = m->max_col;
  g->format = m->format;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_export(tbsla_hip_matrix_t m, int *h_a, int *h_b, double *h_v, void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  hipStream_t s = as_stream(stream);
  auto d2h = [&](void *h, const void *d, size_t bytes) -> int {
    if (!h || !d || bytes == 0) return TBSLA_HIP_SUCCESS;
    TBSLA_HIP_CHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
    return TBSLA_HIP_SUCCESS;
  };
  const size_t st = (size_t)std::max(0LL, m->stored);
  switch (m->format) {
    case TBSLA_HIP_CSR:
      if (m->rowptr) TBSLA_TRY(d2h(h_a, m->rowptr, ((size_t)m->ln_row + 1) * sizeof(int)));
      else if (h_a) std::memset(h_a, 0, ((size_t)m->ln_row + 1) * sizeof(int));
      TBSLA_TRY(d2h(h_b, m->colidx, st * sizeof(int)));
      TBSLA_TRY(d2h(h_v, m->values, st * sizeof(double)));
      break;
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO:
      TBSLA_TRY(d2h(h_a, m->rowidx, st * sizeof(int)));
      TBSLA_TRY(d2h(h_b, m->colidx, st * sizeof(int)));
      TBSLA_TRY(d2h(h_v, m->values, st * sizeof(double)));
      break;
    case TBSLA_HIP_ELL: {
      // column-major on the device, row-major (reference layout) on the host
      if (!m->has_arrays) break;
      const size_t rows = (size_t)m->ln_row, w = (size_t)m->max_col, slots = rows * w;
      std::vector<int> cols(slots);
      std::vector<double> vals(slots);
      TBSLA_TRY(d2h(cols.data(), m->colidx, slots * sizeof(int)));
      TBSLA_TRY(d2h(vals.data(), m->values, slots * sizeof(double)));
      TBSLA_HIP_CHECK(hipStreamSynchronize(s));
      for (size_t i = 0; i < rows; i++)
        for (size_t j = 0; j < w; j++) {
          if (h_a) h_a[i * w + j] = cols[j * rows + i];
          if (h_v) h_v[i * w + j] = vals[j * rows + i];
        }
      break;
    }
    case TBSLA_HIP_DENSE:
      if (m->has_arrays)
        TBSLA_TRY(d2h(h_v, m->values, (size_t)m->ln_row * (size_t)m->ln_col * sizeof(double)));
      break;
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_import(tbsla_hip_matrix_t m, int format, const tbsla_hip_geometry *g,
                     const int *h_a, const int *h_b, const double *h_v, void *stream) {
  if (!m || !g) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (format < TBSLA_HIP_COO || format > TBSLA_HIP_DENSE)
    return fail(TBSLA_HIP_ERR_INVALID, "unknown format");
  if (g->ln_row < 0 || g->ln_col < 0 || g->stored < 0)
    return fail(TBSLA_HIP_ERR_INVALID, "bad geometry");
  hipStream_t s = as_stream(stream);
  // every index must lie inside the block: kernels index x, y and the
  // scatter targets with them unchecked
  {
    const long long st = g->stored;
    const long long r0 = g->f_row, r1 = (long long)g->f_row + g->ln_row;
    const long long c0 = g->f_col, c1 = (long long)g->f_col + g->ln_col;
    if (format == TBSLA_HIP_COO || format == TBSLA_HIP_SCOO) {
      if (st > 0 && (!h_a || !h_b || !h_v)) return fail(TBSLA_HIP_ERR_INVALID, "missing arrays");
      const bool coo = format == TBSLA_HIP_COO;
      for (long long i = 0; i < st; i++) {
        const long long r = h_a[i], c = h_b[i];
        if (coo ? (r < 0 || r >= g->n_row || c < 0 || c >= g->n_col)
                : (r < r0 || r >= r1 || c < c0 || c >= c1))
          return fail(TBSLA_HIP_ERR_INVALID, "triplet " + std::to_string(i) + " outside the block");
      }
    } else if (format == TBSLA_HIP_CSR) {
      if (!h_a) return fail(TBSLA_HIP_ERR_INVALID, "CSR import needs rowptr");
      if (h_a[0] != 0) return fail(TBSLA_HIP_ERR_INVALID, "rowptr[0] must be 0");
      for (long long i = 0; i < g->ln_row; i++)
        if (h_a[i + 1] < h_a[i]) return fail(TBSLA_HIP_ERR_INVALID, "rowptr not monotone");
      const long long n = h_a[g->ln_row];
      if (n > 0 && (!h_b || !h_v)) return fail(TBSLA_HIP_ERR_INVALID, "missing arrays");
      for (long long j = 0; j < n; j++)
        if (h_b[j] < c0 || h_b[j] >= c1)
          return fail(TBSLA_HIP_ERR_INVALID, "column " + std::to_string(j) + " outside the block");
    } else if (format == TBSLA_HIP_ELL) {
      const long long slots = (long long)g->ln_row * std::max(0, g->max_col);
      for (long long j = 0; h_a && j < slots; j++) {
        const bool pad = h_a[j] == 0 && (!h_v || h_v[j] == 0.0);
        if (!pad && (h_a[j] < c0 || h_a[j] >= c1))
          return fail(TBSLA_HIP_ERR_INVALID, "column " + std::to_string(j) + " outside the block");
      }
    }
  }
  release_arrays(m);
  m->format = format;
  m->n_row = g->n_row;
  m->n_col = g->n_col;
  m->f_row = g->f_row;
  m->f_col = g->f_col;
  m->ln_row = g->ln_row;
  m->ln_col = g->ln_col;
  m->pr = g->pr;
  m->pc = g->pc;
  m->NR = g->NR;
  m->NC = g->NC;
  m->nnz = g->nnz;
  m->max_col = format == TBSLA_HIP_ELL ? g->max_col : 0;
  m->row_sorted = 1;
  auto h2d = [&](void **d, const void *h, size_t bytes) -> int {
    TBSLA_TRY(device_alloc(d, bytes));
    if (h && bytes) TBSLA_HIP_CHECK(hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, s));
    return TBSLA_HIP_SUCCESS;
  };
  const size_t st = (size_t)g->stored;
  switch (format) {
    case TBSLA_HIP_CSR: {
      if (!h_a) return fail(TBSLA_HIP_ERR_INVALID, "CSR import needs rowptr");
      m->stored = h_a[g->ln_row];
      TBSLA_TRY(h2d((void **)&m->rowptr, h_a, ((size_t)g->ln_row + 1) * sizeof(int)));
      TBSLA_TRY(h2d((void **)&m->colidx, h_b, (size_t)m->stored * sizeof(int)));
      TBSLA_TRY(h2d((void **)&m->values, h_v, (size_t)m->stored * sizeof(double)));
      m->has_arrays = 1;
      break;
    }
    case TBSLA_HIP_COO:
    case TBSLA_HIP_SCOO: {
      m->stored = g->stored;
      for (size_t i = 1; h_a && i < st; i++)
        if (h_a[i] < h_a[i - 1]) {
          m->row_sorted = 0;
          break;
        }
      TBSLA_TRY(h2d((void **)&m->rowidx, h_a, st * sizeof(int)));
      TBSLA_TRY(h2d((void **)&m->colidx, h_b, st * sizeof(int)));
      TBSLA_TRY(h2d((void **)&m->values, h_v, st * sizeof(double)));
      m->has_arrays = st > 0;
      break;
    }
    case TBSLA_HIP_ELL: {
      const size_t rows = (size_t)g->ln_row, w = (size_t)std::max(0, g->max_col);
      const size_t slots = rows * w;
      std::vector<int> cols(slots);
      std::vector<double> vals(slots);
      for (size_t i = 0; i < rows; i++)
        for (size_t j = 0; j < w; j++) {
          cols[j * rows + i] = h_a ? h_a[i * w + j] : 0;
          vals[j * rows + i] = h_v ? h_v[i * w + j] : 0.0;
        }
      m->stored = (long long)slots;
      TBSLA_TRY(h2d((void **)&m->colidx, cols.data(), slots * sizeof(int)));
      TBSLA_TRY(h2d((void **)&m->values, vals.data(), slots * sizeof(double)));
      TBSLA_HIP_CHECK(hipStreamSynchronize(s));
      m->has_arrays = 1;
      break;
    }
    case TBSLA_HIP_DENSE: {
      const size_t cells = (size_t)g->ln_row * (size_t)g->ln_col;
      m->stored = (long long)cells;
      TBSLA_TRY(h2d((void **)&m->values, h_v, cells * sizeof(double)));
      m->has_arrays = h_v != nullptr;
      break;
    }
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_device_arrays(tbsla_hip_matrix_t m, void **d_a, void **d_b, void **d_v) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  if (d_a) *d_a = m->format == TBSLA_HIP_CSR ? (void *)m->rowptr : (void *)m->rowidx;
  if (d_b) *d_b = m->colidx;
  if (d_v) *d_v = m->values;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_Ax(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int vect_incr,
                 void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  return spmv_launch(m, d_r, d_v, vect_incr, false, as_stream(stream));
}

int tbsla_hip_spmv(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int vect_incr,
                   void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  return spmv_launch(m, d_r, d_v, vect_incr, true, as_stream(stream));
}

// Matrix::a_axpx_ (Matrix.cpp:5-14): tmp = spmv(v); tmp[i] += v[i] over
// ln_col entries; r = spmv(tmp).
int tbsla_hip_a_axpx(tbsla_hip_matrix_t m, double *d_r, const double *d_v, double *d_tmp,
                     int vect_incr, void *stream) {
  if (!m || !d_tmp) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  hipStream_t s = as_stream(stream);
  TBSLA_TRY(spmv_launch(m, d_tmp, d_v, vect_incr, true, s));
  TBSLA_TRY(launch_axpy_inplace(d_tmp, d_v, m->ln_col, s));
  return spmv_launch(m, d_r, d_tmp, vect_incr, true, s);
}

// Matrix::AAxpAx (Matrix.cpp:16-23).
int tbsla_hip_AAxpAx(tbsla_hip_matrix_t m, double *d_r, const double *d_v, double *d_buf,
                     int vect_incr, void *stream) {
  if (!m || !d_buf) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  hipStream_t s = as_stream(stream);
  TBSLA_TRY(spmv_launch(m, d_buf, d_v, vect_incr, false, s));
  TBSLA_TRY(launch_axpy_inplace(d_buf, d_v, m->ln_col, s));
  return spmv_launch(m, d_r, d_buf, vect_incr, false, s);
}

// Matrix::AAxpAxpx (Matrix.cpp:25-31).
int tbsla_hip_AAxpAxpx(tbsla_hip_matrix_t m, double *d_r, const double *d_v, double *d_buf,
                       int vect_incr, void *stream) {
  TBSLA_TRY(tbsla_hip_AAxpAx(m, d_r, d_v, d_buf, vect_incr, stream));
  return launch_axpy_inplace(d_r, d_v, m->ln_col, as_stream(stream));
}

int tbsla_hip_time_spmv(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int iters,
                        float *ms, void *stream) {
  if (!m || !ms || iters < 1) return fail(TBSLA_HIP_ERR_INVALID, "bad argument");
  hipStream_t s = as_stream(stream);
  hipEvent_t e0, e1;
  TBSLA_HIP_CHECK(hipEventCreate(&e0));
  TBSLA_HIP_CHECK(hipEventCreate(&e1));
  TBSLA_HIP_CHECK(hipEventRecord(e0, s));
  int st = TBSLA_HIP_SUCCESS;
  for (int i = 0; i < iters && st == TBSLA_HIP_SUCCESS; i++)
    st = spmv_launch(m, d_r, d_v, 0, true, s);
  TBSLA_HIP_CHECK(hipEventRecord(e1, s));
  TBSLA_HIP_CHECK(hipEventSynchronize(e1));
  TBSLA_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

int tbsla_hip_set_csr_schedule(tbsla_hip_matrix_t m, int lanes, int rows) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  if (lanes != 0 && (lanes < 1 || lanes > 64 || (lanes & (lanes - 1))))
    return fail(TBSLA_HIP_ERR_INVALID, "lanes must be 0 or a power of two in 1..64");
  if (rows != 0 && rows != 1 && rows != 2 && rows != 4)
    return fail(TBSLA_HIP_ERR_INVALID, "rows must be 0, 1, 2 or 4");
  m->lanes_per_row = lanes;
  m->rows_per_group = rows;
  m->csr_vec = 0;
  m->csr_nt = -1;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_get_csr_schedule(tbsla_hip_matrix_t m, int *lanes, int *rows) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  int L = pick_csr_lanes(m);
  if (lanes) *lanes = L;
  if (rows) *rows = pick_csr_rows(m, L);
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_set_csr_kernel(tbsla_hip_matrix_t m, int lanes, int rows, int vec,
                             int nontemporal) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  if (!csr_rows_valid(vec, lanes, rows, nontemporal))
    return fail(TBSLA_HIP_ERR_INVALID,
                "no CSR kernel for this (lanes, rows, vec, nontemporal): vec 1 takes lanes 1..64 "
                "and rows 1, 2, 4; vec 2 lanes 4..32, vec 4 lanes 4..16, both rows 1, 2");
  m->lanes_per_row = lanes;
  m->rows_per_group = rows;
  m->csr_vec = vec;
  m->csr_nt = nontemporal;
  return TBSLA_HIP_SUCCESS;
}

int tbsla_hip_get_csr_kernel(tbsla_hip_matrix_t m, int *lanes, int *rows, int *vec,
                             int *nontemporal) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null matrix handle");
  const int L = pick_csr_lanes(m);
  if (lanes) *lanes = L;
  if (rows) *rows = pick_csr_rows(m, L);
  if (vec) *vec = pick_csr_vec(m);
  if (nontemporal) *nontemporal = pick_csr_nt(m);
  return TBSLA_HIP_SUCCESS;
}

// diagonal layout: offsets searched for dense diagonals (wide enough for a
// 3-D stencil's +-n^2 diagonals up to n = 256)
constexpr int kDiagWindow = 65536;
constexpr float kDerivedGain = 0.97f;

int tbsla_hip_autotune(tbsla_hip_matrix_t m, double *d_r, const double *d_v, int iters,
                       float *best_us, void *stream) {
  if (!m || iters < 1) return fail(TBSLA_HIP_ERR_INVALID, "bad argument");
  float ms = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool coo = m->format == TBSLA_HIP_COO || m->format == TBSLA_HIP_SCOO;
  if ((m->format == TBSLA_HIP_ELL && m->has_arrays && m->max_col > 0 && m->ln_row > 0) ||
      (coo && m->has_arrays && m->stored > 0 && m->row_sorted)) {
    // ELL: its column-major kernel; COO / SCOO: the row kernel over their
    // triplets' row pointer; or the split or diagonal layout (8 or 16 slices)
    split_release(m);
    TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, 1, &ms, stream));
    TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, iters, &ms, stream));
    if (coo && !m->coo_rowptr) {  // the timing found the triplets unsorted
      if (best_us) *best_us = ms * 1000.0f / iters;
      return TBSLA_HIP_SUCCESS;
    }
    float best = ms;
    const float own_ms = ms;  // the format's own kernel
    int best_slices = 0, best_layout = 0, best_window = 0;
    for (int layout : {TBSLA_HIP_CSR_SPLIT, TBSLA_HIP_CSR_DIAG})
      for (int slices : {16, 8}) {
        if (layout == TBSLA_HIP_CSR_DIAG && (coo ? (m->format == TBSLA_HIP_COO ? m->n_col : m->ln_col)
                                                 : m->ln_col) == 0)
          break;
        const int window = layout == TBSLA_HIP_CSR_DIAG ? kDiagWindow : 64;
        TBSLA_TRY(split_build(m, window, slices, s, layout));
        TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, 1, &ms, stream));
        TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, iters, &ms, stream));
        if (ms < best && ms < own_ms * kDerivedGain) {
          best = ms;
          best_slices = slices;
          best_layout = layout;
          best_window = window;
        }
        // the LDS-slice form does not depend on the slice count
        if (layout == TBSLA_HIP_CSR_DIAG && split_lds_block_rows(m) > 0) break;
      }
    split_release(m);
    if (best_slices) TBSLA_TRY(split_build(m, best_window, best_slices, s, best_layout));
    if (best_us) *best_us = best * 1000.0f / iters;
    return TBSLA_HIP_SUCCESS;
  }
  if (m->format != TBSLA_HIP_CSR) {
    TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, iters, &ms, stream));
    if (best_us) *best_us = ms * 1000.0f / iters;
    return TBSLA_HIP_SUCCESS;
  }
  // candidates: every (lanes, rows) schedule on the row layout, then on the
  // split layout with 8 and 16 column slices (when the block has entries
  // far from its diagonal)
  struct Best {
    int layout = -1, window = 0, slices = 0, L = 0, R = 0, V = 1, NT = 1;
    float ms = 0;
  } best;
  // a derived layout (a second copy of the entries) is kept only when it
  // beats the row layout's best time by kDerivedGain
  float rows_ms = 0;
  std::vector<Best> rows_trials;  // every row-layout kernel's short trial
  auto derived_wins = [&](float t) { return t < best.ms && t < rows_ms * kDerivedGain; };
  auto sweep = [&](int layout, int window, int slices) -> int {
    for (int V : {1, 2, 4})
      for (int L = 1; L <= 64; L <<= 1)
        for (int R : {1, 2, 4})
          for (int NT : {1, 0}) {
            if (!csr_rows_valid(V, L, R, NT)) continue;
            m->lanes_per_row = L;
            m->rows_per_group = R;
            m->csr_vec = V;
            m->csr_nt = NT;
            TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, 1, &ms, stream));  // warm
            TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, iters, &ms, stream));
            const bool win = layout == TBSLA_HIP_CSR_ROWS ? (best.layout < 0 || ms < best.ms)
                                                          : derived_wins(ms);
            if (win) best = Best{layout, window, slices, L, R, V, NT, ms};
            if (layout == TBSLA_HIP_CSR_ROWS) rows_trials.push_back(Best{layout, 0, 0, L, R, V, NT, ms});
          }
    return TBSLA_HIP_SUCCESS;
  };
  split_release(m);
  TBSLA_TRY(sweep(TBSLA_HIP_CSR_ROWS, 0, 0));
  // the three fastest row kernels again with three times the launches (the
  // short trials differ by a few percent from run to run)
  {
    std::stable_sort(rows_trials.begin(), rows_trials.end(),
                     [](const Best &x, const Best &y) { return x.ms < y.ms; });
    if (rows_trials.size() > 3) rows_trials.resize(3);
    for (auto &t : rows_trials) {
      m->lanes_per_row = t.L;
      m->rows_per_group = t.R;
      m->csr_vec = t.V;
      m->csr_nt = t.NT;
      TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, 3 * iters, &ms, stream));
      t.ms = ms / 3.0f;
    }
    if (!rows_trials.empty()) {
      best = rows_trials[0];
      for (const auto &t : rows_trials)
        if (t.ms < best.ms) best = t;
    }
  }
  rows_ms = best.ms;
  constexpr int kWindow = 64;
  for (int slices : {8, 16}) {
    if (!m->has_arrays || m->ln_row == 0) break;
    TBSLA_TRY(split_build(m, kWindow, slices, s));
    long long remote = 0;
    TBSLA_TRY(tbsla_hip_get_csr_layout(m, nullptr, nullptr, nullptr, &remote));
    if (remote == 0) break;
    TBSLA_TRY(sweep(TBSLA_HIP_CSR_SPLIT, kWindow, slices));
  }
  // the diagonal layout: its own kernels, one timing per slice count (one
  // only for the LDS-slice form, which does not depend on it)
  for (int slices : {16, 8}) {
    if (!m->has_arrays || m->ln_row == 0 || m->ln_col == 0) break;
    TBSLA_TRY(split_build(m, kDiagWindow, slices, s, TBSLA_HIP_CSR_DIAG));
    TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, 1, &ms, stream));
    TBSLA_TRY(tbsla_hip_time_spmv(m, d_r, d_v, iters, &ms, stream));
    if (derived_wins(ms))
      best = Best{TBSLA_HIP_CSR_DIAG, kDiagWindow, slices, best.L, best.R, best.V, best.NT, ms};
    if (split_lds_block_rows(m) > 0) break;
  }
  split_release(m);
  if (best.layout == TBSLA_HIP_CSR_SPLIT || best.layout == TBSLA_HIP_CSR_DIAG)
    TBSLA_TRY(split_build(m, best.window, best.slices, s, best.layout));
  m->lanes_per_row = best.L;
  m->rows_per_group = best.R;
  m->csr_vec = best.V;
  m->csr_nt = best.NT;
  if (best_us) *best_us = best.ms * 1000.0f / iters;
  return TBSLA_HIP_SUCCESS;
}

}  // extern "C"
