// spgemm.hip -- C = A B for two CSR blocks with short rows (the product
// matrix A G of the factored CG's product form).
//
// One thread per row of A: the products a_ic b_ck of the row (at most
// kMaxProducts; more sets the overflow flag and the caller uses another
// method) are gathered in generation order (A's entries, then B's row c),
// sorted by column with a stable insertion sort, and equal columns summed in
// that order with separate multiply and add -- so C is deterministic.  Two
// passes: counts, then (after the row-pointer scan) the entries.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

constexpr int kMaxProducts = 64;

template <bool FILL>
__global__ __launch_bounds__(kBlock) void k_product_rows(
    long long n, int a_fcol, const int *__restrict__ a_rp, const int *__restrict__ a_ci,
    const double *__restrict__ a_v, int b_fcol, const int *__restrict__ b_rp,
    const int *__restrict__ b_ci, const double *__restrict__ b_v, long long b_rows,
    int *__restrict__ cnt, const int *__restrict__ c_rp, int *__restrict__ c_ci,
    double *__restrict__ c_v, int *__restrict__ overflow) {
  TBSLA_NO_FMA
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    int col[kMaxProducts];
    double val[kMaxProducts];
    int m = 0;
    bool over = false;
    for (int ja = a_rp[i]; ja < a_rp[i + 1]; ja++) {
      const long long c = (long long)a_ci[ja] - a_fcol;
      if (c < 0 || c >= b_rows) continue;  // outside B's rows: no products
      const double av = a_v[ja];
      for (int jb = b_rp[c]; jb < b_rp[c + 1]; jb++) {
        if (m == kMaxProducts) {
          over = true;
          break;
        }
        col[m] = b_ci[jb] - b_fcol;
        val[m] = av * b_v[jb];
        m++;
      }
    }
    if (over) {
      *overflow = 1;
      if (!FILL) cnt[i] = 0;
      continue;
    }
    // stable insertion sort by column
    for (int p = 1; p < m; p++) {
      const int kc = col[p];
      const double kv = val[p];
      int q = p - 1;
      while (q >= 0 && col[q] > kc) {
        col[q + 1] = col[q];
        val[q + 1] = val[q];
        q--;
      }
      col[q + 1] = kc;
      val[q + 1] = kv;
    }
    if (!FILL) {
      int u = 0;
      for (int p = 0; p < m; p++) u += (p == 0 || col[p] != col[p - 1]);
      cnt[i] = u;
    } else {
      int o = c_rp[i] - 1;
      for (int p = 0; p < m; p++) {
        if (p == 0 || col[p] != col[p - 1]) {
          o++;
          c_ci[o] = col[p];
          c_v[o] = val[p];
        } else {
          c_v[o] = c_v[o] + val[p];
        }
      }
    }
  }
}

unsigned pgrid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 65536));
}

}  // namespace

int csr_product_arrays(const tbsla_hip_matrix_s *A, const tbsla_hip_matrix_s *B, int **rowptr,
                       int **colidx, double **values, long long *nnz, bool *fits,
                       hipStream_t s) {
  const long long n = A->ln_row;
  *rowptr = *colidx = nullptr;
  *values = nullptr;
  *nnz = 0;
  *fits = false;
  int *cnt = nullptr, *flag = nullptr, *rp = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(s);
    hipFree(cnt);
    hipFree(flag);
  };
  int st;
  if ((st = device_alloc((void **)&cnt, (size_t)(n + 1) * 4)) ||
      (st = device_alloc((void **)&flag, 4)) ||
      (st = device_alloc((void **)&rp, (size_t)(n + 1) * 4))) {
    cleanup();
    hipFree(rp);
    return st;
  }
  TBSLA_HIP_CHECK(hipMemsetAsync(flag, 0, 4, s));
  k_product_rows<false><<<pgrid(n), kBlock, 0, s>>>(
      n, A->f_col, A->rowptr, A->colidx, A->values, B->f_col, B->rowptr, B->colidx, B->values,
      B->ln_row, cnt, nullptr, nullptr, nullptr, flag);
  int over = 0;
  TBSLA_HIP_CHECK(hipMemcpyAsync(&over, flag, 4, hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  if (over) {
    cleanup();
    hipFree(rp);
    return TBSLA_HIP_SUCCESS;  // *fits = false: rows with more than kMaxProducts products
  }
  long long total = 0;
  st = counts_to_rowptr(cnt, rp, n, &total, s);
  int *ci = nullptr;
  double *v = nullptr;
  if (!st && ((st = device_alloc((void **)&ci, (size_t)std::max(total, 1LL) * 4)) ||
              (st = device_alloc((void **)&v, (size_t)std::max(total, 1LL) * 8)))) {
  }
  if (!st) {
    k_product_rows<true><<<pgrid(n), kBlock, 0, s>>>(
        n, A->f_col, A->rowptr, A->colidx, A->values, B->f_col, B->rowptr, B->colidx,
        B->values, B->ln_row, nullptr, rp, ci, v, flag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) st = fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  }
  cleanup();
  if (st) {
    hipFree(rp);
    hipFree(ci);
    hipFree(v);
    return st;
  }
  *rowptr = rp;
  *colidx = ci;
  *values = v;
  *nnz = total;
  *fits = true;
  return TBSLA_HIP_SUCCESS;
}

}  // namespace tbsla_hip
