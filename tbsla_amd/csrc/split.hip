// split.hip -- a derived execution layout for CSR SpMV on MI355X: the same
// entries, split into a near-diagonal part and a column-sliced remote part.
//
// Why: a row's entries near its diagonal gather x from lines that the next
// rows reuse, but entries far from it (e.g. cqmat's rewired columns,
// MatrixCSR.cpp:373-486 with q > 0) gather uniformly over all of x, which is
// several times one XCD's 4 MB L2; every such gather then pulls a whole line
// from the MALL or HBM for 8 useful bytes.  Here the remote entries are
// sorted by (column slice, row) and multiplied by a kernel whose workgroups
// for slice s all run on XCD s mod 8 (workgroups are dispatched to the eight
// XCDs round-robin), so each XCD keeps its slice of x in its own L2.  The
// products go to an array P in that order; the row kernel then adds, per row,
// its near-diagonal products (streamed CSR) and its remote products P[pos]
// (row-major list, contiguous per row block).  Every sum is taken in a fixed
// order, so results are deterministic; they differ from the row-major CSR
// sum only by rounding.
//
// Layout (device, owned by the handle):
//   lrp[ln_row+1], lci[nl], lv[nl]     near-diagonal entries |c - r| <= window
//   rptr[ln_row+1], rpos[nr]           remote entries per row -> index in P
//   sc[nr], sv[nr]                     remote entries in (slice, row) order
//   P[nr]                              their products, rewritten every SpMV
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "csr_rows.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

struct SplitCsr {
  int window = 0, slices = 0;
  long long nl = 0, nr = 0;
  int *lrp = nullptr, *lci = nullptr, *rptr = nullptr, *rpos = nullptr, *sc = nullptr;
  double *lv = nullptr, *sv = nullptr, *P = nullptr;
  std::vector<long long> soff;  // slice boundaries (host)
  ~SplitCsr() {
    hipDeviceSynchronize();
    hipFree(lrp);
    hipFree(lci);
    hipFree(lv);
    hipFree(rptr);
    hipFree(rpos);
    hipFree(sc);
    hipFree(sv);
    hipFree(P);
  }
};

namespace {

constexpr int kRemoteChunk = 2048;  // remote entries per workgroup

__device__ inline bool is_local(long long c, long long r, int window) {
  return c - r <= window && r - c <= window;
}

// Entries of row r: CSR rp[r]..rp[r+1], or, for ELL (ell_w > 0, column-major
// with n rows), the ell_w slots j * n + r (padding slots included: they hold
// (0, 0.0) and the reference multiplies them too, MatrixELL.cpp:181-205).
struct RowSrc {
  const int *rp;
  int ell_w;
  long long n;
  __device__ int count(long long r) const { return ell_w > 0 ? ell_w : rp[r + 1] - rp[r]; }
  __device__ long long at(long long r, int k) const {
    return ell_w > 0 ? (long long)k * n + r : (long long)rp[r] + k;
  }
};

__global__ __launch_bounds__(kBlock) void k_split_count(long long n, int f_col, int window,
                                                        RowSrc src, const int *__restrict__ ci,
                                                        int *__restrict__ lcnt,
                                                        int *__restrict__ rcnt) {
  for (long long r = (long long)blockIdx.x * kBlock + threadIdx.x; r < n;
       r += (long long)gridDim.x * kBlock) {
    int l = 0, q = 0;
    const int cnt = src.count(r);
    for (int k = 0; k < cnt; k++) {
      const long long c = (long long)max(ci[src.at(r, k)] - f_col, 0);
      if (is_local(c, r, window)) l++;
      else q++;
    }
    lcnt[r] = l;
    rcnt[r] = q;
  }
}

// local entries in row order; remote entries in row-major order with their
// (slice, row) sort key
__global__ __launch_bounds__(kBlock) void k_split_fill(
    long long n, long long n_col, int f_col, int window, int slices, int rb, RowSrc src,
    const int *__restrict__ ci, const double *__restrict__ v,
    const int *__restrict__ lrp, const int *__restrict__ rptr, int *__restrict__ lci,
    double *__restrict__ lv, int *__restrict__ rc, double *__restrict__ rv,
    unsigned long long *__restrict__ keys, int *__restrict__ payload,
    int *__restrict__ scount) {
  for (long long r = (long long)blockIdx.x * kBlock + threadIdx.x; r < n;
       r += (long long)gridDim.x * kBlock) {
    int pl = lrp[r], pr = rptr[r];
    const int cnt = src.count(r);
    for (int k = 0; k < cnt; k++) {
      const long long j = src.at(r, k);
      // ELL clamps a negative index to 0 (MatrixELL.cpp:198-199)
      const long long c = (long long)max(ci[j] - f_col, 0);
      if (is_local(c, r, window)) {
        lci[pl] = (int)c;
        lv[pl] = v[j];
        pl++;
      } else {
        const unsigned long long s = (unsigned long long)(c * slices / n_col);
        rc[pr] = (int)c;
        rv[pr] = v[j];
        keys[pr] = (s << rb) | (unsigned long long)r;
        payload[pr] = pr;
        atomicAdd(scount + s, 1);
        pr++;
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_split_place(long long nr,
                                                        const int *__restrict__ payload,
                                                        const int *__restrict__ rc,
                                                        const double *__restrict__ rv,
                                                        int *__restrict__ sc,
                                                        double *__restrict__ sv,
                                                        int *__restrict__ rpos) {
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < nr;
       e += (long long)gridDim.x * kBlock) {
    const int j = payload[e];
    sc[e] = rc[j];
    sv[e] = rv[j];
    rpos[j] = (int)e;
  }
}

// P[e] = sv[e] x[sc[e]] for the slices of one pass: workgroup b serves slice
// pass * 8 + b % 8 (so all of a slice's workgroups land on one XCD), chunk b / 8.
__global__ __launch_bounds__(kBlock) void k_remote_products(const long long *__restrict__ soff,
                                                            int pass, int slices,
                                                            const int *__restrict__ sc,
                                                            const double *__restrict__ sv,
                                                            const double *__restrict__ x,
                                                            double *__restrict__ P,
                                                            const int *__restrict__ done) {
  if (done && *done) return;
  const int s = pass * 8 + (int)(blockIdx.x % 8);
  if (s >= slices) return;
  const long long b = soff[s] + (long long)(blockIdx.x / 8) * kRemoteChunk;
  const long long e = min(soff[s + 1], b + kRemoteChunk);
  for (long long k = b + threadIdx.x; k < e; k += kBlock) P[k] = ldg_stream(sv + k) * x[ldg_stream(sc + k)];
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

}  // namespace

int split_params(const tbsla_hip_matrix_s *m, int *window, int *slices) {
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  if (!sp) return 0;
  *window = sp->window;
  *slices = sp->slices;
  return 1;
}

void split_release(tbsla_hip_matrix_s *m) {
  if (m->split) {
    delete static_cast<SplitCsr *>(m->split);
    m->split = nullptr;
  }
}

// device copy of the slice offsets lives in the handle's scratch head
int split_launch(tbsla_hip_matrix_s *m, double *y, const double *x, const int *done,
                 hipStream_t s) {
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  const long long *d_soff = reinterpret_cast<const long long *>(sp->sc == nullptr ? nullptr : sp->P + sp->nr);
  if (sp->nr > 0) {
    for (int pass = 0; pass * 8 < sp->slices; pass++) {
      long long chunks = 0;
      for (int k = pass * 8; k < std::min(sp->slices, pass * 8 + 8); k++)
        chunks = std::max(chunks, (sp->soff[k + 1] - sp->soff[k] + kRemoteChunk - 1) / kRemoteChunk);
      if (chunks == 0) continue;
      k_remote_products<<<(unsigned)(8 * chunks), kBlock, 0, s>>>(d_soff, pass, sp->slices, sp->sc,
                                                                  sp->sv, x, sp->P, done);
    }
  }
  // the near-diagonal entries with the block's CSR schedule, then each
  // row's remote products (lanes of the group take them in turn)
  CsrRowsArgs a{m->ln_row, sp->lrp, sp->lci, sp->lv, x, 0, y, done, sp->rptr, sp->rpos, sp->P};
  const int L = pick_csr_lanes(m);
  csr_rows_launch<true>(a, pick_csr_vec(m), L, pick_csr_rows(m, L), pick_csr_nt(m), s);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

int split_build(tbsla_hip_matrix_s *m, int window, int slices, hipStream_t s) {
  split_release(m);
  const long long n = m->ln_row, nc = std::max(m->ln_col, 1);
  auto *sp = new SplitCsr();
  sp->window = window;
  sp->slices = slices;
  m->split = sp;
  int *lcnt = nullptr, *rcnt = nullptr, *rc = nullptr, *pay = nullptr, *scount = nullptr;
  double *rv = nullptr;
  unsigned long long *keys = nullptr;
  void *ws = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(s);
    hipFree(lcnt);
    hipFree(rcnt);
    hipFree(rc);
    hipFree(pay);
    hipFree(scount);
    hipFree(rv);
    hipFree(keys);
    hipFree(ws);
  };
  auto failed = [&](int st) {
    cleanup();
    split_release(m);
    return st;
  };
  int st;
  if ((st = device_alloc((void **)&lcnt, (size_t)(n + 1) * 4)) ||
      (st = device_alloc((void **)&rcnt, (size_t)(n + 1) * 4)) ||
      (st = device_alloc((void **)&sp->lrp, (size_t)(n + 1) * 4)) ||
      (st = device_alloc((void **)&sp->rptr, (size_t)(n + 1) * 4)) ||
      (st = device_alloc((void **)&scount, (size_t)(slices + 1) * 4)) ||
      (st = device_alloc(&ws, scan_workspace_bytes(n + 1) + 64)))
    return failed(st);
  long long *d_total = reinterpret_cast<long long *>(static_cast<char *>(ws) + scan_workspace_bytes(n + 1));
  RowSrc src{m->rowptr, m->format == TBSLA_HIP_ELL ? m->max_col : 0, n};
  k_split_count<<<sgrid(n), kBlock, 0, s>>>(n, m->f_col, window, src, m->colidx, lcnt, rcnt);
  long long tot[2] = {0, 0};
  int *cnts[2] = {lcnt, rcnt}, *ptrs[2] = {sp->lrp, sp->rptr};
  for (int k = 0; k < 2; k++) {
    if ((st = scan<int>(cnts[k], ptrs[k], n, false, d_total, ws, s))) return failed(st);
    if (hipMemcpyAsync(&tot[k], d_total, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return failed(fail(TBSLA_HIP_ERR_HIP, "split layout: scan"));
    const int t = (int)tot[k];
    hipMemcpyAsync(ptrs[k] + n, &t, 4, hipMemcpyHostToDevice, s);
  }
  sp->nl = tot[0];
  sp->nr = tot[1];
  if ((st = device_alloc((void **)&sp->lci, (size_t)std::max(sp->nl, 1LL) * 4)) ||
      (st = device_alloc((void **)&sp->lv, (size_t)std::max(sp->nl, 1LL) * 8)) ||
      (st = device_alloc((void **)&sp->rpos, (size_t)std::max(sp->nr, 1LL) * 4)) ||
      (st = device_alloc((void **)&sp->sc, (size_t)std::max(sp->nr, 1LL) * 4)) ||
      (st = device_alloc((void **)&sp->sv, (size_t)std::max(sp->nr, 1LL) * 8)) ||
      // P, followed by the device copy of the slice offsets
      (st = device_alloc((void **)&sp->P, (size_t)std::max(sp->nr, 1LL) * 8 + (size_t)(slices + 1) * 8)) ||
      (st = device_alloc((void **)&rc, (size_t)std::max(sp->nr, 1LL) * 4)) ||
      (st = device_alloc((void **)&rv, (size_t)std::max(sp->nr, 1LL) * 8)) ||
      (st = device_alloc((void **)&keys, (size_t)std::max(sp->nr, 1LL) * 8)) ||
      (st = device_alloc((void **)&pay, (size_t)std::max(sp->nr, 1LL) * 4)))
    return failed(st);
  hipMemsetAsync(scount, 0, (size_t)(slices + 1) * 4, s);
  const int rb = bits_of(std::max(n, 1LL));
  k_split_fill<<<sgrid(n), kBlock, 0, s>>>(n, nc, m->f_col, window, slices, rb, src, m->colidx,
                                           m->values, sp->lrp, sp->rptr, sp->lci, sp->lv, rc, rv, keys,
                                           pay, scount);
  if (sp->nr > 1 && (st = radix_sort_pairs(keys, pay, sp->nr, rb + bits_of(slices), s))) return failed(st);
  if (sp->nr > 0) k_split_place<<<sgrid(sp->nr), kBlock, 0, s>>>(sp->nr, pay, rc, rv, sp->sc, sp->sv, sp->rpos);
  std::vector<int> hc(slices + 1, 0);
  if (hipMemcpyAsync(hc.data(), scount, (size_t)slices * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return failed(fail(TBSLA_HIP_ERR_HIP, "split layout: slice counts"));
  sp->soff.assign(slices + 1, 0);
  for (int k = 0; k < slices; k++) sp->soff[k + 1] = sp->soff[k] + hc[k];
  long long *d_soff = reinterpret_cast<long long *>(sp->P + sp->nr);
  if (hipMemcpyAsync(d_soff, sp->soff.data(), (size_t)(slices + 1) * 8, hipMemcpyHostToDevice, s) != hipSuccess)
    return failed(fail(TBSLA_HIP_ERR_HIP, "split layout: offsets"));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return failed(fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e)));
  cleanup();
  return TBSLA_HIP_SUCCESS;
}

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_set_csr_layout(tbsla_hip_matrix_t m, int layout, int window, int slices,
                                        void *stream) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null handle");
  if (m->format != TBSLA_HIP_CSR && m->format != TBSLA_HIP_ELL)
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "CSR and ELL blocks only");
  if (layout == TBSLA_HIP_CSR_ROWS) {
    split_release(m);
    return TBSLA_HIP_SUCCESS;
  }
  if (layout != TBSLA_HIP_CSR_SPLIT) return fail(TBSLA_HIP_ERR_INVALID, "unknown layout");
  if (window < 0 || slices < 1 || slices > 64 || slices % 8)
    return fail(TBSLA_HIP_ERR_INVALID, "window >= 0, slices a multiple of 8 up to 64");
  if (!m->has_arrays || m->ln_row == 0 || (m->format == TBSLA_HIP_ELL && m->max_col == 0))
    return fail(TBSLA_HIP_ERR_INVALID, "empty block");
  return split_build(m, window, slices, static_cast<hipStream_t>(stream));
}

extern "C" int tbsla_hip_get_csr_layout(tbsla_hip_matrix_t m, int *layout, int *window,
                                        int *slices, long long *remote_entries) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null handle");
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  if (layout) *layout = sp ? TBSLA_HIP_CSR_SPLIT : TBSLA_HIP_CSR_ROWS;
  if (window) *window = sp ? sp->window : 0;
  if (slices) *slices = sp ? sp->slices : 0;
  if (remote_entries) *remote_entries = sp ? sp->nr : 0;
  return TBSLA_HIP_SUCCESS;
}
