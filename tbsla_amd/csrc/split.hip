// split.hip -- a derived execution layout for CSR SpMV on MI355X: the same
// entries, split into a near-diagonal part and a column-sliced remote part.
//
// Why: a row's entries near its diagonal gather x from lines that the next
// rows reuse, but entries far from it (e.g. cqmat's rewired columns,
// MatrixCSR.cpp:373-486 with q > 0) gather uniformly over all of x, which is
// several times one XCD's 4 MB L2; every such gather then pulls a whole
// 128-byte line from the MALL or HBM for 8 useful bytes.
//
// Two forms of the remote part:
//
// * LDS slices (the diagonal layout, whenever every block of rows fits):
//   the remote entries are grouped by column slice of kSliceCols columns
//   (128 KB of x, held in one workgroup's LDS) and, within a slice, by
//   block of H rows.  k_slice_products gives each slice a workgroup that
//   loads its piece of x into LDS once and multiplies the slice's entries
//   (2-byte column offsets and row-block indices), gathering x from LDS; it
//   writes the products
//   to P ordered by (row block, slice, row), so each row block's products
//   are one contiguous run.  k_dia_rows_remote gives each row block one
//   workgroup, which scatters the block's products into LDS in row-major
//   order (each carries its 2-byte position), sums each row's products in
//   its CSR order and adds the total to the row's diagonal sum.  No x
//   gather leaves LDS, and no row total goes through HBM.
// * L2 slices (the split layout, and the diagonal layout's fallback): the
//   remote entries are sorted by (column slice, row) and multiplied by a
//   kernel whose workgroups for slice s all run on XCD s mod 8 (workgroups
//   are dispatched to the eight XCDs round-robin), so each XCD keeps its
//   slice of x in its own L2.  The products go to an array P in that
//   order.  A combine kernel then gives each block of kCombineRows rows one
//   workgroup, which reads that block's contiguous run of P in every slice
//   (slice by slice, in order) and writes the row totals R[row]; the row
//   kernel streams the near-diagonal entries and stores y[row] = its sum +
//   R[row].
//
// Every sum is taken in a fixed order, so results are deterministic; they
// differ from the row-major CSR sum only by rounding.
//
// Layout (device, owned by the handle):
//   lrp[ln_row+1], lci[nl], lv[nl]     near-diagonal entries |c - r| <= window
//   sr[nr], sc[nr], sv[nr]             remote entries in (slice, row) order
//   boff[slices][nblk+1]               first entry of row block b in slice s
//   P[nr], R[ln_row]                   products and row totals, rewritten
//                                      every SpMV
//   doff[ndiag], dv[tiles]             diagonal layout: stored diagonals
//   c16[nr], q16[nr], sv[nr]           LDS slices: column offset in the slice,
//                                      row block and value, (slice, row) order
//   sst[S R + 1], pst[S R]             LDS slices: run (slice f, block q) at
//                                      sst[f R + q] (entries), pst[f R + q] (P)
//   d16[nr]                            LDS slices, P order: position of the
//                                      product in its block's row-major order
//   pb[R + 1], rs16[R (H + 1)]         LDS slices: first product of each row
//                                      block; row starts within the block
#include <algorithm>
#include <atomic>
#include <climits>
#include <vector>

#include "common.hpp"
#include "csr_rows.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

constexpr int kCombineRows = 1024;  // rows per combine workgroup (LDS: 8 KB)
constexpr int kMaxDiagWindow = 1 << 20;  // diagonal layout: offsets searched
constexpr int kSliceCols = 16384;  // LDS slices: columns per slice (128 KB of x)
constexpr int kBlockCap = 14336;   // LDS slices: products per row block (112 KB)
constexpr int kBigBlock = 1024;    // threads of the LDS-slice kernels

struct SplitCsr {
  int layout = 0, window = 0, slices = 0;
  unsigned long long build_id = 0;  // unique per build in the process
  long long nl = 0, nr = 0, nblk = 0;
  int ndiag = 0;  // diagonal layout: stored diagonals
  int *lrp = nullptr, *lci = nullptr, *sr = nullptr, *sc = nullptr, *boff = nullptr;
  int *doff = nullptr;  // their offsets c - r, ascending
  double *lv = nullptr, *sv = nullptr, *P = nullptr, *R = nullptr;
  double *dv = nullptr;  // dv[k n + r]: row r's entry on diagonal k, or 0
  std::vector<long long> soff;  // slice boundaries (host)
  // LDS slices (lds_h > 0): rows per block H, slices S, blocks R, workgroups
  // per slice G
  int lds_h = 0, lds_s = 0, lds_r = 0, lds_g = 0;
  unsigned short *c16 = nullptr, *q16 = nullptr, *d16 = nullptr, *rs16 = nullptr;
  int *sst = nullptr, *pst = nullptr, *pb = nullptr;
  ~SplitCsr() {
    (void)hipDeviceSynchronize();
    (void)hipFree(c16);
    (void)hipFree(q16);
    (void)hipFree(d16);
    (void)hipFree(rs16);
    (void)hipFree(sst);
    (void)hipFree(pst);
    (void)hipFree(pb);
    (void)hipFree(lrp);
    (void)hipFree(lci);
    (void)hipFree(lv);
    (void)hipFree(sr);
    (void)hipFree(sc);
    (void)hipFree(sv);
    (void)hipFree(boff);
    (void)hipFree(P);
    (void)hipFree(R);
    (void)hipFree(doff);
    (void)hipFree(dv);
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

// The rows a derived layout is built over and launched on: a CSR or ELL
// block's own rows, or a row-sorted COO / SCOO block's triplet rows through
// its derived row pointer (rows coo_row0 ..; COO indexes x by the global
// column, SCOO by the column less f_col, as their SpMVs do).  Diagonals are
// offsets c - (r + r0) of the x index c from the layout row r shifted by r0,
// the row of x that layout row 0 lines up with: 0 for CSR and ELL, the
// global row coo_row0 for COO, coo_row0 - f_row for SCOO, so that a grid
// block's diagonals are the whole matrix's.
struct LayoutRows {
  long long n;   // rows
  long long nc;  // entries of x
  int f_col;     // subtracted from a column index to index x
  const int *rp;
  int ell_w;
  long long r0;
};
LayoutRows layout_rows(const tbsla_hip_matrix_s *m) {
  if (m->format == TBSLA_HIP_COO || m->format == TBSLA_HIP_SCOO) {
    const bool coo = m->format == TBSLA_HIP_COO;
    return {m->coo_nrows, coo ? m->n_col : m->ln_col, coo ? 0 : m->f_col, m->coo_rowptr, 0,
            coo ? m->coo_row0 : (long long)m->coo_row0 - m->f_row};
  }
  return {m->ln_row, m->ln_col, m->f_col, m->rowptr, m->format == TBSLA_HIP_ELL ? m->max_col : 0, 0};
}

// Where an entry of (shifted) row r goes.  Split layout (dmap null): near-diagonal
// (returns 0) or remote (-1).  Diagonal layout: the slot of diagonal k
// (returns k + 1) when c - r is a stored diagonal, the value is not zero and
// the row's slot on it is still free (`used`), else remote; so a zero slot
// means "no entry" and the DIA kernel skips it.
__device__ inline int place_of(long long c, long long r, double v, int window,
                               const int *__restrict__ dmap, unsigned long long &used) {
  if (!dmap) return is_local(c, r, window) ? 0 : -1;
  const long long d = c - r;
  if (d < -window || d > window || !(v != 0.0)) return -1;
  const int k = dmap[d + window];
  if (k < 0 || ((used >> k) & 1ULL)) return -1;
  used |= 1ULL << k;
  return k + 1;
}

__global__ __launch_bounds__(kBlock) void k_split_count(long long n, int f_col, long long r0,
                                                        int window, RowSrc src, const int *__restrict__ ci,
                                                        const double *__restrict__ v,
                                                        const int *__restrict__ dmap,
                                                        int *__restrict__ lcnt,
                                                        int *__restrict__ rcnt) {
  for (long long r = (long long)blockIdx.x * kBlock + threadIdx.x; r < n;
       r += (long long)gridDim.x * kBlock) {
    int l = 0, q = 0;
    unsigned long long used = 0;
    const int cnt = src.count(r);
    for (int k = 0; k < cnt; k++) {
      const long long j = src.at(r, k);
      const long long c = (long long)max(ci[j] - f_col, 0);
      if (place_of(c, r + r0, v[j], window, dmap, used) >= 0) l++;
      else q++;
    }
    lcnt[r] = l;
    rcnt[r] = q;
  }
}

// entries per diagonal offset c - (r + r0) in [-window, window] (nonzero values)
// over the sampled rows r = 0, step, 2 step, ...; offsets within
// kLdsHalf of the diagonal are counted in LDS and added once per workgroup
constexpr int kLdsHalf = 4095;
__global__ __launch_bounds__(kBlock) void k_diag_hist(long long n, long long step, int f_col,
                                                      long long r0, int window, RowSrc src,
                                                      const int *__restrict__ ci,
                                                      const double *__restrict__ v,
                                                      int *__restrict__ hist) {
  __shared__ int lh[2 * kLdsHalf + 1];
  for (int i = threadIdx.x; i < 2 * kLdsHalf + 1; i += kBlock) lh[i] = 0;
  __syncthreads();
  const long long ns = (n + step - 1) / step;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < ns;
       i += (long long)gridDim.x * kBlock) {
    const long long r = i * step;
    const int cnt = src.count(r);
    for (int k = 0; k < cnt; k++) {
      const long long j = src.at(r, k);
      const long long d = (long long)max(ci[j] - f_col, 0) - (r + r0);
      if (d < -window || d > window || !(v[j] != 0.0)) continue;
      if (d >= -kLdsHalf && d <= kLdsHalf) atomicAdd(lh + (d + kLdsHalf), 1);
      else atomicAdd(hist + (d + window), 1);
    }
  }
  __syncthreads();
  const int lo = max(-window, -kLdsHalf), hi = min(window, kLdsHalf);
  for (int d = lo + (int)threadIdx.x; d <= hi; d += kBlock)
    if (lh[d + kLdsHalf]) atomicAdd(hist + (d + window), lh[d + kLdsHalf]);
}

// y[r] = sum over the stored diagonals k (ascending offset) of
// dv[k n + r] x[r + doff[k]] (zero slots skipped) + R[r] (when add is not
// null).  A wavefront takes two tiles of kWave rows, each lane two adjacent
// rows of one of them, so every load of a diagonal's values is 16 bytes per
// lane (one 512-byte piece per half-wavefront).  The diagonals go in chunks
// of eight whose loads are issued together: a last chunk's missing
// diagonals reread the last one and add nothing.  x is read at the clamped
// column for every slot (in bounds), and a zero slot's product is not
// added, so an Inf or NaN in x reaches only the rows that hold an entry on
// its column.
__global__ __launch_bounds__(kBlock) void k_dia_rows(long long n, long long ncol, int ndiag,
                                                     const int *__restrict__ doff,
                                                     const double *__restrict__ dv,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ add,
                                                     double *__restrict__ y,
                                                     const int *__restrict__ done) {
  if (done && *done) return;
  constexpr int U = 8;
  // lane k holds diagonal k's offset (ndiag <= kWave): read by readlane,
  // so no scalar load waits inside the loop
  const int lane = (int)threadIdx.x & (kWave - 1);
  const int od = lane < ndiag ? doff[lane] : 0;
  const long long last = ncol - 1;
  const long long ntiles = (n + kWave - 1) / kWave;
  const long long npairs = (ntiles + 1) / 2;
  const long long wstep = (long long)gridDim.x * (kBlock / kWave);
  for (long long pr = (long long)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; pr < npairs;
       pr += wstep) {
    const long long t = 2 * pr + lane / (kWave / 2);  // this lane's tile
    const bool tile_ok = t < ntiles;
    const long long r = t * kWave + 2 * (lane % (kWave / 2));  // rows r, r + 1
    // rows r, r + 1 on diagonal k: 16 bytes at tile + k kWave (a missing
    // second tile reads tile 0 and stores nothing)
    const double *tile = dv + (tile_ok ? t : 0) * kWave * (long long)ndiag + 2 * (lane % (kWave / 2));
    typedef double v2 __attribute__((ext_vector_type(2)));
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < ndiag; k += U) {
      v2 v[U];
      double x0[U], x1[U];
#pragma unroll
      for (int i = 0; i < U; i++) {
        const int kk = min(k + i, ndiag - 1);
        const int o = __builtin_amdgcn_readlane(od, kk);
        v[i] = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(tile + kk * kWave));
        x0[i] = x[min(max(r + o, 0LL), last)];
        x1[i] = x[min(max(r + 1 + o, 0LL), last)];
      }
#pragma unroll
      for (int i = 0; i < U; i++) {
        const bool on = k + i < ndiag;
        s0 = (on && v[i].x != 0.0) ? fma(v[i].x, x0[i], s0) : s0;
        s1 = (on && v[i].y != 0.0) ? fma(v[i].y, x1[i], s1) : s1;
      }
    }
    if (tile_ok && r < n) y[r] = add ? s0 + add[r] : s0;
    if (tile_ok && r + 1 < n) y[r + 1] = add ? s1 + add[r + 1] : s1;
  }
}

// local entries in row order; remote entries in row-major order with their
// (slice, row) sort key: slice c / slice_cols when slice_cols > 0 (LDS
// slices), else c slices / n_col
__global__ __launch_bounds__(kBlock) void k_split_fill(
    long long n, long long n_col, int f_col, long long r0, int window, int slices, int slice_cols,
    int rb, RowSrc src,
    const int *__restrict__ ci, const double *__restrict__ v, const int *__restrict__ dmap,
    int ndiag, const int *__restrict__ lrp, const int *__restrict__ rptr, int *__restrict__ lci,
    double *__restrict__ lv, double *__restrict__ dv, int *__restrict__ rc,
    double *__restrict__ rv, unsigned long long *__restrict__ keys, int *__restrict__ payload,
    int *__restrict__ scount) {
  for (long long r = (long long)blockIdx.x * kBlock + threadIdx.x; r < n;
       r += (long long)gridDim.x * kBlock) {
    int pl = dmap ? 0 : lrp[r], pr = rptr[r];
    unsigned long long used = 0;
    const int cnt = src.count(r);
    for (int k = 0; k < cnt; k++) {
      const long long j = src.at(r, k);
      // ELL clamps a negative index to 0 (MatrixELL.cpp:198-199)
      const long long c = (long long)max(ci[j] - f_col, 0);
      const double val = v[j];
      const int at = place_of(c, r + r0, val, window, dmap, used);
      if (at > 0) {
        dv[dia_at(r, at - 1, ndiag)] = val;
      } else if (at == 0) {
        lci[pl] = (int)c;
        lv[pl] = val;
        pl++;
      } else {
        const unsigned long long s =
            (unsigned long long)(slice_cols > 0 ? c / slice_cols : c * slices / n_col);
        rc[pr] = (int)c;
        rv[pr] = val;
        keys[pr] = (s << rb) | (unsigned long long)r;
        payload[pr] = pr;
        atomicAdd(scount + s, 1);
        pr++;
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_split_place(long long nr,
                                                        const unsigned long long *__restrict__ keys,
                                                        int rb, const int *__restrict__ payload,
                                                        const int *__restrict__ rc,
                                                        const double *__restrict__ rv,
                                                        int *__restrict__ sr,
                                                        int *__restrict__ sc,
                                                        double *__restrict__ sv) {
  const unsigned long long mask = (1ULL << rb) - 1;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < nr;
       e += (long long)gridDim.x * kBlock) {
    const int j = payload[e];
    sr[e] = (int)(keys[e] & mask);
    sc[e] = rc[j];
    sv[e] = rv[j];
  }
}

// boff[s (nblk + 1) + b] = the first entry of slice s whose row is at least
// b kCombineRows (rows ascend within a slice); b = nblk gives the slice end
__global__ __launch_bounds__(kBlock) void k_block_offsets(const long long *__restrict__ soff,
                                                          int slices, long long nblk,
                                                          const int *__restrict__ sr,
                                                          int *__restrict__ boff) {
  const long long total = (long long)slices * (nblk + 1);
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (long long)gridDim.x * kBlock) {
    const int s = (int)(i / (nblk + 1));
    const long long row = (i % (nblk + 1)) * kCombineRows;
    long long lo = soff[s], hi = soff[s + 1];
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if (sr[mid] < row) lo = mid + 1;
      else hi = mid;
    }
    boff[i] = (int)lo;
  }
}

// R[row] for the rows of block b: per slice in order, each run of one row's
// products (contiguous, rows ascend) is summed by the thread at its first
// entry; runs of one slice have distinct rows, so their LDS adds never meet.
__global__ __launch_bounds__(kBlock) void k_remote_combine(long long n, int slices, long long nblk,
                                                           const int *__restrict__ boff,
                                                           const int *__restrict__ sr,
                                                           const double *__restrict__ P,
                                                           double *__restrict__ R,
                                                           const int *__restrict__ done) {
  __shared__ double acc[kCombineRows];
  if (done && *done) return;
  const long long b = blockIdx.x;
  const long long r0 = b * kCombineRows;
  for (int i = threadIdx.x; i < kCombineRows; i += kBlock) acc[i] = 0.0;
  __syncthreads();
  for (int s = 0; s < slices; s++) {
    const int e0 = boff[(long long)s * (nblk + 1) + b], e1 = boff[(long long)s * (nblk + 1) + b + 1];
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlock) {
      const int row = sr[e];
      if (e == e0 || sr[e - 1] != row) {
        double t = P[e];
        for (int k = e + 1; k < e1 && sr[k] == row; k++) t += P[k];
        acc[row - r0] += t;
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < kCombineRows && r0 + i < n; i += kBlock) R[r0 + i] = acc[i];
}

// P[e] = sv[e] x[sc[e]] for the slices of one pass: workgroup b serves slice
// pass * 8 + b % 8 (so all of a slice's workgroups land on one XCD), chunk
// b / 8.  Each thread issues its kRemoteChunk / kBlock loads before the
// gathers that depend on them.
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
  constexpr int U = kRemoteChunk / kBlock;
  int c[U];
  double v[U];
#pragma unroll
  for (int i = 0; i < U; i++) {
    const long long k = b + threadIdx.x + (long long)i * kBlock;
    c[i] = k < e ? ldg_stream(sc + k) : 0;
    v[i] = k < e ? ldg_stream(sv + k) : 0.0;
  }
#pragma unroll
  for (int i = 0; i < U; i++) {
    const long long k = b + threadIdx.x + (long long)i * kBlock;
    if (k < e) P[k] = v[i] * x[c[i]];
  }
}

// ---- LDS slices ------------------------------------------------------------

// out[i] = rp[min(i h, n)] for i in [0, nb]: the first remote entry of each
// block of h rows (rp: row-major remote offsets, n + 1 values)
__global__ __launch_bounds__(kBlock) void k_block_starts(long long n, int h, long long nb,
                                                         const int *__restrict__ rp,
                                                         int *__restrict__ out) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i <= nb;
       i += (long long)gridDim.x * kBlock)
    out[i] = rp[min(i * h, n)];
}

// (slice, row) order: the column's offset in its slice, the row's block of
// h rows and the value
__global__ __launch_bounds__(kBlock) void k_slice_place(long long nr,
                                                        const unsigned long long *__restrict__ keys,
                                                        int rb, int h, const int *__restrict__ payload,
                                                        const int *__restrict__ rc,
                                                        const double *__restrict__ rv,
                                                        unsigned short *__restrict__ c16,
                                                        unsigned short *__restrict__ q16,
                                                        double *__restrict__ sv) {
  const unsigned long long mask = (1ULL << rb) - 1;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < nr;
       e += (long long)gridDim.x * kBlock) {
    const int j = payload[e];
    const unsigned long long k = keys[e];
    const long long f = (long long)(k >> rb);
    c16[e] = (unsigned short)(rc[j] - f * kSliceCols);
    q16[e] = (unsigned short)((long long)(k & mask) / h);
    sv[e] = rv[j];
  }
}

// sst[f R + q] = the first entry (in (slice, row) order) of slice f whose row
// is at least q h; sst[S R] = nr
__global__ __launch_bounds__(kBlock) void k_slice_starts(long long nr, int rb, int S, int R, int h,
                                                         const unsigned long long *__restrict__ keys,
                                                         int *__restrict__ sst) {
  const long long total = (long long)S * R;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i <= total;
       i += (long long)gridDim.x * kBlock) {
    if (i == total) {
      sst[i] = (int)nr;
      continue;
    }
    const unsigned long long want =
        ((unsigned long long)(i / R) << rb) | (unsigned long long)((i % R) * (long long)h);
    long long lo = 0, hi = nr;
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if (keys[mid] < want) lo = mid + 1;
      else hi = mid;
    }
    sst[i] = (int)lo;
  }
}

// pst[f R + q]: where run (slice f, block q) starts in P, whose order is
// (block, slice, row): block q's runs follow each other in slice order from
// pb[q]
__global__ __launch_bounds__(kBlock) void k_slice_pstarts(int S, int R, const int *__restrict__ sst,
                                                          const int *__restrict__ pb,
                                                          int *__restrict__ pst) {
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < R; q += gridDim.x * kBlock) {
    int acc = pb[q];
    for (int f = 0; f < S; f++) {
      const long long run = (long long)f * R + q;
      pst[run] = acc;
      acc += sst[run + 1] - sst[run];
    }
  }
}

// d16, in P order: the product's position in its block's row-major order
// (payload = the entry's row-major remote index)
__global__ __launch_bounds__(kBlock) void k_slice_dest(long long nr, int rb, int R, int h,
                                                       const unsigned long long *__restrict__ keys,
                                                       const int *__restrict__ payload,
                                                       const int *__restrict__ sst,
                                                       const int *__restrict__ pst,
                                                       const int *__restrict__ pb,
                                                       unsigned short *__restrict__ d16) {
  const unsigned long long mask = (1ULL << rb) - 1;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < nr;
       e += (long long)gridDim.x * kBlock) {
    const unsigned long long k = keys[e];
    const long long f = (long long)(k >> rb);
    const int q = (int)((long long)(k & mask) / h);
    const long long run = f * R + q;
    const long long p = pst[run] + (e - sst[run]);
    d16[p] = (unsigned short)(payload[e] - pb[q]);
  }
}

// rs16[q (h + 1) + i] = the first product of row q h + i within block q
// (rows past n: the block's end)
__global__ __launch_bounds__(kBlock) void k_row_starts16(long long n, int h, int R,
                                                         const int *__restrict__ rp,
                                                         const int *__restrict__ pb,
                                                         unsigned short *__restrict__ rs16) {
  const long long total = (long long)R * (h + 1);
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (long long)gridDim.x * kBlock) {
    const long long q = i / (h + 1), j = i % (h + 1);
    rs16[i] = (unsigned short)(rp[min(q * h + j, n)] - pb[q]);
  }
}

// P[pst[run] + j] = sv[sst[run] + j] x[f kSliceCols + c16[sst[run] + j]] for
// the runs (f, q) of slice f = blockIdx / G over the blocks q of group
// blockIdx % G (at most kRunsCap).  Those runs are one contiguous range of
// entries, which the workgroup streams with every load independent of the
// run tables (kEntriesInFlight entries a lane, a wavefront's loads
// contiguous); each entry carries its block q (q16), whose P offset the
// workgroup holds in LDS with the slice's piece of x.
constexpr int kRunsCap = 2048;
constexpr int kEntriesInFlight = 8;
__global__ __launch_bounds__(kBigBlock) void k_slice_products(
    long long ncol, int S, int R, int G, const int *__restrict__ sst, const int *__restrict__ pst,
    const unsigned short *__restrict__ c16, const unsigned short *__restrict__ q16,
    const double *__restrict__ sv, const double *__restrict__ x, double *__restrict__ P,
    const int *__restrict__ done) {
  __shared__ double xs[kSliceCols];
  __shared__ int dl[kRunsCap];  // P position - entry position, per run
  if (done && *done) return;
  const int f = (int)(blockIdx.x / G), g = (int)(blockIdx.x % G);
  if (f >= S) return;
  const int q0 = (int)((long long)R * g / G), q1 = (int)((long long)R * (g + 1) / G);
  const int nq = q1 - q0;  // <= kRunsCap (G is chosen so)
  const long long run0 = (long long)f * R + q0;
  const long long c0 = (long long)f * kSliceCols;
  constexpr int UX = kSliceCols / kBigBlock;
  double t[UX];
#pragma unroll
  for (int i = 0; i < UX; i++) {
    const long long c = c0 + threadIdx.x + (long long)i * kBigBlock;
    t[i] = c < ncol ? x[c] : 0.0;
  }
  for (int i = (int)threadIdx.x; i < nq; i += kBigBlock) dl[i] = pst[run0 + i] - sst[run0 + i];
#pragma unroll
  for (int i = 0; i < UX; i++) xs[threadIdx.x + i * kBigBlock] = t[i];
  const int e0 = sst[run0], e1 = sst[run0 + nq];
  __syncthreads();
  constexpr int K = kEntriesInFlight;
  for (int base = e0 + (int)threadIdx.x; base < e1; base += kBigBlock * K) {
    int c[K], q[K];
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int e = base + k * kBigBlock;
      const bool in = e < e1;
      c[k] = in ? (int)c16[e] : 0;
      q[k] = in ? (int)q16[e] - q0 : 0;
      v[k] = in ? sv[e] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int e = base + k * kBigBlock;
      if (e < e1) P[e + dl[q[k]]] = v[k] * xs[c[k]];
    }
  }
}

// PageRank's update of one row (Matrix.cpp:123-141, as k_pr_update), each
// operation rounded on its own
__device__ __forceinline__ double pr_value(const PrFuse &pf, double y, double S, double t_sum) {
  TBSLA_NO_FMA
  double v = pf.beta * y + t_sum;
  return v / S;
}

// y[r] = the diagonal sum of row r (as k_dia_rows) + the sum of its remote
// products, for the h rows of block blockIdx.  The block's products
// P[pb[q] .. pb[q + 1]) are first scattered into LDS in row-major order
// (d16); each row then sums its own run in CSR order.  Sixteen wavefronts
// take 2048 rows per pass, two tiles of kWave rows each, a lane two
// adjacent rows.  With PR, PageRank's update of the row replaces the store
// of y: b[r] = (beta y[r] + S (1 - beta) / n_col) / S with b_t = x and S its
// sum, and the workgroup's partials of |b - b_t| and of sum(b) go to
// p_err / p_sum[blockIdx.x] (lanes' partials summed by the wavefront's fixed
// tree, then the wavefronts in order).
template <bool PR>
__global__ __launch_bounds__(kBigBlock) void k_dia_rows_remote(
    long long n, long long ncol, int ndiag, int h, const int *__restrict__ doff,
    const double *__restrict__ dv, const double *__restrict__ x, const int *__restrict__ pb,
    const unsigned short *__restrict__ d16, const unsigned short *__restrict__ rs16,
    const double *__restrict__ P, double *__restrict__ y, const int *__restrict__ done,
    PrFuse pf) {
  __shared__ double buf[kBlockCap];
  if (done && *done) return;
  double S = 0.0, t_sum = 0.0, pe = 0.0, ps = 0.0;
  if constexpr (PR) {
    TBSLA_NO_FMA
    S = *pf.S;
    t_sum = S * ((1 - pf.beta) / pf.n_col);
  }
  const long long q = blockIdx.x;
  const long long r0 = q * h;
  {
    const int p0 = pb[q], cnt = pb[q + 1] - p0;
    constexpr int K = (kBlockCap + kBigBlock - 1) / kBigBlock;
    double pv[K];
    int dd[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int i = (int)threadIdx.x + k * kBigBlock;
      const bool ok = i < cnt;
      pv[k] = ok ? P[p0 + i] : 0.0;
      dd[k] = ok ? (int)d16[p0 + i] : 0;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
      if ((int)threadIdx.x + k * kBigBlock < cnt) buf[dd[k]] = pv[k];
  }
  __syncthreads();
  constexpr int U = 8;
  const int lane = (int)threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
  const int od = lane < ndiag ? doff[lane] : 0;
  const long long last = ncol - 1;
  const long long ntiles = (n + kWave - 1) / kWave;
  const unsigned short *rs = rs16 + q * (h + 1);
  typedef double v2 __attribute__((ext_vector_type(2)));
  for (int pass = 0; pass * 2 * kBigBlock < h; pass++) {
    // local row (even) of this lane's pair
    const int lr = pass * 2 * kBigBlock + wave * 2 * kWave + (lane / (kWave / 2)) * kWave +
                   2 * (lane % (kWave / 2));
    const long long r = r0 + lr;
    const long long t = r / kWave;
    const bool tile_ok = lr < h && t < ntiles;
    const double *tile = dv + (tile_ok ? t : 0) * kWave * (long long)ndiag + 2 * (lane % (kWave / 2));
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < ndiag; k += U) {
      v2 v[U];
      double x0[U], x1[U];
#pragma unroll
      for (int i = 0; i < U; i++) {
        const int kk = min(k + i, ndiag - 1);
        const int o = __builtin_amdgcn_readlane(od, kk);
        v[i] = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(tile + kk * kWave));
        x0[i] = x[min(max(r + o, 0LL), last)];
        x1[i] = x[min(max(r + 1 + o, 0LL), last)];
      }
#pragma unroll
      for (int i = 0; i < U; i++) {
        const bool on = k + i < ndiag;
        s0 = (on && v[i].x != 0.0) ? fma(v[i].x, x0[i], s0) : s0;
        s1 = (on && v[i].y != 0.0) ? fma(v[i].y, x1[i], s1) : s1;
      }
    }
    if (tile_ok) {
      const int a = rs[lr], b = rs[lr + 1], c = rs[lr + 2];
      double t0 = 0.0, t1 = 0.0;
      for (int i = a; i < b; i++) t0 += buf[i];
      for (int i = b; i < c; i++) t1 += buf[i];
      if constexpr (PR) {
        if (r < n) {
          const double v = pr_value(pf, s0 + t0, S, t_sum);
          pf.b[r] = v;
          pe += fabs(v - x[r]);
          ps += v;
        }
        if (r + 1 < n) {
          const double v = pr_value(pf, s1 + t1, S, t_sum);
          pf.b[r + 1] = v;
          pe += fabs(v - x[r + 1]);
          ps += v;
        }
      } else {
        if (r < n) y[r] = s0 + t0;
        if (r + 1 < n) y[r + 1] = s1 + t1;
      }
    }
  }
  if constexpr (PR) {
    pe = group_sum<kWave>(pe);
    ps = group_sum<kWave>(ps);
    __syncthreads();  // every row is done with buf
    if (lane == 0) {
      buf[wave] = pe;
      buf[kBigBlock / kWave + wave] = ps;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double te = 0.0, ts = 0.0;
      for (int w = 0; w < kBigBlock / kWave; w++) {
        te += buf[w];
        ts += buf[kBigBlock / kWave + w];
      }
      pf.p_err[blockIdx.x] = te;
      pf.p_sum[blockIdx.x] = ts;
    }
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

}  // namespace

int split_pr_blocks(const tbsla_hip_matrix_s *m) {
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  const bool own_rows = m->format == TBSLA_HIP_CSR || m->format == TBSLA_HIP_ELL;
  return sp && own_rows && sp->lds_h > 0 && m->ln_row > 0 ? sp->lds_r : 0;
}

int split_launch_pr(tbsla_hip_matrix_s *m, const double *bt, const PrFuse &pf, const int *done,
                    hipStream_t s) {
  if (split_pr_blocks(m) == 0) return fail(TBSLA_HIP_ERR_INVALID, "no LDS-slice diagonal layout");
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  const LayoutRows lr = layout_rows(m);
  const long long n = lr.n, ncol = std::max(lr.nc, 1LL);
  k_slice_products<<<(unsigned)(sp->lds_s * sp->lds_g), kBigBlock, 0, s>>>(
      lr.nc, sp->lds_s, sp->lds_r, sp->lds_g, sp->sst, sp->pst, sp->c16, sp->q16, sp->sv, bt,
      sp->P, done);
  k_dia_rows_remote<true><<<(unsigned)sp->lds_r, kBigBlock, 0, s>>>(
      n, ncol, sp->ndiag, sp->lds_h, sp->doff, sp->dv, bt, sp->pb, sp->d16, sp->rs16, sp->P,
      nullptr, done, pf);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

int split_params(const tbsla_hip_matrix_s *m, int *window, int *slices, int *layout) {
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  if (!sp) return 0;
  *window = sp->window;
  *slices = sp->slices;
  if (layout) *layout = sp->layout;
  return 1;
}

int split_lds_block_rows(const tbsla_hip_matrix_s *m) {
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  return sp ? sp->lds_h : 0;
}

int split_diag_only(const tbsla_hip_matrix_s *m, const int **doff, const double **dv, int *ndiag,
                    long long *dv_len, unsigned long long *build_id) {
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  if (!sp || sp->layout != TBSLA_HIP_CSR_DIAG || sp->nr != 0 || sp->ndiag == 0) return 0;
  *build_id = sp->build_id;
  *doff = sp->doff;
  *dv = sp->dv;
  *ndiag = sp->ndiag;
  *dv_len = (long long)sp->ndiag * ((layout_rows(m).n + kWave - 1) / kWave) * kWave;
  return 1;
}

void split_release(tbsla_hip_matrix_s *m) {
  if (m->split) {
    delete static_cast<SplitCsr *>(m->split);
    m->split = nullptr;
  }
}

// the device copy of the slice offsets follows P
int split_launch(tbsla_hip_matrix_s *m, double *y, const double *x, const int *done,
                 hipStream_t s) {
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  const LayoutRows lr = layout_rows(m);
  const long long n = lr.n, ncol = std::max(lr.nc, 1LL);
  if (sp->lds_h > 0) {
    k_slice_products<<<(unsigned)(sp->lds_s * sp->lds_g), kBigBlock, 0, s>>>(
        lr.nc, sp->lds_s, sp->lds_r, sp->lds_g, sp->sst, sp->pst, sp->c16, sp->q16, sp->sv, x,
        sp->P, done);
    k_dia_rows_remote<false><<<(unsigned)sp->lds_r, kBigBlock, 0, s>>>(
        n, ncol, sp->ndiag, sp->lds_h, sp->doff, sp->dv, x, sp->pb,
        sp->d16, sp->rs16, sp->P, y, done, PrFuse{});
    TBSLA_HIP_LAUNCH_CHECK();
    return TBSLA_HIP_SUCCESS;
  }
  const long long *d_soff = reinterpret_cast<const long long *>(sp->P + std::max(sp->nr, 1LL));
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
  if (sp->layout == TBSLA_HIP_CSR_DIAG) {
    // no remote entries: nothing to combine
    if (sp->nr > 0)
      k_remote_combine<<<(unsigned)sp->nblk, kBlock, 0, s>>>(n, sp->slices, sp->nblk,
                                                             sp->boff, sp->sr, sp->P, sp->R, done);
    k_dia_rows<<<sgrid(n / 2), kBlock, 0, s>>>(n, ncol, sp->ndiag,
                                                   sp->doff, sp->dv, x,
                                                   sp->nr > 0 ? sp->R : nullptr, y, done);
    TBSLA_HIP_LAUNCH_CHECK();
    return TBSLA_HIP_SUCCESS;
  }
  k_remote_combine<<<(unsigned)sp->nblk, kBlock, 0, s>>>(n, sp->slices, sp->nblk, sp->boff,
                                                         sp->sr, sp->P, sp->R, done);
  // the near-diagonal entries with the block's CSR schedule, plus R[row]
  CsrRowsArgs a{n, sp->lrp, sp->lci, sp->lv, x, 0, y, done, sp->R};
  const int L = pick_csr_lanes(m);
  csr_rows_launch<true>(a, pick_csr_vec(m), L, pick_csr_rows(m, L), pick_csr_nt(m), s);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

// Rows per block for the LDS-slice form: the largest power of two from 4096
// down to 64 whose every block of rows holds at most kBlockCap remote
// entries, or 0 (rp: row-major remote offsets, n + 1 values, on the device)
static int pick_block_rows(const int *rp, long long n, hipStream_t s, int *h_out) {
  *h_out = 0;
  const long long nb = (n + kWave - 1) / kWave;
  int *d = nullptr;
  int st;
  if ((st = device_alloc((void **)&d, (size_t)(nb + 1) * 4))) return st;
  std::vector<int> hb(nb + 1);
  k_block_starts<<<sgrid(nb + 1), kBlock, 0, s>>>(n, kWave, nb, rp, d);
  const bool ok = hipGetLastError() == hipSuccess &&
                  hipMemcpyAsync(hb.data(), d, (size_t)(nb + 1) * 4, hipMemcpyDeviceToHost, s) ==
                      hipSuccess &&
                  hipStreamSynchronize(s) == hipSuccess;
  (void)hipFree(d);
  if (!ok) return fail(TBSLA_HIP_ERR_HIP, "diagonal layout: block sizes");
  for (int h = 4096; h >= kWave && *h_out == 0; h /= 2) {
    const long long step = h / kWave;
    bool fits = true;
    for (long long b = 0; b * step < nb && fits; b++)
      fits = hb[std::min((b + 1) * step, nb)] - hb[b * step] <= kBlockCap;
    if (fits) *h_out = h;
  }
  return TBSLA_HIP_SUCCESS;
}

int split_build(tbsla_hip_matrix_s *m, int window, int slices, hipStream_t s, int layout) {
  split_release(m);
  const LayoutRows lr = layout_rows(m);
  const long long n = lr.n, nc = std::max(lr.nc, 1LL);
  const int f_col = lr.f_col;
  const bool diag = layout == TBSLA_HIP_CSR_DIAG;
  if (diag && (lr.r0 + window > INT_MAX || lr.r0 - window < INT_MIN))
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "diagonal layout: offsets beyond 32 bits");
  const int s_lds = (int)((nc + kSliceCols - 1) / kSliceCols);  // LDS slices of x
  static std::atomic<unsigned long long> builds{0};
  auto *sp = new SplitCsr();
  sp->build_id = ++builds;
  sp->layout = layout;
  sp->window = window;
  sp->slices = slices;
  sp->nblk = (n + kCombineRows - 1) / kCombineRows;
  m->split = sp;
  int *lcnt = nullptr, *rcnt = nullptr, *rptr = nullptr, *rc = nullptr, *pay = nullptr,
      *scount = nullptr, *hist = nullptr, *dmap = nullptr;
  double *rv = nullptr;
  unsigned long long *keys = nullptr;
  void *ws = nullptr;
  int tot32[2] = {0, 0};  // rptr[n] and lrp[n], copied from here before the stream syncs
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    (void)hipFree(lcnt);
    (void)hipFree(rcnt);
    (void)hipFree(rptr);
    (void)hipFree(rc);
    (void)hipFree(pay);
    (void)hipFree(scount);
    (void)hipFree(rv);
    (void)hipFree(keys);
    (void)hipFree(ws);
    (void)hipFree(hist);
    (void)hipFree(dmap);
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
      (st = device_alloc((void **)&rptr, (size_t)(n + 1) * 4)) ||
      (st = device_alloc((void **)&scount, (size_t)(std::max(slices, s_lds) + 1) * 4)) ||
      (st = device_alloc(&ws, scan_workspace_bytes(n + 1) + 64)))
    return failed(st);
  long long *d_total = reinterpret_cast<long long *>(static_cast<char *>(ws) + scan_workspace_bytes(n + 1));
  RowSrc src{lr.rp, lr.ell_w, n};
  if (diag) {
    // the stored diagonals: offsets within the window holding at least a
    // quarter of the rows' worth of nonzero entries (at most 64, the
    // fullest), ascending
    const int nb = 2 * window + 1;
    std::vector<int> h(nb), map(nb, -1);
    if ((st = device_alloc((void **)&hist, (size_t)nb * 4)) ||
        (st = device_alloc((void **)&dmap, (size_t)nb * 4)))
      return failed(st);
    (void)hipMemsetAsync(hist, 0, (size_t)nb * 4, s);
    const long long step = std::max(1LL, n / 262144);  // rows sampled: at most 262144
    k_diag_hist<<<256, kBlock, 0, s>>>(n, step, f_col, lr.r0, window, src, m->colidx, m->values,
                                       hist);
    if (hipMemcpyAsync(h.data(), hist, (size_t)nb * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return failed(fail(TBSLA_HIP_ERR_HIP, "diagonal layout: histogram"));
    std::vector<int> cand;
    const long long sampled = (n + step - 1) / step;
    for (int d = 0; d < nb; d++)
      if ((long long)h[d] * 4 >= sampled && h[d] > 0) cand.push_back(d);
    std::stable_sort(cand.begin(), cand.end(), [&h](int a, int b) { return h[a] > h[b]; });
    if (cand.size() > 64) cand.resize(64);
    std::sort(cand.begin(), cand.end());
    sp->ndiag = (int)cand.size();
    // whole tiles of kWave rows
    const size_t dia_bytes =
        (size_t)std::max(sp->ndiag, 1) * (size_t)((n + kWave - 1) / kWave) * kWave * 8;
    std::vector<int> offs(std::max<size_t>(cand.size(), 1), 0);
    for (size_t k = 0; k < cand.size(); k++) {
      map[cand[k]] = (int)k;
      offs[k] = (int)(cand[k] - window + lr.r0);  // the kernels read x at r + offs[k]
    }
    if ((st = device_alloc((void **)&sp->doff, offs.size() * 4)) ||
        (st = device_alloc((void **)&sp->dv, dia_bytes)))
      return failed(st);
    (void)hipMemcpyAsync(dmap, map.data(), (size_t)nb * 4, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(sp->doff, offs.data(), offs.size() * 4, hipMemcpyHostToDevice, s);
    (void)hipMemsetAsync(sp->dv, 0, dia_bytes, s);
  }
  k_split_count<<<sgrid(n), kBlock, 0, s>>>(n, f_col, lr.r0, window, src, m->colidx, m->values,
                                            dmap, lcnt, rcnt);
  long long tot[2] = {0, 0};
  int *cnts[2] = {lcnt, rcnt}, *ptrs[2] = {sp->lrp, rptr};
  for (int k = 0; k < 2; k++) {
    if ((st = scan<int>(cnts[k], ptrs[k], n, false, d_total, ws, s))) return failed(st);
    if (hipMemcpyAsync(&tot[k], d_total, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return failed(fail(TBSLA_HIP_ERR_HIP, "split layout: scan"));
    tot32[k] = (int)tot[k];
    (void)hipMemcpyAsync(ptrs[k] + n, &tot32[k], 4, hipMemcpyHostToDevice, s);
  }
  sp->nl = tot[0];
  sp->nr = tot[1];
  const long long nr1 = std::max(sp->nr, 1LL);
  // the diagonal layout's remote entries through LDS slices when every
  // block of at most 4096 rows fits and the run tables stay small
  int H = 0;
  if (diag && sp->nr > 0) {
    if ((st = pick_block_rows(rptr, n, s, &H))) return failed(st);
    // block indices are 2-byte
    const long long nblocks = H > 0 ? (n + H - 1) / H : 0;
    if (H > 0 && (nblocks > 65536 || (long long)s_lds * nblocks > (1LL << 26))) H = 0;
  }
  const long long nl_store = diag ? 1 : std::max(sp->nl, 1LL);  // DIA entries live in dv
  if ((st = device_alloc((void **)&sp->lci, (size_t)nl_store * 4)) ||
      (st = device_alloc((void **)&sp->lv, (size_t)nl_store * 8)) ||
      (st = device_alloc((void **)&sp->sv, (size_t)nr1 * 8)) ||
      (st = device_alloc((void **)&rc, (size_t)nr1 * 4)) ||
      (st = device_alloc((void **)&rv, (size_t)nr1 * 8)) ||
      (st = device_alloc((void **)&keys, (size_t)nr1 * 8)) ||
      (st = device_alloc((void **)&pay, (size_t)nr1 * 4)))
    return failed(st);
  (void)hipMemsetAsync(scount, 0, (size_t)(std::max(slices, s_lds) + 1) * 4, s);
  const int rb = bits_of(std::max(n, 1LL));
  k_split_fill<<<sgrid(n), kBlock, 0, s>>>(n, nc, f_col, lr.r0, window, slices,
                                           H > 0 ? kSliceCols : 0, rb, src, m->colidx, m->values, dmap, sp->ndiag, sp->lrp,
                                           rptr, sp->lci, sp->lv, sp->dv, rc, rv, keys, pay, scount);
  if (H > 0) {
    const int S = s_lds, R = (int)((n + H - 1) / H);
    sp->lds_h = H;
    sp->lds_s = S;
    sp->lds_r = R;
    // at least 256 workgroups, at most kRunsCap blocks each
    sp->lds_g = std::max(1, std::min(R, std::max((256 + S - 1) / S, (R + kRunsCap - 1) / kRunsCap)));
    const long long runs = (long long)S * R;
    if ((st = device_alloc((void **)&sp->c16, (size_t)nr1 * 2)) ||
        (st = device_alloc((void **)&sp->q16, (size_t)nr1 * 2)) ||
        (st = device_alloc((void **)&sp->d16, (size_t)nr1 * 2)) ||
        (st = device_alloc((void **)&sp->P, (size_t)nr1 * 8)) ||
        (st = device_alloc((void **)&sp->sst, (size_t)(runs + 1) * 4)) ||
        (st = device_alloc((void **)&sp->pst, (size_t)runs * 4)) ||
        (st = device_alloc((void **)&sp->pb, (size_t)(R + 1) * 4)) ||
        (st = device_alloc((void **)&sp->rs16, (size_t)R * (size_t)(H + 1) * 2)))
      return failed(st);
    if (sp->nr > 1 && (st = radix_sort_pairs(keys, pay, sp->nr, rb + bits_of(S), s))) return failed(st);
    k_slice_place<<<sgrid(sp->nr), kBlock, 0, s>>>(sp->nr, keys, rb, H, pay, rc, rv, sp->c16,
                                                   sp->q16, sp->sv);
    k_block_starts<<<sgrid(R + 1), kBlock, 0, s>>>(n, H, R, rptr, sp->pb);
    k_slice_starts<<<sgrid(runs + 1), kBlock, 0, s>>>(sp->nr, rb, S, R, H, keys, sp->sst);
    k_slice_pstarts<<<sgrid(R), kBlock, 0, s>>>(S, R, sp->sst, sp->pb, sp->pst);
    k_slice_dest<<<sgrid(sp->nr), kBlock, 0, s>>>(sp->nr, rb, R, H, keys, pay, sp->sst, sp->pst,
                                                  sp->pb, sp->d16);
    k_row_starts16<<<sgrid((long long)R * (H + 1)), kBlock, 0, s>>>(n, H, R, rptr, sp->pb, sp->rs16);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return failed(fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e)));
    cleanup();
    return TBSLA_HIP_SUCCESS;
  }
  if ((st = device_alloc((void **)&sp->R, (size_t)n * 8)) ||
      (st = device_alloc((void **)&sp->boff, (size_t)slices * (size_t)(sp->nblk + 1) * 4)) ||
      (st = device_alloc((void **)&sp->sr, (size_t)nr1 * 4)) ||
      (st = device_alloc((void **)&sp->sc, (size_t)nr1 * 4)) ||
      // P, followed by the device copy of the slice offsets
      (st = device_alloc((void **)&sp->P, (size_t)nr1 * 8 + (size_t)(slices + 1) * 8)))
    return failed(st);
  if (sp->nr > 1 && (st = radix_sort_pairs(keys, pay, sp->nr, rb + bits_of(slices), s))) return failed(st);
  if (sp->nr > 0)
    k_split_place<<<sgrid(sp->nr), kBlock, 0, s>>>(sp->nr, keys, rb, pay, rc, rv, sp->sr, sp->sc,
                                                   sp->sv);
  std::vector<int> hc(slices + 1, 0);
  if (hipMemcpyAsync(hc.data(), scount, (size_t)slices * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return failed(fail(TBSLA_HIP_ERR_HIP, "split layout: slice counts"));
  sp->soff.assign(slices + 1, 0);
  for (int k = 0; k < slices; k++) sp->soff[k + 1] = sp->soff[k] + hc[k];
  long long *d_soff = reinterpret_cast<long long *>(sp->P + nr1);
  if (hipMemcpyAsync(d_soff, sp->soff.data(), (size_t)(slices + 1) * 8, hipMemcpyHostToDevice, s) != hipSuccess)
    return failed(fail(TBSLA_HIP_ERR_HIP, "split layout: offsets"));
  k_block_offsets<<<sgrid((long long)slices * (sp->nblk + 1)), kBlock, 0, s>>>(
      d_soff, slices, sp->nblk, sp->sr, sp->boff);
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
  if (m->format == TBSLA_HIP_DENSE)
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "CSR, ELL, COO and SCOO blocks only");
  if (layout == TBSLA_HIP_CSR_ROWS) {
    split_release(m);
    return TBSLA_HIP_SUCCESS;
  }
  const bool coo = m->format == TBSLA_HIP_COO, scoo = m->format == TBSLA_HIP_SCOO;
  if (coo || scoo) {
    // over the triplets' rows through their row pointer (as their SpMV)
    const long long n = (scoo && m->nnz == 0) ? 0 : m->stored;
    if (!m->has_arrays || n <= 0) return fail(TBSLA_HIP_ERR_INVALID, "empty block");
    bool ok = false;
    if (m->row_sorted) TBSLA_TRY(coo_rowptr(m, n, &ok, static_cast<hipStream_t>(stream)));
    if (!ok) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "triplets not sorted by row");
  }
  if (layout != TBSLA_HIP_CSR_SPLIT && layout != TBSLA_HIP_CSR_DIAG)
    return fail(TBSLA_HIP_ERR_INVALID, "unknown layout");
  if (window < 0 || slices < 1 || slices > 64 || slices % 8)
    return fail(TBSLA_HIP_ERR_INVALID, "window >= 0, slices a multiple of 8 up to 64");
  if (layout == TBSLA_HIP_CSR_DIAG && window > kMaxDiagWindow)
    return fail(TBSLA_HIP_ERR_INVALID, "diagonal layout: window at most 1048576");
  const LayoutRows lr = layout_rows(m);
  if (layout == TBSLA_HIP_CSR_DIAG && lr.nc == 0)  // its kernel reads x for every slot
    return fail(TBSLA_HIP_ERR_INVALID, "diagonal layout: block without columns");
  if (!m->has_arrays || lr.n == 0 || (m->format == TBSLA_HIP_ELL && m->max_col == 0))
    return fail(TBSLA_HIP_ERR_INVALID, "empty block");
  return split_build(m, window, slices, static_cast<hipStream_t>(stream), layout);
}

extern "C" int tbsla_hip_get_csr_layout(tbsla_hip_matrix_t m, int *layout, int *window,
                                        int *slices, long long *remote_entries) {
  if (!m) return fail(TBSLA_HIP_ERR_INVALID, "null handle");
  const SplitCsr *sp = static_cast<const SplitCsr *>(m->split);
  if (layout) *layout = sp ? sp->layout : TBSLA_HIP_CSR_ROWS;
  if (window) *window = sp ? sp->window : 0;
  if (slices) *slices = sp ? sp->slices : 0;
  if (remote_entries) *remote_entries = sp ? sp->nr : 0;
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_get_csr_remote_form(tbsla_hip_matrix_t m, int *block_rows) {
  if (!m || !block_rows) return fail(TBSLA_HIP_ERR_INVALID, "null handle or output");
  *block_rows = split_lds_block_rows(m);
  return TBSLA_HIP_SUCCESS;
}
