// spmm.hip -- C = A B (MatrixCSR::dense_multiply) and personalised PageRank
// over k teleport vectors as one SpMM per iteration.
//
// One wavefront per row, lane t owning columns t, t+64, ... of the dense
// operands (row-major n x k).  A row's (column, value) pairs are
// wave-uniform and come through the scalar cache; each nonzero reads one
// contiguous k-double row of B.  Products are added in the row's storage
// order with separate multiply and add (no FMA), so C is bit-identical to
// the reference's loop (src/mpi/matrix/MatrixCSR.cpp:11-30).
#include <algorithm>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

constexpr int kMaxTiles = 4;  // columns per launch row: 256; wider B in chunks (grid y)

template <int KT>
__global__ __launch_bounds__(kBlock) void k_csr_spmm(long long n_rows, int k, int f_col, int ln_col,
                                                     const int *__restrict__ rowptr,
                                                     const int *__restrict__ colidx,
                                                     const double *__restrict__ vals,
                                                     const double *__restrict__ B,
                                                     double *__restrict__ C) {
  TBSLA_NO_FMA
  const long long row = __builtin_amdgcn_readfirstlane(
      (int)(((long long)blockIdx.x * kBlock + threadIdx.x) / kWave));
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= n_rows) return;
  const int k0 = blockIdx.y * (kMaxTiles * kWave);  // this launch row's column chunk
  double acc[KT];
#pragma unroll
  for (int t = 0; t < KT; t++) acc[t] = 0.0;
  const int beg = rowptr[row], end = rowptr[row + 1];
  int j = beg;
  // E entries per step: their B rows' loads are in flight together; the
  // products are still added in entry order
  constexpr int E = 4;
  for (; j + E <= end; j += E) {
    bool in[E];
    double v[E];
    const double *bp[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int c = colidx[j + e] - f_col;
      in[e] = c >= 0 && c < ln_col;  // block filter (:19)
      v[e] = vals[j + e];
      bp[e] = B + (size_t)(in[e] ? c : 0) * k;
    }
    double xv[E][KT];
#pragma unroll
    for (int e = 0; e < E; e++)
#pragma unroll
      for (int t = 0; t < KT; t++) {
        const int col = k0 + lane + t * kWave;
        xv[e][t] = col < k ? bp[e][col] : 0.0;
      }
#pragma unroll
    for (int e = 0; e < E; e++)
#pragma unroll
      for (int t = 0; t < KT; t++)
        if (in[e]) acc[t] = acc[t] + v[e] * xv[e][t];
  }
  for (; j < end; j++) {
    const int c = colidx[j] - f_col;
    if (c >= 0 && c < ln_col) {
      const double v = vals[j];
      const double *b = B + (size_t)c * k;
#pragma unroll
      for (int t = 0; t < KT; t++) {
        const int col = k0 + lane + t * kWave;
        if (col < k) acc[t] = acc[t] + v * b[col];
      }
    }
  }
  double *out = C + (size_t)row * k;
#pragma unroll
  for (int t = 0; t < KT; t++) {
    const int col = k0 + lane + t * kWave;
    if (col < k) out[col] = acc[t];
  }
}

struct PprState {
  int iters;
  int done;
  double max_err;
};

// Row slots: a workgroup holds kPprSlots rows at a time (kPprLanes lanes
// per row, each lane owning 4 of the 64 columns).  Workgroup b runs on XCD
// b mod 8 (workgroups are dispatched round-robin over the eight XCDs): XCD x
// takes the x-th contiguous eighth of the rows and its slots take those rows
// in turn (slot v of the XCD's W slots: rows v, v + W, ...), so at any
// moment an XCD works on a window of about W consecutive rows, whose
// near-diagonal X rows stay in its L2.  Other grid sizes: every slot of the
// grid in turn.
constexpr int kPprLanes = 16;                  // lanes per row
constexpr int kPprSlots = kBlock / kPprLanes;  // rows in flight per workgroup
constexpr int kPprDepth = 8;                   // entries whose X rows load together
constexpr int kPprGrid = 2048;                 // workgroups of a step

struct RowWalk {
  long long first, stride, end;
};
__device__ inline RowWalk xcd_walk(long long n_rows, int nb, int slot) {
  const int b = blockIdx.x;
  if (nb % 8 != 0) return RowWalk{(long long)b * kPprSlots + slot, (long long)nb * kPprSlots, n_rows};
  const long long per = (n_rows + 7) / 8;
  const long long x0 = min(n_rows, (long long)(b % 8) * per);
  return RowWalk{x0 + (long long)(b / 8) * kPprSlots + slot, (long long)(nb / 8) * kPprSlots,
                 min(n_rows, x0 + per)};
}

constexpr int kPprCols = kWave / kPprLanes;  // columns per lane

// kPprCols consecutive doubles (aligned to their size) in one load / store
template <int N>
struct DVec {
  typedef double type __attribute__((ext_vector_type(N)));
};
template <>
struct DVec<1> {
  typedef double type;
};
typedef DVec<kPprCols>::type pprvec_t;
template <typename VT>
__device__ inline double vget(const VT &v, int q) {
  if constexpr (std::is_same<VT, double>::value) return v;
  else return v[q];
}
template <typename VT>
__device__ inline void vset(VT &v, int q, double x) {
  if constexpr (std::is_same<VT, double>::value) v = x;
  else v[q] = x;
}

// ---- personalised PageRank step: tiles of rows, the band window of X in LDS ----
//
// One step for k = 64 columns fused with the SpMM:
// Y[i,:] = A[i,:] X; X'[i,k] = (beta Y[i,k] + S_k (1-beta) [i == src_k]) / S_k;
// per-workgroup partials of |X' - X| and of X' per column -- the page_rank
// update (src/mpi/matrix/Matrix.cpp:123-141) with a one-hot teleport per
// column, on the SpMM of src/mpi/matrix/MatrixCSR.cpp:11-30.
//
// A workgroup takes a contiguous run of tiles of kPprTile consecutive rows.
// For a tile starting at row r0 the X rows [r0 + dlo, r0 + dlo + kPprWin) --
// the band the tile's near-diagonal entries read (dlo from a histogram of
// the entries' offsets) -- are in LDS, 512 bytes per row, served to every
// entry of the band at LDS bandwidth.  The LDS holds them as a ring of
// kPprWin rows (X row j in slot j mod kPprWin): the next tile's window is
// this one shifted by kPprTile rows, so only its last kPprTile rows are new;
// they are loaded into registers while this tile is summed and written over
// the departing rows at the tile boundary.  Each X row is read from HBM once
// per step (plus the first window of each workgroup), and a row's old value
// X[r, :] comes from the ring too when the tile's own rows lie in its window.
// Each entry's X row comes from LDS when it lies in the window ("band") or
// from HBM ("remote": rewired entries, about 10% in a cqmat with q = 0.1);
// a row's sum is its band entries in storage order, then its remote entries
// in storage order, each a fused multiply-add (k_ppr_quad below).
constexpr int kPprBlock = 256;                  // threads of a tile workgroup
constexpr int kPprWaves = kPprBlock / kWave;    // 4 wavefronts
constexpr int kPprTile = 64;                    // rows per tile
constexpr int kPprRowsPerWave = kPprTile / kPprWaves;
constexpr int kPprWin = 96;                     // X rows staged per tile (48 KB)
constexpr int kPprReach = kPprWin - kPprTile;   // band offsets [dlo, dlo + 32] always staged
constexpr int kPprHist = 128;                   // offsets [-64, 63] in the histogram

struct PprArgs {
  int n;  // rows = columns of the square block (int, as the reference's ln_row)
  int f_col;
  int old_lds;  // every tile's own rows lie in its window: old X values from the ring
  const int *rowptr, *colidx;
  const double *vals;
  const double *X;
  double *Xn;
  const int *src;
  const double *S;
  double beta;
  const PprState *st;
  double *p_err, *p_sum;
  int dlo;
  int ntiles;
  int last;  // index of the last stored entry (a dummy entry when the block has none)
};

// one product added to a row sum (a fused multiply-add: one VALU
// instruction per entry and column instead of two; the step is within
// rounding of the unfused sum)
__device__ __forceinline__ double ppr_madd(double acc, double v, double x) { return fma(v, x, acc); }

// the ring slot of X row j (j >= -kPprWin)
__device__ __forceinline__ int ring_slot(int j) { return (int)((unsigned)(j + kPprWin) % kPprWin); }

// the window of row r's tile in the ring: X rows [wlo, whi)
struct PprWindow {
  int wlo, whi;
};
__device__ __forceinline__ PprWindow ppr_window(const PprArgs &a, int r) {
  const int wbase = (r / kPprTile) * kPprTile + a.dlo;
  PprWindow w;
  w.wlo = min(max(wbase, 0), a.n);
  w.whi = max(w.wlo, min(wbase + kPprWin, a.n));
  return w;
}

// The rows of one wavefront: its workgroup's tiles t0, t0 + tstep, ... below
// t1; row q of the wavefront is row wave + kPprWaves (q mod kPprRowsPerWave)
// of its (q / kPprRowsPerWave)-th tile.
struct PprWalk {
  int t0, t1, tstep;
  int wave;
};
__device__ __forceinline__ int ppr_row_at(const PprArgs &a, const PprWalk &wk, int q) {
  const int t = wk.t0 + (q / kPprRowsPerWave) * wk.tstep;
  const int r = t * kPprTile + wk.wave + kPprWaves * (q % kPprRowsPerWave);
  return (t < wk.t1 && r < a.n) ? r : -1;
}

// Distances of the stages ahead of the row being summed: P at D0 = NS - 1,
// E at DE, R at DR.
template <int NS, int DE, int DR>
struct PprPipe {
  static constexpr int D0 = NS - 1;
  static_assert(D0 > DE && DE > DR && DR >= 1, "each stage after the one it consumes");
};

// The ring's loads: kPprRingLoads 16-byte pieces per thread cover the
// kPprTile rows a tile adds to the ring.  Rows [j0, j0 + kPprTile) of tile t
// are loaded and stored when t < t1 and the row lies in [0, min(n, jhi)); a
// missing row reads X's first piece and is not stored.
constexpr int kPprPieces = kWave / 2;  // 16-byte pieces per X row
constexpr int kPprRingLoads = kPprTile * kPprPieces / kPprBlock;
static_assert(kPprTile * kPprPieces % kPprBlock == 0, "whole pieces per thread");

// one 16-byte piece of an X row (a vector type, so the prefetched pieces
// stay in registers)
typedef double pprpiece_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bool ring_row(const PprArgs &a, bool tile_ok, int j, int jhi) {
  return tile_ok && j >= 0 && j < a.n && j < jhi;
}
__device__ __forceinline__ void ring_load(const PprArgs &a, bool tile_ok, int j0, int jhi,
                                          pprpiece_t (&pre)[kPprRingLoads]) {
  const pprpiece_t *X2 = reinterpret_cast<const pprpiece_t *>(a.X);
#pragma unroll
  for (int k = 0; k < kPprRingLoads; k++) {
    const int p = (int)threadIdx.x + k * kPprBlock;
    const int j = j0 + p / kPprPieces;
    pre[k] = X2[ring_row(a, tile_ok, j, jhi) ? (size_t)j * kPprPieces + p % kPprPieces : 0];
  }
}
__device__ __forceinline__ void ring_store(const PprArgs &a, bool tile_ok, int j0, int jhi,
                                           const pprpiece_t (&pre)[kPprRingLoads], double *win) {
  pprpiece_t *w2 = reinterpret_cast<pprpiece_t *>(win);
#pragma unroll
  for (int k = 0; k < kPprRingLoads; k++) {
    const int p = (int)threadIdx.x + k * kPprBlock;
    const int j = j0 + p / kPprPieces;
    if (ring_row(a, tile_ok, j, jhi)) w2[ring_slot(j) * kPprPieces + p % kPprPieces] = pre[k];
  }
}

// ---- four rows per wavefront ----------------------------------------------
//
// The same step with a wavefront taking four rows at a time, one per group
// of kQuadLanes = 16 lanes (a DPP row), lane t of a group holding columns
// 4t .. 4t + 3.  A row's first 32 entries sit two per lane (entries t and
// t + 16); entry K reaches its group's lanes by DPP row_newbcast (one
// v_mov_dpp per 32-bit half, no readlane and no scalar register), so a band
// entry costs two 16-byte LDS reads, three broadcasts, one address and four
// fused multiply-adds for the four rows' 64 columns at once.  Remote
// entries are found per group from the wavefront's ballots and fetched by
// lane shuffles.  A workgroup takes a contiguous run of tiles; each
// wavefront runs its rows through four pipeline stages (row extents,
// entries, remote X rows, sum) so that a row's loads are in flight while
// earlier rows are summed, each step running its sum first
// (quad_group_step); entries past a row's first 32 go sixteen at a time,
// unpipelined.
constexpr int kQuadLanes = 16;
constexpr int kQuadRows = kWave / kQuadLanes;  // rows per wavefront step
constexpr int kQuadEnt = 2 * kQuadLanes;       // entries per row in the pipeline

// lane K of each group of 16, in every lane of the group (row_newbcast reads
// a lane of the lane's own row, so every lane has a source; bound_ctrl set
// lets the compiler give the broadcast straight to the instruction using it,
// as the address add does)
template <int K>
__device__ __forceinline__ int qbc(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xf, 0xf, true);
}
template <int K>
__device__ __forceinline__ double qbc_d(double v) {
  return __hiloint2double(qbc<K>(__double2hiint(v)), qbc<K>(__double2loint(v)));
}

// a quad slot: four rows (row q of each group), this lane's part
template <int S>
struct QuadSlot {
  int beg, end;            // the group's row extent
  int col[2];              // entries t, t + 16 of the row: column, -1 past the row
  double val[2];
  pprpiece_t old[2];       // X[r, 4t .. 4t + 3] (when not read from the ring)
  pprpiece_t xr[S][2];     // the group's first S remote entries: X rows, this lane's columns
  double vr[S];            // and values (0 past the count)
};

// row of group g at quad step qs: the row q = 4 qs + g of the wavefront's sequence
__device__ __forceinline__ int quad_row(const PprArgs &a, const PprWalk &wk, int qs, int g) {
  return ppr_row_at(a, wk, kQuadRows * qs + g);
}

// stage P
template <int S>
__device__ __forceinline__ void quad_sp(const PprArgs &a, int r, QuadSlot<S> &s) {
  const int rr = max(r, 0);
  const int b = a.rowptr[rr], e = a.rowptr[rr + 1];
  s.beg = b;
  s.end = r < 0 ? b : e;
}

// stage E: the first 32 entries (two per lane)
template <int S, bool OLD>
__device__ __forceinline__ void quad_se(const PprArgs &a, int r, int t, QuadSlot<S> &s) {
  const int len = s.end - s.beg;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int i = t + h * kQuadLanes;
    const int j = min(s.beg + min(i, max(len - 1, 0)), a.last);
    const int c = a.colidx[j];
    const double v = a.vals[j];
    s.col[h] = i < len ? c - a.f_col : -1;
    s.val[h] = i < len ? v : 0.0;
  }
}

// the group's remote entries among the lane's two (bit t: entry t, bit
// 16 + t: entry 16 + t), from the wavefront's ballots
__device__ __forceinline__ unsigned quad_remote_mask(const PprArgs &a, int c0, int c1, int wlo,
                                                     int whi, int g) {
  const unsigned long long m0 = __ballot(c0 >= 0 && c0 < a.n && (c0 < wlo || c0 >= whi));
  const unsigned long long m1 = __ballot(c1 >= 0 && c1 < a.n && (c1 < wlo || c1 >= whi));
  return (unsigned)((m0 >> (kQuadLanes * g)) & 0xFFFFu) |
         ((unsigned)((m1 >> (kQuadLanes * g)) & 0xFFFFu) << 16);
}

// entry idx (0 .. 31) of the group's row: column and value, by shuffles
__device__ __forceinline__ void quad_pick(const int (&col)[2], const double (&val)[2], int g,
                                          int idx, int &c, double &v) {
  const int src = kQuadLanes * g + (idx & (kQuadLanes - 1));
  const int c0 = __shfl(col[0], src), c1 = __shfl(col[1], src);
  const double v0 = __shfl(val[0], src), v1 = __shfl(val[1], src);
  c = idx < kQuadLanes ? c0 : c1;
  v = idx < kQuadLanes ? v0 : v1;
}

// stage R: the X rows of the group's first S remote entries and, unless OLD
// (every row's old values in the ring), the row's old values
template <int S, bool OLD>
__device__ __forceinline__ void quad_sr(const PprArgs &a, int r, int t, int g, QuadSlot<S> &s) {
  if constexpr (!OLD) {
    const pprpiece_t *X2 = reinterpret_cast<const pprpiece_t *>(a.X);
    const size_t o = (size_t)max(r, 0) * kPprPieces + 2 * t;
    s.old[0] = X2[o];
    s.old[1] = X2[o + 1];
  }
  const PprWindow w = ppr_window(a, max(r, 0));
  unsigned m = quad_remote_mask(a, s.col[0], s.col[1], w.wlo, w.whi, g);
  const pprpiece_t *X2 = reinterpret_cast<const pprpiece_t *>(a.X);
#pragma unroll
  for (int k = 0; k < S; k++) {
    const bool has = m != 0;
    const int idx = has ? __builtin_ctz(m) : 0;
    m &= m - 1;
    int c;
    double v;
    quad_pick(s.col, s.val, g, idx, c, v);
    const size_t o = (size_t)(has ? c : 0) * kPprPieces + 2 * t;  // unused: X row 0
    s.xr[k][0] = X2[o];
    s.xr[k][1] = X2[o + 1];
    s.vr[k] = has ? v : 0.0;
  }
}

// the band walk over one lane's entries (h = 0: entries 0 .. 15, h = 1:
// 16 .. 31 of each group), K = 0 .. 15 in storage order
template <int K>
__device__ __forceinline__ void quad_walk(int off, double v, const char *win_b, int t,
                                          double (&acc)[4]) {
  const int o = qbc<K>(off) + 32 * t;
  const double vk = qbc_d<K>(v);
  const pprpiece_t x0 = *reinterpret_cast<const pprpiece_t *>(win_b + o);
  const pprpiece_t x1 = *reinterpret_cast<const pprpiece_t *>(win_b + o + 16);
  acc[0] = ppr_madd(acc[0], vk, x0.x);
  acc[1] = ppr_madd(acc[1], vk, x0.y);
  acc[2] = ppr_madd(acc[2], vk, x1.x);
  acc[3] = ppr_madd(acc[3], vk, x1.y);
  if constexpr (K + 1 < kQuadLanes) quad_walk<K + 1>(off, v, win_b, t, acc);
}

// byte offset in the ring of entry column c's X row (the zero row outside
// the window), and the value the walk multiplies
__device__ __forceinline__ void quad_band(int c, double val, int wlo, int whi, int &off,
                                          double &v) {
  const bool band = c >= wlo && c < whi;
  off = (band ? ring_slot(c) : kPprWin) * kWave * 8;
  v = band ? val : 0.0;
}

// remote products of entries whose group bits are in m, one at a time
__device__ __forceinline__ void quad_remote_rest(const PprArgs &a, unsigned m, const int (&col)[2],
                                                 const double (&val)[2], int t, int g,
                                                 double (&acc)[4]) {
  const pprpiece_t *X2 = reinterpret_cast<const pprpiece_t *>(a.X);
  while (__ballot(m != 0)) {
    const bool has = m != 0;
    const int idx = has ? __builtin_ctz(m) : 0;
    m &= m - 1;
    int c;
    double v;
    quad_pick(col, val, g, idx, c, v);
    const size_t o = (size_t)(has ? c : 0) * kPprPieces + 2 * t;
    const pprpiece_t x0 = X2[o], x1 = X2[o + 1];
    v = has ? v : 0.0;
    acc[0] = ppr_madd(acc[0], v, x0.x);
    acc[1] = ppr_madd(acc[1], v, x0.y);
    acc[2] = ppr_madd(acc[2], v, x1.x);
    acc[3] = ppr_madd(acc[3], v, x1.y);
  }
}

// stage C: the four rows' sums -- band entries of the first 32 in storage
// order, the first S remote ones, the other remote ones, then entries past
// the first 32 sixteen at a time (band, then remote) -- and their update
template <int S, bool OLD>
__device__ __forceinline__ void quad_sc(const PprArgs &a, int r, int t, int g, const QuadSlot<S> &s,
                                        const double *__restrict__ win, double (&e)[4],
                                        double (&sm)[4]) {
  TBSLA_NO_FMA
  const PprWindow w = ppr_window(a, max(r, 0));
  const char *win_b = reinterpret_cast<const char *>(win);
  const int len = s.end - s.beg;
  // the longest of the four rows (every lane)
  int ml = len;
  ml = max(ml, __shfl_xor(ml, kQuadLanes));
  ml = max(ml, __shfl_xor(ml, 2 * kQuadLanes));
  ml = __builtin_amdgcn_readfirstlane(ml);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  {
    int off;
    double v;
    quad_band(s.col[0], s.val[0], w.wlo, w.whi, off, v);
    quad_walk<0>(off, v, win_b, t, acc);
    if (ml > kQuadLanes) {
      quad_band(s.col[1], s.val[1], w.wlo, w.whi, off, v);
      quad_walk<0>(off, v, win_b, t, acc);
    }
  }
#pragma unroll
  for (int k = 0; k < S; k++) {
    acc[0] = ppr_madd(acc[0], s.vr[k], s.xr[k][0].x);
    acc[1] = ppr_madd(acc[1], s.vr[k], s.xr[k][0].y);
    acc[2] = ppr_madd(acc[2], s.vr[k], s.xr[k][1].x);
    acc[3] = ppr_madd(acc[3], s.vr[k], s.xr[k][1].y);
  }
  {
    unsigned m = quad_remote_mask(a, s.col[0], s.col[1], w.wlo, w.whi, g);
#pragma unroll
    for (int k = 0; k < S; k++) m &= m - 1;  // the ones stage R fetched
    quad_remote_rest(a, m, s.col, s.val, t, g, acc);
  }
  // entries past the first 32, sixteen per chunk
  for (int c0 = kQuadEnt; c0 < ml; c0 += kQuadLanes) {
    const int i = c0 + t;
    const int j = min(s.beg + min(i, max(len - 1, 0)), a.last);
    const int cj = a.colidx[j];
    const double vj = a.vals[j];
    const int col[2] = {i < len ? cj - a.f_col : -1, -1};
    const double val[2] = {i < len ? vj : 0.0, 0.0};
    int off;
    double v;
    quad_band(col[0], val[0], w.wlo, w.whi, off, v);
    quad_walk<0>(off, v, win_b, t, acc);
    const unsigned m = quad_remote_mask(a, col[0], col[1], w.wlo, w.whi, g);
    quad_remote_rest(a, m, col, val, t, g, acc);
  }
  if (r < 0) return;
  double old[4];
  if constexpr (OLD) {
    const pprpiece_t *o2 = reinterpret_cast<const pprpiece_t *>(win + ring_slot(r) * kWave);
    const pprpiece_t p0 = o2[2 * t], p1 = o2[2 * t + 1];
    old[0] = p0.x;
    old[1] = p0.y;
    old[2] = p1.x;
    old[3] = p1.y;
  } else {
    old[0] = s.old[0].x;
    old[1] = s.old[0].y;
    old[2] = s.old[1].x;
    old[3] = s.old[1].y;
  }
  // the columns' sums and sources (read here, not held: registers)
  const pprpiece_t *S2 = reinterpret_cast<const pprpiece_t *>(a.S);
  const pprpiece_t sa = S2[2 * t], sb = S2[2 * t + 1];
  const int4 sr = reinterpret_cast<const int4 *>(a.src)[t];
  const double S4[4] = {sa.x, sa.y, sb.x, sb.y};
  const int src4[4] = {sr.x, sr.y, sr.z, sr.w};
  double tv[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double u = a.beta * acc[i] + ((r + a.f_col == src4[i]) ? S4[i] * (1 - a.beta) : 0.0);
    u = u / S4[i];
    tv[i] = u;
    e[i] += fabs(u - old[i]);
    sm[i] += u;
  }
  pprpiece_t *Xn2 = reinterpret_cast<pprpiece_t *>(a.Xn);
  const size_t o = (size_t)r * kPprPieces + 2 * t;
  Xn2[o] = pprpiece_t{tv[0], tv[1]};
  Xn2[o + 1] = pprpiece_t{tv[2], tv[3]};
}

// One group of NS quad steps (NS quad steps = one tile when NS = 4): for
// step qs0 + K, stage C, then R of step + DR, E of step + DE, P of step + D0.
template <int NS, int DE, int DR, int S, bool OLD, int K>
__device__ __forceinline__ void quad_group_step(const PprArgs &a, const PprWalk &wk, int qs0,
                                                QuadSlot<S> (&st)[NS],
                                                const double *__restrict__ win, int t, int g,
                                                double (&e)[4], double (&sm)[4]) {
  constexpr int D0 = PprPipe<NS, DE, DR>::D0;
  const int qs = qs0 + K;
  // C first: the remote rows R fetches next are then live only until the
  // next step's C, whose band walk covers their latency
  quad_sc<S, OLD>(a, quad_row(a, wk, qs, g), t, g, st[K], win, e, sm);
  quad_sr<S, OLD>(a, quad_row(a, wk, qs + DR, g), t, g, st[(K + DR) % NS]);
  quad_se<S, OLD>(a, quad_row(a, wk, qs + DE, g), t, st[(K + DE) % NS]);
  quad_sp(a, quad_row(a, wk, qs + D0, g), st[(K + D0) % NS]);
  if constexpr (K + 1 < NS)
    quad_group_step<NS, DE, DR, S, OLD, K + 1>(a, wk, qs0, st, win, t, g, e, sm);
}

template <int NS, int DE, int DR, int S, bool OLD, int K>
__device__ __forceinline__ void quad_prologue(const PprArgs &a, const PprWalk &wk,
                                              QuadSlot<S> (&st)[NS], int t, int g) {
  constexpr int D0 = PprPipe<NS, DE, DR>::D0;
  if constexpr (K < D0) {
    quad_sp(a, quad_row(a, wk, K, g), st[K]);
  } else if constexpr (K < D0 + DE) {
    quad_se<S, OLD>(a, quad_row(a, wk, K - D0, g), t, st[K - D0]);
  } else if constexpr (K < D0 + DE + DR) {
    quad_sr<S, OLD>(a, quad_row(a, wk, K - D0 - DE, g), t, g, st[K - D0 - DE]);
  }
  if constexpr (K + 1 < D0 + DE + DR) quad_prologue<NS, DE, DR, S, OLD, K + 1>(a, wk, st, t, g);
}

constexpr int kQuadStepsPerTile = kPprRowsPerWave / kQuadRows;

template <int NS, int DE, int DR, int S, bool OLD>
__global__ __launch_bounds__(kPprBlock, 3) void k_ppr_quad(PprArgs a) {
  TBSLA_NO_FMA
  static_assert(kQuadStepsPerTile % NS == 0, "a tile starts a group");
  static_assert(kPprWin > kPprTile && kPprWin <= 2 * kPprTile, "the first window is two loads");
  if (a.st->done) return;
  __shared__ __attribute__((aligned(16))) double win[(kPprWin + 1) * kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int t = lane & (kQuadLanes - 1), g = lane / kQuadLanes;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
  if (threadIdx.x < kWave) win[kPprWin * kWave + threadIdx.x] = 0.0;  // before the first barrier
  double e[4] = {0.0, 0.0, 0.0, 0.0}, sm[4] = {0.0, 0.0, 0.0, 0.0};
  const int nb = gridDim.x, b = blockIdx.x;
  const int per = (a.ntiles + nb - 1) / nb;
  PprWalk wk;
  wk.wave = wave;
  wk.t0 = min(a.ntiles, b * per);
  wk.t1 = min(a.ntiles, wk.t0 + per);
  wk.tstep = 1;
  const int nqs = (wk.t1 - wk.t0) * kQuadStepsPerTile;
  const int base0 = wk.t0 * kPprTile + a.dlo;
  const bool any = wk.t0 < wk.t1;
  pprpiece_t pre[kPprRingLoads];
  ring_load(a, any, base0, base0 + kPprWin, pre);
  ring_store(a, any, base0, base0 + kPprWin, pre, win);
  ring_load(a, any, base0 + kPprTile, base0 + kPprWin, pre);
  ring_store(a, any, base0 + kPprTile, base0 + kPprWin, pre, win);
  ring_load(a, wk.t0 + 1 < wk.t1, base0 + kPprWin, base0 + kPprWin + kPprTile, pre);
  __syncthreads();
  QuadSlot<S> st[NS];
  quad_prologue<NS, DE, DR, S, OLD, 0>(a, wk, st, t, g);
  for (int qs0 = 0; qs0 < nqs; qs0 += NS) {
    if (qs0 > 0 && qs0 % kQuadStepsPerTile == 0) {
      const int tt = wk.t0 + qs0 / kQuadStepsPerTile;
      const int jn = tt * kPprTile + a.dlo + (kPprWin - kPprTile);  // tile tt's new rows
      __syncthreads();  // tile tt - 1's rows are done with the departing rows
      ring_store(a, true, jn, jn + kPprTile, pre, win);
      __syncthreads();
      ring_load(a, tt + 1 < wk.t1, jn + kPprTile, jn + 2 * kPprTile, pre);
    }
    quad_group_step<NS, DE, DR, S, OLD, 0>(a, wk, qs0, st, win, t, g, e, sm);
  }
  // per-column partials: the four groups (fixed xor tree), then the
  // wavefronts in order
#pragma unroll
  for (int i = 0; i < 4; i++) {
    e[i] += __shfl_xor(e[i], kQuadLanes);
    e[i] += __shfl_xor(e[i], 2 * kQuadLanes);
    sm[i] += __shfl_xor(sm[i], kQuadLanes);
    sm[i] += __shfl_xor(sm[i], 2 * kQuadLanes);
  }
  __syncthreads();
  if (g == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      win[wave * kWave + 4 * t + i] = e[i];
      win[(kPprWaves + wave) * kWave + 4 * t + i] = sm[i];
    }
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    double te = 0.0, ts = 0.0;
    for (int v = 0; v < kPprWaves; v++) {
      te += win[v * kWave + lane];
      ts += win[(kPprWaves + v) * kWave + lane];
    }
    a.p_err[(size_t)b * kWave + lane] = te;
    a.p_sum[(size_t)b * kWave + lane] = ts;
  }
}

// Histogram of the entries' offsets col - row in [-64, 63] over every 16th
// row (the window's placement): per-workgroup bins in LDS, then added to the
// global bins.
__global__ __launch_bounds__(kBlock) void k_ppr_offsets(long long n, int f_col,
                                                        const int *__restrict__ rowptr,
                                                        const int *__restrict__ colidx,
                                                        unsigned int *__restrict__ bins) {
  __shared__ unsigned int h[kPprHist];
  if (threadIdx.x < kPprHist) h[threadIdx.x] = 0;
  __syncthreads();
  for (long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) * 16; i < n;
       i += (long long)gridDim.x * kBlock * 16) {
    for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
      const long long o = (long long)(colidx[j] - f_col) - i;
      if (o >= -kPprHist / 2 && o < kPprHist / 2) atomicAdd(&h[o + kPprHist / 2], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x < kPprHist && h[threadIdx.x]) atomicAdd(&bins[threadIdx.x], h[threadIdx.x]);
}

// Sum over nb per-block rows of 64 partials, in block order within each of
// kBlock / kWave quarters, the quarters then in order (a fixed order);
// result for column threadIdx.x (< 64) in thread threadIdx.x.
__device__ inline double block_order_sum(const double *__restrict__ part, int nb, double *lds) {
  const int col = threadIdx.x & (kWave - 1), qt = threadIdx.x / kWave;
  constexpr int kQ = kBlock / kWave;
  const int per = (nb + kQ - 1) / kQ;
  const int b0 = min(nb, qt * per), b1 = min(nb, b0 + per);
  double t = 0.0;
#pragma unroll 8
  for (int b = b0; b < b1; b++) t += part[(size_t)b * kWave + col];
  lds[qt * kWave + col] = t;
  __syncthreads();
  double tot = 0.0;
  if (qt == 0)
    for (int k = 0; k < kQ; k++) tot += lds[k * kWave + col];
  return tot;
}

// Column totals in a fixed order; converged when every column's change is
// below epsilon.
__global__ __launch_bounds__(kBlock) void k_ppr_finalize(PprState *__restrict__ st,
                                                         const double *__restrict__ p_err,
                                                         const double *__restrict__ p_sum, int nb,
                                                         double *__restrict__ S, double epsilon) {
  if (st->done) return;
  __shared__ double lds[2][kBlock];
  const double e = block_order_sum(p_err, nb, lds[0]);
  const double sm = block_order_sum(p_sum, nb, lds[1]);
  if (threadIdx.x >= kWave) return;
  S[threadIdx.x] = sm;
  double m = e;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
  if (threadIdx.x == 0) {
    st->max_err = m;
    st->iters += 1;
    if (m < epsilon) st->done = 1;
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_d(double *__restrict__ a, long long n, double v) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    a[i] = v;
}

// per-column sums of the final iterate, then X[:, k] /= sum_k
// (pagerank_norma_end, Matrix.cpp:149-171)
__global__ __launch_bounds__(kBlock) void k_col_partial(const double *__restrict__ X, long long n_rows,
                                                        double *__restrict__ part) {
  __shared__ double lds[kPprSlots][kWave];
  const int lane = threadIdx.x & (kPprLanes - 1), slot = threadIdx.x / kPprLanes;
  const int c0 = kPprCols * lane;
  double sm[kPprCols];
#pragma unroll
  for (int q = 0; q < kPprCols; q++) sm[q] = 0.0;
  const RowWalk wk = xcd_walk(n_rows, gridDim.x, slot);
  for (long long r = wk.first; r < wk.end; r += wk.stride) {
    const pprvec_t v = *reinterpret_cast<const pprvec_t *>(X + (size_t)r * kWave + c0);
#pragma unroll
    for (int q = 0; q < kPprCols; q++) sm[q] += vget(v, q);
  }
#pragma unroll
  for (int q = 0; q < kPprCols; q++) lds[slot][c0 + q] = sm[q];
  __syncthreads();
  if (threadIdx.x < kWave) {
    double t = 0.0;
    for (int w = 0; w < kPprSlots; w++) t += lds[w][threadIdx.x];
    part[(size_t)blockIdx.x * kWave + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_col_total(const double *__restrict__ part, int nb,
                                                      double *__restrict__ tot) {
  __shared__ double lds[kBlock];
  const double t = block_order_sum(part, nb, lds);
  if (threadIdx.x < kWave) tot[threadIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_col_divide(double *__restrict__ X, long long n,
                                                       const double *__restrict__ tot) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    X[i] /= tot[i & (kWave - 1)];
}

unsigned stride_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 65536));
}

}  // namespace

}  // namespace tbsla_hip

using namespace tbsla_hip;

extern "C" int tbsla_hip_spmm(tbsla_hip_matrix_t m, const double *d_B, double *d_C, int k,
                              void *stream) {
  if (!m || !d_B || !d_C) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_CSR) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "SpMM needs a CSR block");
  if (k < 1) return fail(TBSLA_HIP_ERR_INVALID, "k must be at least 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long rows = m->ln_row;
  if (rows == 0) return TBSLA_HIP_SUCCESS;
  if (m->ln_col == 0 || !m->has_arrays || m->stored == 0)  // no products: C = 0
    return launch_fill(d_C, rows * (long long)k, 0.0, s);
  const int chunk = kMaxTiles * kWave;
  const dim3 grid(grid_for(rows * kWave), (unsigned)((k + chunk - 1) / chunk));
  const int tiles = k > chunk ? kMaxTiles : (k + kWave - 1) / kWave;
  switch (tiles) {
    case 1: k_csr_spmm<1><<<grid, kBlock, 0, s>>>(rows, k, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values, d_B, d_C); break;
    case 2: k_csr_spmm<2><<<grid, kBlock, 0, s>>>(rows, k, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values, d_B, d_C); break;
    case 3: k_csr_spmm<3><<<grid, kBlock, 0, s>>>(rows, k, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values, d_B, d_C); break;
    default: k_csr_spmm<4><<<grid, kBlock, 0, s>>>(rows, k, m->f_col, m->ln_col, m->rowptr, m->colidx, m->values, d_B, d_C); break;
  }
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_personalized_page_rank(tbsla_hip_matrix_t m, double beta, double epsilon,
                                                int max_iterations, const int *h_sources,
                                                double *d_X, int *nb_iterations,
                                                double *max_error, void *stream) {
  if (!m || !h_sources || !d_X) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  if (m->format != TBSLA_HIP_CSR)
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "personalised PageRank needs a CSR block");
  if (m->ln_row != m->ln_col || m->f_row != m->f_col)
    return fail(TBSLA_HIP_ERR_INVALID, "personalised PageRank needs a square diagonal block");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = m->ln_row;
  if (n < 1) return fail(TBSLA_HIP_ERR_INVALID, "empty block");
  if (!m->rowptr) return fail(TBSLA_HIP_ERR_INVALID, "the block has no row pointer");
  if (m->stored > INT_MAX) return fail(TBSLA_HIP_ERR_RANGE, "more than 2^31 - 1 entries");
  // tile workgroups: three per CU (48 KB of LDS each, 12 wavefronts) on 256
  // CUs, each taking a contiguous run of tiles; fewer when the block has
  // fewer tiles
  const long long ntiles = (n + kPprTile - 1) / kPprTile;
  const unsigned nb = (unsigned)std::min<long long>(768, ntiles);
  // partials of the column sums at the end: the row-slot grid of k_col_partial
  const unsigned nbc = n >= (long long)kPprSlots * kPprGrid
                           ? (unsigned)kPprGrid
                           : (unsigned)std::max<long long>(1, (n + kPprSlots - 1) / kPprSlots);
  const unsigned nparts = std::max(nb, nbc);
  double *buf = nullptr, *part = nullptr, *S = nullptr;
  int *src = nullptr;
  unsigned int *bins = nullptr;
  PprState *st = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    (void)hipFree(part);
    (void)hipFree(S);
    (void)hipFree(src);
    (void)hipFree(st);
    (void)hipFree(bins);
  };
  // the first failing HIP call of the set-up and iteration loop
  hipError_t herr = hipSuccess;
  auto note = [&](hipError_t e) {
    if (herr == hipSuccess && e != hipSuccess) herr = e;
  };
  auto failed = [&](const char *where) {
    cleanup();
    return fail(TBSLA_HIP_ERR_HIP, std::string("personalised PageRank, ") + where + ": " +
                                       hipGetErrorString(herr));
  };
  int rc;
  if ((rc = device_alloc((void **)&buf, (size_t)n * kWave * 8)) ||
      (rc = device_alloc((void **)&part, (size_t)2 * nparts * kWave * 8)) ||
      (rc = device_alloc((void **)&S, kWave * 8)) || (rc = device_alloc((void **)&src, kWave * 4)) ||
      (rc = device_alloc((void **)&st, sizeof(PprState))) ||
      (rc = device_alloc((void **)&bins, kPprHist * sizeof(unsigned int)))) {
    cleanup();
    return rc;
  }
  // where the band lies: the offsets col - row of every 16th row; the window
  // covers the kPprReach + 1 consecutive offsets holding the most entries
  std::vector<unsigned int> hb(kPprHist, 0);
  note(hipMemsetAsync(bins, 0, kPprHist * sizeof(unsigned int), s));
  if (m->stored > 0)
    k_ppr_offsets<<<grid_for((n + 15) / 16), kBlock, 0, s>>>(n, m->f_col, m->rowptr, m->colidx,
                                                             bins);
  note(hipGetLastError());
  note(hipMemcpyAsync(hb.data(), bins, kPprHist * sizeof(unsigned int), hipMemcpyDeviceToHost, s));
  note(hipStreamSynchronize(s));
  if (herr != hipSuccess) return failed("band histogram");
  int dlo = 0;
  unsigned long long best = 0;
  for (int lo = 0; lo + kPprReach < kPprHist; lo++) {
    unsigned long long c = 0;
    for (int o = lo; o <= lo + kPprReach; o++) c += hb[o];
    if (c > best) {
      best = c;
      dlo = lo - kPprHist / 2;
    }
  }
  std::vector<double> hS(kWave, (double)n);  // X starts at 1: column sums n
  PprState hst{0, 0, 0.0};
  note(hipMemcpyAsync(S, hS.data(), kWave * 8, hipMemcpyHostToDevice, s));
  note(hipMemcpyAsync(src, h_sources, kWave * 4, hipMemcpyHostToDevice, s));
  note(hipMemcpyAsync(st, &hst, sizeof(hst), hipMemcpyHostToDevice, s));
  k_fill_d<<<stride_grid(n * kWave), kBlock, 0, s>>>(d_X, n * kWave, 1.0);
  note(hipGetLastError());
  if (herr != hipSuccess) return failed("set-up");
  double *Xb[2] = {d_X, buf};
  PprArgs pa{};
  pa.n = (int)n;
  pa.f_col = m->f_col;
  pa.rowptr = m->rowptr;
  // the kernel reads entry min(j, last) whatever the row's length: with no
  // entries stored it reads a dummy one (src[0], S[0]; never summed)
  const bool any = m->stored > 0 && m->colidx && m->values;
  pa.colidx = any ? m->colidx : src;
  pa.vals = any ? m->values : S;
  pa.last = any ? (int)(m->stored - 1) : 0;
  pa.src = src;
  pa.S = S;
  pa.beta = beta;
  pa.st = st;
  pa.p_err = part;
  pa.p_sum = part + (size_t)nb * kWave;
  pa.dlo = dlo;
  pa.old_lds = dlo <= 0 && dlo + kPprWin >= kPprTile;
  pa.ntiles = (int)ntiles;
  int launched = 0;
  while (launched < max_iterations) {
    const int batch = std::min(16, max_iterations - launched);
    for (int i = 0; i < batch; i++, launched++) {
      pa.X = Xb[launched % 2];
      pa.Xn = Xb[(launched + 1) % 2];
      // old values from the ring (three remote rows ahead per row), or from
      // X (two, for the registers the old values take)
      if (pa.old_lds) k_ppr_quad<4, 2, 1, 3, true><<<nb, kPprBlock, 0, s>>>(pa);
      else k_ppr_quad<4, 2, 1, 2, false><<<nb, kPprBlock, 0, s>>>(pa);
      k_ppr_finalize<<<1, kBlock, 0, s>>>(st, part, part + (size_t)nb * kWave, (int)nb, S, epsilon);
    }
    note(hipGetLastError());
    note(hipMemcpyAsync(&hst, st, sizeof(hst), hipMemcpyDeviceToHost, s));
    note(hipStreamSynchronize(s));
    if (herr != hipSuccess) return failed("iterations");
    if (hst.done) break;
  }
  double *X = Xb[hst.iters % 2];
  k_col_partial<<<nbc, kBlock, 0, s>>>(X, n, part);
  k_col_total<<<1, kBlock, 0, s>>>(part, (int)nbc, S);
  k_col_divide<<<stride_grid(n * kWave), kBlock, 0, s>>>(X, n * kWave, S);
  note(hipGetLastError());
  if (X != d_X) note(hipMemcpyAsync(d_X, X, (size_t)n * kWave * 8, hipMemcpyDeviceToDevice, s));
  note(hipStreamSynchronize(s));
  if (herr != hipSuccess) return failed("normalisation");
  cleanup();
  if (nb_iterations) *nb_iterations = hst.iters;
  if (max_error) *max_error = hst.max_err;
  return TBSLA_HIP_SUCCESS;
}
