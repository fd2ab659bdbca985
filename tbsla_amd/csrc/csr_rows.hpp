// csr_rows.hpp -- the CSR row kernel (MatrixCSR::Ax, MatrixCSR.cpp:209-224)
// shared by the plain row layout and by the near-diagonal part of the split
// layout.
//
// A wave holds kWave / L lane groups; group g takes R rows (base + r G + g).
// Each lane reads V consecutive entries per trip (one 4 V-byte load of
// column indices and one 8 V-byte load of values), starting from the row's
// first entry rounded down to a multiple of V, so the vector loads stay
// aligned; entries of the previous row that such a load covers are
// multiplied by zero.  Lanes add their entries in trip order and the group
// adds its lanes by a fixed xor tree, so results are deterministic and
// within rounding of the sequential row sum.
//
// NT selects the nontemporal load (the line is not kept in L2 for later
// use): it pays when every line is read whole by one instruction, and
// costs a second fetch when a row's lines are split over trips.  The
// autotuner picks V, L, R and NT per matrix.
#pragma once

#include "common.hpp"

namespace tbsla_hip {

struct CsrRowsArgs {
  long long n_rows;
  const int *rowptr;
  const int *colidx;
  const double *vals;
  const double *x;
  int f_col;
  double *y;
  const int *done;
  // split layout: y[row] = the row's sum + add[row] (its remote products)
  const double *add;
  // optional row segments (instances of a batch): rows [seg_offs[k],
  // seg_offs[k+1]); a wavefront whose rows all lie in one segment with
  // seg_done[k] != 0 returns at once (its rows of y are left as they are)
  const int *seg_offs = nullptr;
  int nseg = 0;
  const int *seg_done = nullptr;
};

// the segment holding row: the largest k with seg_offs[k] <= row
__device__ inline int seg_of_row(const int *offs, int nseg, long long row) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (offs[mid] <= row) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int V, int NT>
__device__ inline void csr_load(const int *ci, const double *va, int j, int (&c)[V],
                                double (&v)[V]) {
  typedef int vi __attribute__((ext_vector_type(V)));
  typedef double vd __attribute__((ext_vector_type(V)));
  if (V == 1) {
    c[0] = NT ? __builtin_nontemporal_load(ci + j) : ci[j];
    v[0] = NT ? __builtin_nontemporal_load(va + j) : va[j];
  } else {
    const vi cc = NT ? __builtin_nontemporal_load(reinterpret_cast<const vi *>(ci + j))
                     : *reinterpret_cast<const vi *>(ci + j);
    const vd vv = NT ? __builtin_nontemporal_load(reinterpret_cast<const vd *>(va + j))
                     : *reinterpret_cast<const vd *>(va + j);
#pragma unroll
    for (int k = 0; k < V; k++) {
      c[k] = cc[k];
      v[k] = vv[k];
    }
  }
}

template <int V, int L, int R, int NT, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_csr_rows(CsrRowsArgs a) {
  if (a.done && *a.done) return;
  constexpr int G = kWave / L;
  const int lane = threadIdx.x & (L - 1);
  const int g = (threadIdx.x & (kWave - 1)) / L;
  const long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) / kWave;
  const long long base = wave * (G * R);
  if (base >= a.n_rows) return;
  if (a.seg_offs) {  // wave-uniform
    const long long last = min(base + G * R, a.n_rows) - 1;
    const int k0 = seg_of_row(a.seg_offs, a.nseg, base);
    if (a.seg_done[k0] && (last < a.seg_offs[k0 + 1])) return;
  }
  int beg[R], end[R];
  int len = 0;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const long long row = base + r * G + g;
    if (row < a.n_rows) {
      beg[r] = a.rowptr[row];
      end[r] = a.rowptr[row + 1];
    } else {
      beg[r] = end[r] = 0;
    }
    len = max(len, end[r] - (beg[r] & ~(V - 1)));
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off, kWave));
  const int trips = (len + V * L - 1) / (V * L);
  double sum[R];
#pragma unroll
  for (int r = 0; r < R; r++) sum[r] = 0.0;
  for (int t = 0; t < trips; t++) {
    int c[R][V];
    double v[R][V];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int j = (beg[r] & ~(V - 1)) + V * (t * L + lane);
      if (j + V <= end[r]) {
        csr_load<V, NT>(a.colidx, a.vals, j, c[r], v[r]);
#pragma unroll
        for (int k = 0; k < V; k++) {
          c[r][k] -= a.f_col;
          if (V > 1 && j + k < beg[r]) v[r][k] = 0.0;
        }
      } else {
#pragma unroll
        for (int k = 0; k < V; k++) {
          const bool in = j + k < end[r] && j + k >= beg[r];
          c[r][k] = in ? a.colidx[j + k] - a.f_col : 0;
          v[r][k] = in ? a.vals[j + k] : 0.0;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int k = 0; k < V; k++) sum[r] = fma(v[r][k], a.x[c[r][k]], sum[r]);
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const double s = group_sum<L>(sum[r]);
    const long long row = base + r * G + g;
    if (lane == 0 && row < a.n_rows) a.y[row] = SPLIT ? s + a.add[row] : s;
  }
}

// Schedules with a compiled kernel: V = 1 with L = 1..64 and R = 1, 2, 4;
// V = 2 with L = 4..32 and V = 4 with L = 4..16, both with R = 1, 2.
inline bool csr_rows_valid(int V, int L, int R, int NT) {
  if (NT != 0 && NT != 1) return false;
  if (L < 1 || L > 64 || (L & (L - 1))) return false;
  if (V == 1) return R == 1 || R == 2 || R == 4;
  if (R != 1 && R != 2) return false;
  if (V == 2) return L >= 4 && L <= 32;
  if (V == 4) return L >= 4 && L <= 16;
  return false;
}

template <bool SPLIT, int V, int L, int R, int NT>
void csr_rows_go(const CsrRowsArgs &a, hipStream_t s) {
  const long long rows_per_wave = (long long)(kWave / L) * R;
  const long long waves = (a.n_rows + rows_per_wave - 1) / rows_per_wave;
  k_csr_rows<V, L, R, NT, SPLIT><<<grid_for(waves * kWave), kBlock, 0, s>>>(a);
}

template <bool SPLIT, int V, int L, int R>
void csr_rows_nt(const CsrRowsArgs &a, int NT, hipStream_t s) {
  if (NT) csr_rows_go<SPLIT, V, L, R, 1>(a, s);
  else csr_rows_go<SPLIT, V, L, R, 0>(a, s);
}

template <bool SPLIT, int V, int L>
void csr_rows_r(const CsrRowsArgs &a, int R, int NT, hipStream_t s) {
  if (R == 1) csr_rows_nt<SPLIT, V, L, 1>(a, NT, s);
  else if (V > 1 || R == 2) csr_rows_nt<SPLIT, V, L, 2>(a, NT, s);
  else csr_rows_nt<SPLIT, V, L, (V == 1 ? 4 : 2)>(a, NT, s);
}

// Launches the kernel for a schedule that csr_rows_valid accepts.
template <bool SPLIT>
void csr_rows_launch(const CsrRowsArgs &a, int V, int L, int R, int NT, hipStream_t s) {
  if (V == 1) {
    switch (L) {
      case 1: csr_rows_r<SPLIT, 1, 1>(a, R, NT, s); break;
      case 2: csr_rows_r<SPLIT, 1, 2>(a, R, NT, s); break;
      case 4: csr_rows_r<SPLIT, 1, 4>(a, R, NT, s); break;
      case 8: csr_rows_r<SPLIT, 1, 8>(a, R, NT, s); break;
      case 16: csr_rows_r<SPLIT, 1, 16>(a, R, NT, s); break;
      case 32: csr_rows_r<SPLIT, 1, 32>(a, R, NT, s); break;
      default: csr_rows_r<SPLIT, 1, 64>(a, R, NT, s); break;
    }
  } else if (V == 2) {
    switch (L) {
      case 4: csr_rows_r<SPLIT, 2, 4>(a, R, NT, s); break;
      case 8: csr_rows_r<SPLIT, 2, 8>(a, R, NT, s); break;
      case 16: csr_rows_r<SPLIT, 2, 16>(a, R, NT, s); break;
      default: csr_rows_r<SPLIT, 2, 32>(a, R, NT, s); break;
    }
  } else {
    switch (L) {
      case 4: csr_rows_r<SPLIT, 4, 4>(a, R, NT, s); break;
      case 8: csr_rows_r<SPLIT, 4, 8>(a, R, NT, s); break;
      default: csr_rows_r<SPLIT, 4, 16>(a, R, NT, s); break;
    }
  }
}

}  // namespace tbsla_hip
