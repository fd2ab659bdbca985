// precond_net.hip -- the learned factored preconditioner: a 3-layer
// message-passing network (width 64) that maps an SPD CSR matrix A to a
// lower-triangular G on tril(A)'s pattern; CG then uses M = G G^T
// (tbsla_hip_cg with TBSLA_HIP_PRECOND_FACTORED).  No counterpart in the
// reference, whose CG (src/mpi/matrix/CG.cpp:32-73) is unpreconditioned.
//
// With s_i = A_ii^-1/2 and A^ = S A S (unit diagonal), edges e_ij = A^_ij:
//   x_i  = [1, sum_j |e_ij|, max_j |e_ij|, deg_i / 8]
//   h1   = relu([x, E x] W0 + b0)                              (8 -> 64, FP32)
//   U    = h_l W_l (BF16 MFMA, FP32 accumulation), l = 1, 2
//   h_l+1= relu(U[:, :64] + E U[:, 64:] + b_l)
//   P    = h3 V (BF16 MFMA)
//   g^_ij= wo . relu(P_i[:64] + P_j[64:] + e_ij ve + c) + bo     (j < i)
//   g^_ii= exp(wd . h3_i + bd)
//   G    = S G^  (so G^T A G = G^^T A^ G^; zero heads give G^ = I: Jacobi)
// Training minimises |Z - G^^T A^ G^ Z|_F^2 / (n k) over Rademacher probes Z
// with hand-written backward kernels and Adam.  Every reduction writes
// per-workgroup partials that one kernel sums in a fixed order, so a step is
// deterministic.
//
// Layouts: activations row-major n x 64 (one wavefront per node, lane =
// channel); GEMM operands BF16 row-major; weights kept FP32 (master) with
// BF16 copies in both orientations.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

constexpr int H = 64;       // hidden width
constexpr int F = 4;        // node features
constexpr int H2 = 2 * H;   // transform width (self | neighbour)

// parameter layout (FP32), in this order
constexpr int oW0 = 0;                   // (2F) x H
constexpr int ob0 = oW0 + 2 * F * H;     // H
constexpr int oW1 = ob0 + H;             // H x 2H
constexpr int ob1 = oW1 + H * H2;        // H
constexpr int oW2 = ob1 + H;             // H x 2H
constexpr int ob2 = oW2 + H * H2;        // H
constexpr int oV = ob2 + H;              // H x 2H
constexpr int ove = oV + H * H2;         // H
constexpr int oc = ove + H;              // H
constexpr int owo = oc + H;              // H
constexpr int obo = owo + H;             // 1
constexpr int owd = obo + 1;             // H
constexpr int obd = owd + H;             // 1
constexpr int kParams = obd + 1;
static_assert(ob1 % 4 == 0 && ob2 % 4 == 0 && ove % 4 == 0 && oc % 4 == 0 && owo % 4 == 0,
              "offsets loaded as 16-byte vectors");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;  // raw BF16 storage

__device__ __host__ inline bf16_t to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u = u + 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (bf16_t)(u >> 16);
}

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

template <int W>
__device__ inline float group_sumf(float v) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
  return v;
}

// ---------------------------------------------------------------------------
// graph preparation
// ---------------------------------------------------------------------------

// diag, s = diag^-1/2, off-diagonal degree, lower count (columns < row)
__global__ __launch_bounds__(kBlock) void k_prep_rows(long long n, int f_col,
                                                      const int *__restrict__ rp,
                                                      const int *__restrict__ ci,
                                                      const double *__restrict__ v,
                                                      double *__restrict__ s, int *__restrict__ deg,
                                                      int *__restrict__ nlow, int *__restrict__ gcnt,
                                                      int *__restrict__ tcnt) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    double d = 0.0;
    int dg = 0, lo = 0;
    for (int j = rp[i]; j < rp[i + 1]; j++) {
      const long long c = ci[j] - f_col;
      if (c == i) {
        d += v[j];
      } else {
        dg++;
        lo += c < i;
      }
    }
    s[i] = d > 0.0 ? 1.0 / sqrt(d) : 1.0;
    deg[i] = dg;
    nlow[i] = lo;
    gcnt[i] = lo + 1;
    tcnt[i] = dg - lo + 1;
  }
}

// edges (off-diagonal, column order), features, G and G^T column indices
__global__ __launch_bounds__(kBlock) void k_prep_fill(
    long long n, int f_col, const int *__restrict__ rp, const int *__restrict__ ci,
    const double *__restrict__ v, const double *__restrict__ s, const int *__restrict__ erp,
    const int *__restrict__ nlow, const int *__restrict__ grp, const int *__restrict__ trp,
    int *__restrict__ eci, float *__restrict__ ev, float *__restrict__ x,
    int *__restrict__ gci, int *__restrict__ tci) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    int pe = erp[i];
    float sa = 0.f, mx = 0.f;
    for (int j = rp[i]; j < rp[i + 1]; j++) {
      const long long c = ci[j] - f_col;
      if (c == i) continue;
      const float e = (float)(v[j] * (s[i] * s[c]));  // same bits for (i,j) and (j,i)
      eci[pe] = (int)c;
      ev[pe] = e;
      pe++;
      sa += fabsf(e);
      mx = fmaxf(mx, fabsf(e));
    }
    const int dg = pe - erp[i];
    x[i * F + 0] = 1.f;
    x[i * F + 1] = sa;
    x[i * F + 2] = mx;
    x[i * F + 3] = (float)dg * 0.125f;
    // G row i: lower columns, then i; G^T row i: i, then upper columns
    const int lo = nlow[i];
    int pg = grp[i], pt = trp[i];
    for (int k = 0; k < lo; k++) gci[pg++] = eci[erp[i] + k];
    gci[pg] = (int)i;
    tci[pt++] = (int)i;
    for (int k = lo; k < dg; k++) tci[pt++] = eci[erp[i] + k];
  }
}

// position in G's values of every G^T entry: (i, j >= i) -> G entry (j, i)
__global__ __launch_bounds__(kBlock) void k_prep_tmap(long long n, const int *__restrict__ grp,
                                                      const int *__restrict__ gci,
                                                      const int *__restrict__ nlow,
                                                      const int *__restrict__ trp,
                                                      const int *__restrict__ tci,
                                                      int *__restrict__ tmap) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int t0 = trp[i];
    tmap[t0] = grp[i] + nlow[i];  // diagonal
    for (int t = t0 + 1; t < trp[i + 1]; t++) {
      const int j = tci[t];
      // column i among the (sorted) lower columns of G row j
      int lo = grp[j], hi = grp[j] + nlow[j];
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (gci[mid] < i) lo = mid + 1;
        else hi = mid;
      }
      tmap[t] = lo;
    }
  }
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------

// Layer 0: a workgroup takes 64 nodes.  Threads 0..63 form each node's
// 8 inputs [x_i, sum_j e_ij x_j] (into LDS and z0); then every thread
// computes 16 of the 64 x 64 outputs, channel = thread % 64, with its W0
// column in registers (pre = b0 + sum_k z_k W0[k][c], k in order).
__global__ __launch_bounds__(kBlock) void k_layer0(long long n, const int *__restrict__ erp,
                                                   const int *__restrict__ eci,
                                                   const float *__restrict__ ev,
                                                   const float *__restrict__ x,
                                                   const float *__restrict__ th,
                                                   float *__restrict__ z0, float *__restrict__ h,
                                                   bf16_t *__restrict__ hb) {
  __shared__ float zs[64][2 * F];
  const long long n0 = (long long)blockIdx.x * 64;
  if (threadIdx.x < 64) {
    const long long i = n0 + threadIdx.x;
    float z[2 * F];
#pragma unroll
    for (int k = 0; k < 2 * F; k++) z[k] = 0.f;
    if (i < n) {
#pragma unroll
      for (int k = 0; k < F; k++) z[k] = x[i * F + k];
      for (int j = erp[i]; j < erp[i + 1]; j++) {
        const float e = ev[j];
        const long long c = eci[j];
#pragma unroll
        for (int k = 0; k < F; k++) z[F + k] = fmaf(e, x[c * F + k], z[F + k]);
      }
#pragma unroll
      for (int k = 0; k < 2 * F; k++) z0[i * 2 * F + k] = z[k];
    }
#pragma unroll
    for (int k = 0; k < 2 * F; k++) zs[threadIdx.x][k] = z[k];
  }
  __syncthreads();
  const int c = threadIdx.x & (kWave - 1);
  float w[2 * F];
#pragma unroll
  for (int k = 0; k < 2 * F; k++) w[k] = th[oW0 + k * H + c];
  const float b0 = th[ob0 + c];
  for (int nd = threadIdx.x / kWave; nd < 64; nd += kBlock / kWave) {
    const long long i = n0 + nd;
    if (i >= n) break;
    float pre = b0;
#pragma unroll
    for (int k = 0; k < 2 * F; k++) pre = fmaf(zs[nd][k], w[k], pre);
    const float hv = fmaxf(pre, 0.f);
    h[i * H + c] = hv;
    hb[i * H + c] = to_bf16(hv);
  }
}

typedef float floatq __attribute__((ext_vector_type(4)));

// h' = relu(U[:, :64] + E U[:, 64:] + b); the last layer also gives the
// diagonal head g^_ii = exp(wd . h + bd).  Four nodes per wavefront, 16
// lanes per node, four channels per lane (16-byte loads).
__global__ __launch_bounds__(kBlock) void k_aggregate(long long n, const int *__restrict__ erp,
                                                      const int *__restrict__ eci,
                                                      const float *__restrict__ ev,
                                                      const float *__restrict__ U,
                                                      const float *__restrict__ th, int ob,
                                                      float *__restrict__ h, bf16_t *__restrict__ hb,
                                                      float *__restrict__ gdiag) {
  const int q = threadIdx.x & 15;  // channels 4q .. 4q+3
  const long long groups = (long long)gridDim.x * (kBlock / 16);
  const floatq bias = *reinterpret_cast<const floatq *>(th + ob + 4 * q);
  // wd starts at an offset that is not a multiple of four floats: scalar loads
  const floatq wd = {th[owd + 4 * q], th[owd + 4 * q + 1], th[owd + 4 * q + 2], th[owd + 4 * q + 3]};
  const float bd = th[obd];
  for (long long i = (long long)blockIdx.x * (kBlock / 16) + threadIdx.x / 16; i < n; i += groups) {
    floatq pre = *reinterpret_cast<const floatq *>(U + i * H2 + 4 * q) + bias;
    for (int j = erp[i]; j < erp[i + 1]; j++) {
      const float e = ev[j];
      const floatq u = *reinterpret_cast<const floatq *>(U + (long long)eci[j] * H2 + H + 4 * q);
      pre.x = fmaf(e, u.x, pre.x);
      pre.y = fmaf(e, u.y, pre.y);
      pre.z = fmaf(e, u.z, pre.z);
      pre.w = fmaf(e, u.w, pre.w);
    }
    floatq hv;
    hv.x = fmaxf(pre.x, 0.f);
    hv.y = fmaxf(pre.y, 0.f);
    hv.z = fmaxf(pre.z, 0.f);
    hv.w = fmaxf(pre.w, 0.f);
    *reinterpret_cast<floatq *>(h + i * H + 4 * q) = hv;
    const uint32_t lo = (uint32_t)to_bf16(hv.x) | ((uint32_t)to_bf16(hv.y) << 16);
    const uint32_t hi = (uint32_t)to_bf16(hv.z) | ((uint32_t)to_bf16(hv.w) << 16);
    *reinterpret_cast<uint2 *>(hb + i * H + 4 * q) = make_uint2(lo, hi);
    if (gdiag) {
      float d = wd.x * hv.x + wd.y * hv.y + wd.z * hv.z + wd.w * hv.w;
      d = group_sumf<16>(d);
      if (q == 0) gdiag[i] = expf(d + bd);
    }
  }
}

// C[n x N] (+)= A[n x K] B[K x N] with A row-major BF16 and B given as Bt
// (N x K, row-major BF16), FP32 result.  Workgroup: 64 rows, 4 waves of 16
// rows; each wave holds N/16 16x16 tiles (v_mfma_f32_16x16x32_bf16:
// lane l holds A[l%16][8(l/16)..+8], Bt[l%16][8(l/16)..+8] and
// D[4(l/16)..+4][l%16]).  Epilogue: add the existing C, then zero where
// mask <= 0 (the ReLU derivative).
template <int K, int N>
__global__ __launch_bounds__(kBlock) void k_gemm_rows(long long n, const bf16_t *__restrict__ A,
                                                      const bf16_t *__restrict__ Bt,
                                                      float *__restrict__ C, int add,
                                                      const float *__restrict__ mask) {
  constexpr int NT = N / 16;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const long long r0 = (long long)blockIdx.x * 64 + wave * 16;
  if (r0 >= n) return;
  floatx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  const long long ar = r0 + (lane & 15);
  const int kq = (lane >> 4) * 8;
#pragma unroll
  for (int kb = 0; kb < K; kb += 32) {
    bf16x8 a;
    if (ar < n) {
      a = *reinterpret_cast<const bf16x8 *>(A + ar * K + kb + kq);
    } else {
      a = bf16x8{};
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const bf16x8 b = *reinterpret_cast<const bf16x8 *>(Bt + (long long)(t * 16 + (lane & 15)) * K + kb + kq);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const long long r = r0 + (lane >> 4) * 4 + q;
      if (r >= n) continue;
      const long long o = r * N + t * 16 + (lane & 15);
      float v = acc[t][q];
      if (add) v += C[o];
      if (mask && !(mask[o] > 0.f)) v = 0.f;
      C[o] = v;
    }
  }
}

// Split-K weight gradient: part[blk] = A[rows]^T B[rows] (M = 64, N = 128)
// over the workgroup's chunk of rows, staged through LDS (transposed) in
// steps of 32 rows.  Each wave owns 16 rows of the 64 x 128 result.
constexpr int kWgRows = 512;

__global__ __launch_bounds__(kBlock) void k_gemm_wgrad(long long n, const bf16_t *__restrict__ A,
                                                       const bf16_t *__restrict__ B,
                                                       float *__restrict__ part) {
  __shared__ bf16_t At[H][32 + 8];   // At[m][k]
  __shared__ bf16_t Bs[H2][32 + 8];  // Bs[nn][k]
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const long long row0 = (long long)blockIdx.x * kWgRows;
  floatx4 acc[8];
#pragma unroll
  for (int t = 0; t < 8; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < kWgRows; ks += 32) {
    // stage 32 rows: A (32 x 64) and B (32 x 128), transposed
    for (int e = threadIdx.x; e < 32 * H; e += kBlock) {
      const int k = e / H, m = e % H;
      const long long r = row0 + ks + k;
      At[m][k] = r < n ? A[r * H + m] : (bf16_t)0;
    }
    for (int e = threadIdx.x; e < 32 * H2; e += kBlock) {
      const int k = e / H2, c = e % H2;
      const long long r = row0 + ks + k;
      Bs[c][k] = r < n ? B[r * H2 + c] : (bf16_t)0;
    }
    __syncthreads();
    const int kq = (lane >> 4) * 8;
    bf16x8 a;
#pragma unroll
    for (int q = 0; q < 8; q++) a[q] = __builtin_bit_cast(__bf16, At[wave * 16 + (lane & 15)][kq + q]);
#pragma unroll
    for (int t = 0; t < 8; t++) {
      bf16x8 b;
#pragma unroll
      for (int q = 0; q < 8; q++) b[q] = __builtin_bit_cast(__bf16, Bs[t * 16 + (lane & 15)][kq + q]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
    __syncthreads();
  }
  float *out = part + (long long)blockIdx.x * H * H2;
#pragma unroll
  for (int t = 0; t < 8; t++)
#pragma unroll
    for (int q = 0; q < 4; q++)
      out[(wave * 16 + (lane >> 4) * 4 + q) * H2 + t * 16 + (lane & 15)] = acc[t][q];
}

// edge head: four rows per wavefront, 16 lanes per row, four channels per
// lane; the row's lower edges one at a time
__global__ __launch_bounds__(kBlock) void k_edge_head(long long n, const int *__restrict__ erp,
                                                      const int *__restrict__ eci,
                                                      const float *__restrict__ ev,
                                                      const int *__restrict__ nlow,
                                                      const int *__restrict__ grp,
                                                      const float *__restrict__ P,
                                                      const float *__restrict__ th,
                                                      const float *__restrict__ gdiag,
                                                      const double *__restrict__ s,
                                                      float *__restrict__ ghat,
                                                      double *__restrict__ gval) {
  const int q = threadIdx.x & 15;
  const long long groups = (long long)gridDim.x * (kBlock / 16);
  const floatq ve = *reinterpret_cast<const floatq *>(th + ove + 4 * q);
  const floatq cc = *reinterpret_cast<const floatq *>(th + oc + 4 * q);
  const floatq wo = *reinterpret_cast<const floatq *>(th + owo + 4 * q);
  const float bo = th[obo];
  for (long long i = (long long)blockIdx.x * (kBlock / 16) + threadIdx.x / 16; i < n; i += groups) {
    const floatq pi = *reinterpret_cast<const floatq *>(P + i * H2 + 4 * q);
    const int lo = nlow[i], g0 = grp[i];
    for (int k = 0; k < lo; k++) {
      const int j = eci[erp[i] + k];
      const float e = ev[erp[i] + k];
      const floatq pj = *reinterpret_cast<const floatq *>(P + (long long)j * H2 + H + 4 * q);
      const floatq sv = pi + pj + e * ve + cc;
      float g = wo.x * fmaxf(sv.x, 0.f) + wo.y * fmaxf(sv.y, 0.f) + wo.z * fmaxf(sv.z, 0.f) +
                wo.w * fmaxf(sv.w, 0.f);
      g = group_sumf<16>(g) + bo;
      if (q == 0) {
        ghat[g0 + k] = g;
        if (gval) gval[g0 + k] = s[i] * (double)g;
      }
    }
    if (q == 0) {
      ghat[g0 + lo] = gdiag[i];
      if (gval) gval[g0 + lo] = s[i] * (double)gdiag[i];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_gather_t(long long m, const int *__restrict__ tmap,
                                                     const double *__restrict__ gval,
                                                     double *__restrict__ tval) {
  for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < m;
       t += (long long)gridDim.x * kBlock)
    tval[t] = gval[tmap[t]];
}

// ---------------------------------------------------------------------------
// loss |Z - G^^T A^ G^ Z|^2 / (n k) and dL/dg^
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_probes(long long nk, uint64_t seed_hash,
                                                   float *__restrict__ Z) {
  for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < nk;
       t += (long long)gridDim.x * kBlock)
    Z[t] = unit_value(seed_hash, (uint64_t)t) < 0.5 ? -1.f : 1.f;
}

// out[i, p] = sum over G row i of g^ in[col, p]   (KP lanes per row)
template <int KP>
__global__ __launch_bounds__(kBlock) void k_spmm_g(long long n, const int *__restrict__ rp,
                                                   const int *__restrict__ ci,
                                                   const float *__restrict__ val,
                                                   const int *__restrict__ vmap,
                                                   const float *__restrict__ in,
                                                   float *__restrict__ out) {
  const int p = threadIdx.x % KP;
  for (long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) / KP; i < n;
       i += (long long)gridDim.x * (kBlock / KP)) {
    float acc = 0.f;
    for (int j = rp[i]; j < rp[i + 1]; j++) {
      const float g = vmap ? val[vmap[j]] : val[j];
      acc = fmaf(g, in[(long long)ci[j] * KP + p], acc);
    }
    out[i * KP + p] = acc;
  }
}

// out = A^ in = in + E in
template <int KP>
__global__ __launch_bounds__(kBlock) void k_spmm_a(long long n, const int *__restrict__ erp,
                                                   const int *__restrict__ eci,
                                                   const float *__restrict__ ev,
                                                   const float *__restrict__ in,
                                                   float *__restrict__ out) {
  const int p = threadIdx.x % KP;
  for (long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) / KP; i < n;
       i += (long long)gridDim.x * (kBlock / KP)) {
    float acc = in[i * KP + p];
    for (int j = erp[i]; j < erp[i + 1]; j++) acc = fmaf(ev[j], in[(long long)eci[j] * KP + p], acc);
    out[i * KP + p] = acc;
  }
}

// R = Z - W; per-workgroup partial of sum R^2
__global__ __launch_bounds__(kBlock) void k_residual_probe(long long nk, const float *__restrict__ Z,
                                                           const float *__restrict__ W,
                                                           float *__restrict__ R,
                                                           double *__restrict__ part) {
  __shared__ double lds[kBlock / kWave];
  double acc = 0.0;
  for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < nk;
       t += (long long)gridDim.x * kBlock) {
    const float r = Z[t] - W[t];
    R[t] = r;
    acc += (double)r * r;
  }
  acc = group_sum<kWave>(acc);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; w++) t += lds[w];
    part[blockIdx.x] = t;
  }
}

// dg^_ij = c sum_p (V_ip R_jp + Y_ip Z_jp) over the entries of G row i
template <int KP>
__global__ __launch_bounds__(kBlock) void k_dghat(long long n, const int *__restrict__ grp,
                                                  const int *__restrict__ gci,
                                                  const float *__restrict__ V,
                                                  const float *__restrict__ R,
                                                  const float *__restrict__ Y,
                                                  const float *__restrict__ Z, float c,
                                                  float *__restrict__ dg) {
  const int p = threadIdx.x % KP;
  for (long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) / KP; i < n;
       i += (long long)gridDim.x * (kBlock / KP)) {
    const float vi = V[i * KP + p], yi = Y[i * KP + p];
    for (int j = grp[i]; j < grp[i + 1]; j++) {
      const long long cj = gci[j];
      float t = vi * R[cj * KP + p] + yi * Z[cj * KP + p];
      t = group_sumf<KP>(t);
      if (p == 0) dg[j] = c * t;
    }
  }
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------

// Head, row pass (wave per row i): diagonal head, and the lower edges (i, j)
// for dP_i[:64]; parameter partials for wd, bd, wo, bo, ve, c per workgroup.
// Partial layout per workgroup: [wd 64 | bd | wo 64 | bo | ve 64 | c 64].
constexpr int kHeadPart = 4 * H + 2;

__global__ __launch_bounds__(kBlock) void k_head_bwd_rows(
    long long n, const int *__restrict__ erp, const int *__restrict__ eci,
    const float *__restrict__ ev, const int *__restrict__ nlow, const int *__restrict__ grp,
    const float *__restrict__ P, const float *__restrict__ h3, const float *__restrict__ th,
    const float *__restrict__ ghat, const float *__restrict__ dg, float *__restrict__ dh3,
    float *__restrict__ dP, bf16_t *__restrict__ dPb, float *__restrict__ part) {
  __shared__ float lds[kBlock / kWave][kHeadPart];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const long long waves = (long long)gridDim.x * (kBlock / kWave);
  const float ve = th[ove + lane], cc = th[oc + lane], wo = th[owo + lane], wd = th[owd + lane];
  float a_wd = 0.f, a_bd = 0.f, a_wo = 0.f, a_bo = 0.f, a_ve = 0.f, a_c = 0.f;
  for (long long i = (long long)blockIdx.x * (kBlock / kWave) + wave; i < n; i += waves) {
    const int lo = nlow[i], g0 = grp[i];
    const float dq = dg[g0 + lo] * ghat[g0 + lo];
    a_wd = fmaf(dq, h3[i * H + lane], a_wd);
    a_bd += dq;
    dh3[i * H + lane] = dq * wd;
    const float pi = P[i * H2 + lane];
    float dpi = 0.f;
    for (int k = 0; k < lo; k++) {
      const int j = eci[erp[i] + k];
      const float e = ev[erp[i] + k];
      const float sv = pi + P[(long long)j * H2 + H + lane] + e * ve + cc;
      const float d = dg[g0 + k];
      a_wo = fmaf(d, fmaxf(sv, 0.f), a_wo);
      a_bo += d;
      const float ds = sv > 0.f ? d * wo : 0.f;
      a_ve = fmaf(ds, e, a_ve);
      a_c += ds;
      dpi += ds;
    }
    dP[i * H2 + lane] = dpi;
    dPb[i * H2 + lane] = to_bf16(dpi);
  }
  // workgroup partials (waves in order)
  lds[wave][lane] = a_wd;
  lds[wave][H + 1 + lane] = a_wo;
  lds[wave][2 * H + 2 + lane] = a_ve;
  lds[wave][3 * H + 2 + lane] = a_c;
  // a_bd and a_bo hold the same per-node / per-edge sums in every lane
  const float bd = __shfl(a_bd, 0, kWave), bo = __shfl(a_bo, 0, kWave);
  if (lane == 0) {
    lds[wave][H] = bd;
    lds[wave][2 * H + 1] = bo;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kHeadPart; e += kBlock) {
    float t = 0.f;
    for (int w = 0; w < kBlock / kWave; w++) t += lds[w][e];
    part[(long long)blockIdx.x * kHeadPart + e] = t;
  }
}

// Head, column pass (wave per row j): edges (i, j), i > j, are the upper
// entries of row j; they give dP_j[64:].
__global__ __launch_bounds__(kBlock) void k_head_bwd_cols(
    long long n, const int *__restrict__ trp, const int *__restrict__ tci,
    const int *__restrict__ tmap, const int *__restrict__ erp, const int *__restrict__ eci,
    const float *__restrict__ ev, const int *__restrict__ nlow, const float *__restrict__ P,
    const float *__restrict__ th, const float *__restrict__ dg, float *__restrict__ dP,
    bf16_t *__restrict__ dPb) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long waves = (long long)gridDim.x * (kBlock / kWave);
  const float ve = th[ove + lane], cc = th[oc + lane], wo = th[owo + lane];
  for (long long j = (long long)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; j < n;
       j += waves) {
    const float pj = P[j * H2 + H + lane];
    float dpj = 0.f;
    // G^T row j: the diagonal, then upper columns i (the same order as the
    // upper part of E's row j, whose values are e_ji = e_ij)
    const int t0 = trp[j];
    const int eu = erp[j] + nlow[j];
    for (int t = t0 + 1; t < trp[j + 1]; t++) {
      const long long i = tci[t];
      const float e = ev[eu + (t - t0 - 1)];
      const float sv = P[i * H2 + lane] + pj + e * ve + cc;
      const float d = dg[tmap[t]];
      dpj += sv > 0.f ? d * wo : 0.f;
    }
    dP[j * H2 + H + lane] = dpj;
    dPb[j * H2 + H + lane] = to_bf16(dpj);
  }
}

// dU = [dpre, E dpre] (BF16) and per-workgroup partials of db = sum dpre
__global__ __launch_bounds__(kBlock) void k_aggregate_bwd(long long n, const int *__restrict__ erp,
                                                          const int *__restrict__ eci,
                                                          const float *__restrict__ ev,
                                                          const float *__restrict__ dpre,
                                                          bf16_t *__restrict__ dUb,
                                                          float *__restrict__ part) {
  __shared__ float lds[kBlock / kWave][H];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const long long waves = (long long)gridDim.x * (kBlock / kWave);
  float ab = 0.f;
  for (long long i = (long long)blockIdx.x * (kBlock / kWave) + wave; i < n; i += waves) {
    const float d = dpre[i * H + lane];
    ab += d;
    float nb = 0.f;
    for (int j = erp[i]; j < erp[i + 1]; j++) nb = fmaf(ev[j], dpre[(long long)eci[j] * H + lane], nb);
    dUb[i * H2 + lane] = to_bf16(d);
    dUb[i * H2 + H + lane] = to_bf16(nb);
  }
  lds[wave][lane] = ab;
  __syncthreads();
  if (threadIdx.x < H) {
    float t = 0.f;
    for (int w = 0; w < kBlock / kWave; w++) t += lds[w][threadIdx.x];
    part[(long long)blockIdx.x * H + threadIdx.x] = t;
  }
}

// layer 0: partials of dW0 = z0^T dpre1 (8 x 64) and db0 = sum dpre1
constexpr int kL0Part = 2 * F * H + H;

__global__ __launch_bounds__(kBlock) void k_layer0_bwd(long long n, const float *__restrict__ z0,
                                                       const float *__restrict__ dpre,
                                                       float *__restrict__ part) {
  __shared__ float lds[kBlock / kWave][kL0Part];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const long long waves = (long long)gridDim.x * (kBlock / kWave);
  float aw[2 * F], ab = 0.f;
#pragma unroll
  for (int k = 0; k < 2 * F; k++) aw[k] = 0.f;
  for (long long i = (long long)blockIdx.x * (kBlock / kWave) + wave; i < n; i += waves) {
    const float d = dpre[i * H + lane];
    ab += d;
#pragma unroll
    for (int k = 0; k < 2 * F; k++) aw[k] = fmaf(z0[i * 2 * F + k], d, aw[k]);
  }
#pragma unroll
  for (int k = 0; k < 2 * F; k++) lds[wave][k * H + lane] = aw[k];
  lds[wave][2 * F * H + lane] = ab;
  __syncthreads();
  for (int e = threadIdx.x; e < kL0Part; e += kBlock) {
    float t = 0.f;
    for (int w = 0; w < kBlock / kWave; w++) t += lds[w][e];
    part[(long long)blockIdx.x * kL0Part + e] = t;
  }
}

// out[j] (+)= sum over nb workgroups of part[b * len + j], in workgroup order
__global__ __launch_bounds__(kBlock) void k_reduce_parts(const float *__restrict__ part, int nb,
                                                         int len, float *__restrict__ out, int add) {
  for (int j = blockIdx.x * kBlock + threadIdx.x; j < len; j += gridDim.x * kBlock) {
    float t = 0.f;
    for (int b = 0; b < nb; b++) t += part[(long long)b * len + j];
    out[j] = add ? out[j] + t : t;
  }
}

// scatter the head partial sums into the gradient vector
__global__ void k_head_scatter(const float *__restrict__ red, float *__restrict__ grad) {
  for (int e = threadIdx.x; e < kHeadPart; e += blockDim.x) {
    const float v = red[e];
    if (e < H) grad[owd + e] = v;
    else if (e == H) grad[obd] = v;
    else if (e < 2 * H + 1) grad[owo + e - H - 1] = v;
    else if (e == 2 * H + 1) grad[obo] = v;
    else if (e < 3 * H + 2) grad[ove + e - 2 * H - 2] = v;
    else grad[oc + e - 3 * H - 2] = v;
  }
}

__global__ void k_sum_double(const double *__restrict__ part, int nb, double *__restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double t = 0.0;
    for (int b = 0; b < nb; b++) t += part[b];
    *out = t;
  }
}

// Adam with global-norm clipping (one workgroup computes the norm, every
// workgroup applies it identically)
__global__ __launch_bounds__(kBlock) void k_grad_norm(const float *__restrict__ g, int len,
                                                      float *__restrict__ out) {
  __shared__ float lds[kBlock / kWave];
  float a = 0.f;
  for (int j = threadIdx.x; j < len; j += kBlock) a = fmaf(g[j], g[j], a);
  a = wave_sum(a);
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < kBlock / kWave; w++) t += lds[w];
    *out = sqrtf(t);
  }
}

__global__ __launch_bounds__(kBlock) void k_adam(float *__restrict__ th, float *__restrict__ m,
                                                 float *__restrict__ v,
                                                 const float *__restrict__ g, int len,
                                                 const float *__restrict__ gnorm, float lr,
                                                 float clip, float b1, float b2, float eps,
                                                 float c1, float c2) {
  const float nrm = *gnorm;
  const float sc = (clip > 0.f && nrm > clip) ? clip / nrm : 1.f;
  for (int j = blockIdx.x * kBlock + threadIdx.x; j < len; j += gridDim.x * kBlock) {
    const float gj = g[j] * sc;
    const float mj = b1 * m[j] + (1.f - b1) * gj;
    const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    th[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
  }
}

// BF16 copies of W1, W2, V in both orientations: wt (2H x H, = W^T, for
// U = h W) and w (H x 2H, for dh = dU W^T)
__global__ __launch_bounds__(kBlock) void k_weights_bf16(const float *__restrict__ th,
                                                         bf16_t *__restrict__ wt,
                                                         bf16_t *__restrict__ w) {
  const int offs[3] = {oW1, oW2, oV};
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < 3 * H * H2; e += gridDim.x * kBlock) {
    const int l = e / (H * H2), r = e % (H * H2);
    const int k = r / H2, c = r % H2;  // W[k][c]
    const bf16_t b = to_bf16(th[offs[l] + r]);
    w[l * H * H2 + r] = b;
    wt[l * H * H2 + c * H + k] = b;
  }
}

unsigned sgrid(long long n, int per = kBlock) {
  long long g = (n + per - 1) / per;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 65536));
}

constexpr unsigned kRowGrid = 1024;  // workgroups of the wave-per-row kernels (partials)

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  int ensure(size_t count) {
    if (p && count <= n) return TBSLA_HIP_SUCCESS;
    if (p) {
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    TBSLA_TRY(device_alloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)));
    n = count;
    return TBSLA_HIP_SUCCESS;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

}  // namespace tbsla_hip

using namespace tbsla_hip;

struct tbsla_hip_precond_net_s {
  DBuf<float> th, m, v, grad, gnorm;
  DBuf<bf16_t> wt, w;  // BF16 weight copies
  int step = 0;
  // graph of the current matrix
  long long n = 0, ne = 0, ng = 0;
  DBuf<double> s;
  DBuf<int> deg, nlow, gcnt, tcnt, erp, grp, trp, eci, gci, tci, tmap;
  DBuf<float> ev, x, z0;
  DBuf<char> scan_ws;
  // activations and gradients
  DBuf<float> h1, h2, h3, U, P, gdiag, ghat, dg, dh, dP;
  DBuf<bf16_t> hb1, hb2, hb3, Ub, dPb;
  DBuf<float> Z, T1, Vv, W, R, T2, Y;
  DBuf<double> lpart, lsum;
  DBuf<float> part, red;
};

namespace {

int scan_rowptr(tbsla_hip_precond_net_s *net, const int *cnt, int *rowptr, long long n,
                long long *total, hipStream_t s) {
  TBSLA_TRY(net->scan_ws.ensure(scan_workspace_bytes(n + 1) + 64));
  long long *d_total = reinterpret_cast<long long *>(net->scan_ws.p + scan_workspace_bytes(n + 1));
  TBSLA_TRY(scan<int>(cnt, rowptr, n, false, d_total, net->scan_ws.p, s));
  long long t = 0;
  TBSLA_HIP_CHECK(hipMemcpyAsync(&t, d_total, 8, hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  if (t > INT32_MAX) return fail(TBSLA_HIP_ERR_RANGE, "pattern beyond int32");
  const int ti = (int)t;
  TBSLA_HIP_CHECK(hipMemcpyAsync(rowptr + n, &ti, 4, hipMemcpyHostToDevice, s));
  *total = t;
  return TBSLA_HIP_SUCCESS;
}

// structure and scaled edges of A (every call: values may differ)
int prepare(tbsla_hip_precond_net_s *net, tbsla_hip_matrix_s *A, hipStream_t s) {
  if (A->format != TBSLA_HIP_CSR) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "the network reads a CSR block");
  if (A->ln_row != A->ln_col || A->f_row != A->f_col)
    return fail(TBSLA_HIP_ERR_INVALID, "the network needs a square diagonal block");
  const long long n = A->ln_row;
  if (n < 1) return fail(TBSLA_HIP_ERR_INVALID, "empty block");
  net->n = n;
  TBSLA_TRY(net->s.ensure(n));
  DBuf<int> *ib[] = {&net->deg, &net->nlow, &net->gcnt, &net->tcnt};
  for (auto *b : ib) TBSLA_TRY(b->ensure(n + 1));
  TBSLA_TRY(net->erp.ensure(n + 1));
  TBSLA_TRY(net->grp.ensure(n + 1));
  TBSLA_TRY(net->trp.ensure(n + 1));
  k_prep_rows<<<sgrid(n), kBlock, 0, s>>>(n, A->f_col, A->rowptr, A->colidx, A->values, net->s.p,
                                          net->deg.p, net->nlow.p, net->gcnt.p, net->tcnt.p);
  TBSLA_HIP_LAUNCH_CHECK();
  long long ne = 0, ng = 0, nt = 0;
  TBSLA_TRY(scan_rowptr(net, net->deg.p, net->erp.p, n, &ne, s));
  TBSLA_TRY(scan_rowptr(net, net->gcnt.p, net->grp.p, n, &ng, s));
  TBSLA_TRY(scan_rowptr(net, net->tcnt.p, net->trp.p, n, &nt, s));
  if (nt != ng) return fail(TBSLA_HIP_ERR_INVALID, "the matrix pattern is not symmetric");
  net->ne = ne;
  net->ng = ng;
  TBSLA_TRY(net->eci.ensure(ne));
  TBSLA_TRY(net->ev.ensure(ne));
  TBSLA_TRY(net->x.ensure(n * F));
  TBSLA_TRY(net->gci.ensure(ng));
  TBSLA_TRY(net->tci.ensure(ng));
  TBSLA_TRY(net->tmap.ensure(ng));
  k_prep_fill<<<sgrid(n), kBlock, 0, s>>>(n, A->f_col, A->rowptr, A->colidx, A->values, net->s.p,
                                          net->erp.p, net->nlow.p, net->grp.p, net->trp.p,
                                          net->eci.p, net->ev.p, net->x.p, net->gci.p, net->tci.p);
  k_prep_tmap<<<sgrid(n), kBlock, 0, s>>>(n, net->grp.p, net->gci.p, net->nlow.p, net->trp.p,
                                          net->tci.p, net->tmap.p);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

unsigned row_grid(long long n) {
  return (unsigned)std::min<long long>(kRowGrid, std::max<long long>(1, (n + 3) / 4));
}

int forward(tbsla_hip_precond_net_s *net, double *gval, hipStream_t s) {
  const long long n = net->n;
  DBuf<float> *fb[] = {&net->h1, &net->h2, &net->h3};
  for (auto *b : fb) TBSLA_TRY(b->ensure(n * H));
  DBuf<bf16_t> *bb[] = {&net->hb1, &net->hb2, &net->hb3};
  for (auto *b : bb) TBSLA_TRY(b->ensure(n * H));
  TBSLA_TRY(net->z0.ensure(n * 2 * F));
  TBSLA_TRY(net->U.ensure(n * H2));
  TBSLA_TRY(net->P.ensure(n * H2));
  TBSLA_TRY(net->gdiag.ensure(n));
  TBSLA_TRY(net->ghat.ensure(net->ng));
  const unsigned gg = (unsigned)((n + 63) / 64);
  const unsigned qg = (unsigned)std::min<long long>(4096, (n + 15) / 16);  // 16 nodes per workgroup
  const float *th = net->th.p;
  k_layer0<<<(unsigned)((n + 63) / 64), kBlock, 0, s>>>(n, net->erp.p, net->eci.p, net->ev.p,
                                                         net->x.p, th, net->z0.p, net->h1.p,
                                                         net->hb1.p);
  k_gemm_rows<H, H2><<<gg, kBlock, 0, s>>>(n, net->hb1.p, net->wt.p, net->U.p, 0, nullptr);
  k_aggregate<<<qg, kBlock, 0, s>>>(n, net->erp.p, net->eci.p, net->ev.p, net->U.p, th, ob1,
                                    net->h2.p, net->hb2.p, nullptr);
  k_gemm_rows<H, H2><<<gg, kBlock, 0, s>>>(n, net->hb2.p, net->wt.p + H * H2, net->U.p, 0, nullptr);
  k_aggregate<<<qg, kBlock, 0, s>>>(n, net->erp.p, net->eci.p, net->ev.p, net->U.p, th, ob2,
                                    net->h3.p, net->hb3.p, net->gdiag.p);
  k_gemm_rows<H, H2><<<gg, kBlock, 0, s>>>(n, net->hb3.p, net->wt.p + 2 * H * H2, net->P.p, 0,
                                           nullptr);
  k_edge_head<<<qg, kBlock, 0, s>>>(n, net->erp.p, net->eci.p, net->ev.p, net->nlow.p, net->grp.p,
                                    net->P.p, th, net->gdiag.p, net->s.p, net->ghat.p, gval);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

template <int KP>
int loss_and_dghat_k(tbsla_hip_precond_net_s *net, unsigned long long seed, hipStream_t s) {
  const long long n = net->n, nk = n * KP;
  DBuf<float> *pb[] = {&net->Z, &net->T1, &net->Vv, &net->W, &net->R, &net->T2, &net->Y};
  for (auto *b : pb) TBSLA_TRY(b->ensure(nk));
  TBSLA_TRY(net->dg.ensure(net->ng));
  TBSLA_TRY(net->lpart.ensure(1024));
  TBSLA_TRY(net->lsum.ensure(1));
  const unsigned g = sgrid(nk);
  k_probes<<<g, kBlock, 0, s>>>(nk, mix64(seed), net->Z.p);
  k_spmm_g<KP><<<g, kBlock, 0, s>>>(n, net->grp.p, net->gci.p, net->ghat.p, nullptr, net->Z.p, net->T1.p);
  k_spmm_a<KP><<<g, kBlock, 0, s>>>(n, net->erp.p, net->eci.p, net->ev.p, net->T1.p, net->Vv.p);
  k_spmm_g<KP><<<g, kBlock, 0, s>>>(n, net->trp.p, net->tci.p, net->ghat.p, net->tmap.p, net->Vv.p,
                                    net->W.p);
  const unsigned lg = std::min<unsigned>(1024, sgrid(nk));
  k_residual_probe<<<lg, kBlock, 0, s>>>(nk, net->Z.p, net->W.p, net->R.p, net->lpart.p);
  k_sum_double<<<1, 64, 0, s>>>(net->lpart.p, (int)lg, net->lsum.p);
  k_spmm_g<KP><<<g, kBlock, 0, s>>>(n, net->grp.p, net->gci.p, net->ghat.p, nullptr, net->R.p, net->T2.p);
  k_spmm_a<KP><<<g, kBlock, 0, s>>>(n, net->erp.p, net->eci.p, net->ev.p, net->T2.p, net->Y.p);
  const float c = (float)(-2.0 / ((double)n * KP));
  k_dghat<KP><<<g, kBlock, 0, s>>>(n, net->grp.p, net->gci.p, net->Vv.p, net->R.p, net->Y.p,
                                   net->Z.p, c, net->dg.p);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

int loss_and_dghat(tbsla_hip_precond_net_s *net, int k, unsigned long long seed, hipStream_t s) {
  switch (k) {
    case 1: return loss_and_dghat_k<1>(net, seed, s);
    case 2: return loss_and_dghat_k<2>(net, seed, s);
    case 4: return loss_and_dghat_k<4>(net, seed, s);
    case 8: return loss_and_dghat_k<8>(net, seed, s);
    case 16: return loss_and_dghat_k<16>(net, seed, s);
  }
  return fail(TBSLA_HIP_ERR_INVALID, "probes must be 1, 2, 4, 8 or 16");
}

// dL/dtheta into net->grad
int backward(tbsla_hip_precond_net_s *net, hipStream_t s) {
  const long long n = net->n;
  const unsigned rg = row_grid(n);
  const unsigned gg = (unsigned)((n + 63) / 64);
  const unsigned wg = (unsigned)((n + kWgRows - 1) / kWgRows);
  TBSLA_TRY(net->dh.ensure(n * H));
  TBSLA_TRY(net->dP.ensure(n * H2));
  TBSLA_TRY(net->dPb.ensure(n * H2));
  TBSLA_TRY(net->Ub.ensure(n * H2));
  TBSLA_TRY(net->grad.ensure(kParams));
  const size_t part_len = std::max<size_t>({(size_t)rg * kHeadPart, (size_t)rg * kL0Part,
                                            (size_t)wg * H * H2, (size_t)rg * H});
  TBSLA_TRY(net->part.ensure(part_len));
  TBSLA_TRY(net->red.ensure(kHeadPart));
  const float *th = net->th.p;
  float *grad = net->grad.p;
  // head
  k_head_bwd_rows<<<rg, kBlock, 0, s>>>(n, net->erp.p, net->eci.p, net->ev.p, net->nlow.p,
                                        net->grp.p, net->P.p, net->h3.p, th, net->ghat.p, net->dg.p,
                                        net->dh.p, net->dP.p, net->dPb.p, net->part.p);
  k_reduce_parts<<<1, kBlock, 0, s>>>(net->part.p, (int)rg, kHeadPart, net->red.p, 0);
  k_head_scatter<<<1, kBlock, 0, s>>>(net->red.p, grad);
  k_head_bwd_cols<<<rg, kBlock, 0, s>>>(n, net->trp.p, net->tci.p, net->tmap.p, net->erp.p,
                                        net->eci.p, net->ev.p, net->nlow.p, net->P.p, th, net->dg.p,
                                        net->dP.p, net->dPb.p);
  // V: dV = h3^T dP; dh3 += dP V^T, masked -> dpre3
  k_gemm_wgrad<<<wg, kBlock, 0, s>>>(n, net->hb3.p, net->dPb.p, net->part.p);
  k_reduce_parts<<<sgrid(H * H2), kBlock, 0, s>>>(net->part.p, (int)wg, H * H2, grad + oV, 0);
  k_gemm_rows<H2, H><<<gg, kBlock, 0, s>>>(n, net->dPb.p, net->w.p + 2 * H * H2, net->dh.p, 1,
                                           net->h3.p);
  // layer 2: pre3 = U2[:, :64] + E U2[:, 64:] + b2
  const struct {
    const bf16_t *hb;
    const float *h;
    int oW, ob, wl;
  } L[2] = {{net->hb2.p, net->h2.p, oW2, ob2, 1}, {net->hb1.p, net->h1.p, oW1, ob1, 0}};
  for (int l = 0; l < 2; l++) {
    k_aggregate_bwd<<<rg, kBlock, 0, s>>>(n, net->erp.p, net->eci.p, net->ev.p, net->dh.p,
                                          net->Ub.p, net->part.p);
    k_reduce_parts<<<1, kBlock, 0, s>>>(net->part.p, (int)rg, H, grad + L[l].ob, 0);
    k_gemm_wgrad<<<wg, kBlock, 0, s>>>(n, L[l].hb, net->Ub.p, net->part.p);
    k_reduce_parts<<<sgrid(H * H2), kBlock, 0, s>>>(net->part.p, (int)wg, H * H2, grad + L[l].oW, 0);
    // dh_l = dU W_l^T, masked by h_l > 0 -> dpre_l (overwrites dh)
    k_gemm_rows<H2, H><<<gg, kBlock, 0, s>>>(n, net->Ub.p, net->w.p + L[l].wl * H * H2, net->dh.p, 0,
                                             L[l].h);
  }
  k_layer0_bwd<<<rg, kBlock, 0, s>>>(n, net->z0.p, net->dh.p, net->part.p);
  k_reduce_parts<<<sgrid(kL0Part), kBlock, 0, s>>>(net->part.p, (int)rg, kL0Part, grad + oW0, 0);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

int refresh_weights(tbsla_hip_precond_net_s *net, hipStream_t s) {
  k_weights_bf16<<<sgrid(3 * H * H2), kBlock, 0, s>>>(net->th.p, net->wt.p, net->w.p);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

// initial parameters: N(0,1) draws by Box-Muller on V(seed, .) scaled per
// tensor; heads at zero
void init_params(unsigned long long seed, std::vector<float> &th) {
  th.assign(kParams, 0.f);
  const uint64_t h = mix64(seed);
  struct Seg {
    int off, len;
    double scale;
  } segs[] = {{oW0, 2 * F * H, std::sqrt(2.0 / (2 * F))}, {ob0, H, 0.0}, {oW1, H * H2, std::sqrt(2.0 / H)},
              {ob1, H, 0.0}, {oW2, H * H2, std::sqrt(2.0 / H)}, {ob2, H, 0.0},
              {oV, H * H2, std::sqrt(1.0 / H)}, {ove, H, 0.1}, {oc, H, 0.0}, {owo, H, 0.0},
              {obo, 1, 0.0}, {owd, H, 0.0}, {obd, 1, 0.0}};
  for (const Seg &sg : segs)
    for (int t = 0; t < sg.len; t++) {
      const uint64_t idx = (uint64_t)(sg.off + t);
      const double u1 = unit_value(h, 2 * idx), u2 = unit_value(h, 2 * idx + 1);
      const double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
      th[sg.off + t] = (float)(z * sg.scale);
    }
}

}  // namespace

extern "C" int tbsla_hip_precond_net_param_count(void) { return kParams; }

extern "C" int tbsla_hip_precond_net_create(tbsla_hip_precond_net_t *out,
                                            unsigned long long seed) {
  if (!out) return fail(TBSLA_HIP_ERR_INVALID, "null output");
  auto *net = new tbsla_hip_precond_net_s();
  std::vector<float> th;
  init_params(seed, th);
  int st;
  if ((st = net->th.ensure(kParams)) || (st = net->m.ensure(kParams)) ||
      (st = net->v.ensure(kParams)) || (st = net->grad.ensure(kParams)) ||
      (st = net->gnorm.ensure(1)) || (st = net->wt.ensure(3 * H * H2)) ||
      (st = net->w.ensure(3 * H * H2))) {
    delete net;
    return st;
  }
  hipError_t e = hipMemcpy(net->th.p, th.data(), kParams * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(net->m.p, 0, kParams * 4);
  if (e == hipSuccess) e = hipMemset(net->v.p, 0, kParams * 4);
  if (e != hipSuccess) {
    delete net;
    return fail(TBSLA_HIP_ERR_HIP, hipGetErrorString(e));
  }
  st = refresh_weights(net, nullptr);
  if (st == TBSLA_HIP_SUCCESS && hipDeviceSynchronize() != hipSuccess)
    st = fail(TBSLA_HIP_ERR_HIP, "initialising the network");
  if (st) {
    delete net;
    return st;
  }
  *out = net;
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_precond_net_destroy(tbsla_hip_precond_net_t net) {
  if (!net) return TBSLA_HIP_SUCCESS;
  (void)hipDeviceSynchronize();
  delete net;
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_precond_net_get_params(tbsla_hip_precond_net_t net, float *h_theta) {
  if (!net || !h_theta) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  TBSLA_HIP_CHECK(hipMemcpy(h_theta, net->th.p, kParams * 4, hipMemcpyDeviceToHost));
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_precond_net_set_params(tbsla_hip_precond_net_t net, const float *h_theta) {
  if (!net || !h_theta) return fail(TBSLA_HIP_ERR_INVALID, "null argument");
  TBSLA_HIP_CHECK(hipMemcpy(net->th.p, h_theta, kParams * 4, hipMemcpyHostToDevice));
  TBSLA_HIP_CHECK(hipMemset(net->m.p, 0, kParams * 4));
  TBSLA_HIP_CHECK(hipMemset(net->v.p, 0, kParams * 4));
  net->step = 0;
  TBSLA_TRY(refresh_weights(net, nullptr));
  TBSLA_HIP_CHECK(hipDeviceSynchronize());
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_precond_net_apply(tbsla_hip_precond_net_t net, tbsla_hip_matrix_t A,
                                           tbsla_hip_matrix_t G, tbsla_hip_matrix_t GT,
                                           void *stream) {
  if (!net || !A || !G || G == A || (GT && (GT == A || GT == G)))
    return fail(TBSLA_HIP_ERR_INVALID, "bad handles");
  if (G->format != TBSLA_HIP_CSR || (GT && GT->format != TBSLA_HIP_CSR))
    return fail(TBSLA_HIP_ERR_UNSUPPORTED, "G and G^T are CSR blocks");
  if (A->f_col != 0 || A->f_row != 0) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "the network needs a block at (0, 0)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TBSLA_TRY(prepare(net, A, s));
  const long long n = net->n, ng = net->ng;
  // G: the lower pattern (with the diagonal) of A, FP64 values
  // arrays of a previous call of the same size are reused (no allocation on
  // the solve path, and graphs captured on them stay valid)
  auto shape = [&](tbsla_hip_matrix_s *M, const int *rp, const int *ci, int **rpo, int **cio,
                   double **vo) -> int {
    const bool reuse = M->format == TBSLA_HIP_CSR && M->has_arrays && M->ln_row == n &&
                       M->stored == ng && M->rowptr && M->colidx && M->values;
    if (reuse) {
      TBSLA_HIP_CHECK(hipMemcpyAsync(M->rowptr, rp, (size_t)(n + 1) * 4, hipMemcpyDeviceToDevice, s));
      TBSLA_HIP_CHECK(hipMemcpyAsync(M->colidx, ci, (size_t)ng * 4, hipMemcpyDeviceToDevice, s));
      *rpo = M->rowptr;
      *cio = M->colidx;
      *vo = M->values;
      return TBSLA_HIP_SUCCESS;
    }
    release_arrays(M);
    M->n_row = M->n_col = M->ln_row = M->ln_col = (int)n;
    M->f_row = M->f_col = A->f_row;
    M->pr = M->pc = 0;
    M->NR = M->NC = 1;
    M->max_col = 0;
    M->row_sorted = 1;
    TBSLA_TRY(device_alloc((void **)&M->rowptr, (size_t)(n + 1) * 4));
    TBSLA_TRY(device_alloc((void **)&M->colidx, (size_t)ng * 4));
    TBSLA_TRY(device_alloc((void **)&M->values, (size_t)ng * 8));
    TBSLA_HIP_CHECK(hipMemcpyAsync(M->rowptr, rp, (size_t)(n + 1) * 4, hipMemcpyDeviceToDevice, s));
    TBSLA_HIP_CHECK(hipMemcpyAsync(M->colidx, ci, (size_t)ng * 4, hipMemcpyDeviceToDevice, s));
    M->nnz = M->stored = ng;
    M->has_arrays = 1;
    *rpo = M->rowptr;
    *cio = M->colidx;
    *vo = M->values;
    return TBSLA_HIP_SUCCESS;
  };
  int *rp, *ci;
  double *gv, *tv = nullptr;
  TBSLA_TRY(shape(G, net->grp.p, net->gci.p, &rp, &ci, &gv));
  if (GT) TBSLA_TRY(shape(GT, net->trp.p, net->tci.p, &rp, &ci, &tv));
  TBSLA_TRY(forward(net, gv, s));
  if (GT) {
    k_gather_t<<<sgrid(ng), kBlock, 0, s>>>(ng, net->tmap.p, gv, tv);
    TBSLA_HIP_LAUNCH_CHECK();
  }
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_precond_net_loss_grad(tbsla_hip_precond_net_t net, tbsla_hip_matrix_t A,
                                               int probes, unsigned long long probe_seed,
                                               double *loss, float *h_grad, void *stream) {
  if (!net || !A) return fail(TBSLA_HIP_ERR_INVALID, "bad handles");
  if (A->f_col != 0) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "the network needs f_col = 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TBSLA_TRY(prepare(net, A, s));
  TBSLA_TRY(forward(net, nullptr, s));
  TBSLA_TRY(loss_and_dghat(net, probes, probe_seed, s));
  TBSLA_TRY(backward(net, s));
  double sum = 0;
  TBSLA_HIP_CHECK(hipMemcpyAsync(&sum, net->lsum.p, 8, hipMemcpyDeviceToHost, s));
  if (h_grad) TBSLA_HIP_CHECK(hipMemcpyAsync(h_grad, net->grad.p, kParams * 4, hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  if (loss) *loss = sum / ((double)net->n * probes);
  return TBSLA_HIP_SUCCESS;
}

extern "C" int tbsla_hip_precond_net_train_step(tbsla_hip_precond_net_t net, tbsla_hip_matrix_t A,
                                                int probes, unsigned long long probe_seed,
                                                float lr, float clip, double *loss, void *stream) {
  if (!net || !A) return fail(TBSLA_HIP_ERR_INVALID, "bad handles");
  if (A->f_col != 0) return fail(TBSLA_HIP_ERR_UNSUPPORTED, "the network needs f_col = 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  TBSLA_TRY(prepare(net, A, s));
  TBSLA_TRY(forward(net, nullptr, s));
  TBSLA_TRY(loss_and_dghat(net, probes, probe_seed, s));
  TBSLA_TRY(backward(net, s));
  net->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)net->step), c2 = 1.f - std::pow(b2, (float)net->step);
  k_grad_norm<<<1, kBlock, 0, s>>>(net->grad.p, kParams, net->gnorm.p);
  k_adam<<<sgrid(kParams), kBlock, 0, s>>>(net->th.p, net->m.p, net->v.p, net->grad.p, kParams,
                                           net->gnorm.p, lr, clip, b1, b2, 1e-8f, c1, c2);
  TBSLA_TRY(refresh_weights(net, s));
  double sum = 0;
  TBSLA_HIP_CHECK(hipMemcpyAsync(&sum, net->lsum.p, 8, hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  if (loss) *loss = sum / ((double)net->n * probes);
  return TBSLA_HIP_SUCCESS;
}
