// common.hpp -- shared device helpers and the internal matrix record of libtbsla_hip.
//
// gfx950 only: 64-lane wavefronts, 256-thread workgroups unless a kernel says
// otherwise.  Error handling: every entry point returns a tbsla_hip status
// and records a message retrievable with tbsla_hip_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "tbsla_hip.h"

namespace tbsla_hip {

constexpr int kWave = 64;
constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int status, const std::string &msg);

#define TBSLA_HIP_CHECK(expr)                                                             \
  do {                                                                                    \
    hipError_t err__ = (expr);                                                            \
    if (err__ != hipSuccess)                                                              \
      return ::tbsla_hip::fail(TBSLA_HIP_ERR_HIP, std::string(#expr) + ": " +             \
                                                      hipGetErrorString(err__));          \
  } while (0)

#define TBSLA_HIP_LAUNCH_CHECK() TBSLA_HIP_CHECK(hipGetLastError())

#define TBSLA_TRY(expr)            \
  do {                             \
    int st__ = (expr);             \
    if (st__ != TBSLA_HIP_SUCCESS) \
      return st__;                 \
  } while (0)

// Diagonal layout of a CSR block: values stored in tiles of kWave rows; row
// r's entry on stored diagonal k sits at (r / kWave) kWave ndiag + k kWave +
// r % kWave, so one wavefront's loads of a diagonal are one 512-byte piece
// and its whole slab of diagonals is contiguous.
__host__ __device__ inline long long dia_at(long long r, int k, int ndiag) {
  return (r / kWave) * (long long)kWave * ndiag + (long long)k * kWave + (r % kWave);
}

inline unsigned grid_for(long long n, int per_block = kBlock) {
  long long g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// generators (device and host)
// ---------------------------------------------------------------------------

// glibc rand_r (stdlib/rand_r.c): three LCG steps, 11 + 10 + 10 bits.
__host__ __device__ inline int rand_r32(uint32_t &seed) {
  uint32_t next = seed;
  next = next * 1103515245u + 12345u;
  int result = (int)((next >> 16) & 2047u);
  next = next * 1103515245u + 12345u;
  result = (result << 10) ^ (int)((next >> 16) & 1023u);
  next = next * 1103515245u + 12345u;
  result = (result << 10) ^ (int)((next >> 16) & 1023u);
  seed = next;
  return result;
}

// values_generation::cqmat_value (src/cpp/utils/values_generation.cpp:28-47):
// generator index k -> (row, col).  c is already min(c, nc).
struct CqmatParams {
  uint64_t nr, nc, c, nfull;  // nfull = min(nc - c + 1, nr)
  double q;
  uint32_t seed_mult;
};

__host__ __device__ inline void cqmat_decode(const CqmatParams &p, uint64_t k, int64_t &row,
                                             int64_t &col) {
  uint32_t seed = (uint32_t)k;
  if (p.seed_mult > 0) seed = (uint32_t)((uint64_t)p.seed_mult * k);
  if (k < p.c * p.nfull) {
    uint64_t r = k / p.c;
    int draw = rand_r32(seed);
    row = (int64_t)r;
    if ((double)draw / 2147483647.0 < p.q)
      col = (int64_t)((uint64_t)rand_r32(seed) % p.nc);
    else
      col = (int64_t)(r + k % p.c);
  } else {
    uint64_t it = k - p.c * p.nfull;
    uint64_t cur = 1;
    while (cur < p.nr - p.nfull && it >= p.c - cur) {
      it -= p.c - cur;
      cur++;
    }
    cur += p.nfull - 1;
    int draw = rand_r32(seed);
    row = (int64_t)cur;
    // integer division, as the reference (:45): true for every draw < RAND_MAX
    if ((double)(draw / 2147483647) < p.q)
      col = (int64_t)((uint64_t)rand_r32(seed) % p.nc);
    else
      col = (int64_t)(cur + it);
  }
}

// values_generation::cdiag_value (values_generation.cpp:6-26).
__host__ __device__ inline void cdiag_decode(uint64_t i, int64_t nc, int64_t cd, int64_t &row,
                                             int64_t &col) {
  int64_t ii = (int64_t)i;
  if (cd == 0) {
    row = ii;
    col = ii;
    return;
  }
  int64_t lim1 = (nc - cd) < cd ? (nc - cd) : cd;
  if (lim1 < 0) lim1 = 0;
  if (ii < lim1) {
    row = ii;
    col = ii + cd;
    return;
  }
  int64_t lim2 = cd + 2 * (nc - 2 * cd);
  if (lim2 < 0) lim2 = 0;
  if (ii < lim2) {
    int64_t it = (int64_t)((i + (uint64_t)cd) / 2);
    row = it;
    col = (i % 2 == 0) ? it - cd : it + cd;
    return;
  }
  int64_t it = ii - (nc - 2 * cd);
  if (cd > nc) it -= cd - nc;
  row = it;
  col = it - cd;
}

// Project value generator V(seed, k): top 53 bits of
// splitmix64(splitmix64(seed) ^ k) * 2^-53 (splitmix64: Steele, Lea & Flood,
// OOPSLA 2014).  Integer-only, identical on host and device.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline double unit_value(uint64_t seed_hash, uint64_t k) {
  uint64_t h = mix64(seed_hash ^ k);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// range::lnv / range::pflv (src/cpp/utils/range.cpp:7-26).
inline long long lnv(long long size, int pos, int npos) {
  long long n = size / npos;
  long long mod = size % npos;
  if (pos < mod) n++;
  return n;
}
inline long long pflv(long long size, int pos, int npos) {
  long long mod = size % npos;
  long long n = lnv(size, pos, npos) * pos;
  if (pos >= mod) n += mod;
  return n;
}

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
template <int W>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
  return v;
}

// Inside a kernel that must round a*b+c twice, as the reference's loops do
// (HIP compiles with fp-contract=fast, which would fuse them).
#define TBSLA_NO_FMA _Pragma("clang fp contract(off)")

template <typename T>
__device__ inline T ldg_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}

}  // namespace tbsla_hip

// ---------------------------------------------------------------------------
// The matrix record behind tbsla_hip_matrix_t.
// ---------------------------------------------------------------------------
struct tbsla_hip_matrix_s {
  int format = TBSLA_HIP_CSR;
  // block geometry, as tbsla::cpp::Matrix (Matrix.hpp:54-55)
  int n_row = 0, n_col = 0, f_row = 0, f_col = 0, ln_row = 0, ln_col = 0;
  int pr = 0, pc = 0, NR = 1, NC = 1;
  long long nnz = 0;     // the reference's nnz field
  long long stored = 0;  // entries held: CSR rowptr[ln_row]; COO/SCOO triplets
  int max_col = 0;       // ELL
  int has_arrays = 0;    // the reference allocated its arrays (fills can skip it)
  int row_sorted = 1;    // COO/SCOO triplets in nondecreasing row order
  // device arrays (owned)
  int *rowptr = nullptr;     // CSR: ln_row + 1
  int *colidx = nullptr;     // CSR: stored;  ELL: max_col * ln_row, column-major
  int *rowidx = nullptr;     // COO/SCOO: stored
  double *values = nullptr;  // CSR/COO/SCOO: stored; ELL: column-major; DENSE: row-major
  // SpMV schedule
  int lanes_per_row = 0;     // CSR: 0 = pick from the mean row length
  int rows_per_group = 0;    // CSR: rows each lane group keeps in flight (0 = pick)
  int csr_vec = 0;           // CSR: entries per lane load (0 = 1)
  int csr_nt = -1;           // CSR: non-temporal loads of the entries (-1 = on)
  // scratch owned by the record (reused across calls; grows on demand)
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // CG work vectors and captured graphs kept between solves (cg.hip)
  void *cg_cache = nullptr;
  // CSR split execution layout (split.hip), when selected
  void *split = nullptr;
  // COO/SCOO with row-sorted triplets: row pointer over the rows
  // [coo_row0, coo_row0 + coo_nrows) into the triplet arrays, built on the
  // first SpMV (spmv.hip)
  int *coo_rowptr = nullptr;
  int coo_row0 = 0, coo_nrows = 0;
  // identity of the handle and of its values: copies derived from the
  // values elsewhere (the CG's FP32 copy) are current while both match
  unsigned long long uid = 0;
  unsigned long long values_version = 0;
};

namespace tbsla_hip {
int ensure_scratch(tbsla_hip_matrix_s *m, size_t bytes);
void release_arrays(tbsla_hip_matrix_s *m);
void cg_cache_release(tbsla_hip_matrix_s *m);
void split_release(tbsla_hip_matrix_s *m);
int device_alloc(void **p, size_t bytes);
}  // namespace tbsla_hip
