// scan.hip -- device prefix sums (rowptr construction, sort offsets).
//
// Reduce-then-scan in three launches: per-tile sums, one workgroup scanning
// the tile sums in order, per-tile rescan with the tile's offset.  Tiles are
// 256 threads x 8 consecutive items; accumulation is 64-bit, so the totals
// of int32 inputs are exact.  Integer work: the result is bit-exact and
// independent of the launch shape.
#include "common.hpp"
#include "kernels.hpp"

namespace tbsla_hip {

namespace {

constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;

// Inclusive scan of one value per thread across the 256-thread workgroup.
__device__ long long block_inclusive_scan(long long v, long long *lds_waves) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    long long o = __shfl_up(v, off, kWave);
    if (lane >= off) v += o;
  }
  if (lane == kWave - 1) lds_waves[wave] = v;
  __syncthreads();
  long long add = 0;
  for (int w = 0; w < wave; w++) add += lds_waves[w];
  __syncthreads();
  return v + add;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_tile_sums(const T *__restrict__ in, long long n,
                                                      long long *__restrict__ sums) {
  __shared__ long long lds[kBlock / kWave];
  long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kItems;
  long long s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++)
    if (base + i < n) s += (long long)in[base + i];
  s = block_inclusive_scan(s, lds);
  if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = s;
}

// One workgroup: exclusive scan of the tile sums, in chunks, in order.
__global__ __launch_bounds__(kBlock) void k_scan_tile_sums(long long *__restrict__ sums,
                                                           long long ntiles,
                                                           long long *__restrict__ total) {
  __shared__ long long lds[kBlock / kWave];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (long long base = 0; base < ntiles; base += kBlock) {
    long long i = base + threadIdx.x;
    long long v = i < ntiles ? sums[i] : 0;
    long long inc = block_inclusive_scan(v, lds);
    long long c = carry;
    if (i < ntiles) sums[i] = c + inc - v;
    __syncthreads();
    if (threadIdx.x == kBlock - 1) carry = c + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_tile_scan(const T *__restrict__ in, T *__restrict__ out,
                                                      long long n,
                                                      const long long *__restrict__ offsets,
                                                      int inclusive) {
  __shared__ long long lds[kBlock / kWave];
  long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kItems;
  long long vals[kItems];
  long long s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    vals[i] = base + i < n ? (long long)in[base + i] : 0;
    s += vals[i];
  }
  long long run = block_inclusive_scan(s, lds) - s + offsets[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    long long next = run + vals[i];
    if (base + i < n) out[base + i] = (T)(inclusive ? next : run);
    run = next;
  }
}

}  // namespace

template <typename T>
int scan(const T *in, T *out, long long n, bool inclusive, long long *d_total, void *workspace,
         hipStream_t s) {
  long long ntiles = (n + kTile - 1) / kTile;
  if (ntiles < 1) ntiles = 1;
  long long *sums = static_cast<long long *>(workspace);
  k_tile_sums<T><<<(unsigned)ntiles, kBlock, 0, s>>>(in, n, sums);
  TBSLA_HIP_LAUNCH_CHECK();
  k_scan_tile_sums<<<1, kBlock, 0, s>>>(sums, ntiles, d_total);
  TBSLA_HIP_LAUNCH_CHECK();
  k_tile_scan<T><<<(unsigned)ntiles, kBlock, 0, s>>>(in, out, n, sums, inclusive ? 1 : 0);
  TBSLA_HIP_LAUNCH_CHECK();
  return TBSLA_HIP_SUCCESS;
}

size_t scan_workspace_bytes(long long n) {
  long long ntiles = (n + kTile - 1) / kTile;
  if (ntiles < 1) ntiles = 1;
  return (size_t)ntiles * sizeof(long long);
}

template int scan<int>(const int *, int *, long long, bool, long long *, void *, hipStream_t);
template int scan<long long>(const long long *, long long *, long long, bool, long long *, void *,
                             hipStream_t);

}  // namespace tbsla_hip
