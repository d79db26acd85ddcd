// Element-wise / reduction kernels of libxtpgpu (packing, weights, synthetic data,
// BF16 splitting, Davidson vector updates).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "gemm_kernels.h"
#include "xtpgpu_internal.h"

namespace xg {

template <class F>
__global__ void map_kernel(size_t n, F f) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) f(i);
}

template <class F>
inline void launch_map(hipStream_t st, size_t n, F f) {
  if (n == 0) return;
  size_t blocks = std::min<size_t>((n + 255) / 256, (size_t)1 << 16);
  hipLaunchKernelGGL(map_kernel<F>, dim3((unsigned)blocks), dim3(256), 0, st, n, f);
  XG_HIP(hipGetLastError());
}

// splitmix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number generators", 2014)
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Standard-normal variate as an Irwin-Hall sum of 12 uniform 16-bit integers, shifted
// to zero mean and unit variance. Pure integer arithmetic plus one exact conversion, so
// a host restatement reproduces it bit for bit.
__host__ __device__ inline double synth_normal(uint64_t seed, uint64_t p, uint64_t i, uint64_t j) {
  uint64_t h = splitmix64(seed);
  h = splitmix64(h ^ (p * 0x100000001B3ull));
  h = splitmix64(h ^ (i * 0xC2B2AE3D27D4EB4Full));
  h = splitmix64(h ^ (j * 0x165667B19E3779F9ull));
  uint64_t s = 0;
  for (int r = 0; r < 3; ++r) {
    h = splitmix64(h);
    s += (h & 0xFFFFull) + ((h >> 16) & 0xFFFFull) + ((h >> 32) & 0xFFFFull) + ((h >> 48) & 0xFFFFull);
  }
  // s in [0, 12*65535]; (s + 6) / 65536 - 6 has mean 0 and variance 12*(65536^2-1)/12/65536^2 ~ 1
  return ((double)s + 6.0) * (1.0 / 65536.0) - 6.0;
}

// Split a double into bf16 hi + bf16 lo (bit patterns): hi = rn(x), lo = rn(x - hi).
__device__ __forceinline__ void split_bf16(double x, unsigned short& hi, unsigned short& lo) {
  float f = (float)x;
  hi = bf16_bits_rn(f);
  lo = bf16_bits_rn((float)(x - (double)bf16_bits_to_f(hi)));
}

// Deterministic column norms: one workgroup per column, fixed reduction order.
__global__ __launch_bounds__(256) void colnorm_kernel(const double* __restrict__ X, long ld, long n,
                                                      double* __restrict__ out) {
  __shared__ double red[256];
  const double* x = X + (long)blockIdx.x * ld;
  double s = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) s += x[i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt(red[0]);
}

// Deterministic column norms in two passes for tall columns: pass 1 gives partial sums of
// squares over `nb` row chunks per column (grid nb x ncol), pass 2 reduces each column's nb
// partials with one wave in a fixed butterfly order and takes the square root.
__global__ __launch_bounds__(256) void colnorm_partial_kernel(const double* __restrict__ X, long ld, long n,
                                                              double* __restrict__ part) {
  __shared__ double red[256];
  const int nb = gridDim.x;
  const long chunk = (n + nb - 1) / nb;
  const long lo = (long)blockIdx.x * chunk, hi = min(n, lo + chunk);
  const double* x = X + (long)blockIdx.y * ld;
  double s = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += 256) s += x[i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)blockIdx.y * nb + blockIdx.x] = red[0];
}

__device__ __forceinline__ double wave_sum(double s) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// out[c] = sqrt(sum_b part[c*nb + b]); one wave per column, 4 columns per block.
__global__ __launch_bounds__(256) void colnorm_finish_kernel(const double* __restrict__ part, int nb, long ncol,
                                                             double* __restrict__ out) {
  const long c = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= ncol) return;
  double s = 0.0;
  for (int b = lane; b < nb; b += 64) s += part[c * nb + b];
  s = wave_sum(s);
  if (lane == 0) out[c] = sqrt(s);
}

// Sum of split-K partials, one wave per output (4 per block): lanes stride over the splits and
// a fixed butterfly combines them, so the order is deterministic.
__global__ __launch_bounds__(256) void reduce_partials_wave_kernel(const double* __restrict__ part, long stride,
                                                                   int nsplit, long n, double* __restrict__ out) {
  const long i = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  double s = 0.0;
  for (int k = lane; k < nsplit; k += 64) s += part[(long)k * stride + i];
  s = wave_sum(s);
  if (lane == 0) out[i] = s;
}

// Sum of split-K partials: out[i] = sum_s part[s*stride + i] (fixed order).
__global__ void reduce_partials_kernel(const double* __restrict__ part, long stride, int nsplit,
                                       long n, double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += part[(long)k * stride + i];
    out[i] = s;
  }
}

}  // namespace xg
