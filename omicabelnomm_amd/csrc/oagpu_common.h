// Shared device helpers for the gfx950 scan kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oa {

enum { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16X3 = 2 };
enum { DT_F64 = 0, DT_F32 = 1, DT_F16 = 2 };

// Largest coefficient count a pair can have (l + r), and the stride of a
// coefficient vector in the fallback list.
constexpr int PMAX = 40;

// Relative size below which the closed-form RSS is replaced by an explicit
// residual pass: RSS < RSS_GUARD * y'y.
constexpr double RSS_GUARD = 1e-4;

typedef __attribute__((ext_vector_type(4))) double d4;
typedef __attribute__((ext_vector_type(4))) float f4;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;

__device__ __forceinline__ double load_elem(const void* base, int dt, size_t idx) {
  if (dt == DT_F64) return reinterpret_cast<const double*>(base)[idx];
  if (dt == DT_F32) return (double)reinterpret_cast<const float*>(base)[idx];
  return (double)(float)reinterpret_cast<const _Float16*>(base)[idx];
}

// Four consecutive elements idx..idx+3.  vec: idx is a multiple of 4 and the
// base is 16-byte aligned, so one vector load covers them.
__device__ __forceinline__ void load_elem4(const void* base, int dt, bool vec, size_t idx, double* o) {
  if (vec) {
    if (dt == DT_F32) {
      const float4 v = *reinterpret_cast<const float4*>(reinterpret_cast<const float*>(base) + idx);
      o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else if (dt == DT_F64) {
      const double2 a = *reinterpret_cast<const double2*>(reinterpret_cast<const double*>(base) + idx);
      const double2 b = *reinterpret_cast<const double2*>(reinterpret_cast<const double*>(base) + idx + 2);
      o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    } else {
      const uint2 u = *reinterpret_cast<const uint2*>(reinterpret_cast<const _Float16*>(base) + idx);
      _Float16 h[4];
      __builtin_memcpy(h, &u, sizeof(h));
#pragma unroll
      for (int e = 0; e < 4; ++e) o[e] = (double)(float)h[e];
    }
    return;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) o[e] = load_elem(base, dt, idx + e);
}

// float -> bf16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t bf16_bits_rn(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_bits_to_float(uint16_t b) {
  return __uint_as_float(((uint32_t)b) << 16);
}
// Two-term split x ~= hi + lo with hi = bf16(x), lo = bf16(x - hi).
__device__ __forceinline__ void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
  hi = bf16_bits_rn(x);
  lo = bf16_bits_rn(x - bf16_bits_to_float(hi));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum for blockDim.x == 256 (4 waves).  All threads get the sum.
__device__ __forceinline__ double block_sum256(double v, double* sh /*[4]*/) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = sh[0] + sh[1] + sh[2] + sh[3];
  return s;
}
// Block-wide sums of N values at once (blockDim.x == 256): one pair of
// barriers for all N, and the N shuffle chains interleave.  Each sum is
// formed in the same order as block_sum256's.
template <int N>
__device__ __forceinline__ void block_sum256_n(double* v, double* sh /*[4 * N]*/) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int e = 0; e < N; ++e) v[e] += __shfl_xor(v[e], off, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int e = 0; e < N; ++e) sh[w * N + e] = v[e];
  __syncthreads();
#pragma unroll
  for (int e = 0; e < N; ++e) v[e] = sh[e] + sh[N + e] + sh[2 * N + e] + sh[3 * N + e];
}
__device__ __forceinline__ long long block_sum256_ll(long long v, long long* sh) {
  v = wave_sum_ll(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  return sh[0] + sh[1] + sh[2] + sh[3];
}

// Packed upper-triangular index of (row i, col j), i <= j.
__host__ __device__ __forceinline__ int upk(int i, int j) { return j * (j + 1) / 2 + i; }

__device__ __forceinline__ uint64_t mix64a(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t mix64b(uint64_t x) {
  x = (x + 0x632BE59BD9B4E019ull) * 0xD6E8FEB86659FD93ull;
  x = (x ^ (x >> 32)) * 0xD6E8FEB86659FD93ull;
  return x ^ (x >> 32);
}

}  // namespace oa
