// Device helpers shared by the kernels: amplitude types, index-bit insertion
// and the reference's complex arithmetic order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qcs {

template <typename T>
struct Amp;
template <>
struct Amp<double> {
  typedef double2 type;
};
template <>
struct Amp<float> {
  typedef float2 type;
};

// A 2x2 gate in the precision of the state: {m00, m01, m10, m11} as re/im.
template <typename T>
struct Mat2 {
  T v[8];
};

// Non-temporal (streaming) loads and stores of one amplitude.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double2 ld_stream(const double2 *p) {
  const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 ld_stream(const float2 *p) {
  const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v *>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(double2 *p, double2 x) {
  const d2v v = {x.x, x.y};
  __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(p));
}
__device__ __forceinline__ void st_stream(float2 *p, float2 x) {
  const f2v v = {x.x, x.y};
  __builtin_nontemporal_store(v, reinterpret_cast<f2v *>(p));
}

// XCD-aware work-group order.  The hardware places work-group b on XCD b % 8;
// this maps b to a position so that each XCD's work-groups cover one
// contiguous range of the grid (a bijection of [0, nb) for any nb).
__device__ __forceinline__ uint64_t xcd_order(uint64_t b, uint64_t nb) {
  const uint64_t x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Insert a zero bit at position b of x (pairs of a gate on qubit b).
__device__ __forceinline__ uint64_t insert0(uint64_t x, int b) {
  const uint64_t lo = x & ((uint64_t(1) << b) - 1);
  return ((x >> b) << (b + 1)) | lo;
}

// c_mul of the reference (src/complex.c:47-57):
//   re = ar*br - ai*bi, im = ar*bi + ai*br
// The library is built with -ffp-contract=off so these stay separate
// multiplies and adds, as gcc -std=c89 compiles the reference.
template <typename T>
__device__ __forceinline__ void cmul(T ar, T ai, T br, T bi, T &rr, T &ri) {
  rr = ar * br - ai * bi;
  ri = ar * bi + ai * br;
}

// out = c_add(c_mul(g0, v0), c_mul(g1, v1))   (src/q_gates.c:140-141)
template <typename T>
__device__ __forceinline__ void row2(T g0r, T g0i, T g1r, T g1i, T v0r, T v0i,
                                     T v1r, T v1i, T &orr, T &oi) {
  T ar, ai, br, bi;
  cmul(g0r, g0i, v0r, v0i, ar, ai);
  cmul(g1r, g1i, v1r, v1i, br, bi);
  orr = ar + br;
  oi = ai + bi;
}

// |a|^2 = re*re + im*im in FP64 (c_norm_sq, src/complex.c:89-93)
template <typename A>
__device__ __forceinline__ double norm2(const A &a) {
  const double re = a.x, im = a.y;
  return re * re + im * im;
}

// Wave-level (64 lanes) and block-level FP64 sums with a fixed association,
// so every reduction is deterministic run to run.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Inclusive prefix sum over the 64 lanes of a wave (Hillis-Steele order).
__device__ __forceinline__ double wave_incl_scan(double x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// Sum of v over the work-group; `lds` holds blockDim.x/64 doubles.
// The result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) r += lds[w];
  }
  __syncthreads();
  return r;
}

}  // namespace qcs
