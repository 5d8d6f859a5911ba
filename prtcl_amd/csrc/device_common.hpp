// Device-side building blocks shared by the SPH kernels: vector types,
// the cubic spline kernel in the reference's operation order, the exact
// neighbour predicate, and deterministic FP64 block reductions for CDNA4's
// 64-wide wavefronts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace prtcl_amd {

constexpr int kWave = 64;

template <class T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};
template <class T>
using vec4_t = typename Vec4<T>::type;

template <class V>
__device__ __forceinline__ V make4(decltype(V().x) x, decltype(V().x) y,
                                   decltype(V().x) z, decltype(V().x) w) {
  V r;
  r.x = x;
  r.y = y;
  r.z = z;
  r.w = w;
  return r;
}

// Squared distance with the reference's association d0*d0 + (d1*d1 + d2*d2)
// and no FMA contraction, so the neighbour predicate decides exactly as the
// FP32 CPU code does (grouped_uniform_grid.hpp:415-418).
template <class T>
__device__ __forceinline__ T sqdist_exact(T d0, T d1, T d2) {
#pragma clang fp contract(off)
  return d0 * d0 + (d1 * d1 + d2 * d2);
}

template <class T>
__device__ __forceinline__ T dot3_exact(T a0, T a1, T a2, T b0, T b1, T b2) {
#pragma clang fp contract(off)
  return a0 * b0 + (a1 * b1 + a2 * b2);
}

// Cubic spline kernel with support 2h (Monaghan 1992), evaluated in the
// operation order of cubic_spline_kernel.hpp:43-74 and
// kernel_facade.hpp:204-229: q = |r|/h, W = alpha*((2-q)^3 - 4(1-q)^3)/h^3,
// dW/dr = -|alpha*(12(1-q)^2 - 3(2-q)^2)|/h^4, gradient (dW/dr / r) * x_ij,
// zero for r below sqrt(eps) (constants.hpp:75-78). Division and sqrt are
// IEEE-correctly rounded (HIP's default for FP32 division and sqrt).
template <class T>
struct CubicSpline {
  T h;      // smoothing scale
  T alpha;  // 1/(4 pi) rounded to T
  T eps;    // sqrt(epsilon<T>)

  __device__ __forceinline__ T h3() const {
#pragma clang fp contract(off)
    return (h * h) * h;
  }
  __device__ __forceinline__ T h4() const {
#pragma clang fp contract(off)
    return ((h * h) * h) * h;
  }
  __device__ __forceinline__ T evalq(T q) const {
#pragma clang fp contract(off)
    T result = 0;
    if (q < T(1)) {
      T b = T(1) - q;
      result -= T(4) * ((b * b) * b);
    }
    if (q < T(2)) {
      T b = T(2) - q;
      result += (b * b) * b;
    }
    return alpha * result;
  }
  __device__ __forceinline__ T evaldq(T q) const {
#pragma clang fp contract(off)
    T result = 0;
    if (q < T(1)) {
      T b = T(1) - q;
      result += T(12) * (b * b);
    }
    if (q < T(2)) {
      T b = T(2) - q;
      result -= T(3) * (b * b);
    }
    return alpha * result;
  }
  // W from the squared distance
  __device__ __forceinline__ T w(T r2) const {
    T r = sqrt(r2);
    return evalq(fabs(r) / h) / h3();
  }
  // factor f with grad W = f * x_ij
  __device__ __forceinline__ T gradf(T r2) const {
#pragma clang fp contract(off)
    T r = sqrt(r2);
    if (r < eps)
      return T(0);
    T dr = -copysign(evaldq(fabs(r) / h) / h4(), r);
    return dr / r;
  }
};

// ---------------------------------------------------------------------------
// deterministic reductions (fixed tree: wave shuffles then one wave over the
// per-wave partials), valid in thread 0 of the block
// ---------------------------------------------------------------------------

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1)
    v += __shfl_down(v, o, kWave);
  return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T u = __shfl_down(v, o, kWave);
    v = u > v ? u : v;
  }
  return v;
}

template <int BS, class T>
__device__ __forceinline__ T block_sum(T v) {
  static_assert(BS % kWave == 0 && BS / kWave <= kWave, "block size");
  __shared__ T ws[BS / kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  v = wave_sum(v);
  if (lane == 0)
    ws[w] = v;
  __syncthreads();
  T r = T(0);
  if (w == 0) {
    r = lane < BS / kWave ? ws[lane] : T(0);
    r = wave_sum(r);
  }
  __syncthreads();
  return r;
}

template <int BS, class T>
__device__ __forceinline__ T block_max(T v) {
  __shared__ T ws[BS / kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  v = wave_max(v);
  if (lane == 0)
    ws[w] = v;
  __syncthreads();
  T r = v;
  if (w == 0) {
    r = lane < BS / kWave ? ws[lane] : ws[0];
    r = wave_max(r);
  }
  __syncthreads();
  return r;
}

// A thread's share of the block partials p[0, n): the terms k = threadIdx.x,
// threadIdx.x + BS, ... in that order, summed (max: their maximum, at least
// m0). The loads of a batch of 16 terms are all issued before the first add,
// so a thread waits for one memory round trip per batch and not per term
// (the partials come from other XCDs, a miss in this XCD's L2). The order of
// the adds is that of the plain strided loop, so the sums are bit-identical.
template <int BS, class T>
__device__ __forceinline__ T strided_sum(const T *p, unsigned n) {
  constexpr unsigned kBatch = 16;
  T acc = T(0);
  for (unsigned k0 = threadIdx.x; k0 < n; k0 += kBatch * BS) {
    T t[kBatch];
#pragma unroll
    for (unsigned u = 0; u < kBatch; ++u) {
      const unsigned k = k0 + u * BS;
      t[u] = p[k < n ? k : n - 1];
    }
#pragma unroll
    for (unsigned u = 0; u < kBatch; ++u)
      if (k0 + u * BS < n)
        acc += t[u];
  }
  return acc;
}

template <int BS>
__device__ __forceinline__ double strided_max(const double *p, unsigned n, double m0) {
  constexpr unsigned kBatch = 16;
  double m = m0;
  for (unsigned k0 = threadIdx.x; k0 < n; k0 += kBatch * BS) {
    double t[kBatch];
#pragma unroll
    for (unsigned u = 0; u < kBatch; ++u) {
      const unsigned k = k0 + u * BS;
      t[u] = p[k < n ? k : n - 1];
    }
#pragma unroll
    for (unsigned u = 0; u < kBatch; ++u)
      if (k0 + u * BS < n)
        m = t[u] > m ? t[u] : m;
  }
  return m;
}

}  // namespace prtcl_amd
