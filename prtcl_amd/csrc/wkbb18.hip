// WKBB18 implicit viscosity (Weiler et al. 2018) on the device: the 3x3
// block-diagonal system matrix and its inverse (wkbb18.hpp:194-270) and the
// preconditioned CG of cg_openmp.hpp:32-215 with the matrix-free product of
// wkbb18.hpp:342-390.
//
// Two product kernels share the CG engine:
//   * exact: walks the 27-cell grid with the reference predicate and
//     evaluates every pair term in the reference's FP32 operation order
//     (used by the wkbb18.* procedures);
//   * list: walks the step's fluid neighbour list with fused arithmetic and
//     applies the boundary part as a precomputed symmetric 3x3 matrix per
//     particle, B_i = sum_b k_ib gradW_ib x_ib^T (used by prtcl_hip_step).
// The CG scalars (dot products, alpha, beta, thresholds) are FP64 on the
// device; vectors are the scene's real type. The iteration logic follows the
// reference exactly: r = A x - b, p0 = -y0, break on |p.q| < tol and on
// |r.y|_prev < tol (absolute), stop when |r|^2 < max(tol^2 |b|^2, FLT_MIN),
// at most max_k + 1 iterations, the returned count is the loop counter.
#include "wkbb18.hpp"
#include "cg_ctl.hpp"
#include "exact_walk.hpp"

namespace prtcl_amd {

namespace {

template <class T>
__device__ __forceinline__ T cof3(const T *m, int i, int j) {
#pragma clang fp contract(off)
  const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return m[i1 + 3 * j1] * m[i2 + 3 * j2] - m[i1 + 3 * j2] * m[i2 + 3 * j1];
}

// inverse by adjugate / determinant, column-major (element (r,c) at r + 3c)
template <class T>
__device__ __forceinline__ void invert3(const T *m, T *out) {
#pragma clang fp contract(off)
  const T c00 = cof3(m, 0, 0), c10 = cof3(m, 1, 0), c20 = cof3(m, 2, 0);
  const T det = c00 * m[0] + (c10 * m[1] + c20 * m[2]);
  const T invdet = T(1) / det;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      out[r + 3 * c] = cof3(m, c, r) * invdet;
}

template <class T>
__device__ __forceinline__ T step_l(T eps, T x) {
  return (eps < x) ? T(1) : T(0);
}

template <class T>
struct RealEps;
template <>
struct RealEps<float> {
  static constexpr double value = 1.1920928955078125e-07;  // 2^-23
};
template <>
struct RealEps<double> {
  static constexpr double value = 2.220446049250313e-16;  // 2^-52
};

// Moore-Penrose pseudo-inverse of a 3x3 matrix m (column-major) into out, as
// invert_pm (eigen.hpp:266-276). The reference takes it from Eigen's complete
// orthogonal decomposition, whose rank counts the pivots above 3 eps of the
// largest; here a one-sided Jacobi SVD in FP64 (Hestenes rotations of the
// column pairs until they are orthogonal to working precision, at most 12
// sweeps) gives A V = B with orthogonal columns b_k = sigma_k u_k, and
// pinv(A) = sum over sigma_k > 3 eps sigma_max of v_k b_k^T / sigma_k^2.
//
// Fast path: sigma_min / sigma_max >= |det| / sigma_max^3 >= |det| / |A|_F^3,
// so |det| > 2 (3 eps) |A|_F^3 proves the matrix full rank under the
// threshold, and then the pseudo-inverse is the inverse (adjugate / det in
// FP64). Only the other matrices take the SVD.
template <class T>
__device__ void pinv3(const T *m, T *out) {
  double b[9], v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
  for (int k = 0; k < 9; ++k)
    b[k] = static_cast<double>(m[k]);
  {
    double cof[9], fro2 = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        cof[r + 3 * c] = b[r1 + 3 * c1] * b[r2 + 3 * c2] - b[r1 + 3 * c2] * b[r2 + 3 * c1];
        fro2 += b[r + 3 * c] * b[r + 3 * c];
      }
    const double det = b[0] * cof[0] + b[1] * cof[1] + b[2] * cof[2];
    if (fabs(det) > 6.0 * RealEps<T>::value * fro2 * sqrt(fro2)) {
      const double id = 1.0 / det;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
          out[r + 3 * c] = static_cast<T>(cof[c + 3 * r] * id);  // adj(r, c) = cof(c, r)
      return;
    }
  }
  for (int sweep = 0; sweep < 12; ++sweep) {
    bool rotated = false;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      double al = 0, be = 0, ga = 0;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        al += b[r + 3 * p] * b[r + 3 * p];
        be += b[r + 3 * q] * b[r + 3 * q];
        ga += b[r + 3 * p] * b[r + 3 * q];
      }
      if (!(fabs(ga) > 1e-17 * sqrt(al * be)))
        continue;
      rotated = true;
      const double zeta = (be - al) / (2.0 * ga);
      const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double bp = b[r + 3 * p], bq = b[r + 3 * q];
        b[r + 3 * p] = c * bp - sn * bq;
        b[r + 3 * q] = sn * bp + c * bq;
        const double vp = v[r + 3 * p], vq = v[r + 3 * q];
        v[r + 3 * p] = c * vp - sn * vq;
        v[r + 3 * q] = sn * vp + c * vq;
      }
    }
    if (!rotated)
      break;
  }
  double s2[3], smax2 = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s2[k] = b[3 * k] * b[3 * k] + b[3 * k + 1] * b[3 * k + 1] + b[3 * k + 2] * b[3 * k + 2];
    smax2 = s2[k] > smax2 ? s2[k] : smax2;
  }
  const double thr = 3.0 * RealEps<T>::value;
  const double thr2 = thr * thr * smax2;
  double inv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    inv[k] = (s2[k] > thr2 && s2[k] > 0.0) ? 1.0 / s2[k] : 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      out[r + 3 * c] = static_cast<T>(v[r] * b[c] * inv[0] + v[r + 3] * b[c + 3] * inv[1] +
                                      v[r + 6] * b[c + 6] * inv[2]);
}

// ---- correction.identity_gradient_correction (correction.hpp:133-158)
template <class T>
__global__ __launch_bounds__(kBlock) void k_corr_identity(SceneView<T> S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nf)
    return;
  T *L = S.gc + 9 * static_cast<size_t>(i);
#pragma unroll
  for (int k = 0; k < 9; ++k)
    L[k] = (k % 4 == 0) ? T(1) : T(0);
}

// ---- correction.compute_gradient_correction (correction.hpp:161-217):
// L_i = pinv(sum_f (m_j / rho_i) gradW_ij (x_j - x_i)^T), exact walk in the
// reference's FP32 operation order, or over the step's neighbour list
template <class T, bool kList>
__device__ __forceinline__ void corr_body(const SceneView<T> &S, const NbrList &L,
                                          unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i];
  const T rhoi = S.rho[i];
  T M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  auto add = [&](const vec4_t<T> &pj, T f, T d0, T d1, T d2) {
    const T g[3] = {f * d0, f * d1, f * d2};
    const T e[3] = {pj.x - xi.x, pj.y - xi.y, pj.z - xi.z};
    const T k = pj.w / rhoi;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r)
        M[r + 3 * c] += k * (g[r] * e[c]);
  };
  const uint32_t info = kList ? L.info[i] : kListOverflow;
  if (kList && !(info & kListOverflow)) {
    list_visit(L, Spans::of(L.blk[i / kBlock]), static_cast<size_t>(S.nf), i, info,
               [&](uint32_t j) {
                 const vec4_t<T> pj = S.xm[j];
                 const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
                 add(pj, S.K.gradf(sqdist_exact(d0, d1, d2)), d0, d1, d2);
               });
  } else {
    nbrs_exact(S, xi, S.fcell[i], S.fstart, S.xm,
               [&](uint32_t, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                 add(pj, S.K.gradf(r2), d0, d1, d2);
               });
  }
  pinv3(M, S.gc + 9 * static_cast<size_t>(i));
}

template <class T, bool kList>
__global__ __launch_bounds__(kBlock) void k_corr(SceneView<T> S, NbrList L) {
  corr_body<T, kList>(S, L, xcd_block(blockIdx.x, gridDim.x));
}

// L_i g with Eigen's fixed-size product order per row: l0 g0 + (l1 g1 + l2 g2)
template <class T>
__device__ __forceinline__ void apply_gc(const T *L, const T g[3], T out[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int r = 0; r < 3; ++r)
    out[r] = L[r] * g[0] + (L[r + 3] * g[1] + L[r + 6] * g[2]);
}

// ---- wkbb18.compute_diagonal, exact (wkbb18.hpp:194-270); with kGC the
// gradient is L_i gradW (wkbb18_gc.hpp:236-262)
template <class T, bool kGC>
__global__ __launch_bounds__(kBlock) void k_diag_exact(SceneView<T> S) {
#pragma clang fp contract(off)
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i];
  const T rhoi = S.rho[i];
  const uint32_t cell = S.fcell[i];
  T Li[9];
  if (kGC) {
#pragma unroll
    for (int k = 0; k < 9; ++k)
      Li[k] = S.gc[9 * static_cast<size_t>(i) + k];
  }
  auto grad = [&](T f, T d0, T d1, T d2, T g[3]) {
    const T g0[3] = {f * d0, f * d1, f * d2};
    if (kGC) {
      apply_gc(Li, g0, g);
    } else {
      g[0] = g0[0];
      g[1] = g0[1];
      g[2] = g0[2];
    }
  };
  T D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  nbrs_exact(S, xi, cell, S.fstart, S.xm, [&](uint32_t j, T d0, T d1, T d2, T r2,
                                                const vec4_t<T> &pj) {
    const T f = S.K.gradf(r2);
    T g[3];
    grad(f, d0, d1, d2, g);
    const T d[3] = {d0, d1, d2};
    const T k = ((T(10) * (S.mu_f / rhoi)) * (pj.w / S.rho[j])) / (r2 + S.hh001);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r)
        D[r + 3 * c] += k * (g[r] * d[c]);
  });
  nbrs_exact(S, xi, cell, S.bstart, S.bxV, [&](uint32_t, T d0, T d1, T d2, T r2,
                                                 const vec4_t<T> &pj) {
    const T f = S.K.gradf(r2);
    T g[3];
    grad(f, d0, d1, d2, g);
    const T d[3] = {d0, d1, d2};
    const T k = ((T(10) * (S.mu_b / rhoi)) * ((S.rho0 * pj.w) / rhoi)) / (r2 + S.hh001);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r)
        D[r + 3 * c] += k * (g[r] * d[c]);
  });
  T M[9];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      M[r + 3 * c] = ((r == c) ? T(1) : T(0)) - S.dt * D[r + 3 * c];
  invert3(M, S.sp + 9 * static_cast<size_t>(i));
}

// ---- product_s, exact (wkbb18.hpp:342-390): out = w - dt * l(w); with kGC
// every term is (s L_i) gradW (wkbb18_gc.hpp:362-390)
template <class T, bool kGC>
__device__ __forceinline__ void product_exact(const SceneView<T> &S, int i,
                                              const vec4_t<T> *__restrict__ w, int ws,
                                              T out[3]) {
#pragma clang fp contract(off)
  const vec4_t<T> xi = S.xm[i];
  const vec4_t<T> wi = w[static_cast<size_t>(ws) * i];
  const T rhoi = S.rho[i];
  const uint32_t cell = S.fcell[i];
  T l0 = 0, l1 = 0, l2 = 0;
  T Li[9];
  if (kGC) {
#pragma unroll
    for (int k = 0; k < 9; ++k)
      Li[k] = S.gc[9 * static_cast<size_t>(i) + k];
  }
  auto acc = [&](T s, T f, T d0, T d1, T d2) {
    if (kGC) {
      const T g[3] = {f * d0, f * d1, f * d2};
      T sl[9];
#pragma unroll
      for (int k = 0; k < 9; ++k)
        sl[k] = s * Li[k];
      T t[3];
      apply_gc(sl, g, t);
      l0 += t[0];
      l1 += t[1];
      l2 += t[2];
    } else {
      l0 += s * (f * d0);
      l1 += s * (f * d1);
      l2 += s * (f * d2);
    }
  };
  nbrs_exact(S, xi, cell, S.fstart, S.xm, [&](uint32_t j, T d0, T d1, T d2, T r2,
                                                const vec4_t<T> &pj) {
    const vec4_t<T> wj = w[static_cast<size_t>(ws) * j];
    const T f = S.K.gradf(r2);
    const T s = (((T(10) * (S.mu_f / rhoi)) * (pj.w / S.rho[j])) *
                 dot3_exact(wi.x - wj.x, wi.y - wj.y, wi.z - wj.z, d0, d1, d2)) /
                (r2 + S.hh001);
    acc(s, f, d0, d1, d2);
  });
  nbrs_exact(S, xi, cell, S.bstart, S.bxV, [&](uint32_t, T d0, T d1, T d2, T r2,
                                                 const vec4_t<T> &pj) {
    const T f = S.K.gradf(r2);
    const T s = (((T(10) * (S.mu_b / rhoi)) * ((S.rho0 * pj.w) / rhoi)) *
                 dot3_exact(wi.x, wi.y, wi.z, d0, d1, d2)) /
                (r2 + S.hh001);
    acc(s, f, d0, d1, d2);
  });
  out[0] = wi.x - S.dt * l0;
  out[1] = wi.y - S.dt * l1;
  out[2] = wi.z - S.dt * l2;
}

// ---- fast path: diagonal over the list, B_i, and (x, m/rho) packing
template <class T>
__device__ __forceinline__ T gradf_fast(const SceneView<T> &S, T r2) {
  // same closed form as the fused IISPH kernels: 3 alpha/h^4 (u^2 - t^2)/r
  const T h = S.K.h;
  const T inv_h = T(1) / h;
  const T rinv = T(1) / sqrt(r2);
  const T q = r2 * rinv * inv_h;
  const T t = T(2) - q;
  T u = T(2) - T(2) * q;
  u = u > T(0) ? u : T(0);
  const T ag3 = T(3) * S.K.alpha / (h * h * h * h);
  return r2 < S.K.eps * S.K.eps ? T(0) : ag3 * (u * u - t * t) * rinv;
}

// with kGC: M = I - dt L_i (D + B), and the product applies L_i to the summed
// vector (equal in real arithmetic to applying it per pair)
template <class T, bool kGC>
__global__ __launch_bounds__(kBlock) void k_diag_list(SceneView<T> S, NbrList L) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i];
  const T rhoi = S.rho[i];
  const size_t n = static_cast<size_t>(S.nf);
  T D[6] = {0, 0, 0, 0, 0, 0};  // symmetric: xx yy zz xy xz yz
  auto add = [&](T k, T d0, T d1, T d2) {
    D[0] += k * d0 * d0;
    D[1] += k * d1 * d1;
    D[2] += k * d2 * d2;
    D[3] += k * d0 * d1;
    D[4] += k * d0 * d2;
    D[5] += k * d1 * d2;
  };
  const T cf = T(10) * S.mu_f / rhoi;
  auto fluid = [&](uint32_t j) {
    const vec4_t<T> pj = S.xm[j];
    const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
    const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
    add(cf * pj.w / S.rho[j] * gradf_fast(S, r2) / (r2 + S.hh001), d0, d1, d2);
  };
  const uint32_t info = L.info[i];
  if (!(info & kListOverflow)) {
    list_visit(L, Spans::of(L.blk[i / kBlock]), n, i, info, fluid);
  } else {
    nbrs_exact(S, xi, S.fcell[i], S.fstart, S.xm,
               [&](uint32_t j, T, T, T, T, const vec4_t<T> &) {
                 if (j != static_cast<uint32_t>(i))
                   fluid(j);
               });
  }
  T B[6] = {0, 0, 0, 0, 0, 0};
  const T cb = T(10) * S.mu_b / rhoi * S.rho0 / rhoi;
  nbrs_exact(S, xi, S.fcell[i], S.bstart, S.bxV, [&](uint32_t, T d0, T d1, T d2, T r2e,
                                                       const vec4_t<T> &pj) {
    (void)r2e;
    const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
    const T k = cb * pj.w * gradf_fast(S, r2) / (r2 + S.hh001);
    B[0] += k * d0 * d0;
    B[1] += k * d1 * d1;
    B[2] += k * d2 * d2;
    B[3] += k * d0 * d1;
    B[4] += k * d0 * d2;
    B[5] += k * d1 * d2;
  });
  T M[9];
  const T dt = S.dt;
  const T Dt[6] = {D[0] + B[0], D[1] + B[1], D[2] + B[2], D[3] + B[3], D[4] + B[4], D[5] + B[5]};
  if (kGC) {
    const T Ds[9] = {Dt[0], Dt[3], Dt[4], Dt[3], Dt[1], Dt[5], Dt[4], Dt[5], Dt[2]};
    const T *Li = S.gc + 9 * static_cast<size_t>(i);
    T Lr[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
      Lr[k] = Li[k];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r)
        M[r + 3 * c] = ((r == c) ? T(1) : T(0)) -
                       dt * (Lr[r] * Ds[3 * c] + Lr[r + 3] * Ds[1 + 3 * c] +
                             Lr[r + 6] * Ds[2 + 3 * c]);
  } else {
    M[0] = T(1) - dt * Dt[0];
    M[4] = T(1) - dt * Dt[1];
    M[8] = T(1) - dt * Dt[2];
    M[1] = M[3] = -dt * Dt[3];
    M[2] = M[6] = -dt * Dt[4];
    M[5] = M[7] = -dt * Dt[5];
  }
  invert3(M, S.sp + 9 * static_cast<size_t>(i));
  // B_i for the list product, (x_i, m_i / rho_i) for its gathers
  S.cb[i] = make4<vec4_t<T>>(B[0], B[1], B[2], T(0));
  S.G[i] = make4<vec4_t<T>>(B[3], B[4], B[5], T(0));
  const vec4_t<T> q = make4<vec4_t<T>>(xi.x, xi.y, xi.z, xi.w / rhoi);
  S.xq[i] = q;
  S.cxp[2 * static_cast<size_t>(i)] = q;
}

template <class T>
__device__ __forceinline__ T rsqrt_hw(T x);
template <>
__device__ __forceinline__ float rsqrt_hw<float>(float x) {
  return __builtin_amdgcn_rsqf(x);
}
template <>
__device__ __forceinline__ double rsqrt_hw<double>(double x) {
  return rsqrt(x);
}
template <class T>
__device__ __forceinline__ T rcp_hw(T x);
template <>
__device__ __forceinline__ float rcp_hw<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}
template <>
__device__ __forceinline__ double rcp_hw<double>(double x) {
  return 1.0 / x;
}

// one fluid pair of the list product: l += (m_j / rho_j) ((w_i - w_j) . d)
// f(r) / (r^2 + 0.01 h^2) d with d = x_i - x_j; xq = (x_i, m_i / rho_i),
// qx = (x_j, m_j / rho_j)
template <class T>
struct ProductPair {
  T inv_h, ag3, eps2, hh001;
  __device__ __forceinline__ explicit ProductPair(const SceneView<T> &S)
      : inv_h(T(1) / S.K.h), ag3(T(3) * S.K.alpha / (S.K.h * S.K.h * S.K.h * S.K.h)),
        eps2(S.K.eps * S.K.eps), hh001(S.hh001) {}
  __device__ __forceinline__ void operator()(const vec4_t<T> &xq, const vec4_t<T> &wi,
                                             const vec4_t<T> &qx, const vec4_t<T> &wj, T &l0,
                                             T &l1, T &l2) const {
    const T d0 = xq.x - qx.x, d1 = xq.y - qx.y, d2 = xq.z - qx.z;
    const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
    const T rinv = rsqrt_hw(r2);
    const T qh = r2 * rinv * inv_h;
    const T t = T(2) - qh;
    T u = T(2) - T(2) * qh;
    u = u > T(0) ? u : T(0);
    const T f = r2 < eps2 ? T(0) : ag3 * (u * u - t * t) * rinv;
    const T s = qx.w * ((wi.x - wj.x) * d0 + (wi.y - wj.y) * d1 + (wi.z - wj.z) * d2) * f *
                rcp_hw(r2 + hh001);
    l0 += s * d0;
    l1 += s * d1;
    l2 += s * d2;
  }
};

// out = w_i - dt (10 mu_f / rho_i l + B_i w_i) from the fluid pair sum l
// and B_i = (bd: xx yy zz, bo: xy xz yz); with kGC L_i applies to the bracket
template <class T, bool kGC>
__device__ __forceinline__ void product_finish(const SceneView<T> &S, int i,
                                               const vec4_t<T> &wi, T l0, T l1, T l2,
                                               const vec4_t<T> &bd, const vec4_t<T> &bo,
                                               T out[3]) {
  const T cf = T(10) * S.mu_f / S.rho[i];
  const T b0 = bd.x * wi.x + bo.x * wi.y + bo.y * wi.z;
  const T b1 = bo.x * wi.x + bd.y * wi.y + bo.z * wi.z;
  const T b2 = bo.y * wi.x + bo.z * wi.y + bd.z * wi.z;
  if (kGC) {
    const T *Li = S.gc + 9 * static_cast<size_t>(i);
    const T u0 = cf * l0 + b0, u1 = cf * l1 + b1, u2 = cf * l2 + b2;
    out[0] = wi.x - S.dt * (Li[0] * u0 + Li[3] * u1 + Li[6] * u2);
    out[1] = wi.y - S.dt * (Li[1] * u0 + Li[4] * u1 + Li[7] * u2);
    out[2] = wi.z - S.dt * (Li[2] * u0 + Li[5] * u1 + Li[8] * u2);
  } else {
    out[0] = wi.x - S.dt * (cf * l0 + b0);
    out[1] = wi.y - S.dt * (cf * l1 + b1);
    out[2] = wi.z - S.dt * (cf * l2 + b2);
  }
}

template <class T, bool kGC>
__device__ __forceinline__ void product_list(const SceneView<T> &S, const NbrList &L, int i,
                                             const vec4_t<T> *__restrict__ w, int ws,
                                             T out[3]) {
  const vec4_t<T> xq = S.xq[i];  // (x_i, m_i / rho_i)
  const vec4_t<T> wi = w[static_cast<size_t>(ws) * i];
  const size_t n = static_cast<size_t>(S.nf);
  const ProductPair<T> pair(S);
  T l0 = 0, l1 = 0, l2 = 0;
  struct XW {
    vec4_t<T> x, w;  // (x_j, m_j / rho_j) and w_j
  };
  auto use = [&](const XW &q) { pair(xq, wi, q.x, q.w, l0, l1, l2); };
  const uint32_t info = L.info[i];
  if (!(info & kListOverflow)) {
    list_gather<4>(L, Spans::of(L.blk[i / kBlock]), n, i, info,
                   [&](uint32_t j) {
                     return XW{S.cxp[2 * static_cast<size_t>(j)], w[static_cast<size_t>(ws) * j]};
                   },
                   use);
  } else {
    nbrs_exact(S, xq, S.fcell[i], S.fstart, S.xq,
               [&](uint32_t j, T, T, T, T, const vec4_t<T> &) {
                 if (j != static_cast<uint32_t>(i))
                   use(XW{S.xq[j], w[static_cast<size_t>(ws) * j]});
               });
  }
  product_finish<T, kGC>(S, i, wi, l0, l1, l2, S.cb[i], S.G[i], out);
}

template <class T>
__device__ __forceinline__ T dot3r(const vec4_t<T> &a, const vec4_t<T> &b) {
#pragma clang fp contract(off)
  return a.x * b.x + (a.y * b.y + a.z * b.z);  // sum(cmul(l, r)) per item
}

template <class T>
__device__ __forceinline__ vec4_t<T> precond(const T *M, const vec4_t<T> &w) {
#pragma clang fp contract(off)
  vec4_t<T> y;
  y.x = M[0] * w.x + (M[3] * w.y + M[6] * w.z);
  y.y = M[1] * w.x + (M[4] * w.y + M[7] * w.z);
  y.z = M[2] * w.x + (M[5] * w.y + M[8] * w.z);
  y.w = T(0);
  return y;
}

// setup: x = v (guess), r = A x - b with b = v (the rhs of wkbb18.hpp:296-334
// is v - dt a_b with a_b = 0 because the boundary velocity is zero),
// y = P r, p = -y; partials of b.b, r.r, r.y
template <class T, bool kExact, bool kGC>
__global__ __launch_bounds__(kBlock) void k_cg_setup(SceneView<T> S, NbrList L, double *part) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < S.nf) {
    const vec4_t<T> v = S.v[i];
    T ax[3];
    if (kExact)
      product_exact<T, kGC>(S, i, S.v, 1, ax);
    else
      product_list<T, kGC>(S, L, i, S.v, 1, ax);
    vec4_t<T> b = v;
    b.w = T(0);
    vec4_t<T> r;
    r.x = ax[0] - b.x;
    r.y = ax[1] - b.y;
    r.z = ax[2] - b.z;
    r.w = T(0);
    const vec4_t<T> y = precond(S.sp + 9 * static_cast<size_t>(i), r);
    S.cx[i] = b;
    S.cr[i] = r;
    S.cy[i] = y;
    S.cxp[2 * static_cast<size_t>(i) + 1] = make4<vec4_t<T>>(-y.x, -y.y, -y.z, T(0));
    acc[0] = static_cast<double>(dot3r<T>(b, b));
    acc[1] = static_cast<double>(dot3r<T>(r, r));
    acc[2] = static_cast<double>(dot3r<T>(r, y));
  }
  store_partials<3>(part, gridDim.x, acc, lb);
}

// k_diag_list and the list form of k_cg_setup in one walk over the list
// (fused steps): every pair term is the expression of the two kernels, in the
// same order, with (x_j, m_j / rho_j) formed from x_j, m_j and rho_j as the
// diagonal pass writes it, so the results equal those of the two kernels
// (lb: the block's logical index, nb: the scene's block count)
template <class T, bool kGC>
__device__ __forceinline__ void diag_setup_list_body(const SceneView<T> &S, const NbrList &L,
                                                     double *part, unsigned lb, unsigned nb) {
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < S.nf) {
    const vec4_t<T> xi = S.xm[i];
    const T rhoi = S.rho[i];
    const vec4_t<T> wi = S.v[i];  // the guess x0 = v
    const vec4_t<T> xqi = make4<vec4_t<T>>(xi.x, xi.y, xi.z, xi.w / rhoi);
    const size_t n = static_cast<size_t>(S.nf);
    T D[6] = {0, 0, 0, 0, 0, 0};  // symmetric: xx yy zz xy xz yz
    const T cf = T(10) * S.mu_f / rhoi;
    const ProductPair<T> pair(S);
    T l0 = 0, l1 = 0, l2 = 0;
    struct P {
      vec4_t<T> x;  // (x_j, m_j)
      T rho;
      vec4_t<T> w;  // v_j
    };
    auto both = [&](const P &q) {
      const T d0 = xi.x - q.x.x, d1 = xi.y - q.x.y, d2 = xi.z - q.x.z;
      const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const T k = cf * q.x.w / q.rho * gradf_fast(S, r2) / (r2 + S.hh001);
      D[0] += k * d0 * d0;
      D[1] += k * d1 * d1;
      D[2] += k * d2 * d2;
      D[3] += k * d0 * d1;
      D[4] += k * d0 * d2;
      D[5] += k * d1 * d2;
      pair(xqi, wi, make4<vec4_t<T>>(q.x.x, q.x.y, q.x.z, q.x.w / q.rho), q.w, l0, l1, l2);
    };
    const uint32_t info = L.info[i];
    if (!(info & kListOverflow)) {
      list_gather<4>(L, Spans::of(L.blk[i / kBlock]), n, i, info,
                     [&](uint32_t j) { return P{S.xm[j], S.rho[j], S.v[j]}; }, both);
    } else {
      nbrs_exact(S, xi, S.fcell[i], S.fstart, S.xm,
                 [&](uint32_t j, T, T, T, T, const vec4_t<T> &) {
                   if (j != static_cast<uint32_t>(i))
                     both(P{S.xm[j], S.rho[j], S.v[j]});
                 });
    }
    T B[6] = {0, 0, 0, 0, 0, 0};
    const T cb = T(10) * S.mu_b / rhoi * S.rho0 / rhoi;
    nbrs_exact(S, xi, S.fcell[i], S.bstart, S.bxV, [&](uint32_t, T d0, T d1, T d2, T r2e,
                                                         const vec4_t<T> &pj) {
      (void)r2e;
      const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const T k = cb * pj.w * gradf_fast(S, r2) / (r2 + S.hh001);
      B[0] += k * d0 * d0;
      B[1] += k * d1 * d1;
      B[2] += k * d2 * d2;
      B[3] += k * d0 * d1;
      B[4] += k * d0 * d2;
      B[5] += k * d1 * d2;
    });
    T M[9];
    const T dt = S.dt;
    const T Dt[6] = {D[0] + B[0], D[1] + B[1], D[2] + B[2],
                     D[3] + B[3], D[4] + B[4], D[5] + B[5]};
    if (kGC) {
      const T Ds[9] = {Dt[0], Dt[3], Dt[4], Dt[3], Dt[1], Dt[5], Dt[4], Dt[5], Dt[2]};
      const T *Li = S.gc + 9 * static_cast<size_t>(i);
      T Lr[9];
#pragma unroll
      for (int k = 0; k < 9; ++k)
        Lr[k] = Li[k];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
          M[r + 3 * c] = ((r == c) ? T(1) : T(0)) -
                         dt * (Lr[r] * Ds[3 * c] + Lr[r + 3] * Ds[1 + 3 * c] +
                               Lr[r + 6] * Ds[2 + 3 * c]);
    } else {
      M[0] = T(1) - dt * Dt[0];
      M[4] = T(1) - dt * Dt[1];
      M[8] = T(1) - dt * Dt[2];
      M[1] = M[3] = -dt * Dt[3];
      M[2] = M[6] = -dt * Dt[4];
      M[5] = M[7] = -dt * Dt[5];
    }
    T *spi = S.sp + 9 * static_cast<size_t>(i);
    invert3(M, spi);
    const vec4_t<T> bd = make4<vec4_t<T>>(B[0], B[1], B[2], T(0));
    const vec4_t<T> bo = make4<vec4_t<T>>(B[3], B[4], B[5], T(0));
    S.cb[i] = bd;
    S.G[i] = bo;
    S.xq[i] = xqi;
    S.cxp[2 * static_cast<size_t>(i)] = xqi;
    // the CG set-up (k_cg_setup)
    T ax[3];
    product_finish<T, kGC>(S, i, wi, l0, l1, l2, bd, bo, ax);
    vec4_t<T> b = wi;
    b.w = T(0);
    vec4_t<T> r;
    r.x = ax[0] - b.x;
    r.y = ax[1] - b.y;
    r.z = ax[2] - b.z;
    r.w = T(0);
    const vec4_t<T> y = precond(spi, r);
    S.cx[i] = b;
    S.cr[i] = r;
    S.cy[i] = y;
    S.cxp[2 * static_cast<size_t>(i) + 1] = make4<vec4_t<T>>(-y.x, -y.y, -y.z, T(0));
    acc[0] = static_cast<double>(dot3r<T>(b, b));
    acc[1] = static_cast<double>(dot3r<T>(r, r));
    acc[2] = static_cast<double>(dot3r<T>(r, y));
  }
  store_partials<3>(part, nb, acc, lb);
}

template <class T, bool kGC>
__global__ __launch_bounds__(kBlock) void k_diag_setup_list(SceneView<T> S, NbrList L,
                                                            double *part) {
  diag_setup_list_body<T, kGC>(S, L, part, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}

// q = A p, partial p.q
template <class T, bool kExact, bool kGC>
__device__ __forceinline__ void cg_spmv_body(const SceneView<T> &S, const NbrList &L,
                                             double *part, unsigned lb, unsigned nb) {
  if (S.ctl->cg_done)
    return;
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double acc[1] = {0.0};
  if (i < S.nf) {
    T q[3];
    if (kExact)
      product_exact<T, kGC>(S, i, S.cxp + 1, 2, q);
    else
      product_list<T, kGC>(S, L, i, S.cxp + 1, 2, q);
    const vec4_t<T> qv = make4<vec4_t<T>>(q[0], q[1], q[2], T(0));
    S.cq[i] = qv;
    acc[0] = static_cast<double>(dot3r<T>(S.cxp[2 * static_cast<size_t>(i) + 1], qv));
  }
  store_partials<1>(part, nb, acc, lb);
}

template <class T, bool kExact, bool kGC>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(SceneView<T> S, NbrList L, double *part) {
  cg_spmv_body<T, kExact, kGC>(S, L, part, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}

// x += alpha p, r += alpha q, y = P r; partials r.r, r.y
template <class T>
__device__ __forceinline__ void cg_update_body(const SceneView<T> &S, T alpha, double *part,
                                               unsigned lb, unsigned nb) {
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double acc[2] = {0.0, 0.0};
  if (i < S.nf) {
    const vec4_t<T> p = S.cxp[2 * static_cast<size_t>(i) + 1], q = S.cq[i];
    vec4_t<T> x = S.cx[i], r = S.cr[i];
    x.x += alpha * p.x;
    x.y += alpha * p.y;
    x.z += alpha * p.z;
    r.x += alpha * q.x;
    r.y += alpha * q.y;
    r.z += alpha * q.z;
    const vec4_t<T> y = precond(S.sp + 9 * static_cast<size_t>(i), r);
    S.cx[i] = x;
    S.cr[i] = r;
    S.cy[i] = y;
    acc[0] = static_cast<double>(dot3r<T>(r, r));
    acc[1] = static_cast<double>(dot3r<T>(r, y));
  }
  store_partials<2>(part, nb, acc, lb);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_update(SceneView<T> S, double *part) {
  if (S.ctl->cg_done)
    return;
  cg_update_body(S, static_cast<T>(S.ctl->cg_alpha), part, blockIdx.x, gridDim.x);
}

// p = -y + beta p
template <class T>
__device__ __forceinline__ void cg_direction_body(const SceneView<T> &S, T beta, unsigned lb) {
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> y = S.cy[i];
  vec4_t<T> p = S.cxp[2 * static_cast<size_t>(i) + 1];
  p.x = -y.x + beta * p.x;
  p.y = -y.y + beta * p.y;
  p.z = -y.z + beta * p.z;
  S.cxp[2 * static_cast<size_t>(i) + 1] = p;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_direction(SceneView<T> S) {
  if (S.ctl->cg_done)
    return;
  cg_direction_body(S, static_cast<T>(S.ctl->cg_beta), blockIdx.x);
}

// The iteration without its two one-block kernels (cg_alpha_in_block,
// cg_beta_in_block): k_cg_alpha with k_cg_update, k_cg_beta with
// k_cg_direction
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_update_alpha(SceneView<T> S, const double *pq_part,
                                                            double *part) {
  if (cg_done_for_block(S.ctl))
    return;
  double alpha;
  if (!cg_alpha_in_block(S.ctl, pq_part, gridDim.x, &alpha))
    return;
  cg_update_body(S, static_cast<T>(alpha), part, blockIdx.x, gridDim.x);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_direction_beta(SceneView<T> S, const double *part) {
  if (cg_done_for_block(S.ctl))
    return;
  double beta;
  if (!cg_beta_in_block(S.ctl, part, gridDim.x, &beta))
    return;
  cg_direction_body(S, static_cast<T>(beta), blockIdx.x);
}

// apply (wkbb18.hpp:401-413): vd = (x - v) H(t - t_birth - dt_fade), a = vd / dt
template <class T>
__device__ __forceinline__ void cg_apply_body(const SceneView<T> &S, unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> x = S.cx[i], v = S.v[i];
  const T H = step_l(T(0), (S.t - v.w) - S.fade);
  const vec4_t<T> vd = make4<vec4_t<T>>((x.x - v.x) * H, (x.y - v.y) * H, (x.z - v.z) * H, T(0));
  S.vd[i] = vd;
  S.a[i] = make4<vec4_t<T>>(vd.x / S.dt, vd.y / S.dt, vd.z / S.dt, T(0));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_apply(SceneView<T> S) {
  cg_apply_body(S, blockIdx.x);
}

// ---- batched WKBB18 (prtcl_hip_step over several scenes): the kernels above
// over the blocks of every scene of a table (map: block map, entry of scene k
// in bits 20.., its local block below), each block with its scene's view,
// list, CG partials and block count, so every scene's sums and decisions are
// those of its single-scene step. A scene with cg_fold derives alpha and
// beta in its vector kernels, otherwise its one-block kernels do.
template <class T>
__device__ __forceinline__ NbrList entry_list(const BatchEntry<T> &E) {
  return NbrList{E.nidx, E.ninfo, E.nblk};
}

template <class T, bool kGC>
__global__ __launch_bounds__(kBlock) void k_diag_setup_b(const BatchEntry<T> *tab,
                                                         const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  diag_setup_list_body<T, kGC>(E.S, entry_list(E), E.cg_part, u & kBatchLocalMask, E.nblocks);
}

// the gradient correction of wkbb18_gc scenes over the list
template <class T>
__global__ __launch_bounds__(kBlock) void k_corr_b(const BatchEntry<T> *tab,
                                                   const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  corr_body<T, true>(E.S, entry_list(E), u & kBatchLocalMask);
}

// one block per scene
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_setup_fin_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  cg_setup_fin_body(E.S.ctl, E.cg_part, E.nblocks, E.cg_tol, E.cg_max_k, E.cg_seq);
}

template <class T, bool kGC>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_b(const BatchEntry<T> *tab,
                                                      const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  cg_spmv_body<T, false, kGC>(E.S, entry_list(E), E.cg_part, u & kBatchLocalMask, E.nblocks);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_alpha_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  if (!E.cg_fold)
    cg_alpha_body(E.S.ctl, E.cg_part, E.nblocks);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_update_b(const BatchEntry<T> *tab,
                                                        const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  const unsigned lb = u & kBatchLocalMask, nb = E.nblocks;
  Ctl *c = E.S.ctl;
  if (E.cg_fold) {  // p.q partials at [0, nb), r.r and r.y at [nb, 3 nb)
    if (cg_done_for_block(c))
      return;
    double alpha;
    if (!cg_alpha_in_block(c, E.cg_part, nb, &alpha, lb))
      return;
    cg_update_body(E.S, static_cast<T>(alpha), E.cg_part + nb, lb, nb);
  } else {
    if (c->cg_done)
      return;
    cg_update_body(E.S, static_cast<T>(c->cg_alpha), E.cg_part, lb, nb);
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_beta_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  if (!E.cg_fold)
    cg_beta_body(E.S.ctl, E.cg_part, E.nblocks);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_direction_b(const BatchEntry<T> *tab,
                                                           const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  const unsigned lb = u & kBatchLocalMask, nb = E.nblocks;
  Ctl *c = E.S.ctl;
  if (E.cg_fold) {
    if (cg_done_for_block(c))
      return;
    double beta;
    if (!cg_beta_in_block(c, E.cg_part + nb, nb, &beta, lb))
      return;
    cg_direction_body(E.S, static_cast<T>(beta), lb);
  } else {
    if (c->cg_done)
      return;
    cg_direction_body(E.S, static_cast<T>(c->cg_beta), lb);
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_apply_b(const BatchEntry<T> *tab,
                                                       const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  cg_apply_body(E.S, u & kBatchLocalMask);
}

}  // namespace

template <class T>
void Wkbb18Ops<T>::diagonal(hipStream_t st, const SceneView<T> &S, const NbrList *L, bool gc) {
  if (S.nf <= 0)
    return;
  const dim3 gr(grid_blocks(S.nf)), bl(kBlock);
  if (L) {
    if (gc)
      hipLaunchKernelGGL((k_diag_list<T, true>), gr, bl, 0, st, S, *L);
    else
      hipLaunchKernelGGL((k_diag_list<T, false>), gr, bl, 0, st, S, *L);
  } else {
    if (gc)
      hipLaunchKernelGGL((k_diag_exact<T, true>), gr, bl, 0, st, S);
    else
      hipLaunchKernelGGL((k_diag_exact<T, false>), gr, bl, 0, st, S);
  }
}

template <class T>
void Wkbb18Ops<T>::setup(hipStream_t st, const SceneView<T> &S, const NbrList *L, bool gc,
                         double *part, double tol, int max_k, unsigned seq) {
  const unsigned nb = grid_blocks(S.nf);
  if (S.nf > 0) {
    const dim3 gr(nb), bl(kBlock);
    const NbrList Lv = L ? *L : NbrList{};
    if (L && gc)
      hipLaunchKernelGGL((k_cg_setup<T, false, true>), gr, bl, 0, st, S, Lv, part);
    else if (L)
      hipLaunchKernelGGL((k_cg_setup<T, false, false>), gr, bl, 0, st, S, Lv, part);
    else if (gc)
      hipLaunchKernelGGL((k_cg_setup<T, true, true>), gr, bl, 0, st, S, Lv, part);
    else
      hipLaunchKernelGGL((k_cg_setup<T, true, false>), gr, bl, 0, st, S, Lv, part);
  }
  hipLaunchKernelGGL(k_cg_setup_fin, dim3(1), dim3(kBlock), 0, st, S.ctl, part, nb, tol, max_k,
                     seq);
}

template <class T>
void Wkbb18Ops<T>::diagonal_setup(hipStream_t st, const SceneView<T> &S, const NbrList &L,
                                  bool gc, double *part, double tol, int max_k, unsigned seq) {
  const unsigned nb = grid_blocks(S.nf);
  if (S.nf > 0) {
    if (gc)
      hipLaunchKernelGGL((k_diag_setup_list<T, true>), dim3(nb), dim3(kBlock), 0, st, S, L, part);
    else
      hipLaunchKernelGGL((k_diag_setup_list<T, false>), dim3(nb), dim3(kBlock), 0, st, S, L,
                         part);
  }
  hipLaunchKernelGGL(k_cg_setup_fin, dim3(1), dim3(kBlock), 0, st, S.ctl, part, nb, tol, max_k,
                     seq);
}

template <class T>
void Wkbb18Ops<T>::iteration_product(hipStream_t st, const SceneView<T> &S, const NbrList *L,
                                     bool gc, double *part) {
  if (S.nf <= 0)
    return;
  const dim3 gr(grid_blocks(S.nf)), bl(kBlock);
  const NbrList Lv = L ? *L : NbrList{};
  if (L && gc)
    hipLaunchKernelGGL((k_cg_spmv<T, false, true>), gr, bl, 0, st, S, Lv, part);
  else if (L)
    hipLaunchKernelGGL((k_cg_spmv<T, false, false>), gr, bl, 0, st, S, Lv, part);
  else if (gc)
    hipLaunchKernelGGL((k_cg_spmv<T, true, true>), gr, bl, 0, st, S, Lv, part);
  else
    hipLaunchKernelGGL((k_cg_spmv<T, true, false>), gr, bl, 0, st, S, Lv, part);
}

template <class T>
void Wkbb18Ops<T>::iteration_vectors(hipStream_t st, const SceneView<T> &S, double *part,
                                     bool split) {
  if (S.nf <= 0)
    return;
  const unsigned nb = grid_blocks(S.nf);
  const dim3 gr(nb), bl(kBlock);
  if (!split && nb <= kFusedCgBlocks) {
    // p.q partials at part[0, nb), r.r and r.y at part[nb, 3 nb)
    hipLaunchKernelGGL(k_cg_update_alpha<T>, gr, bl, 0, st, S, part, part + nb);
    hipLaunchKernelGGL(k_cg_direction_beta<T>, gr, bl, 0, st, S, part + nb);
    return;
  }
  hipLaunchKernelGGL(k_cg_alpha, dim3(1), bl, 0, st, S.ctl, part, nb);
  hipLaunchKernelGGL(k_cg_update<T>, gr, bl, 0, st, S, part);
  hipLaunchKernelGGL(k_cg_beta, dim3(1), bl, 0, st, S.ctl, part, nb);
  hipLaunchKernelGGL(k_cg_direction<T>, gr, bl, 0, st, S);
}

template <class T>
void Wkbb18Ops<T>::correction(hipStream_t st, const SceneView<T> &S, const NbrList *L) {
  if (S.nf <= 0)
    return;
  const dim3 gr(grid_blocks(S.nf)), bl(kBlock);
  if (L)
    hipLaunchKernelGGL((k_corr<T, true>), gr, bl, 0, st, S, *L);
  else
    hipLaunchKernelGGL((k_corr<T, false>), gr, bl, 0, st, S, NbrList{});
}

template <class T>
void Wkbb18Ops<T>::identity_correction(hipStream_t st, const SceneView<T> &S) {
  if (S.nf > 0)
    hipLaunchKernelGGL(k_corr_identity<T>, dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st, S);
}

template <class T>
void Wkbb18Ops<T>::apply(hipStream_t st, const SceneView<T> &S) {
  if (S.nf > 0)
    hipLaunchKernelGGL(k_cg_apply<T>, dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st, S);
}

template <class T>
void Wkbb18Ops<T>::correction_batch(hipStream_t st, const BatchEntry<T> *tab,
                                    const BatchMaps &m) {
  if (m.n_gc > 0)
    hipLaunchKernelGGL(k_corr_b<T>, dim3(m.n_gc), dim3(kBlock), 0, st, tab, m.gc);
}

template <class T>
void Wkbb18Ops<T>::diagonal_setup_batch(hipStream_t st, const BatchEntry<T> *tab,
                                        const BatchMaps &m, unsigned nscenes) {
  if (nscenes == 0)
    return;
  if (m.n_plain > 0)
    hipLaunchKernelGGL((k_diag_setup_b<T, false>), dim3(m.n_plain), dim3(kBlock), 0, st, tab,
                       m.plain);
  if (m.n_gc > 0)
    hipLaunchKernelGGL((k_diag_setup_b<T, true>), dim3(m.n_gc), dim3(kBlock), 0, st, tab, m.gc);
  hipLaunchKernelGGL(k_cg_setup_fin_b<T>, dim3(nscenes), dim3(kBlock), 0, st, tab);
}

template <class T>
void Wkbb18Ops<T>::iteration_batch(hipStream_t st, const BatchEntry<T> *tab, const BatchMaps &m,
                                   unsigned nscenes, bool any_split) {
  if (m.n_all == 0)
    return;
  const dim3 gr(m.n_all), bl(kBlock);
  if (m.n_plain > 0)
    hipLaunchKernelGGL((k_cg_spmv_b<T, false>), dim3(m.n_plain), bl, 0, st, tab, m.plain);
  if (m.n_gc > 0)
    hipLaunchKernelGGL((k_cg_spmv_b<T, true>), dim3(m.n_gc), bl, 0, st, tab, m.gc);
  if (any_split)
    hipLaunchKernelGGL(k_cg_alpha_b<T>, dim3(nscenes), bl, 0, st, tab);
  hipLaunchKernelGGL(k_cg_update_b<T>, gr, bl, 0, st, tab, m.all);
  if (any_split)
    hipLaunchKernelGGL(k_cg_beta_b<T>, dim3(nscenes), bl, 0, st, tab);
  hipLaunchKernelGGL(k_cg_direction_b<T>, gr, bl, 0, st, tab, m.all);
}

template <class T>
void Wkbb18Ops<T>::apply_batch(hipStream_t st, const BatchEntry<T> *tab, const BatchMaps &m) {
  if (m.n_all > 0)
    hipLaunchKernelGGL(k_cg_apply_b<T>, dim3(m.n_all), dim3(kBlock), 0, st, tab, m.all);
}

template struct Wkbb18Ops<float>;
template struct Wkbb18Ops<double>;

}  // namespace prtcl_amd
