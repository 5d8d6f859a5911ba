// Fused step kernels: the reference's step (density, gravity, explicit
// viscosity, integrate velocity, IISPH setup, IISPH iterations, integrate
// velocity / position) regrouped into the fewest neighbour passes.
//
// Regrouping (exact in real arithmetic; FP32 results differ from the
// reference's per-term evaluation by rounding only):
//   G_f = sum_f m_j gradW_ij, G_b = sum_b V_b gradW_ib,
//   c_i  = -(G_f + 1.4 rho0 G_b) / rho_i^2                    (iisph.hpp:197-249)
//   AA_i = dt^2 rho0/rho_i (c_i . (G_f + rho0 G_b) - m_i/rho_i^2 sum_f m_j |gradW|^2)
//   s_i  = (rho0 - rho_i) - dt rho0/rho_i (sum_f m_j (v_i - v_j).gradW + rho0 v_i.G_b)
//   a_i  = p_i c_i - sum_f m_j (p_j/rho_j^2) gradW_ij           (iisph.hpp:358-423)
//   Ap_i = dt^2 rho0/rho_i (a_i . (G_f + rho0 G_b) - sum_f m_j a_j.gradW_ij)
// so boundary particles are visited once per step (in the setup pass) and
// each IISPH iteration is two passes over fluid neighbours only. The setup
// pass tests the 27-cell candidates with the exact predicate and records the
// accepted fluid neighbours (self excluded) in a per-step list
// (slot-interleaved: entry k of particle i at k * n + i, so a wavefront reads
// 64 consecutive words); later passes walk the list. A particle with more
// neighbours than the list holds is walked over the grid instead.
#include "fused.hpp"

namespace prtcl_amd {

namespace {

// neighbours staged per thread in LDS by the setup pass (40 KB per block)
constexpr uint32_t kSetupLdsCap = 40;
static_assert(kSetupLdsCap <= kNbrCap, "the LDS table must fit the list");

// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// works on one contiguous range of cell-ordered particles and its L2 holds
// that range's neighbourhood.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblocks) {
  constexpr unsigned kXcd = 8;
  const unsigned per = nblocks / kXcd, rem = nblocks % kXcd;
  const unsigned x = b % kXcd, k = b / kXcd;
  // XCD x owns logical blocks [x*per + min(x, rem), ...) of length per + (x < rem)
  return x * per + (x < rem ? x : rem) + k;
}

template <class T>
__device__ __forceinline__ T rsqrt_fast(T x);
template <>
__device__ __forceinline__ float rsqrt_fast<float>(float x) {
  return __builtin_amdgcn_rsqf(x);
}
template <>
__device__ __forceinline__ double rsqrt_fast<double>(double x) {
  return rsqrt(x);
}

template <class Fn>
__device__ __forceinline__ void runs(const GridDims &g, uint32_t cellc,
                                     const uint32_t *__restrict__ start, Fn &&fn) {
  const int cx = static_cast<int>(cellc & 0x7FFu);
  const int cy = static_cast<int>((cellc >> 11) & 0x7FFu);
  const int cz = static_cast<int>(cellc >> 22);
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx + 1 < g.nx ? cx + 1 : g.nx - 1;
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= g.nz)
      continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = cy + dy;
      if (y < 0 || y >= g.ny)
        continue;
      const uint32_t row = static_cast<uint32_t>(g.nx) *
                           (static_cast<uint32_t>(y) +
                            static_cast<uint32_t>(g.ny) * static_cast<uint32_t>(z));
      fn(start[row + static_cast<uint32_t>(x0)], start[row + static_cast<uint32_t>(x1) + 1u]);
    }
  }
}

// W and grad-W factor of the cubic spline from r^2 with one reciprocal
// square root: q = r/h, W = aw ((2-q)^3 - 4 (1-q)^3_+),
// dW/dr = ag (12 (1-q)^2_+ - 3 (2-q)^2), gradW = (dW/dr / r) x_ij, and
// gradW = 0 below r = sqrt(eps) as the reference (kernel_facade.hpp:226).
//
// Neighbours satisfy r < 2h, so 2 - q >= 0 up to rounding; with
// t = 2 - q and u = max(2 - 2q, 0):  12 (1-q)^2_+ - 3 (2-q)^2 = 3 (u^2 - t^2).
// gradf_unscaled() returns (u^2 - t^2) / r; the caller multiplies the sum by
// ag3 = 3 alpha / h^4 once.
template <class T>
struct FastSpline {
  T inv_h, aw, ag, ag3, eps2;
  __device__ __forceinline__ void eval(T r2, T &w, T &f) const {
    const T rinv = rsqrt_fast<T>(r2 > T(0) ? r2 : T(1));
    const T r = r2 * rinv;
    const T q = r * inv_h;
    const T b2 = q < T(2) ? T(2) - q : T(0);
    const T b1 = T(1) - q;
    const T b1p = b1 > T(0) ? b1 : T(0);
    w = aw * (b2 * b2 * b2 - T(4) * b1p * b1p * b1p);
    const T dw = ag * (T(12) * b1p * b1p - T(3) * b2 * b2);
    f = r2 < eps2 ? T(0) : dw * rinv;
    if (r2 <= T(0))
      w = aw * T(4);  // W(0) = alpha/h^3 (8 - 4)
  }
  __device__ __forceinline__ T gradf(T r2) const { return ag3 * gradf_unscaled(r2); }
  __device__ __forceinline__ T gradf_unscaled(T r2) const {
    const T rinv = rsqrt_fast<T>(r2);  // r2 = 0 gives inf/NaN, masked below
    const T q = r2 * rinv * inv_h;
    const T t = T(2) - q;
    T u = T(2) - T(2) * q;
    u = u > T(0) ? u : T(0);
    const T d = (u * u - t * t) * rinv;
    return r2 < eps2 ? T(0) : d;
  }
};

template <class T>
__device__ __forceinline__ FastSpline<T> fast_spline(const SceneView<T> &S) {
  FastSpline<T> k;
  const T h = S.K.h;
  k.inv_h = T(1) / h;
  k.aw = S.K.alpha / (h * h * h);
  k.ag = S.K.alpha / (h * h * h * h);
  k.ag3 = T(3) * k.ag;
  k.eps2 = S.K.eps * S.K.eps;
  return k;
}

template <class T>
__device__ __forceinline__ T fade_step(const SceneView<T> &S, T t_birth) {
  return (T(0) < (S.t - t_birth) - S.fade) ? T(1) : T(0);
}

// Calls fn(j) for every fluid neighbour j != i of particle i: from the list
// (four indices loaded ahead so the dependent gathers of four neighbours are
// in flight together) or, for a particle whose list overflowed, over the grid
// with the exact predicate.
template <class T, class Fn>
__device__ __forceinline__ void walk(const SceneView<T> &S, const NbrList &L, int i,
                                     const vec4_t<T> &xi, Fn &&fn) {
  const uint32_t cnt = L.count[i];
  const size_t n = static_cast<size_t>(S.nf);
  const uint32_t *__restrict__ idx = L.idx + i;
  if (cnt <= L.cap) {
    uint32_t k = 0;
    for (; k + 4 <= cnt; k += 4) {
      const uint32_t j0 = idx[(k + 0) * n], j1 = idx[(k + 1) * n];
      const uint32_t j2 = idx[(k + 2) * n], j3 = idx[(k + 3) * n];
      fn(j0);
      fn(j1);
      fn(j2);
      fn(j3);
    }
    for (; k < cnt; ++k)
      fn(idx[k * n]);
  } else {
    runs(S.grid, S.fcell[i], S.fstart, [&](uint32_t b, uint32_t e) {
      for (uint32_t j = b; j < e; ++j) {
        const vec4_t<T> pj = S.xm[j];
        if (j != static_cast<uint32_t>(i) &&
            sqdist_exact(xi.x - pj.x, xi.y - pj.y, xi.z - pj.z) < S.thr2)
          fn(j);
      }
    });
  }
}

// ---------------------------------------------------------------------------
// setup pass: density + neighbour list + IISPH c, AA (+ s when the velocity
// of every neighbour is known without another pass, i.e. no explicit
// viscosity: v*_j = v_j + dt g H_j)
// ---------------------------------------------------------------------------
template <class T, bool kWithSource>
__device__ __forceinline__ void k_setup_body(SceneView<T> S, NbrList L,
                                                  vec4_t<T> *vout, unsigned lb) {
  
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const FastSpline<T> K = fast_spline(S);
  const vec4_t<T> xi = S.xm[i];
  const vec4_t<T> vi = S.v[i];
  const T Hi = fade_step(S, vi.w);
  const T dtg0 = S.dt * S.g0, dtg1 = S.dt * S.g1, dtg2 = S.dt * S.g2;
  const T vs0 = vi.x + dtg0 * Hi, vs1 = vi.y + dtg1 * Hi, vs2 = vi.z + dtg2 * Hi;
  const uint32_t cell = S.fcell[i];
  const int n = S.nf;
  // phase 1: exact predicate over the 27-cell candidates, accepted fluid
  // neighbours appended to the list (the kernel arithmetic of the accepted
  // pairs runs in phase 2 over the compacted list, so the SIMD lanes that
  // reject a candidate do not wait for the ones that accept it)
  // The accepted indices go to this thread's column of an LDS table first
  // (conflict-free: entry k of thread t at lst[k][t]); phase 2 copies them
  // to the list row by row, so every global store is a coalesced row write.
  __shared__ uint32_t lst[kSetupLdsCap][kBlock];
  const unsigned t = threadIdx.x;
  uint32_t cnt = 0;
  runs(S.grid, cell, S.fstart, [&](uint32_t b, uint32_t e) {
    // four candidates per trip, loads issued together (indices clamped to the
    // run; the clamped duplicates are rejected by the bound check)
    for (uint32_t j = b; j < e; j += 4) {
      vec4_t<T> pk[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        pk[u] = S.xm[min(j + u, e - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t ju = j + u;
        if (ju < e && ju != static_cast<uint32_t>(i) &&
            sqdist_exact(xi.x - pk[u].x, xi.y - pk[u].y, xi.z - pk[u].z) < S.thr2) {
          if (cnt < kSetupLdsCap)
            lst[cnt][t] = ju;
          ++cnt;
        }
      }
    }
  });
  L.count[i] = cnt;
  // phase 2: density (self term W(0) = 4 alpha / h^3), G_f, sum |gradW|^2,
  // and the velocity-divergence sum of the source term
  T rho = xi.w * K.aw * T(4), gf0 = 0, gf1 = 0, gf2 = 0, s2 = 0, dsum = 0;
  auto pair = [&](uint32_t j) {
    const vec4_t<T> pj = S.xm[j];
    const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
    T w, f;
    K.eval(d0 * d0 + d1 * d1 + d2 * d2, w, f);
    rho += pj.w * w;
    const T g0 = f * d0, g1 = f * d1, g2 = f * d2;
    gf0 += pj.w * g0;
    gf1 += pj.w * g1;
    gf2 += pj.w * g2;
    s2 += pj.w * (g0 * g0 + g1 * g1 + g2 * g2);
    if (kWithSource) {
      const vec4_t<T> vj = S.v[j];
      const T Hj = fade_step(S, vj.w);
      const T u0 = vs0 - (vj.x + dtg0 * Hj);
      const T u1 = vs1 - (vj.y + dtg1 * Hj);
      const T u2 = vs2 - (vj.z + dtg2 * Hj);
      dsum += pj.w * (u0 * g0 + u1 * g1 + u2 * g2);
    }
  };
  const size_t nn = static_cast<size_t>(n);
  if (cnt <= kSetupLdsCap) {
#pragma unroll 4
    for (uint32_t k = 0; k < cnt; ++k) {
      const uint32_t j = lst[k][t];
      L.idx[k * nn + i] = j;  // row k of the list, written by all lanes together
      pair(j);
    }
  } else {
    // more neighbours than the LDS table holds: walk the grid again in the
    // same candidate order, writing the first L.cap entries
    uint32_t k = 0;
    runs(S.grid, cell, S.fstart, [&](uint32_t b, uint32_t e) {
      for (uint32_t j = b; j < e; ++j) {
        const vec4_t<T> pj = S.xm[j];
        if (j != static_cast<uint32_t>(i) &&
            sqdist_exact(xi.x - pj.x, xi.y - pj.y, xi.z - pj.z) < S.thr2) {
          if (k < L.cap)
            L.idx[k * nn + i] = j;
          ++k;
          pair(j);
        }
      }
    });
  }
  T rhob = 0, gb0 = 0, gb1 = 0, gb2 = 0;
  runs(S.grid, cell, S.bstart, [&](uint32_t b, uint32_t e) {
    for (uint32_t j = b; j < e; j += 4) {
      vec4_t<T> pk[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        pk[u] = S.bxV[min(j + u, e - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const vec4_t<T> pj = pk[u];
        const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
        const T r2 = sqdist_exact(d0, d1, d2);
        if (j + u >= e || !(r2 < S.thr2))
          continue;
        T w, f;
        K.eval(r2, w, f);
        rhob += pj.w * w;
        gb0 += pj.w * (f * d0);
        gb1 += pj.w * (f * d1);
        gb2 += pj.w * (f * d2);
      }
    }
  });
  const T rho0 = S.rho0;
  rho += rho0 * rhob;
  const T irho = T(1) / rho;
  const T irho2 = irho * irho;
  const T k14 = T(1.4) * rho0;
  const T c0 = -(gf0 + k14 * gb0) * irho2;
  const T c1 = -(gf1 + k14 * gb1) * irho2;
  const T c2 = -(gf2 + k14 * gb2) * irho2;
  const T gs0 = gf0 + rho0 * gb0, gs1 = gf1 + rho0 * gb1, gs2 = gf2 + rho0 * gb2;
  const T fac = S.dt * S.dt * rho0 * irho;
  const T AA = fac * ((c0 * gs0 + c1 * gs1 + c2 * gs2) - xi.w * irho2 * s2);
  S.rho[i] = rho;
  S.c[i] = make4<vec4_t<T>>(c0, c1, c2, T(0));
  S.G[i] = make4<vec4_t<T>>(gs0, gs1, gs2, T(0));
  S.AA[i] = AA;
  S.p[i] = T(0);
  S.pr[i] = T(0);
  S.xq[i] = make4<vec4_t<T>>(xi.x, xi.y, xi.z, T(0));
  if (kWithSource) {
    const T src = (rho0 - rho) -
                  S.dt * rho0 * irho * (dsum + rho0 * (vs0 * gb0 + vs1 * gb1 + vs2 * gb2));
    S.s[i] = src;
    vout[i] = make4<vec4_t<T>>(vs0, vs1, vs2, vi.w);
  } else {
    S.vd[i] = make4<vec4_t<T>>(gb0, gb1, gb2, T(0));  // G_b for the source pass
  }
}

template <class T, bool kWithSource>
__global__ __launch_bounds__(kBlock) void k_setup(SceneView<T> S, NbrList L,
                                                  vec4_t<T> *vout) {
  k_setup_body<T, kWithSource>(S, L, vout, xcd_block(blockIdx.x, gridDim.x));
}

// explicit viscosity (viscosity.hpp:171-240) over the list, then
// v* = v + dt (g + a_visc) H   (gravity.hpp:133, symplectic_euler.hpp:172-200)
template <class T>
__global__ __launch_bounds__(kBlock) void k_viscosity_list(SceneView<T> S, NbrList L,
                                                           vec4_t<T> *vout) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const FastSpline<T> K = fast_spline(S);
  const vec4_t<T> xi = S.xm[i], vi = S.v[i];
  const T rhoi = S.rho[i];
  const T cf = T(10) * S.mu_f / rhoi;
  T a0 = 0, a1 = 0, a2 = 0;
  auto pair = [&](uint32_t j) {
    const vec4_t<T> pj = S.xm[j];
    const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
    const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
    const vec4_t<T> vj = S.v[j];
    const T s = cf * (pj.w / S.rho[j]) *
                ((vi.x - vj.x) * d0 + (vi.y - vj.y) * d1 + (vi.z - vj.z) * d2) /
                (r2 + S.hh001) * K.gradf(r2);
    a0 += s * d0;
    a1 += s * d1;
    a2 += s * d2;
  };
  walk(S, L, i, xi, pair);
  const T cb = T(10) * S.mu_b / rhoi * S.rho0 / rhoi;
  runs(S.grid, S.fcell[i], S.bstart, [&](uint32_t b, uint32_t e) {
    for (uint32_t j = b; j < e; ++j) {
      const vec4_t<T> pj = S.bxV[j];
      const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
      const T r2e = sqdist_exact(d0, d1, d2);
      if (!(r2e < S.thr2))
        continue;
      const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const T s = cb * pj.w * (vi.x * d0 + vi.y * d1 + vi.z * d2) / (r2 + S.hh001) *
                  K.gradf(r2);
      a0 += s * d0;
      a1 += s * d1;
      a2 += s * d2;
    }
  });
  const T H = fade_step(S, vi.w);
  vout[i] = make4<vec4_t<T>>(vi.x + S.dt * (S.g0 + a0) * H, vi.y + S.dt * (S.g1 + a1) * H,
                             vi.z + S.dt * (S.g2 + a2) * H, vi.w);
}

// source term s from the integrated velocities (explicit-viscosity path)
template <class T>
__global__ __launch_bounds__(kBlock) void k_source_list(SceneView<T> S, NbrList L) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const FastSpline<T> K = fast_spline(S);
  const vec4_t<T> xi = S.xm[i], vi = S.v[i], gb = S.vd[i];
  const T rhoi = S.rho[i];
  T dsum = 0;
  auto pair = [&](uint32_t j) {
    const vec4_t<T> pj = S.xm[j];
    const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
    const T f = K.gradf(d0 * d0 + d1 * d1 + d2 * d2);
    const vec4_t<T> vj = S.v[j];
    dsum += pj.w * f * ((vi.x - vj.x) * d0 + (vi.y - vj.y) * d1 + (vi.z - vj.z) * d2);
  };
  walk(S, L, i, xi, pair);
  const T rho0 = S.rho0;
  S.s[i] = (rho0 - rhoi) -
           S.dt * rho0 / rhoi * (dsum + rho0 * (vi.x * gb.x + vi.y * gb.y + vi.z * gb.z));
}

// IISPH pressure acceleration over the list: a_i = p_i c_i - sum m_j pr_j gradW
template <class T>
__device__ __forceinline__ void k_accel_list_body(SceneView<T> S, NbrList L, unsigned lb) {
  if (S.ctl->done)
    return;
  
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const FastSpline<T> K = fast_spline(S);
  const vec4_t<T> xi = S.xm[i];
  const vec4_t<T> ci = S.c[i];
  const T pi = S.p[i];
  T a0 = 0, a1 = 0, a2 = 0;
  auto pair = [&](uint32_t j) {
    const vec4_t<T> qj = S.xq[j];  // (x_j, m_j p_j / rho_j^2)
    const T d0 = xi.x - qj.x, d1 = xi.y - qj.y, d2 = xi.z - qj.z;
    const T s = qj.w * K.gradf_unscaled(d0 * d0 + d1 * d1 + d2 * d2);
    a0 += s * d0;
    a1 += s * d1;
    a2 += s * d2;
  };
  walk(S, L, i, xi, pair);
  S.a[i] = make4<vec4_t<T>>(pi * ci.x - K.ag3 * a0, pi * ci.y - K.ag3 * a1,
                            pi * ci.z - K.ag3 * a2, T(0));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_accel_list(SceneView<T> S, NbrList L) {
  k_accel_list_body<T>(S, L, xcd_block(blockIdx.x, gridDim.x));
}

// IISPH relaxed Jacobi update over the list, FP64 aprde reduction and the
// device-side convergence decision (example.lua:409-421)
template <class T>
__device__ __forceinline__ void k_pressure_list_body(SceneView<T> S, NbrList L, unsigned lb) {
  if (S.ctl->done)
    return;
  
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double term = 0.0;
  if (i < S.nf) {
    const FastSpline<T> K = fast_spline(S);
    const vec4_t<T> xi = S.xm[i], ai = S.a[i], gs = S.G[i];
    const T rhoi = S.rho[i];
    T sum = 0;
    auto pair = [&](uint32_t j) {
      const vec4_t<T> pj = S.xm[j];
      const vec4_t<T> aj = S.a[j];
      const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
      sum += pj.w * K.gradf_unscaled(d0 * d0 + d1 * d1 + d2 * d2) *
             (aj.x * d0 + aj.y * d1 + aj.z * d2);
    };
    walk(S, L, i, xi, pair);
    const T Ap = S.dt * S.dt * S.rho0 / rhoi *
                 ((ai.x * gs.x + ai.y * gs.y + ai.z * gs.z) - K.ag3 * sum);
    const T AA = S.AA[i], si = S.s[i];
    const T rcp = (fabs(AA) > T(1e-9)) ? T(1) / AA : T(0);
    const T upd = S.p[i] + S.omega * (si - Ap) * rcp;
    const T pn = ((T(0) < upd) ? upd : T(0)) * ((T(1e-9) < fabs(AA)) ? T(1) : T(0));
    S.Ap[i] = Ap;
    S.p[i] = pn;
    S.pr[i] = pn / (rhoi * rhoi);
    S.xq[i] = make4<vec4_t<T>>(xi.x, xi.y, xi.z, xi.w * pn / (rhoi * rhoi));
    term = static_cast<double>((Ap - si) / S.rho0 * ((T(1e-9) < pn) ? T(1) : T(0)));
  }
  const double bs = block_sum<kBlock>(term);
  if (threadIdx.x == 0)
    S.partials[lb] = bs;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pressure_list(SceneView<T> S, NbrList L) {
  k_pressure_list_body<T>(S, L, xcd_block(blockIdx.x, gridDim.x));
}

// Sum of the block partials in block order (FP64, fixed tree) and the
// convergence decision of example.lua:418-419: stop when it >= min_it and
// aprde / nprde < tol, or at max_it.
__device__ __forceinline__ void k_pressure_finalize_body(Ctl *c, const double *partials,
                                                        unsigned nparts) {
  if (c->done)
    return;
  double acc = 0.0;
  for (unsigned k = threadIdx.x; k < nparts; k += kBlock)
    acc += partials[k];
  acc = block_sum<kBlock>(acc);
  if (threadIdx.x == 0) {
    c->aprde = acc;
    const int it = c->iters + 1;
    c->iters = it;
    const double rel = acc / static_cast<double>(c->nprde);
    c->rel = rel;
    if ((it >= c->min_it && rel < c->tol) || it >= c->max_it) {
      c->done = 1;
      c->converged = (it >= c->min_it && rel < c->tol) ? 1 : 0;
    }
    publish_progress(c);
  }
}

__global__ __launch_bounds__(kBlock) void k_pressure_finalize(Ctl *c, const double *partials,
                                                              unsigned nparts) {
  k_pressure_finalize_body(c, partials, nparts);
}

template <class T>
__device__ __forceinline__ void max_speed_atomic(Ctl *ctl, T s);
template <>
__device__ __forceinline__ void max_speed_atomic<float>(Ctl *ctl, float s) {
  atomicMax(&ctl->max_speed_bits_f, __float_as_uint(s));
}
template <>
__device__ __forceinline__ void max_speed_atomic<double>(Ctl *ctl, double s) {
  atomicMax(&ctl->max_speed_bits_d, static_cast<unsigned long long>(__double_as_longlong(s)));
}

// v += dt a H (pressure acceleration), x += dt v, running max |v|
template <class T>
__device__ __forceinline__ void k_integrate_body(SceneView<T> S, unsigned lb) {
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  T speed = T(0);
  bool bad = false;
  if (i < S.nf) {
    vec4_t<T> v = S.v[i];
    const vec4_t<T> a = S.a[i];
    const T H = fade_step(S, v.w);
    v.x += S.dt * a.x * H;
    v.y += S.dt * a.y * H;
    v.z += S.dt * a.z * H;
    vec4_t<T> x = S.xm[i];
    x.x += S.dt * v.x;
    x.y += S.dt * v.y;
    x.z += S.dt * v.z;
    S.v[i] = v;
    S.xm[i] = x;
    speed = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    bad = !(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(speed));
    if (bad)
      speed = T(0);
  }
  speed = block_max<kBlock>(speed);
  if (threadIdx.x == 0)
    S.partials[lb] = static_cast<double>(speed);
  if (bad)
    S.ctl->nonfinite = 1;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_integrate(SceneView<T> S) {
  k_integrate_body<T>(S, blockIdx.x);
}

// running maximum speed (symplectic_euler.hpp:250-268) from the block maxima
template <class T>
__device__ __forceinline__ void k_speed_finalize_body(Ctl *c, const double *partials,
                                                      unsigned nparts) {
  double m = 0.0;
  for (unsigned k = threadIdx.x; k < nparts; k += kBlock)
    m = partials[k] > m ? partials[k] : m;
  m = block_max<kBlock>(m);
  if (threadIdx.x == 0)
    max_speed_atomic<T>(c, static_cast<T>(m));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_speed_finalize(Ctl *c, const double *partials,
                                                           unsigned nparts) {
  k_speed_finalize_body<T>(c, partials, nparts);
}

// ---- batched wrappers (prtcl_hip_step over several scenes)
template <class T>
__global__ __launch_bounds__(kBlock) void k_setup_b(const BatchEntry<T> *tab,
                                                    const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_setup_body<T, true>(E.S, NbrList{E.nidx, E.ncnt, kNbrCap}, E.vtmp, u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_accel_b(const BatchEntry<T> *tab,
                                                    const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_accel_list_body<T>(E.S, NbrList{E.nidx, E.ncnt, kNbrCap}, u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pressure_b(const BatchEntry<T> *tab,
                                                       const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_pressure_list_body<T>(E.S, NbrList{E.nidx, E.ncnt, kNbrCap}, u & kBatchLocalMask);
}

// one block per scene
template <class T>
__global__ __launch_bounds__(kBlock) void k_pressure_finalize_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  k_pressure_finalize_body(E.S.ctl, E.S.partials, E.nblocks);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_integrate_b(const BatchEntry<T> *tab,
                                                        const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_integrate_body<T>(E.S, u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_speed_finalize_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  k_speed_finalize_body<T>(E.S.ctl, E.S.partials, E.nblocks);
}

}  // namespace

template <class T>
void FusedOps<T>::setup(hipStream_t st, const SceneView<T> &S, const NbrList &L,
                        bool explicit_viscosity, vec4_t<T> *vtmp) {
  if (S.nf <= 0)
    return;
  const dim3 gr(grid_blocks(S.nf)), bl(kBlock);
  if (!explicit_viscosity) {
    hipLaunchKernelGGL((k_setup<T, true>), gr, bl, 0, st, S, L, vtmp);
    return;
  }
  hipLaunchKernelGGL((k_setup<T, false>), gr, bl, 0, st, S, L, vtmp);
  hipLaunchKernelGGL(k_viscosity_list<T>, gr, bl, 0, st, S, L, vtmp);
  // the integrated velocities replace v once every thread has read the old ones
  SceneView<T> S2 = S;
  S2.v = vtmp;
  hipLaunchKernelGGL(k_source_list<T>, gr, bl, 0, st, S2, L);
}

template <class T>
void FusedOps<T>::accel(hipStream_t st, const SceneView<T> &S, const NbrList &L) {
  if (S.nf > 0)
    hipLaunchKernelGGL(k_accel_list<T>, dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st, S, L);
}

template <class T>
void FusedOps<T>::pressure(hipStream_t st, const SceneView<T> &S, const NbrList &L) {
  if (S.nf <= 0)
    return;
  const unsigned nb = grid_blocks(S.nf);
  hipLaunchKernelGGL(k_pressure_list<T>, dim3(nb), dim3(kBlock), 0, st, S, L);
  hipLaunchKernelGGL(k_pressure_finalize, dim3(1), dim3(kBlock), 0, st, S.ctl, S.partials, nb);
}

template <class T>
void FusedOps<T>::integrate(hipStream_t st, const SceneView<T> &S) {
  if (S.nf <= 0)
    return;
  const unsigned nb = grid_blocks(S.nf);
  hipLaunchKernelGGL(k_integrate<T>, dim3(nb), dim3(kBlock), 0, st, S);
  hipLaunchKernelGGL(k_speed_finalize<T>, dim3(1), dim3(kBlock), 0, st, S.ctl, S.partials, nb);
}

template <class T>
void FusedOps<T>::setup_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                              unsigned npblocks) {
  if (npblocks > 0)
    hipLaunchKernelGGL(k_setup_b<T>, dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
}

template <class T>
void FusedOps<T>::accel_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                              unsigned npblocks) {
  if (npblocks > 0)
    hipLaunchKernelGGL(k_accel_b<T>, dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
}

template <class T>
void FusedOps<T>::pressure_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                                 unsigned npblocks, unsigned nscenes) {
  if (npblocks == 0)
    return;
  hipLaunchKernelGGL(k_pressure_b<T>, dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
  hipLaunchKernelGGL(k_pressure_finalize_b<T>, dim3(nscenes), dim3(kBlock), 0, st, tab);
}

template <class T>
void FusedOps<T>::integrate_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                                  unsigned npblocks, unsigned nscenes) {
  if (npblocks == 0)
    return;
  hipLaunchKernelGGL(k_integrate_b<T>, dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
  hipLaunchKernelGGL(k_speed_finalize_b<T>, dim3(nscenes), dim3(kBlock), 0, st, tab);
}

template struct FusedOps<float>;
template struct FusedOps<double>;

}  // namespace prtcl_amd
