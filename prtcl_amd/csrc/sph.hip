// SPH procedure kernels of the reference schemes, one thread per particle in
// cell order, each gathering over the 9 x-runs of the 27-cell neighbourhood
// (3 contiguous cells per run). The per-pair arithmetic follows the
// generated reference code operation by operation (expression trees cited
// per kernel), FP32 without FMA contraction, so individual pair terms equal
// the CPU's; only the order in which neighbours are summed differs.
#include "sph.hpp"

namespace prtcl_amd {

namespace {

template <class Fn>
__device__ __forceinline__ void for_runs(const GridDims &g, uint32_t cellc,
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
      const uint32_t row =
          static_cast<uint32_t>(g.nx) * (static_cast<uint32_t>(y) +
                                         static_cast<uint32_t>(g.ny) * static_cast<uint32_t>(z));
      const uint32_t b = start[row + static_cast<uint32_t>(x0)];
      const uint32_t e = start[row + static_cast<uint32_t>(x1) + 1u];
      fn(b, e);
    }
  }
}

// Visits every neighbour j (squared distance < thr2) of position xi among the
// particles `pts` bucketed by `start`; fn(j, d0, d1, d2, r2, pj) with
// d = xi - xj.
template <class T, class Fn>
__device__ __forceinline__ void for_neighbors(const SceneView<T> &S, const vec4_t<T> &xi,
                                              uint32_t cellc, const uint32_t *start,
                                              const vec4_t<T> *__restrict__ pts, Fn &&fn) {
  for_runs(S.grid, cellc, start, [&](uint32_t b, uint32_t e) {
    for (uint32_t j = b; j < e; ++j) {
      const vec4_t<T> pj = pts[j];
      const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
      const T r2 = sqdist_exact(d0, d1, d2);
      if (r2 < S.thr2)
        fn(j, d0, d1, d2, r2, pj);
    }
  });
}

template <class T>
__device__ __forceinline__ T step_l(T eps, T x) {  // unit_step_l (eigen.hpp:306-314)
  return (eps < x) ? T(1) : T(0);
}

// ---- aiast12.compute_volume (aiast12.hpp:125-175): V_b = 1 / sum_b' W
template <class T>
__global__ __launch_bounds__(kBlock) void k_volume(SceneView<T> S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nb)
    return;
  const vec4_t<T> xi = S.bxV[i];
  T V = 0;
  for_neighbors(S, xi, S.bcell[i], S.bstart, S.bxV,
                [&](uint32_t, T, T, T, T r2, const vec4_t<T> &) { V += S.K.w(r2); });
  vec4_t<T> o = xi;
  o.w = T(1) / V;
  S.bxV[i] = o;
}

// ---- density.compute_density (density.hpp:160-215)
template <class T>
__global__ __launch_bounds__(kBlock) void k_density(SceneView<T> S) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i];
  const uint32_t ci = S.fcell[i];
  T rho = 0;
  for_neighbors(S, xi, ci, S.fstart, S.xm,
                [&](uint32_t, T, T, T, T r2, const vec4_t<T> &pj) { rho += pj.w * S.K.w(r2); });
  for_neighbors(S, xi, ci, S.bstart, S.bxV, [&](uint32_t, T, T, T, T r2, const vec4_t<T> &pj) {
    rho += (pj.w * S.rho0) * S.K.w(r2);
  });
  S.rho[i] = rho;
}

// ---- gravity.initialize_acceleration (gravity.hpp:133)
template <class T>
__global__ __launch_bounds__(kBlock) void k_gravity(SceneView<T> S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < S.nf)
    S.a[i] = make4<vec4_t<T>>(S.g0, S.g1, S.g2, T(0));
}

// ---- gravity.accumulate_acceleration (gravity.hpp:157-186): a += g
template <class T>
__global__ __launch_bounds__(kBlock) void k_gravity_accumulate(SceneView<T> S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < S.nf) {
    vec4_t<T> a = S.a[i];
    a.x += S.g0;
    a.y += S.g1;
    a.z += S.g2;
    S.a[i] = a;
  }
}

// ---- gravity.initialize_axial_acceleration (gravity.hpp:189-215)
template <class T>
__global__ __launch_bounds__(kBlock) void k_gravity_axial(SceneView<T> S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < S.nf) {
    const vec4_t<T> x = S.xm[i];
    S.a[i] = axial_acceleration(S, x.x, x.y, x.z);
  }
}

// ---- viscosity.accumulate_acceleration (viscosity.hpp:171-240)
template <class T>
__global__ __launch_bounds__(kBlock) void k_viscosity(SceneView<T> S) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i], vi = S.v[i];
  const T rhoi = S.rho[i];
  const uint32_t ci = S.fcell[i];
  vec4_t<T> ai = S.a[i];
  for_neighbors(S, xi, ci, S.fstart, S.xm,
                [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                  const vec4_t<T> vj = S.v[j];
                  const T f = S.K.gradf(r2);
                  const T g0 = f * d0, g1 = f * d1, g2 = f * d2;
                  const T s = (((T(10) * (S.mu_f / rhoi)) * (pj.w / S.rho[j])) *
                               dot3_exact(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, d0, d1, d2)) /
                              (r2 + S.hh001);
                  ai.x += s * g0;
                  ai.y += s * g1;
                  ai.z += s * g2;
                });
  for_neighbors(S, xi, ci, S.bstart, S.bxV,
                [&](uint32_t, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                  const T f = S.K.gradf(r2);
                  const T g0 = f * d0, g1 = f * d1, g2 = f * d2;
                  const T s = (((T(10) * (S.mu_b / rhoi)) * ((S.rho0 * pj.w) / rhoi)) *
                               dot3_exact(vi.x, vi.y, vi.z, d0, d1, d2)) /
                              (r2 + S.hh001);
                  ai.x += s * g0;
                  ai.y += s * g1;
                  ai.z += s * g2;
                });
  S.a[i] = ai;
}

// ---- symplectic_euler.integrate_velocity_with_hard_fade
//      (symplectic_euler.hpp:172-200): v += (dt * a) * [0 < (t - t_birth) - fade]
template <class T>
__device__ __forceinline__ void integrate_v_body(const SceneView<T> &S, int i,
                                                 const vec4_t<T> &a) {
#pragma clang fp contract(off)
  vec4_t<T> v = S.v[i];
  const T H = step_l(T(0), (S.t - v.w) - S.fade);
  v.x += (S.dt * a.x) * H;
  v.y += (S.dt * a.y) * H;
  v.z += (S.dt * a.z) * H;
  S.v[i] = v;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_integrate_v(SceneView<T> S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nf)
    return;
  integrate_v_body(S, i, S.a[i]);
}

template <class T>
__device__ __forceinline__ void atomic_max_speed(Ctl *ctl, T s);
template <>
__device__ __forceinline__ void atomic_max_speed<float>(Ctl *ctl, float s) {
  atomicMax(&ctl->max_speed_bits_f, __float_as_uint(s));
}
template <>
__device__ __forceinline__ void atomic_max_speed<double>(Ctl *ctl, double s) {
  atomicMax(&ctl->max_speed_bits_d,
            static_cast<unsigned long long>(__double_as_longlong(s)));
}

// ---- symplectic_euler.integrate_position (symplectic_euler.hpp:236-270):
//      x += dt * v, running max of |v| (bit patterns of non-negative IEEE
//      values order like the values, so an integer atomic max is exact)
template <class T>
__device__ __forceinline__ void integrate_x_body(const SceneView<T> &S, unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  T speed = T(0);
  bool bad = false;
  if (i < S.nf) {
    vec4_t<T> x = S.xm[i];
    const vec4_t<T> v = S.v[i];
    x.x += S.dt * v.x;
    x.y += S.dt * v.y;
    x.z += S.dt * v.z;
    S.xm[i] = x;
    speed = sqrt(sqdist_exact(v.x, v.y, v.z));
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
__global__ __launch_bounds__(kBlock) void k_integrate_x(SceneView<T> S) {
  integrate_x_body(S, blockIdx.x);
}

// ---- iisph.setup pass 1 (iisph.hpp:197-249): helper c
template <class T>
__global__ __launch_bounds__(kBlock) void k_iisph_c(SceneView<T> S) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i];
  const T rhoi = S.rho[i];
  const uint32_t ci = S.fcell[i];
  T c0 = 0, c1 = 0, c2 = 0;
  for_neighbors(S, xi, ci, S.fstart, S.xm,
                [&](uint32_t, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                  const T f = S.K.gradf(r2);
                  const T s = pj.w / (rhoi * rhoi);
                  c0 -= s * (f * d0);
                  c1 -= s * (f * d1);
                  c2 -= s * (f * d2);
                });
  for_neighbors(S, xi, ci, S.bstart, S.bxV,
                [&](uint32_t, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                  const T f = S.K.gradf(r2);
                  const T s = ((T(0.7) * T(2)) * S.rho0) * (pj.w / (rhoi * rhoi));
                  c0 -= s * (f * d0);
                  c1 -= s * (f * d1);
                  c2 -= s * (f * d2);
                });
  S.c[i] = make4<vec4_t<T>>(c0, c1, c2, T(0));
}

// ---- iisph.setup pass 2 (iisph.hpp:252-354): p = 0, source s, diagonal AA
template <class T>
__global__ __launch_bounds__(kBlock) void k_iisph_s_aa(SceneView<T> S) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i], vi = S.v[i], ci = S.c[i];
  const T rhoi = S.rho[i];
  const T mi = xi.w;
  const uint32_t cell = S.fcell[i];
  T s = 0, AA = 0;
  for_neighbors(S, xi, cell, S.fstart, S.xm,
                [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                  const vec4_t<T> vj = S.v[j];
                  const T f = S.K.gradf(r2);
                  const T g0 = f * d0, g1 = f * d1, g2 = f * d2;
                  s += pj.w * dot3_exact(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, g0, g1, g2);
                  AA += pj.w * dot3_exact(ci.x, ci.y, ci.z, g0, g1, g2);
                  AA -= (pj.w * (mi / (rhoi * rhoi))) * dot3_exact(g0, g1, g2, g0, g1, g2);
                });
  for_neighbors(S, xi, cell, S.bstart, S.bxV,
                [&](uint32_t, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                  const T f = S.K.gradf(r2);
                  const T g0 = f * d0, g1 = f * d1, g2 = f * d2;
                  s += (S.rho0 * pj.w) * dot3_exact(vi.x, vi.y, vi.z, g0, g1, g2);
                  AA += (S.rho0 * pj.w) * dot3_exact(ci.x, ci.y, ci.z, g0, g1, g2);
                });
  s = (S.rho0 - rhoi) - ((S.dt * (S.rho0 / rhoi)) * s);
  AA *= (S.dt * S.dt) * (S.rho0 / rhoi);
  S.s[i] = s;
  S.AA[i] = AA;
  S.p[i] = T(0);
  S.pr[i] = T(0);
}

// ---- iisph.iteration_pressure_acceleration (iisph.hpp:358-423)
template <class T, bool kSolve>
__global__ __launch_bounds__(kBlock) void k_iisph_accel(SceneView<T> S) {
#pragma clang fp contract(off)
  if (kSolve && S.ctl->done)
    return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i];
  const T rhoi = S.rho[i], pi = S.p[i];
  const uint32_t cell = S.fcell[i];
  T a0 = 0, a1 = 0, a2 = 0;
  for_neighbors(S, xi, cell, S.fstart, S.xm,
                [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                  const T rhoj = S.rho[j];
                  const T f = S.K.gradf(r2);
                  const T s = pj.w * ((pi / (rhoi * rhoi)) + (S.p[j] / (rhoj * rhoj)));
                  a0 -= s * (f * d0);
                  a1 -= s * (f * d1);
                  a2 -= s * (f * d2);
                });
  for_neighbors(S, xi, cell, S.bstart, S.bxV,
                [&](uint32_t, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                  const T f = S.K.gradf(r2);
                  const T s = ((T(0.7) * S.rho0) * pj.w) * ((T(2) * pi) / (rhoi * rhoi));
                  a0 -= s * (f * d0);
                  a1 -= s * (f * d1);
                  a2 -= s * (f * d2);
                });
  S.a[i] = make4<vec4_t<T>>(a0, a1, a2, T(0));
}

// ---- iisph.iteration_pressure (iisph.hpp:426-516) with the aprde reduction
//      in FP64: block partials in a fixed tree, the last block sums them in
//      block order and takes the convergence decision of example.lua:418-419
//      on the device when kSolve.
template <class T, bool kSolve>
__global__ __launch_bounds__(kBlock) void k_iisph_pressure(SceneView<T> S) {
#pragma clang fp contract(off)
  if (kSolve && S.ctl->done)
    return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double term = 0.0;
  if (i < S.nf) {
    const vec4_t<T> xi = S.xm[i], ai = S.a[i];
    const T rhoi = S.rho[i];
    const uint32_t cell = S.fcell[i];
    T Ap = 0;
    for_neighbors(S, xi, cell, S.fstart, S.xm,
                  [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                    const vec4_t<T> aj = S.a[j];
                    const T f = S.K.gradf(r2);
                    Ap += pj.w * dot3_exact(ai.x - aj.x, ai.y - aj.y, ai.z - aj.z, f * d0,
                                            f * d1, f * d2);
                  });
    for_neighbors(S, xi, cell, S.bstart, S.bxV,
                  [&](uint32_t, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                    const T f = S.K.gradf(r2);
                    Ap += (S.rho0 * pj.w) * dot3_exact(ai.x, ai.y, ai.z, f * d0, f * d1, f * d2);
                  });
    Ap *= (S.dt * S.dt) * (S.rho0 / rhoi);
    const T AA = S.AA[i], si = S.s[i];
    const T rcp = (fabs(AA) > T(1e-9)) ? T(1) / AA : T(0);
    const T upd = S.p[i] + ((S.omega * (si - Ap)) * rcp);
    const T pn = ((T(0) < upd) ? upd : T(0)) * step_l(T(1e-9), fabs(AA));
    S.Ap[i] = Ap;
    S.p[i] = pn;
    S.pr[i] = pn / (rhoi * rhoi);
    term = static_cast<double>(((Ap - si) / S.rho0) * step_l(T(1e-9), pn));
  }
  const double bs = block_sum<kBlock>(term);
  if (threadIdx.x == 0)
    S.partials[blockIdx.x] = bs;
}

// aprde = sum of the block partials in block order (FP64); kSolve also takes
// the convergence decision, otherwise the sum is added to the global as the
// reference's reduction clause does (iisph.hpp:440,505-507)
template <bool kSolve>
__global__ __launch_bounds__(kBlock) void k_pressure_finalize(Ctl *c, const double *partials,
                                                              unsigned nparts) {
  if (kSolve && c->done)
    return;
  const double acc = block_sum<kBlock>(strided_sum<kBlock>(partials, nparts));
  if (threadIdx.x == 0) {
    if (kSolve) {
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
    } else {
      c->aprde = c->aprde + acc;
      c->rel = c->aprde / static_cast<double>(c->nprde);
    }
  }
}

template <class T>
__device__ __forceinline__ void speed_finalize_body(Ctl *c, const double *partials,
                                                    unsigned nparts) {
  const double m = block_max<kBlock>(strided_max<kBlock>(partials, nparts, 0.0));
  if (threadIdx.x == 0) {
    atomic_max_speed<T>(c, static_cast<T>(m));
    publish_step_end(c);
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_speed_finalize(Ctl *c, const double *partials,
                                                           unsigned nparts) {
  speed_finalize_body<T>(c, partials, nparts);
}

// ---- batched forms over the blocks of several scenes (map: block map, entry
// of scene k in bits 20.., its local block below), each block with its
// scene's view: the bodies above, so every scene's result is the same as
// stepped alone. from_xa: a is first taken from the pressure acceleration
// the last IISPH pass wrote (the copy the fused path makes before a
// viscosity stage)
template <class T>
__global__ __launch_bounds__(kBlock) void k_integrate_v_b(const BatchEntry<T> *tab,
                                                          const uint32_t *map, int from_xa) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  const SceneView<T> &S = E.S;
  const int i = static_cast<int>((u & kBatchLocalMask) * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  vec4_t<T> a;
  if (from_xa) {
    a = S.xa[S.xa_a(static_cast<size_t>(i))];
    S.a[i] = a;
  } else {
    a = S.a[i];
  }
  integrate_v_body(S, i, a);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_integrate_x_b(const BatchEntry<T> *tab,
                                                          const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  integrate_x_body(E.S, u & kBatchLocalMask);
}

// one block per scene
template <class T>
__global__ __launch_bounds__(kBlock) void k_speed_finalize_tab(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  speed_finalize_body<T>(E.S.ctl, E.S.partials, E.nblocks);
}

// ---- neighbour export (Neighborhood::CopyNeighbors, neighborhood.cpp:156-167)
template <class T>
__global__ __launch_bounds__(kBlock) void k_count_nbrs(SceneView<T> S, int qgroup, int ngroup,
                                                       uint32_t *cnt) {
  const int nq = qgroup == 1 ? S.nf : S.nb;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nq)
    return;
  const vec4_t<T> xi = qgroup == 1 ? S.xm[i] : S.bxV[i];
  const uint32_t cell = qgroup == 1 ? S.fcell[i] : S.bcell[i];
  uint32_t n = 0;
  for_neighbors(S, xi, cell, ngroup == 1 ? S.fstart : S.bstart, ngroup == 1 ? S.xm : S.bxV,
                [&](uint32_t, T, T, T, T, const vec4_t<T> &) { ++n; });
  cnt[qgroup == 1 ? S.id[i] : S.bid[i]] = n;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_write_nbrs(SceneView<T> S, int qgroup, int ngroup,
                                                       const uint32_t *offsets, uint32_t *out) {
  const int nq = qgroup == 1 ? S.nf : S.nb;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nq)
    return;
  const vec4_t<T> xi = qgroup == 1 ? S.xm[i] : S.bxV[i];
  const uint32_t cell = qgroup == 1 ? S.fcell[i] : S.bcell[i];
  const int32_t *nid = ngroup == 1 ? S.id : S.bid;
  uint32_t o = offsets[qgroup == 1 ? S.id[i] : S.bid[i]];
  for_neighbors(S, xi, cell, ngroup == 1 ? S.fstart : S.bstart, ngroup == 1 ? S.xm : S.bxV,
                [&](uint32_t j, T, T, T, T, const vec4_t<T> &) {
                  out[o++] = static_cast<uint32_t>(nid[j]);
                });
}

}  // namespace

template <class T>
void SphOps<T>::volume(hipStream_t st, const SceneView<T> &S) {
  if (S.nb > 0)
    hipLaunchKernelGGL(k_volume<T>, dim3(grid_blocks(S.nb)), dim3(kBlock), 0, st, S);
}

#define PRTCL_FLUID_LAUNCH(kern)                                                     \
  if (S.nf > 0)                                                                      \
  hipLaunchKernelGGL(kern, dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st, S)

template <class T>
void SphOps<T>::density(hipStream_t st, const SceneView<T> &S) {
  PRTCL_FLUID_LAUNCH(k_density<T>);
}
template <class T>
void SphOps<T>::gravity(hipStream_t st, const SceneView<T> &S) {
  PRTCL_FLUID_LAUNCH(k_gravity<T>);
}
template <class T>
void SphOps<T>::gravity_accumulate(hipStream_t st, const SceneView<T> &S) {
  PRTCL_FLUID_LAUNCH(k_gravity_accumulate<T>);
}
template <class T>
void SphOps<T>::gravity_axial(hipStream_t st, const SceneView<T> &S) {
  PRTCL_FLUID_LAUNCH(k_gravity_axial<T>);
}
template <class T>
void SphOps<T>::viscosity(hipStream_t st, const SceneView<T> &S) {
  PRTCL_FLUID_LAUNCH(k_viscosity<T>);
}
template <class T>
void SphOps<T>::integrate_v(hipStream_t st, const SceneView<T> &S) {
  PRTCL_FLUID_LAUNCH(k_integrate_v<T>);
}
template <class T>
void SphOps<T>::integrate_v_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                  unsigned nblocks, bool from_xa) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_integrate_v_b<T>, dim3(nblocks), dim3(kBlock), 0, st, tab, map,
                       from_xa ? 1 : 0);
}

template <class T>
void SphOps<T>::integrate_x_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                  unsigned nblocks, unsigned nscenes) {
  if (nblocks == 0)
    return;
  hipLaunchKernelGGL(k_integrate_x_b<T>, dim3(nblocks), dim3(kBlock), 0, st, tab, map);
  hipLaunchKernelGGL(k_speed_finalize_tab<T>, dim3(nscenes), dim3(kBlock), 0, st, tab);
}

template <class T>
void SphOps<T>::integrate_x(hipStream_t st, const SceneView<T> &S) {
  if (S.nf <= 0)
    return;
  const unsigned nb = grid_blocks(S.nf);
  hipLaunchKernelGGL(k_integrate_x<T>, dim3(nb), dim3(kBlock), 0, st, S);
  hipLaunchKernelGGL(k_speed_finalize<T>, dim3(1), dim3(kBlock), 0, st, S.ctl, S.partials, nb);
}
template <class T>
void SphOps<T>::iisph_setup(hipStream_t st, const SceneView<T> &S) {
  PRTCL_FLUID_LAUNCH(k_iisph_c<T>);
  PRTCL_FLUID_LAUNCH(k_iisph_s_aa<T>);
}
template <class T>
void SphOps<T>::iisph_accel(hipStream_t st, const SceneView<T> &S, bool solve) {
  if (solve) {
    PRTCL_FLUID_LAUNCH((k_iisph_accel<T, true>));
  } else {
    PRTCL_FLUID_LAUNCH((k_iisph_accel<T, false>));
  }
}
template <class T>
void SphOps<T>::iisph_pressure(hipStream_t st, const SceneView<T> &S, bool solve) {
  if (S.nf <= 0)
    return;
  const unsigned nb = grid_blocks(S.nf);
  if (solve) {
    hipLaunchKernelGGL((k_iisph_pressure<T, true>), dim3(nb), dim3(kBlock), 0, st, S);
    hipLaunchKernelGGL(k_pressure_finalize<true>, dim3(1), dim3(kBlock), 0, st, S.ctl,
                       S.partials, nb);
  } else {
    hipLaunchKernelGGL((k_iisph_pressure<T, false>), dim3(nb), dim3(kBlock), 0, st, S);
    hipLaunchKernelGGL(k_pressure_finalize<false>, dim3(1), dim3(kBlock), 0, st, S.ctl,
                       S.partials, nb);
  }
}
template <class T>
void SphOps<T>::count_neighbors(hipStream_t st, const SceneView<T> &S, int qgroup, int ngroup,
                                uint32_t *cnt) {
  const int nq = qgroup == 1 ? S.nf : S.nb;
  if (nq > 0)
    hipLaunchKernelGGL(k_count_nbrs<T>, dim3(grid_blocks(nq)), dim3(kBlock), 0, st, S, qgroup,
                       ngroup, cnt);
}
template <class T>
void SphOps<T>::write_neighbors(hipStream_t st, const SceneView<T> &S, int qgroup,
                                int ngroup, const uint32_t *offsets, uint32_t *out) {
  const int nq = qgroup == 1 ? S.nf : S.nb;
  if (nq > 0)
    hipLaunchKernelGGL(k_write_nbrs<T>, dim3(grid_blocks(nq)), dim3(kBlock), 0, st, S, qgroup,
                       ngroup, offsets, out);
}

template struct SphOps<float>;
template struct SphOps<double>;

}  // namespace prtcl_amd
