// pt16 vorticity-preserving viscosity (pt16.hpp / pt16.prtcl) on the device.
// Every pair term walks the 27-cell grid with the reference predicate (the
// particle itself included, as CopyNeighbors) and is evaluated in the
// reference's FP32 operation order (this file is built without FMA
// contraction); matrices are 3x3 column-major, Eigen's fixed-size sums of
// three terms are x0 + (x1 + x2). In fused steps the CG products use the
// step's neighbour list instead: the pair coefficients m_j W_ij are written
// once per step and streamed by every product (the self term first, then
// the list order). The two solves run the CG of
// cg_openmp.hpp:32-215 once per vector component, with the FP64 device
// control of cg_ctl.hpp, as the reference's `foreach dimension` loop of
// scalar CGOpenMP<T> solves (pt16.hpp:420-519, 608-760).
#include "cg_ctl.hpp"
#include "exact_walk.hpp"
#include "nbrlist.hpp"
#include "pt16.hpp"

namespace prtcl_amd {

namespace {

// component d (0, 1, 2) of the real4 at p, addressed arithmetically (base +
// 4 d): a select over the vector type's x / y / z accessors inside the
// neighbour loop was compiled without the address of the d = 2 case
template <class T>
__device__ __forceinline__ T comp(const vec4_t<T> *p, int d) {
  return reinterpret_cast<const T *>(p)[d];
}

template <class T>
__device__ __forceinline__ T *comp_ptr(vec4_t<T> *p, int d) {
  return reinterpret_cast<T *>(p) + d;
}

// ---- list form (fused steps)
// index of the coefficient m_j W_ij of list entry e of particle i (chunks of
// 4 entries, so a lane reads 4 with one load and a wavefront's load is
// contiguous); written by the set-up, read by the products and the
// vorticity right-hand side of the same step
__device__ __forceinline__ size_t coef_index(size_t n, int i, uint32_t e) {
  return ((static_cast<size_t>(e / 4) * n + static_cast<size_t>(i)) * 4) + e % 4;
}

constexpr uint32_t kSelfEntry = 0xFFFFFFFEu;  // the particle itself
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;    // a grid-walk neighbour (list overflowed)

// The fluid neighbours of particle i in a fused step: fn(j, d0, d1, d2, r2,
// x_j, e) for the particle itself first (d = 0, e = kSelfEntry), then the
// list entries in list order (e = entry index); a particle whose list
// overflowed walks the grid (itself included, e = kNoEntry).
template <class T, class Fn>
__device__ __forceinline__ void fluid_pairs_list(const SceneView<T> &S, const NbrList &L,
                                                 unsigned lb, int i, const vec4_t<T> &xv,
                                                 Fn &&fn) {
  const uint32_t info = L.info[i];
  if (info & kListOverflow) {
    nbrs_exact(S, xv, S.fcell[i], S.fstart, S.xm,
               [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
                 fn(j, d0, d1, d2, r2, pj, kNoEntry);
               });
    return;
  }
  fn(static_cast<uint32_t>(i), T(0), T(0), T(0), T(0), xv, kSelfEntry);
  list_visit_indexed(L, Spans::of(L.blk[lb]), static_cast<size_t>(S.nf), i, info,
                     [&](uint32_t e, uint32_t j) {
                       const vec4_t<T> pj = S.xm[j];
                       const T d0 = xv.x - pj.x, d1 = xv.y - pj.y, d2 = xv.z - pj.z;
                       fn(j, d0, d1, d2, d0 * d0 + (d1 * d1 + d2 * d2), pj, e);
                     });
}

// ---- pt16.setup; with kList (fused steps) one list pass for the fluid sums
// that also writes the step's pair coefficients m_j W_ij
template <class T, bool kList>
__device__ __forceinline__ void pt16_setup_body(const SceneView<T> &S, const NbrList &L, T xi,
                                                unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> xv = S.xm[i];
  const vec4_t<T> vi = S.v[i];
  const T rhoi = S.rho[i];
  const uint32_t cell = S.fcell[i];
  // same-group density and the smoothed density (pt16.prtcl:53-71), and the
  // velocity gradient sum_f m_j (v_j - v_i) gradW_ij^T / rho_i
  // (pt16.prtcl:76-88)
  T same = 0, prho = 0;
  T vg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  auto grad_term = [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
    const vec4_t<T> vj = S.v[j];
    const T dv[3] = {vj.x - vi.x, vj.y - vi.y, vj.z - vi.z};
    const T f = S.K.gradf(r2);
    const T g[3] = {f * d0, f * d1, f * d2};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r)
        vg[r + 3 * c] += pj.w * (dv[r] * g[c]);
  };
  if (kList) {
    const size_t n = static_cast<size_t>(S.nf);
    fluid_pairs_list(S, L, lb, i, xv,
                     [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj, uint32_t e) {
                       const T w = S.K.w(r2);
                       same += pj.w * w;
                       prho += (pj.w / S.rho0) * w;
                       if (e < kSelfEntry)
                         S.pcoef[coef_index(n, i, e)] = pj.w * w;
                       if (e != kSelfEntry)  // the particle's own term is zero
                         grad_term(j, d0, d1, d2, r2, pj);
                     });
    nbrs_exact(S, xv, cell, S.bstart, S.bxV,
               [&](uint32_t, T, T, T, T r2, const vec4_t<T> &pb) { prho += pb.w * S.K.w(r2); });
    prho *= S.rho0;
  } else {
    nbrs_exact(S, xv, cell, S.fstart, S.xm,
               [&](uint32_t, T, T, T, T r2, const vec4_t<T> &pj) {
                 same += pj.w * S.K.w(r2);
               });
    nbrs_exact(S, xv, cell, S.fstart, S.xm,
               [&](uint32_t, T, T, T, T r2, const vec4_t<T> &pj) {
                 prho += (pj.w / S.rho0) * S.K.w(r2);
               });
    nbrs_exact(S, xv, cell, S.bstart, S.bxV,
               [&](uint32_t, T, T, T, T r2, const vec4_t<T> &pb) { prho += pb.w * S.K.w(r2); });
    prho *= S.rho0;
    nbrs_exact(S, xv, cell, S.fstart, S.xm, grad_term);
  }
#pragma unroll
  for (int k = 0; k < 9; ++k)
    vg[k] /= rhoi;
  // spin, expansion and shear rate (pt16.prtcl:90-105)
  const T div = vg[0] + (vg[4] + vg[8]);
  T R[9], V[9], Sh[9];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      R[r + 3 * c] = (vg[r + 3 * c] - vg[c + 3 * r]) / T(2);
      V[r + 3 * c] = (div / T(3)) * ((r == c) ? T(1) : T(0));
      Sh[r + 3 * c] = ((vg[r + 3 * c] + vg[c + 3 * r]) / T(2)) - V[r + 3 * c];
    }
  // vorticity = 2 (R(2,1), R(0,2), R(1,0)); the target velocity gradient
  // without the spin: V unit_step_r(0, rho - rho0, -div) + xi S
  S.om[i] = make4<vec4_t<T>>(T(2) * R[2 + 3 * 1], T(2) * R[0 + 3 * 2], T(2) * R[1 + 3 * 0], T(0));
  const T step = ((T(0) <= rhoi - S.rho0) || (T(0) <= -div)) ? T(1) : T(0);
  T *G = S.tvg + 9 * static_cast<size_t>(i);
#pragma unroll
  for (int k = 0; k < 9; ++k)
    G[k] = V[k] * step + xi * Sh[k];
  S.psrho[i] = same;
  S.prho[i] = prho;
}

template <class T, bool kList>
__global__ __launch_bounds__(kBlock) void k_pt16_setup(SceneView<T> S, NbrList L, T xi) {
  pt16_setup_body<T, kList>(S, L, xi, kList ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x);
}

// ---- vorticity diffusion rhs; with kList the coefficients m_j W_ij of the
// set-up are streamed and only w_j is gathered (the particle's own term is
// zero)
template <class T, bool kList>
__device__ __forceinline__ void pt16_omega_rhs_body(const SceneView<T> &S, const NbrList &L,
                                                    T xi, unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> xv = S.xm[i];
  const vec4_t<T> wi = S.om[i];
  T a0 = 0, a1 = 0, a2 = 0;
  auto exact = [&] {
    nbrs_exact(S, xv, S.fcell[i], S.fstart, S.xm,
               [&](uint32_t j, T, T, T, T r2, const vec4_t<T> &pj) {
                 const vec4_t<T> wj = S.om[j];
                 const T w = S.K.w(r2);
                 a0 += (pj.w * (wi.x - wj.x)) * w;
                 a1 += (pj.w * (wi.y - wj.y)) * w;
                 a2 += (pj.w * (wi.z - wj.z)) * w;
               });
  };
  if (!kList || (L.info[i] & kListOverflow)) {
    exact();
  } else {
    const size_t n = static_cast<size_t>(S.nf);
    list_visit_indexed(L, Spans::of(L.blk[lb]), n, i, L.info[i], [&](uint32_t e, uint32_t j) {
      const vec4_t<T> wj = S.om[j];
      const T c = S.pcoef[coef_index(n, i, e)];
      a0 += c * (wi.x - wj.x);
      a1 += c * (wi.y - wj.y);
      a2 += c * (wi.z - wj.z);
    });
  }
  S.orhs[i] = make4<vec4_t<T>>(a0 * xi, a1 * xi, a2 * xi, T(0));
}

template <class T, bool kList>
__global__ __launch_bounds__(kBlock) void k_pt16_omega_rhs(SceneView<T> S, NbrList L, T xi) {
  pt16_omega_rhs_body<T, kList>(S, L, xi, kList ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x);
}

// ---- velocity reconstruction rhs
template <class T, bool kList>
__device__ __forceinline__ void pt16_vrec_rhs_body(const SceneView<T> &S, const NbrList &L,
                                                   unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> xv = S.xm[i];
  T Gi[9];
#pragma unroll
  for (int k = 0; k < 9; ++k)
    Gi[k] = S.tvg[9 * static_cast<size_t>(i) + k];
  T a[3] = {0, 0, 0};
  auto term = [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj) {
    const T *Gj = S.tvg + 9 * static_cast<size_t>(j);
    T M[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
      M[k] = Gi[k] + Gj[k];
    const T mh = pj.w * T(0.5);
    const T w = S.K.w(r2);
#pragma unroll
    for (int r = 0; r < 3; ++r)
      a[r] += (mh * (M[r] * d0 + (M[r + 3] * d1 + M[r + 6] * d2))) * w;
  };
  if (kList)
    fluid_pairs_list(S, L, lb, i, xv,
                     [&](uint32_t j, T d0, T d1, T d2, T r2, const vec4_t<T> &pj, uint32_t e) {
                       if (e != kSelfEntry)  // x_ii = 0: the particle's own term is zero
                         term(j, d0, d1, d2, r2, pj);
                     });
  else
    nbrs_exact(S, xv, S.fcell[i], S.fstart, S.xm, term);
  S.vrhs[i] = make4<vec4_t<T>>(a[0], a[1], a[2], T(0));
}

template <class T, bool kList>
__global__ __launch_bounds__(kBlock) void k_pt16_vrec_rhs(SceneView<T> S, NbrList L) {
  pt16_vrec_rhs_body<T, kList>(S, L, kList ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x);
}

// ---- spin: G += cross_product_matrix(0.5 w) (eigen.hpp:330-339)
template <class T>
__device__ __forceinline__ void pt16_spin_body(const SceneView<T> &S, unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> w = S.om[i];
  const T a0 = T(0.5) * w.x, a1 = T(0.5) * w.y, a2 = T(0.5) * w.z;
  T *G = S.tvg + 9 * static_cast<size_t>(i);
  // column-major (r, c) at r + 3c: [[0, -a2, a1], [a2, 0, -a0], [-a1, a0, 0]]
  const T C[9] = {T(0), a2, -a1, -a2, T(0), a0, a1, -a0, T(0)};
#pragma unroll
  for (int k = 0; k < 9; ++k)
    G[k] += C[k];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_spin(SceneView<T> S) {
  pt16_spin_body(S, blockIdx.x);
}

// ---- scalar PCG (one component of one system)
// system diagonal coefficient and preconditioner of the reference:
//   (A x)_i = c_i x_i - sum_f (m_j W_ij) x_j, c = pt16_same_density (which 0)
//   or pt16_density (which 1); P y = y / (pt16_same_density_i - m_i W(0))
template <class T>
__device__ __forceinline__ T product_s(const SceneView<T> &S, int which, int i, const T *w) {
#pragma clang fp contract(off)
  const vec4_t<T> xv = S.xm[i];
  T out = (which == 0 ? S.psrho[i] : S.prho[i]) * w[i];
  nbrs_exact(S, xv, S.fcell[i], S.fstart, S.xm,
             [&](uint32_t j, T, T, T, T r2, const vec4_t<T> &pj) {
               out -= (pj.w * S.K.w(r2)) * w[j];
             });
  return out;
}

template <class T>
__device__ __forceinline__ T product_p(const SceneView<T> &S, int i, T r) {
#pragma clang fp contract(off)
  return r / (S.psrho[i] - S.xm[i].w * S.K.w(T(0)));
}

// ---- list form of the products (fused steps)
// (A w)_i = c_i w_i - m_i W(0) w_i - sum_list (m_j W_ij) w_j with w_j at
// w[ws * j] (ws = 4: one component of a real4 field); the coefficients of a
// chunk of 8 entries and the 8 gathers are issued before the arithmetic
template <class T>
__device__ __forceinline__ T product_s_list(const SceneView<T> &S, const NbrList &L, int which,
                                            int i, unsigned lb, const T *__restrict__ w,
                                            int ws) {
#pragma clang fp contract(off)
  const vec4_t<T> xv = S.xm[i];
  const T wi = w[static_cast<size_t>(ws) * i];
  T out = (which == 0 ? S.psrho[i] : S.prho[i]) * wi;
  const uint32_t info = L.info[i];
  if (info & kListOverflow) {
    nbrs_exact(S, xv, S.fcell[i], S.fstart, S.xm,
               [&](uint32_t j, T, T, T, T r2, const vec4_t<T> &pj) {
                 out -= (pj.w * S.K.w(r2)) * w[static_cast<size_t>(ws) * j];
               });
    return out;
  }
  out -= (xv.w * S.K.w(T(0))) * wi;
  const size_t n = static_cast<size_t>(S.nf);
  const Spans R = Spans::of(L.blk[lb]);
  const uint32_t c0 = info & 0xFFu;
  const uint32_t c01 = c0 + ((info >> 8) & 0xFFu);
  const uint32_t c = list_count(info);
  const uint32_t d1 = R.b1 - R.b0, d2 = R.b2 - R.b1;
  const vec4_t<T> *__restrict__ coef = reinterpret_cast<const vec4_t<T> *>(S.pcoef);
  for (uint32_t g = 0; g * kChunk < c; ++g) {
    const uint4 q = list_chunk(L, n, i, g);
    const vec4_t<T> ca = coef[(2 * static_cast<size_t>(g)) * n + i];
    vec4_t<T> cb = ca;
    if (g * kChunk + 4 < c)
      cb = coef[(2 * static_cast<size_t>(g) + 1) * n + i];
    T wj[kChunk];
#pragma unroll
    for (int k = 0; k < static_cast<int>(kChunk); ++k) {
      const uint32_t e = g * kChunk + k;
      if (e < c) {
        const uint32_t j = R.b0 + chunk_entry(q, k) + (e >= c0 ? d1 : 0u) + (e >= c01 ? d2 : 0u);
        wj[k] = w[static_cast<size_t>(ws) * j];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    const T cf[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
#pragma unroll
    for (int k = 0; k < static_cast<int>(kChunk); ++k)
      if (g * kChunk + k < c)
        out -= cf[k] * wj[k];
  }
  return out;
}

// x = guess, r = A x - b, y = P r, p = -y; partials b.b, r.r, r.y
template <class T, bool kList>
__device__ __forceinline__ void pt16_cg_setup_body(const SceneView<T> &S, const NbrList &L,
                                                   int which, int dim, double *part,
                                                   unsigned lb, unsigned nb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < S.nf) {
    const vec4_t<T> *guess = which == 0 ? S.om : S.v;
    const T b = comp<T>((which == 0 ? S.orhs : S.vrhs) + i, dim);
    const T x0 = comp<T>(guess + i, dim);
    // A x0 over the guess field's component
    T ax;
    if (kList) {
      ax = product_s_list(S, L, which, i, lb, reinterpret_cast<const T *>(guess) + dim, 4);
    } else {
      const vec4_t<T> xv = S.xm[i];
      ax = (which == 0 ? S.psrho[i] : S.prho[i]) * x0;
      nbrs_exact(S, xv, S.fcell[i], S.fstart, S.xm,
                 [&](uint32_t j, T, T, T, T r2, const vec4_t<T> &pj) {
                   ax -= (pj.w * S.K.w(r2)) * comp<T>(guess + j, dim);
                 });
    }
    const T r = ax - b;
    const T y = product_p(S, i, r);
    S.scx[i] = x0;
    S.scr[i] = r;
    S.scy[i] = y;
    S.scp[i] = -y;
    acc[0] = static_cast<double>(b * b);
    acc[1] = static_cast<double>(r * r);
    acc[2] = static_cast<double>(r * y);
  }
  store_partials<3>(part, nb, acc, lb);
}

template <class T, bool kList>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_setup(SceneView<T> S, NbrList L, int which,
                                                          int dim, double *part) {
  pt16_cg_setup_body<T, kList>(S, L, which, dim, part, xcd_block(blockIdx.x, gridDim.x),
                               gridDim.x);
}

// q = A p, partial p.q
template <class T, bool kList>
__device__ __forceinline__ void pt16_cg_spmv_body(const SceneView<T> &S, const NbrList &L,
                                                  int which, double *part, unsigned lb,
                                                  unsigned nb) {
#pragma clang fp contract(off)
  if (S.ctl->cg_done)
    return;
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double acc[1] = {0.0};
  if (i < S.nf) {
    const T q = kList ? product_s_list(S, L, which, i, lb, S.scp, 1)
                      : product_s(S, which, i, S.scp);
    S.scq[i] = q;
    acc[0] = static_cast<double>(S.scp[i] * q);
  }
  store_partials<1>(part, nb, acc, lb);
}

template <class T, bool kList>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_spmv(SceneView<T> S, NbrList L, int which,
                                                         double *part) {
  pt16_cg_spmv_body<T, kList>(S, L, which, part, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}

// x += alpha p, r += alpha q, y = P r; partials r.r, r.y
template <class T>
__device__ __forceinline__ void pt16_update_body(const SceneView<T> &S, T alpha, double *part,
                                                 unsigned lb, unsigned nb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double acc[2] = {0.0, 0.0};
  if (i < S.nf) {
    const T x = S.scx[i] + alpha * S.scp[i];
    const T r = S.scr[i] + alpha * S.scq[i];
    const T y = product_p(S, i, r);
    S.scx[i] = x;
    S.scr[i] = r;
    S.scy[i] = y;
    acc[0] = static_cast<double>(r * r);
    acc[1] = static_cast<double>(r * y);
  }
  store_partials<2>(part, nb, acc, lb);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_update(SceneView<T> S, double *part) {
  if (S.ctl->cg_done)
    return;
  pt16_update_body(S, static_cast<T>(S.ctl->cg_alpha), part, blockIdx.x, gridDim.x);
}

// p = -y + beta p
template <class T>
__device__ __forceinline__ void pt16_direction_body(const SceneView<T> &S, T beta, unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  S.scp[i] = -S.scy[i] + beta * S.scp[i];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_direction(SceneView<T> S) {
  if (S.ctl->cg_done)
    return;
  pt16_direction_body(S, static_cast<T>(S.ctl->cg_beta), blockIdx.x);
}

// the iteration without its two one-block kernels (cg_alpha_in_block,
// cg_beta_in_block)
template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_update_alpha(SceneView<T> S,
                                                                 const double *pq_part,
                                                                 double *part) {
  if (cg_done_for_block(S.ctl))
    return;
  double alpha;
  if (!cg_alpha_in_block(S.ctl, pq_part, gridDim.x, &alpha))
    return;
  pt16_update_body(S, static_cast<T>(alpha), part, blockIdx.x, gridDim.x);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_direction_beta(SceneView<T> S,
                                                                   const double *part) {
  if (cg_done_for_block(S.ctl))
    return;
  double beta;
  if (!cg_beta_in_block(S.ctl, part, gridDim.x, &beta))
    return;
  pt16_direction_body(S, static_cast<T>(beta), blockIdx.x);
}

// apply (pt16.hpp:491-502, 700-717): component dim of the vorticity (which 0)
// or the velocity (which 1) = x unit_step_l(0, (t - t_birth) - dt_fade) +
// old unit_step_r(0, dt_fade - (t - t_birth))
template <class T>
__device__ __forceinline__ void pt16_cg_apply_body(const SceneView<T> &S, int which, int dim,
                                                   unsigned lb) {
#pragma clang fp contract(off)
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const T tb = S.v[i].w;
  const T age = S.t - tb;
  const T on = (T(0) < age - S.fade) ? T(1) : T(0);
  const T keep = (T(0) <= S.fade - age) ? T(1) : T(0);
  T *f = comp_ptr<T>((which == 0 ? S.om : S.v) + i, dim);
  *f = S.scx[i] * on + *f * keep;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_apply(SceneView<T> S, int which, int dim) {
  pt16_cg_apply_body(S, which, dim, blockIdx.x);
}

// ---- batched forms (prtcl_hip_step over several scenes, list forms): the
// bodies above over the blocks of every pt16 scene of a table (map: block
// map, entry of scene k in bits 20.., its local block below), each block with
// its scene's view, list, CG partials and block count. Solve s of a step
// (vorticity components 0-2, velocity components 3-5) runs under sequence
// number cg_seq + s with the tolerance and cap of its system.
template <class T>
__device__ __forceinline__ NbrList pt16_list(const BatchEntry<T> &E) {
  return NbrList{E.nidx, E.ninfo, E.nblk};
}

#define PT16_BATCH_ENTRY                              \
  const uint32_t u = map[blockIdx.x];                 \
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits]; \
  const unsigned lb = u & kBatchLocalMask

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_setup_b(const BatchEntry<T> *tab,
                                                         const uint32_t *map) {
  PT16_BATCH_ENTRY;
  pt16_setup_body<T, true>(E.S, pt16_list(E), static_cast<T>(E.pt16_xi), lb);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_omega_rhs_b(const BatchEntry<T> *tab,
                                                             const uint32_t *map) {
  PT16_BATCH_ENTRY;
  pt16_omega_rhs_body<T, true>(E.S, pt16_list(E), static_cast<T>(E.pt16_xi), lb);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_spin_b(const BatchEntry<T> *tab,
                                                        const uint32_t *map) {
  PT16_BATCH_ENTRY;
  pt16_spin_body(E.S, lb);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_vrec_rhs_b(const BatchEntry<T> *tab,
                                                            const uint32_t *map) {
  PT16_BATCH_ENTRY;
  pt16_vrec_rhs_body<T, true>(E.S, pt16_list(E), lb);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_setup_b(const BatchEntry<T> *tab,
                                                            const uint32_t *map, int which,
                                                            int dim) {
  PT16_BATCH_ENTRY;
  pt16_cg_setup_body<T, true>(E.S, pt16_list(E), which, dim, E.cg_part, lb, E.nblocks);
}

// one block per scene
template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_setup_fin_b(const BatchEntry<T> *tab,
                                                                int which, int solve) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  cg_setup_fin_body(E.S.ctl, E.cg_part, E.nblocks, E.pt16_tol[which], E.pt16_max_k[which],
                    E.cg_seq + static_cast<unsigned>(solve));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_spmv_b(const BatchEntry<T> *tab,
                                                           const uint32_t *map, int which) {
  PT16_BATCH_ENTRY;
  pt16_cg_spmv_body<T, true>(E.S, pt16_list(E), which, E.cg_part, lb, E.nblocks);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_alpha_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  if (!E.cg_fold)
    cg_alpha_body(E.S.ctl, E.cg_part, E.nblocks);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_beta_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  if (!E.cg_fold)
    cg_beta_body(E.S.ctl, E.cg_part, E.nblocks);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_update_b(const BatchEntry<T> *tab,
                                                             const uint32_t *map) {
  PT16_BATCH_ENTRY;
  const unsigned nb = E.nblocks;
  Ctl *c = E.S.ctl;
  if (E.cg_fold) {  // p.q partials at [0, nb), r.r and r.y at [nb, 3 nb)
    if (cg_done_for_block(c))
      return;
    double alpha;
    if (!cg_alpha_in_block(c, E.cg_part, nb, &alpha, lb))
      return;
    pt16_update_body(E.S, static_cast<T>(alpha), E.cg_part + nb, lb, nb);
  } else {
    if (c->cg_done)
      return;
    pt16_update_body(E.S, static_cast<T>(c->cg_alpha), E.cg_part, lb, nb);
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_direction_b(const BatchEntry<T> *tab,
                                                                const uint32_t *map) {
  PT16_BATCH_ENTRY;
  const unsigned nb = E.nblocks;
  Ctl *c = E.S.ctl;
  if (E.cg_fold) {
    if (cg_done_for_block(c))
      return;
    double beta;
    if (!cg_beta_in_block(c, E.cg_part + nb, nb, &beta, lb))
      return;
    pt16_direction_body(E.S, static_cast<T>(beta), lb);
  } else {
    if (c->cg_done)
      return;
    pt16_direction_body(E.S, static_cast<T>(c->cg_beta), lb);
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pt16_cg_apply_b(const BatchEntry<T> *tab,
                                                            const uint32_t *map, int which,
                                                            int dim) {
  PT16_BATCH_ENTRY;
  pt16_cg_apply_body(E.S, which, dim, lb);
}

#undef PT16_BATCH_ENTRY

}  // namespace

template <class T>
void Pt16Ops<T>::setup_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                             unsigned nblocks) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_pt16_setup_b<T>, dim3(nblocks), dim3(kBlock), 0, st, tab, map);
}

template <class T>
void Pt16Ops<T>::omega_rhs_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                 unsigned nblocks) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_pt16_omega_rhs_b<T>, dim3(nblocks), dim3(kBlock), 0, st, tab, map);
}

template <class T>
void Pt16Ops<T>::spin_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                            unsigned nblocks) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_pt16_spin_b<T>, dim3(nblocks), dim3(kBlock), 0, st, tab, map);
}

template <class T>
void Pt16Ops<T>::vrec_rhs_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                unsigned nblocks) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_pt16_vrec_rhs_b<T>, dim3(nblocks), dim3(kBlock), 0, st, tab, map);
}

template <class T>
void Pt16Ops<T>::cg_setup_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                unsigned nblocks, unsigned nscenes, int which, int dim,
                                int solve) {
  if (nscenes == 0)
    return;
  if (nblocks > 0)
    hipLaunchKernelGGL(k_pt16_cg_setup_b<T>, dim3(nblocks), dim3(kBlock), 0, st, tab, map,
                       which, dim);
  hipLaunchKernelGGL(k_pt16_cg_setup_fin_b<T>, dim3(nscenes), dim3(kBlock), 0, st, tab, which,
                     solve);
}

template <class T>
void Pt16Ops<T>::cg_iteration_batch(hipStream_t st, const BatchEntry<T> *tab,
                                    const uint32_t *map, unsigned nblocks, unsigned nscenes,
                                    int which, bool any_split) {
  if (nblocks == 0)
    return;
  const dim3 gr(nblocks), bl(kBlock);
  hipLaunchKernelGGL(k_pt16_cg_spmv_b<T>, gr, bl, 0, st, tab, map, which);
  if (any_split)
    hipLaunchKernelGGL(k_pt16_cg_alpha_b<T>, dim3(nscenes), bl, 0, st, tab);
  hipLaunchKernelGGL(k_pt16_cg_update_b<T>, gr, bl, 0, st, tab, map);
  if (any_split)
    hipLaunchKernelGGL(k_pt16_cg_beta_b<T>, dim3(nscenes), bl, 0, st, tab);
  hipLaunchKernelGGL(k_pt16_cg_direction_b<T>, gr, bl, 0, st, tab, map);
}

template <class T>
void Pt16Ops<T>::cg_apply_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                unsigned nblocks, int which, int dim) {
  if (nblocks > 0)
    hipLaunchKernelGGL(k_pt16_cg_apply_b<T>, dim3(nblocks), dim3(kBlock), 0, st, tab, map, which,
                       dim);
}

template <class T>
void Pt16Ops<T>::setup(hipStream_t st, const SceneView<T> &S, const NbrList *L, T xi) {
  if (S.nf <= 0)
    return;
  if (L)
    hipLaunchKernelGGL((k_pt16_setup<T, true>), dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st, S,
                       *L, xi);
  else
    hipLaunchKernelGGL((k_pt16_setup<T, false>), dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st,
                       S, NbrList{}, xi);
}

template <class T>
void Pt16Ops<T>::omega_rhs(hipStream_t st, const SceneView<T> &S, const NbrList *L, T xi) {
  if (S.nf <= 0)
    return;
  if (L)
    hipLaunchKernelGGL((k_pt16_omega_rhs<T, true>), dim3(grid_blocks(S.nf)), dim3(kBlock), 0,
                       st, S, *L, xi);
  else
    hipLaunchKernelGGL((k_pt16_omega_rhs<T, false>), dim3(grid_blocks(S.nf)), dim3(kBlock), 0,
                       st, S, NbrList{}, xi);
}

template <class T>
void Pt16Ops<T>::vrec_rhs(hipStream_t st, const SceneView<T> &S, const NbrList *L) {
  if (S.nf <= 0)
    return;
  if (L)
    hipLaunchKernelGGL((k_pt16_vrec_rhs<T, true>), dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st,
                       S, *L);
  else
    hipLaunchKernelGGL((k_pt16_vrec_rhs<T, false>), dim3(grid_blocks(S.nf)), dim3(kBlock), 0,
                       st, S, NbrList{});
}

template <class T>
void Pt16Ops<T>::spin(hipStream_t st, const SceneView<T> &S) {
  if (S.nf > 0)
    hipLaunchKernelGGL(k_pt16_spin<T>, dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st, S);
}

template <class T>
void Pt16Ops<T>::cg_setup(hipStream_t st, const SceneView<T> &S, const NbrList *L, int which,
                          int dim, double *part, double tol, int max_k, unsigned seq) {
  const unsigned nb = grid_blocks(S.nf);
  if (S.nf > 0) {
    if (L)
      hipLaunchKernelGGL((k_pt16_cg_setup<T, true>), dim3(nb), dim3(kBlock), 0, st, S, *L,
                         which, dim, part);
    else
      hipLaunchKernelGGL((k_pt16_cg_setup<T, false>), dim3(nb), dim3(kBlock), 0, st, S,
                         NbrList{}, which, dim, part);
  }
  hipLaunchKernelGGL(k_cg_setup_fin, dim3(1), dim3(kBlock), 0, st, S.ctl, part, nb, tol, max_k,
                     seq);
}

template <class T>
void Pt16Ops<T>::cg_iteration(hipStream_t st, const SceneView<T> &S, const NbrList *L,
                              int which, double *part, bool split) {
  if (S.nf <= 0)
    return;
  const unsigned nb = grid_blocks(S.nf);
  const dim3 gr(nb), bl(kBlock);
  if (L)
    hipLaunchKernelGGL((k_pt16_cg_spmv<T, true>), gr, bl, 0, st, S, *L, which, part);
  else
    hipLaunchKernelGGL((k_pt16_cg_spmv<T, false>), gr, bl, 0, st, S, NbrList{}, which, part);
  if (!split && nb <= kFusedCgBlocks) {
    // p.q partials at part[0, nb), r.r and r.y at part[nb, 3 nb)
    hipLaunchKernelGGL(k_pt16_cg_update_alpha<T>, gr, bl, 0, st, S, part, part + nb);
    hipLaunchKernelGGL(k_pt16_cg_direction_beta<T>, gr, bl, 0, st, S, part + nb);
    return;
  }
  hipLaunchKernelGGL(k_cg_alpha, dim3(1), bl, 0, st, S.ctl, part, nb);
  hipLaunchKernelGGL(k_pt16_cg_update<T>, gr, bl, 0, st, S, part);
  hipLaunchKernelGGL(k_cg_beta, dim3(1), bl, 0, st, S.ctl, part, nb);
  hipLaunchKernelGGL(k_pt16_cg_direction<T>, gr, bl, 0, st, S);
}

template <class T>
void Pt16Ops<T>::cg_apply(hipStream_t st, const SceneView<T> &S, int which, int dim) {
  if (S.nf > 0)
    hipLaunchKernelGGL(k_pt16_cg_apply<T>, dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st, S,
                       which, dim);
}

template struct Pt16Ops<float>;
template struct Pt16Ops<double>;

}  // namespace prtcl_amd
