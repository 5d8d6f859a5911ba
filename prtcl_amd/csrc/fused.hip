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
// accepted fluid neighbours (self excluded) in a per-step list of 16-bit
// slab indices (nbrlist.hpp); the IISPH passes stage each slab of their
// block's neighbourhood in LDS and gather the neighbours' fields from there.
// A particle with more neighbours than the list holds is walked over the
// grid instead.
#include <type_traits>

#include "fused.hpp"

namespace prtcl_amd {

namespace {

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

// Exact-predicate walk of particle i's 27-cell candidates in global memory:
// fn(j) for every fluid j != i with |x_i - x_j|^2 < r^2 (particles whose list
// overflowed).
template <class T, class Fn>
__device__ __forceinline__ void grid_walk(const SceneView<T> &S, int i, const vec4_t<T> &xi,
                                          Fn &&fn) {
  runs(S.grid, S.fcell[i], S.fstart, [&](uint32_t b, uint32_t e) {
    for (uint32_t j = b; j < e; ++j) {
      const vec4_t<T> pj = S.xm[j];
      if (j != static_cast<uint32_t>(i) &&
          sqdist_exact(xi.x - pj.x, xi.y - pj.y, xi.z - pj.z) < S.thr2)
        fn(j);
    }
  });
}

// Calls fn(j) for every fluid neighbour j != i of particle i, from the list
// (global slots) or over the grid for a particle whose list overflowed.
template <class T, class Fn>
__device__ __forceinline__ void walk(const SceneView<T> &S, const NbrList &L, int i,
                                     const vec4_t<T> &xi, Fn &&fn) {
  const uint32_t info = L.info[i];
  if (!(info & kListOverflow))
    list_visit(L, Spans::of(L.blk[i / kBlock]), static_cast<size_t>(S.nf), i, info, fn);
  else
    grid_walk(S, i, xi, fn);
}

// the three x-runs of slab s (rows dz = s - 1, dy = -1..1) of a cell, in
// stencil order: fn(b, e); rows outside the box are skipped
template <class Fn>
__device__ __forceinline__ void slab_runs(const GridDims &g, uint32_t cellc,
                                          const uint32_t *__restrict__ start, int s, Fn &&fn) {
  const int cx = static_cast<int>(cellc & 0x7FFu);
  const int cy = static_cast<int>((cellc >> 11) & 0x7FFu);
  const int z = static_cast<int>(cellc >> 22) + s - 1;
  if (z < 0 || z >= g.nz)
    return;
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx + 1 < g.nx ? cx + 1 : g.nx - 1;
#pragma unroll
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

// bounds of the three x-runs of slab s of a cell, loaded together (rows
// outside the box give empty runs)
__device__ __forceinline__ void slab_bounds(const GridDims &g, uint32_t cellc,
                                            const uint32_t *__restrict__ start, int s,
                                            uint32_t rb[3], uint32_t re[3]) {
  const int cx = static_cast<int>(cellc & 0x7FFu);
  const int cy = static_cast<int>((cellc >> 11) & 0x7FFu);
  const int z = static_cast<int>(cellc >> 22) + s - 1;
  const bool zok = z >= 0 && z < g.nz;
  const uint32_t x0 = static_cast<uint32_t>(cx > 0 ? cx - 1 : 0);
  const uint32_t x1 = static_cast<uint32_t>(cx + 1 < g.nx ? cx + 1 : g.nx - 1);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int y = cy + r - 1;
    const bool ok = zok && y >= 0 && y < g.ny;
    const uint32_t row = ok ? static_cast<uint32_t>(g.nx) *
                                  (static_cast<uint32_t>(y) +
                                   static_cast<uint32_t>(g.ny) * static_cast<uint32_t>(z))
                            : 0u;
    rb[r] = ok ? start[row + x0] : 0u;
    re[r] = ok ? start[row + x1 + 1u] : 0u;
  }
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const uint32_t u = __shfl_xor(v, o, kWave);
    v = u < v ? u : v;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const uint32_t u = __shfl_xor(v, o, kWave);
    v = u > v ? u : v;
  }
  return v;
}

// Spans of a block (all threads call it): span s runs from the first start
// to the last end of the slab-s runs of the block's particles. Written to R
// (LDS).
__device__ __forceinline__ void block_spans(const GridDims &g, const uint32_t *fstart,
                                            bool valid, uint32_t cell, BlockRanges &R) {
  __shared__ uint32_t lo[3], hi[3];
  if (threadIdx.x < 3) {
    lo[threadIdx.x] = 0xFFFFFFFFu;
    hi[threadIdx.x] = 0u;
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    uint32_t bm = 0xFFFFFFFFu, em = 0u;
    if (valid)
      slab_runs(g, cell, fstart, s, [&](uint32_t b, uint32_t e) {
        if (e > b) {
          bm = b < bm ? b : bm;
          em = e > em ? e : em;
        }
      });
    bm = wave_min_u32(bm);
    em = wave_max_u32(em);
    if (__lane_id() == 0) {
      atomicMin(&lo[s], bm);
      atomicMax(&hi[s], em);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const uint32_t b = lo[threadIdx.x], e = hi[threadIdx.x];
    R.base[threadIdx.x] = b < e ? b : 0u;
    R.len[threadIdx.x] = b < e ? e - b : 0u;
  }
  __syncthreads();
}

// IISPH iteration 1 of a solve without its neighbour passes (the loop of
// example.lua:409-421 over iisph.hpp:358-516): every pressure is 0 when the
// solve starts (iisph.hpp:282), so the pressure acceleration of iteration 1
// is zero, Ap = 0, and the relaxed Jacobi update and the aprde term need only
// the particle's own s and AA. The expressions are those of
// k_pressure_list_body with p_old = Ap = 0, so p and the aprde term are the
// ones the two passes give, bit for bit. Returns the particle's aprde term.
template <class T>
__device__ __forceinline__ double first_iteration(const SceneView<T> &S, int i,
                                                  const vec4_t<T> &xi, T rhoi, T si, T AA) {
  const T p_old = T(0), Ap = T(0);
  const T rcp = (fabs(AA) > T(1e-9)) ? T(1) / AA : T(0);
  const T upd = p_old + S.omega * (si - Ap) * rcp;
  const T pn = ((T(0) < upd) ? upd : T(0)) * ((T(1e-9) < fabs(AA)) ? T(1) : T(0));
  S.Ap[i] = Ap;
  S.p[i] = pn;
  S.pr[i] = pn / (rhoi * rhoi);
  S.xq[i] = make4<vec4_t<T>>(xi.x, xi.y, xi.z, xi.w * pn / (rhoi * rhoi));
  // the pressure acceleration of iteration 1 (read if the solve stops there)
  S.xa[S.xa_a(static_cast<size_t>(i))] = make4<vec4_t<T>>(T(0), T(0), T(0), T(0));
  return static_cast<double>((Ap - si) / S.rho0 * ((T(1e-9) < pn) ? T(1) : T(0)));
}

// ---------------------------------------------------------------------------
// setup pass: density + neighbour list + IISPH c, AA (+ s when the velocity
// of every neighbour is known without another pass, i.e. no explicit
// viscosity: v*_j = v_j + dt g H_j)
//
// Phase 1 tests the 27-cell candidates with the exact predicate and appends
// the accepted neighbours (entries relative to the slab's span) to this
// thread's column of an LDS table; phase 2 runs the pair sums over the
// compacted entries, so the lanes that reject a candidate do not wait for the
// kernel arithmetic of the ones that accept it. The table is written to the
// list row by row (coalesced). With kStage the block stages each slab's span
// in LDS and runs both phases per slab against the staged copy; otherwise
// the candidates are read from global memory (through L2) and both phases
// run once over all 27 cells.
// ---------------------------------------------------------------------------
template <class T, bool kWithSource, bool kStage>
__device__ __forceinline__ void k_setup_body(const SceneView<T> &S, const NbrList &L,
                                             vec4_t<T> *vout, unsigned lb) {
  constexpr uint32_t kCap = slab_cap<T>();
  __shared__ vec4_t<T> sx[kStage ? kCap + kRunPad : 1];
  // rows kNbrCap + 1 .. kNbrCap + 8 take the writes of a lane past its cap
  __shared__ uint16_t lst[kNbrCap + 9][kBlock];
  __shared__ BlockRanges sR;
  const unsigned t = threadIdx.x;
  const int i = static_cast<int>(lb * kBlock + t);
  const bool valid = i < S.nf;
  const uint32_t cell = valid ? S.fcell[i] : 0u;
  block_spans(S.grid, S.fstart, valid, cell, sR);
  if (t < 6)
    reinterpret_cast<uint32_t *>(L.blk + lb)[t] = t < 3 ? sR.base[t] : sR.len[t - 3];
  const Spans R = Spans::of(sR);
  const uint32_t smax = R.max_len();
  const bool staged = kStage && smax <= kCap;
  const bool encodable = smax <= 0x10000u;

  const FastSpline<T> K = fast_spline(S);
  const vec4_t<T> zero4 = make4<vec4_t<T>>(T(0), T(0), T(0), T(0));
  const vec4_t<T> xi = valid ? S.xm[i] : zero4;
  const vec4_t<T> vi = valid ? S.v[i] : zero4;
  const T Hi = fade_step(S, vi.w);
  const T dtg0 = S.dt * S.g0, dtg1 = S.dt * S.g1, dtg2 = S.dt * S.g2;
  const T vs0 = vi.x + dtg0 * Hi, vs1 = vi.y + dtg1 * Hi, vs2 = vi.z + dtg2 * Hi;
  // density (self term W(0) = 4 alpha / h^3), G_f, sum |gradW|^2 and the
  // velocity-divergence sum of the source term
  T rho = xi.w * K.aw * T(4), gf0 = 0, gf1 = 0, gf2 = 0, s2 = 0, dsum = 0;
  // pair sums of neighbour (x_j, m_j) = pj with velocity vj (read only
  // with kWithSource)
  auto pair2 = [&](const vec4_t<T> &pj, const vec4_t<T> &vj) {
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
      const T Hj = fade_step(S, vj.w);
      const T u0 = vs0 - (vj.x + dtg0 * Hj);
      const T u1 = vs1 - (vj.y + dtg1 * Hj);
      const T u2 = vs2 - (vj.z + dtg2 * Hj);
      dsum += pj.w * (u0 * g0 + u1 * g1 + u2 * g2);
    }
  };
  auto load_v = [&](uint32_t j) { return kWithSource ? S.v[j] : zero4; };
  auto pair = [&](const vec4_t<T> &pj, uint32_t j) { pair2(pj, load_v(j)); };
  uint32_t cnt = 0;
  uint32_t cs[3] = {0u, 0u, 0u};  // entries per slab
  // candidate test of the slab-s runs, positions from global memory
  // (kLds = false) or from the staged span; accepted entries are appended
  auto scan = [&](int s, auto lds_tag) {
    constexpr bool kLds = decltype(lds_tag)::value;
    const uint32_t base = R.base(s);
    uint32_t rb[3], re[3];
    slab_bounds(S.grid, cell, S.fstart, s, rb, re);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const uint32_t b = rb[r], e = re[r];
      const bool own = (s == 1) & (r == 1);  // the run holding particle i
      // eight candidates per trip from consecutive slots, loads issued
      // together. Branch-free append: every candidate is written to row cnt
      // of this thread's column of the LDS table and cnt advances only for
      // an accepted one. The table has 8 rows beyond kNbrCap + 1 and cnt is
      // clamped to kNbrCap + 1 after each trip, so the writes of a trip stay
      // in the table and a count above kNbrCap still reads as overflow. Full
      // trips need no bound test; the last, partial trip of a run reads up
      // to 7 slots past it (the position arrays carry kRunPad slots of
      // padding) and rejects them. Only the run of the particle's own cell
      // can hold the particle itself.
      auto trip = [&](uint32_t j, auto tail_tag) {
        constexpr bool kTail = decltype(tail_tag)::value;
        const vec4_t<T> *__restrict__ src = kLds ? sx + (j - base) : S.xm + j;
        vec4_t<T> pk[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          pk[u] = src[u];
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t lim = e - j;                          // u < lim: inside the run
        const uint32_t self = static_cast<uint32_t>(i) - j;  // u == self: the particle
        const uint32_t ent = j - base;                       // entry of candidate 0
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const bool ok = (!kTail | (static_cast<uint32_t>(u) < lim)) &
                          (!own | (static_cast<uint32_t>(u) != self)) &
                          (sqdist_exact(xi.x - pk[u].x, xi.y - pk[u].y, xi.z - pk[u].z) <
                           S.thr2);
          lst[cnt][t] = static_cast<uint16_t>(ent + u);
          cnt += ok ? 1u : 0u;
        }
        cnt = cnt < kNbrCap + 1u ? cnt : kNbrCap + 1u;
      };
      uint32_t j = b;
      for (; j + 8u <= e; j += 8u)
        trip(j, std::false_type{});
      if (j < e)
        trip(j, std::true_type{});
    }
  };
  bool over = !encodable;
  if (staged) {
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      stage_span(R, s, S.xm, sx);
      __syncthreads();
      if (valid && !over) {
        const uint32_t c0 = cnt;
        scan(s, std::true_type{});
        if (cnt > kNbrCap) {
          over = true;
        } else {
          cs[s] = cnt - c0;
          const uint32_t base = R.base(s);
#pragma unroll 4
          for (uint32_t e = c0; e < cnt; ++e) {
            const uint32_t u = lst[e][t];
            pair(sx[u], base + u);
          }
        }
      }
      __syncthreads();
    }
  } else if (valid && !over) {
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const uint32_t c0 = cnt;
      scan(s, std::false_type{});
      cs[s] = cnt - c0;
    }
    if (cnt > kNbrCap) {
      over = true;
    } else {
      const uint32_t c0 = cs[0], c01 = cs[0] + cs[1];
      const uint32_t d1 = R.b1 - R.b0, d2 = R.b2 - R.b1;
      auto slot = [&](uint32_t e) {
        return R.b0 + lst[e][t] + (e >= c0 ? d1 : 0u) + (e >= c01 ? d2 : 0u);
      };
      // groups of four: every gather of a group ahead of its arithmetic
      uint32_t e = 0;
      for (; e + 4 <= cnt; e += 4) {
        vec4_t<T> px[4], pv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t j = slot(e + k);
          px[k] = S.xm[j];
          pv[k] = load_v(j);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < 4; ++k)
          pair2(px[k], pv[k]);
      }
      for (; e < cnt; ++e) {
        const uint32_t j = slot(e);
        pair2(S.xm[j], load_v(j));
      }
    }
  }
  // the rest is per particle; with S.first_iter the block then sums the
  // aprde terms of IISPH iteration 1 (every thread takes part)
  auto finish = [&]() -> double {
  const size_t nn = static_cast<size_t>(S.nf);
  uint32_t info;
  if (over) {
    // more neighbours than the list holds: all pair sums again over the grid
    rho = xi.w * K.aw * T(4);
    gf0 = gf1 = gf2 = s2 = dsum = T(0);
    grid_walk(S, i, xi, [&](uint32_t j) { pair(S.xm[j], j); });
    info = kListOverflow;
  } else {
    // chunks of 8 entries, one 16-byte store per lane (1 KB per wavefront)
    for (uint32_t g = 0; g * kChunk < cnt; ++g) {
      uint32_t w[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t e0 = g * kChunk + 2 * k;
        const uint32_t lo = e0 < cnt ? lst[e0][t] : 0u;
        const uint32_t hi = e0 + 1 < cnt ? lst[e0 + 1][t] : 0u;
        w[k] = lo | (hi << 16);
      }
      reinterpret_cast<uint4 *>(L.idx)[static_cast<size_t>(g) * nn + static_cast<size_t>(i)] =
          make_uint4(w[0], w[1], w[2], w[3]);
    }
    info = cs[0] | (cs[1] << 8) | (cs[2] << 16);
  }
  L.info[i] = info;
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
  S.xa[S.xa_x(static_cast<size_t>(i))] = xi;
  if (kWithSource) {
    const T src = (rho0 - rho) -
                  S.dt * rho0 * irho * (dsum + rho0 * (vs0 * gb0 + vs1 * gb1 + vs2 * gb2));
    S.s[i] = src;
    vout[i] = make4<vec4_t<T>>(vs0, vs1, vs2, vi.w);
    if (S.first_iter)
      return first_iteration(S, i, xi, rho, src, AA);
  } else {
    S.vd[i] = make4<vec4_t<T>>(gb0, gb1, gb2, T(0));  // G_b for the source pass
  }
  return 0.0;
  };
  const double term = valid ? finish() : 0.0;
  if (kWithSource && S.first_iter) {
    const double bs = block_sum<kBlock>(term);
    if (t == 0)
      S.partials[lb] = bs;
  }
}

template <class T, bool kWithSource, bool kStage>
__global__ __launch_bounds__(kBlock) void k_setup(SceneView<T> S, NbrList L,
                                                  vec4_t<T> *vout, SolveBegin B) {
  k_setup_body<T, kWithSource, kStage>(S, L, vout, xcd_block(blockIdx.x, gridDim.x));
  // no thread of the setup reads the control block
  if (B.seq != 0u && blockIdx.x == 0 && threadIdx.x == 0)
    begin_solve(S.ctl, B);
}

// explicit viscosity (viscosity.hpp:171-240) over the list, then
// v* = v + dt (g + a_visc) H   (gravity.hpp:133, symplectic_euler.hpp:172-200)
template <class T>
__device__ __forceinline__ void k_viscosity_list_body(const SceneView<T> &S, const NbrList &L,
                                                      vec4_t<T> *vout, unsigned lb) {
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
  T g0, g1, g2;
  step_gravity(S, xi.x, xi.y, xi.z, g0, g1, g2);
  vout[i] = make4<vec4_t<T>>(vi.x + S.dt * (g0 + a0) * H, vi.y + S.dt * (g1 + a1) * H,
                             vi.z + S.dt * (g2 + a2) * H, vi.w);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_viscosity_list(SceneView<T> S, NbrList L,
                                                           vec4_t<T> *vout) {
  k_viscosity_list_body(S, L, vout, xcd_block(blockIdx.x, gridDim.x));
}

// v* = v + dt a_axial(x) H (gravity.hpp:189-215, symplectic_euler.hpp:172-200)
// for steps with axial gravity and no explicit viscosity
template <class T>
__device__ __forceinline__ void k_gravity_v_body(const SceneView<T> &S, vec4_t<T> *vout,
                                                 unsigned lb) {
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  if (i >= S.nf)
    return;
  const vec4_t<T> xi = S.xm[i], vi = S.v[i];
  const T H = fade_step(S, vi.w);
  T g0, g1, g2;
  step_gravity(S, xi.x, xi.y, xi.z, g0, g1, g2);
  vout[i] = make4<vec4_t<T>>(vi.x + S.dt * g0 * H, vi.y + S.dt * g1 * H,
                             vi.z + S.dt * g2 * H, vi.w);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_gravity_v(SceneView<T> S, vec4_t<T> *vout) {
  k_gravity_v_body(S, vout, blockIdx.x);
}

// source term s from the integrated velocities (explicit-viscosity path)
template <class T>
__device__ __forceinline__ void k_source_list_body(const SceneView<T> &S, const NbrList &L,
                                                   unsigned lb) {
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  double term = 0.0;
  if (i < S.nf) {
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
    const T src = (rho0 - rhoi) -
                  S.dt * rho0 / rhoi * (dsum + rho0 * (vi.x * gb.x + vi.y * gb.y + vi.z * gb.z));
    S.s[i] = src;
    if (S.first_iter)
      term = first_iteration(S, i, xi, rhoi, src, S.AA[i]);
  }
  if (S.first_iter) {  // IISPH iteration 1: the block's aprde partial
    const double bs = block_sum<kBlock>(term);
    if (threadIdx.x == 0)
      S.partials[lb] = bs;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_source_list(SceneView<T> S, NbrList L) {
  k_source_list_body(S, L, xcd_block(blockIdx.x, gridDim.x));
}

// Walk of the list entries of particle i (all threads of the block call it).
// kStage and spans that fit LDS: per slab, stage(s) fills LDS (all threads),
// then use(lds(u)) for the neighbour at LDS index u. Otherwise
// use(load(j)) over the whole list with global slots, a chunk of 8 gathers
// in flight.
template <bool kStage, uint32_t kCap, int kGroup = 8, class Stage, class Lds, class Load,
          class Use>
__device__ __forceinline__ void list_walk(const NbrList &L, const Spans &R, size_t n, int i,
                                          uint32_t info, Stage &&stage, Lds &&lds, Load &&load,
                                          Use &&use) {
  const uint32_t live = (info & kListOverflow) ? 0u : info;
  if (kStage && R.max_len() <= kCap) {
    uint32_t e = 0;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const uint32_t c = (live >> (8 * s)) & 0xFFu;
      stage(s);
      __syncthreads();
      uint32_t q = 0;
      for (; q + 4 <= c; q += 4) {
        const uint32_t u0 = list_entry(L, n, i, e + q), u1 = list_entry(L, n, i, e + q + 1);
        const uint32_t u2 = list_entry(L, n, i, e + q + 2), u3 = list_entry(L, n, i, e + q + 3);
        use(lds(u0));
        use(lds(u1));
        use(lds(u2));
        use(lds(u3));
      }
      for (; q < c; ++q)
        use(lds(list_entry(L, n, i, e + q)));
      __syncthreads();
      e += c;
    }
  } else {
    list_gather<kGroup>(L, R, n, i, live, load, use);
  }
}

// IISPH pressure acceleration over the list: a_i = p_i c_i - sum m_j pr_j gradW
template <class T, bool kStage>
__device__ __forceinline__ void k_accel_list_body(const SceneView<T> &S, const NbrList &L,
                                                  unsigned lb) {
  if (S.ctl->done)
    return;
  constexpr uint32_t kCap = slab_cap<T>();
  __shared__ vec4_t<T> sq[kStage ? kCap : 1];
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  const bool valid = i < S.nf;
  const Spans R = Spans::of(L.blk[lb]);
  const FastSpline<T> K = fast_spline(S);
  const vec4_t<T> zero4 = make4<vec4_t<T>>(T(0), T(0), T(0), T(0));
  const vec4_t<T> xi = valid ? S.xm[i] : zero4;
  const uint32_t info = valid ? L.info[i] : kListOverflow;
  // own fields loaded ahead of the walk so their latency overlaps it
  const vec4_t<T> ci = valid ? S.c[i] : zero4;
  const T pi = valid ? S.p[i] : T(0);
  T a0 = 0, a1 = 0, a2 = 0;
  auto pair = [&](const vec4_t<T> &qj) {  // (x_j, m_j p_j / rho_j^2)
    const T d0 = xi.x - qj.x, d1 = xi.y - qj.y, d2 = xi.z - qj.z;
    const T s = qj.w * K.gradf_unscaled(d0 * d0 + d1 * d1 + d2 * d2);
    a0 += s * d0;
    a1 += s * d1;
    a2 += s * d2;
  };
  list_walk<kStage, kCap>(
      L, R, static_cast<size_t>(S.nf), i, info, [&](int s) { stage_span(R, s, S.xq, sq); },
      [&](uint32_t u) { return sq[u]; }, [&](uint32_t j) { return S.xq[j]; }, pair);
  if (!valid)
    return;
  if (info & kListOverflow)
    grid_walk(S, i, xi, [&](uint32_t j) { pair(S.xq[j]); });
  const vec4_t<T> ai = make4<vec4_t<T>>(pi * ci.x - K.ag3 * a0, pi * ci.y - K.ag3 * a1,
                                        pi * ci.z - K.ag3 * a2, T(0));
  // the pressure pass gathers it from here; the integration copies it to a
  S.xa[S.xa_a(static_cast<size_t>(i))] = ai;
}

template <class T, bool kStage>
__global__ __launch_bounds__(kBlock) void k_accel_list(SceneView<T> S, NbrList L) {
  k_accel_list_body<T, kStage>(S, L, xcd_block(blockIdx.x, gridDim.x));
}

// IISPH relaxed Jacobi update over the list, FP64 aprde reduction and the
// device-side convergence decision (example.lua:409-421)
template <class T, bool kStage>
__device__ __forceinline__ void k_pressure_list_body(const SceneView<T> &S, const NbrList &L,
                                                     unsigned lb) {
  if (S.ctl->done)
    return;
  constexpr uint32_t kCap = slab_cap<T>();
  struct XA {
    vec4_t<T> x, a;  // (x, m) and (a, 0) of a neighbour
  };
  __shared__ vec4_t<T> sx[kStage ? kCap : 1];
  __shared__ vec4_t<T> sa[kStage ? kCap : 1];
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  const bool valid = i < S.nf;
  const Spans R = Spans::of(L.blk[lb]);
  const FastSpline<T> K = fast_spline(S);
  const vec4_t<T> zero4 = make4<vec4_t<T>>(T(0), T(0), T(0), T(0));
  const vec4_t<T> xi = valid ? S.xm[i] : zero4;
  const uint32_t info = valid ? L.info[i] : kListOverflow;
  // own fields loaded ahead of the walk so their latency overlaps it
  const vec4_t<T> ai = valid ? S.xa[S.xa_a(static_cast<size_t>(i))] : zero4;
  const vec4_t<T> gs = valid ? S.G[i] : zero4;
  const T rhoi = valid ? S.rho[i] : T(1), AA = valid ? S.AA[i] : T(0);
  const T si = valid ? S.s[i] : T(0), p_old = valid ? S.p[i] : T(0);
  T sum = 0;
  auto pair = [&](const XA &q) {
    const T d0 = xi.x - q.x.x, d1 = xi.y - q.x.y, d2 = xi.z - q.x.z;
    sum += q.x.w * K.gradf_unscaled(d0 * d0 + d1 * d1 + d2 * d2) *
           (q.a.x * d0 + q.a.y * d1 + q.a.z * d2);
  };
  list_walk<kStage, kCap, 4>(
      L, R, static_cast<size_t>(S.nf), i, info,
      [&](int s) {
        stage_span(R, s, S.xm, sx);
        stage_span_fn(R, s, [&](size_t g) { return S.xa[S.xa_a(g)]; }, sa);
      },
      [&](uint32_t u) { return XA{sx[u], sa[u]}; },
      [&](uint32_t j) { return XA{S.xa[S.xa_x(j)], S.xa[S.xa_a(j)]}; },
      pair);
  double term = 0.0;
  if (valid) {
    if (info & kListOverflow)
      grid_walk(S, i, xi, [&](uint32_t j) {
        pair(XA{S.xa[S.xa_x(j)], S.xa[S.xa_a(j)]});
      });
    const T Ap = S.dt * S.dt * S.rho0 / rhoi *
                 ((ai.x * gs.x + ai.y * gs.y + ai.z * gs.z) - K.ag3 * sum);
    const T rcp = (fabs(AA) > T(1e-9)) ? T(1) / AA : T(0);
    const T upd = p_old + S.omega * (si - Ap) * rcp;
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

template <class T, bool kStage>
__global__ __launch_bounds__(kBlock) void k_pressure_list(SceneView<T> S, NbrList L) {
  k_pressure_list_body<T, kStage>(S, L, xcd_block(blockIdx.x, gridDim.x));
}

// Sum of the block partials in block order (FP64, fixed tree) and the
// convergence decision of example.lua:418-419: stop when it >= min_it and
// aprde / nprde < tol, or at max_it.
__device__ __forceinline__ void k_pressure_finalize_body(Ctl *c, const double *partials,
                                                        unsigned nparts) {
  if (c->done)
    return;
  const double acc = block_sum<kBlock>(strided_sum<kBlock>(partials, nparts));
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

// v += dt a H (pressure acceleration), x += dt v, running max |v|; seq > 0:
// only once solve seq is done (integrate_now)
template <class T>
__device__ __forceinline__ void k_integrate_body(SceneView<T> S, unsigned lb, unsigned seq) {
  if (seq != 0u && !integrate_now(S.ctl, seq))
    return;
  const int i = static_cast<int>(lb * kBlock + threadIdx.x);
  T speed = T(0);
  bool bad = false;
  if (i < S.nf) {
    vec4_t<T> v = S.v[i];
    const vec4_t<T> a = S.xa[S.xa_a(static_cast<size_t>(i))];
    S.a[i] = a;  // the acceleration field after the step, as the reference's
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
__global__ __launch_bounds__(kBlock) void k_integrate(SceneView<T> S, unsigned seq) {
  k_integrate_body<T>(S, blockIdx.x, seq);
}

// running maximum speed (symplectic_euler.hpp:250-268) from the block maxima;
// seq > 0: after the integration of solve seq's step, which it marks taken
template <class T>
__device__ __forceinline__ void k_speed_finalize_body(Ctl *c, const double *partials,
                                                      unsigned nparts, unsigned seq) {
  if (seq != 0u && !integrate_now(c, seq))
    return;
  const double m = block_max<kBlock>(strided_max<kBlock>(partials, nparts, 0.0));
  if (threadIdx.x == 0) {
    max_speed_atomic<T>(c, static_cast<T>(m));
    if (seq != 0u)
      c->integrated = seq;
    publish_step_end(c);
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_speed_finalize(Ctl *c, const double *partials,
                                                           unsigned nparts, unsigned seq) {
  k_speed_finalize_body<T>(c, partials, nparts, seq);
}

// ---- batched wrappers (prtcl_hip_step over several scenes)
template <class T, bool kStage>
__global__ __launch_bounds__(kBlock) void k_setup_b(const BatchEntry<T> *tab,
                                                    const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_setup_body<T, true, kStage>(E.S, NbrList{E.nidx, E.ninfo, E.nblk}, E.vtmp,
                                u & kBatchLocalMask);
}

// scenes with explicit viscosity or axial gravity: the setup without the
// source term, then v* (explicit viscosity and gravity, or the axial gravity
// alone) into vtmp, then the source term and iteration 1 from v* (the three
// launches of FusedOps::setup for such a scene)
template <class T, bool kStage>
__global__ __launch_bounds__(kBlock) void k_setup_nosrc_b(const BatchEntry<T> *tab,
                                                          const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_setup_body<T, false, kStage>(E.S, NbrList{E.nidx, E.ninfo, E.nblk}, E.vtmp,
                                 u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_vstar_b(const BatchEntry<T> *tab,
                                                    const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  if (E.explicit_visc)
    k_viscosity_list_body(E.S, NbrList{E.nidx, E.ninfo, E.nblk}, E.vtmp, u & kBatchLocalMask);
  else
    k_gravity_v_body(E.S, E.vtmp, u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_source_b(const BatchEntry<T> *tab,
                                                     const uint32_t *map) {
  const uint32_t u = map[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  SceneView<T> S2 = E.S;
  S2.v = E.vtmp;  // the integrated velocities
  k_source_list_body(S2, NbrList{E.nidx, E.ninfo, E.nblk}, u & kBatchLocalMask);
}

template <class T, bool kStage>
__global__ __launch_bounds__(kBlock) void k_accel_b(const BatchEntry<T> *tab,
                                                    const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_accel_list_body<T, kStage>(E.S, NbrList{E.nidx, E.ninfo, E.nblk}, u & kBatchLocalMask);
}

template <class T, bool kStage>
__global__ __launch_bounds__(kBlock) void k_pressure_b(const BatchEntry<T> *tab,
                                                       const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_pressure_list_body<T, kStage>(E.S, NbrList{E.nidx, E.ninfo, E.nblk}, u & kBatchLocalMask);
}

// one block per scene
template <class T>
__global__ __launch_bounds__(kBlock) void k_pressure_finalize_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  k_pressure_finalize_body(E.S.ctl, E.S.partials, E.nblocks);
}

// iteration 1's convergence step for the scenes whose solve it could end
template <class T>
__global__ __launch_bounds__(kBlock) void k_first_finalize_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  if (!E.first_preset)
    k_pressure_finalize_body(E.S.ctl, E.S.partials, E.nblocks);
}

// guarded: each scene's integration only once its solve (E.seq) is done;
// scenes whose step ends with the WKBB18 path are left to it
template <class T>
__global__ __launch_bounds__(kBlock) void k_integrate_b(const BatchEntry<T> *tab,
                                                        const uint32_t *pmap, int guarded) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  if (E.viscous)
    return;
  k_integrate_body<T>(E.S, u & kBatchLocalMask, guarded ? E.seq : 0u);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_speed_finalize_b(const BatchEntry<T> *tab,
                                                             int guarded) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  if (E.viscous)
    return;
  k_speed_finalize_body<T>(E.S.ctl, E.S.partials, E.nblocks, guarded ? E.seq : 0u);
}

}  // namespace

template <class T>
void FusedOps<T>::setup(hipStream_t st, const SceneView<T> &S, const NbrList &L,
                        bool explicit_viscosity, vec4_t<T> *vtmp, bool lds,
                        const SolveBegin &B) {
  if (S.nf <= 0)
    return;
  const dim3 gr(grid_blocks(S.nf)), bl(kBlock);
  // the one-pass setup folds a uniform g into every neighbour's velocity;
  // explicit viscosity or a position-dependent (axial) gravity need the
  // integrated velocities first
  if (!explicit_viscosity && !S.axial) {
    if (lds)
      hipLaunchKernelGGL((k_setup<T, true, true>), gr, bl, 0, st, S, L, vtmp, B);
    else
      hipLaunchKernelGGL((k_setup<T, true, false>), gr, bl, 0, st, S, L, vtmp, B);
    return;
  }
  if (lds)
    hipLaunchKernelGGL((k_setup<T, false, true>), gr, bl, 0, st, S, L, vtmp, B);
  else
    hipLaunchKernelGGL((k_setup<T, false, false>), gr, bl, 0, st, S, L, vtmp, B);
  if (explicit_viscosity)
    hipLaunchKernelGGL(k_viscosity_list<T>, gr, bl, 0, st, S, L, vtmp);
  else
    hipLaunchKernelGGL(k_gravity_v<T>, gr, bl, 0, st, S, vtmp);
  // the integrated velocities replace v once every thread has read the old ones
  SceneView<T> S2 = S;
  S2.v = vtmp;
  hipLaunchKernelGGL(k_source_list<T>, gr, bl, 0, st, S2, L);
}

template <class T>
void FusedOps<T>::accel(hipStream_t st, const SceneView<T> &S, const NbrList &L, bool lds) {
  if (S.nf <= 0)
    return;
  const dim3 gr(grid_blocks(S.nf)), bl(kBlock);
  if (lds)
    hipLaunchKernelGGL((k_accel_list<T, true>), gr, bl, 0, st, S, L);
  else
    hipLaunchKernelGGL((k_accel_list<T, false>), gr, bl, 0, st, S, L);
}

template <class T>
void FusedOps<T>::pressure(hipStream_t st, const SceneView<T> &S, const NbrList &L, bool lds) {
  if (S.nf <= 0)
    return;
  const unsigned nb = grid_blocks(S.nf);
  if (lds)
    hipLaunchKernelGGL((k_pressure_list<T, true>), dim3(nb), dim3(kBlock), 0, st, S, L);
  else
    hipLaunchKernelGGL((k_pressure_list<T, false>), dim3(nb), dim3(kBlock), 0, st, S, L);
  hipLaunchKernelGGL(k_pressure_finalize, dim3(1), dim3(kBlock), 0, st, S.ctl, S.partials, nb);
}

// a = the acceleration of the last pressure-acceleration pass
template <class T>
__global__ __launch_bounds__(kBlock) void k_copy_accel(SceneView<T> S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < S.nf)
    S.a[i] = S.xa[S.xa_a(static_cast<size_t>(i))];
}

template <class T>
void FusedOps<T>::finish_first_iteration(hipStream_t st, const SceneView<T> &S) {
  if (S.nf > 0)
    hipLaunchKernelGGL(k_pressure_finalize, dim3(1), dim3(kBlock), 0, st, S.ctl, S.partials,
                       grid_blocks(S.nf));
}

template <class T>
void FusedOps<T>::finish_first_iteration_batch(hipStream_t st, const BatchEntry<T> *tab,
                                               unsigned nscenes) {
  if (nscenes > 0)
    hipLaunchKernelGGL(k_first_finalize_b<T>, dim3(nscenes), dim3(kBlock), 0, st, tab);
}

template <class T>
void FusedOps<T>::copy_accel(hipStream_t st, const SceneView<T> &S) {
  if (S.nf > 0)
    hipLaunchKernelGGL(k_copy_accel<T>, dim3(grid_blocks(S.nf)), dim3(kBlock), 0, st, S);
}

template <class T>
void FusedOps<T>::integrate(hipStream_t st, const SceneView<T> &S, unsigned seq) {
  if (S.nf <= 0)
    return;
  const unsigned nb = grid_blocks(S.nf);
  hipLaunchKernelGGL(k_integrate<T>, dim3(nb), dim3(kBlock), 0, st, S, seq);
  hipLaunchKernelGGL(k_speed_finalize<T>, dim3(1), dim3(kBlock), 0, st, S.ctl, S.partials, nb,
                     seq);
}

template <class T>
void FusedOps<T>::setup_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                              unsigned npblocks, bool lds) {
  if (npblocks == 0)
    return;
  if (lds)
    hipLaunchKernelGGL((k_setup_b<T, true>), dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
  else
    hipLaunchKernelGGL((k_setup_b<T, false>), dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
}

template <class T>
void FusedOps<T>::setup_split_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                    unsigned nblocks, bool lds) {
  if (nblocks == 0)
    return;
  const dim3 gr(nblocks), bl(kBlock);
  if (lds)
    hipLaunchKernelGGL((k_setup_nosrc_b<T, true>), gr, bl, 0, st, tab, map);
  else
    hipLaunchKernelGGL((k_setup_nosrc_b<T, false>), gr, bl, 0, st, tab, map);
  hipLaunchKernelGGL(k_vstar_b<T>, gr, bl, 0, st, tab, map);
  hipLaunchKernelGGL(k_source_b<T>, gr, bl, 0, st, tab, map);
}

template <class T>
void FusedOps<T>::accel_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                              unsigned npblocks, bool lds) {
  if (npblocks == 0)
    return;
  if (lds)
    hipLaunchKernelGGL((k_accel_b<T, true>), dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
  else
    hipLaunchKernelGGL((k_accel_b<T, false>), dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
}

template <class T>
void FusedOps<T>::pressure_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                                 unsigned npblocks, unsigned nscenes, bool lds) {
  if (npblocks == 0)
    return;
  if (lds)
    hipLaunchKernelGGL((k_pressure_b<T, true>), dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
  else
    hipLaunchKernelGGL((k_pressure_b<T, false>), dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
  hipLaunchKernelGGL(k_pressure_finalize_b<T>, dim3(nscenes), dim3(kBlock), 0, st, tab);
}

template <class T>
void FusedOps<T>::integrate_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                                  unsigned npblocks, unsigned nscenes, bool guarded) {
  if (npblocks == 0)
    return;
  const int g = guarded ? 1 : 0;
  hipLaunchKernelGGL(k_integrate_b<T>, dim3(npblocks), dim3(kBlock), 0, st, tab, pmap, g);
  hipLaunchKernelGGL(k_speed_finalize_b<T>, dim3(nscenes), dim3(kBlock), 0, st, tab, g);
}

template struct FusedOps<float>;
template struct FusedOps<double>;

}  // namespace prtcl_amd
