// Device-resident scene: the fluid and boundary groups of one model, their
// uniform grid and the solver control block, laid out for coalesced access on
// MI355X. Shared by the kernels (nsearch.hip, sph.hip, wkbb18.hip) and the C
// ABI (capi.hip).
//
// Layout in HBM (per scene, slot order = the grid's cell order, rebuilt at
// every neighbourhood update):
//   fluid  xm  real4 (x, y, z, mass)           16/32 B
//          v   real4 (vx, vy, vz, time_of_birth)
//          a, c, G, vd real4; rho, p, pr, s, AA, Ap real; sp real[9]
//          id  int32 host index of the particle in each slot
//          cell uint32 packed cell coordinates (x | y << 11 | z << 22)
//   boundary bxV real4 (x, y, z, volume), bid int32, bcell uint32
//   grid   fstart / bstart uint32[ncells + 1]: exclusive prefix sums of the
//          per-cell particle counts over a dense box of cells of edge r,
//          x fastest, so the three cells x-1..x+1 of a row are one slot range.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "device_common.hpp"

namespace prtcl_amd {

constexpr int kBlock = 256;
constexpr int kMaxAxisCellsXY = 2047;  // 11 bits
constexpr int kMaxAxisCellsZ = 1023;   // 10 bits
constexpr uint32_t kNbrCap = 48;       // fluid neighbours kept per particle and step
// slots of padding after the position arrays (xm, xq): candidate trips of 8
// may read up to 7 slots past the end of a run
constexpr uint32_t kRunPad = 8;

// per-scene control block (device memory)
struct Ctl {
  double aprde;      // iisph_aprde
  double rel;        // aprde / nprde of the last iteration
  double tol;        // IISPH tolerance
  unsigned ticket;   // last-block counter of the pressure reduction
  unsigned ticket2;  // last-block counter of the other reductions
  int nprde;         // iisph_nprde
  int iters;         // iterations done in the current solve
  int done;          // solve finished (converged or capped)
  int converged;
  int min_it, max_it;
  int nonfinite;
  unsigned max_speed_bits_f;       // running max |v| (float bits)
  unsigned long long max_speed_bits_d;  // running max |v| (double bits)
  // host-visible progress word (pinned, mapped): (seq << 32) | (done << 31) | iters,
  // written by the solver's finalize kernel so the host can poll it
  unsigned long long *host_mirror;
  unsigned seq;
  // PCG scalars (FP64)
  double cg_b_nsq, cg_thr, cg_r_nsq, cg_ry, cg_ry_prev, cg_pq, cg_alpha, cg_beta;
  int cg_iter, cg_done, cg_max_k;
  int cg_iter_next;  // cg_iter + 1, set beside alpha (read while cg_iter is written)
  double cg_tol;
  // end-of-step report: steps finished (integrations of the position) and
  // how many of them produced a non-finite position or speed; published to
  // host_mirror[3] as (step_count << 32) | nonfinite_steps
  unsigned step_count;
  unsigned nonfinite_steps;
  // the IISPH solve (seq) whose step the integration kernels have taken
  unsigned integrated;
};

// An integration enqueued before the host knows that IISPH solve `seq` is
// done (right after the predicted iterations) does its work only once that
// solve is done, and only once: a launch that finds the solve still running
// or its step already integrated exits at its first instruction. Every block
// of the launch reads the same words, written by earlier kernels.
__device__ __forceinline__ bool integrate_now(const Ctl *c, unsigned seq) {
  return c->seq == seq && c->done != 0 && c->integrated != seq;
}

// host_mirror words: [0] IISPH progress, [1] its aprde / nprde, [2] CG
// progress, [3] end-of-step report
constexpr int kMirrorStep = 3;

// Counts the finished step (one thread; after the integration kernel of the
// step): folds the non-finite flag of the step into the running count,
// clears it and publishes the report with one system-scope store.
__device__ __forceinline__ void publish_step_end(Ctl *c) {
  c->step_count += 1u;
  c->nonfinite_steps += c->nonfinite ? 1u : 0u;
  c->nonfinite = 0;
  if (c->host_mirror) {
    const unsigned long long w = (static_cast<unsigned long long>(c->step_count) << 32) |
                                 static_cast<unsigned long long>(c->nonfinite_steps);
    __hip_atomic_store(c->host_mirror + kMirrorStep, w, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// device -> host progress of a solve (one 64-bit store at system scope)
__device__ __forceinline__ void publish_progress(Ctl *c) {
  if (c->host_mirror) {
    __hip_atomic_store(c->host_mirror + 1,
                       static_cast<unsigned long long>(__double_as_longlong(c->rel)),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const unsigned long long w = (static_cast<unsigned long long>(c->seq) << 32) |
                                 (static_cast<unsigned long long>(c->done ? 1u : 0u) << 31) |
                                 static_cast<unsigned long long>(static_cast<unsigned>(c->iters));
    __hip_atomic_store(c->host_mirror, w, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// A new solve's control state. iters0 = 1: iteration 1 is taken (by the
// setup pass) and, with min_it >= 2 and max_it >= 2, its convergence step
// could not stop the solve, so the count starts at 1 without that step (its
// aprde is never read: the solve reports the last iteration's)
__device__ __forceinline__ void begin_solve(Ctl *c, double tol, int min_it, int max_it,
                                            unsigned seq, int iters0 = 0) {
  c->aprde = 0.0;
  c->rel = 0.0;
  c->tol = tol;
  c->iters = iters0;
  c->done = (c->nprde >= 1) ? 0 : 1;
  c->converged = 0;
  c->min_it = min_it;
  c->max_it = max_it;
  c->ticket = 0u;
  c->seq = seq;
  publish_progress(c);
}

// the begin of a solve carried by a launch that runs before the solve's
// first iteration (seq 0: none); nprde >= 0 is set first
struct SolveBegin {
  double tol = 0.0;
  int min_it = 0, max_it = 0, nprde = -1, iters0 = 0;
  unsigned seq = 0u;
};

__device__ __forceinline__ void begin_solve(Ctl *c, const SolveBegin &B) {
  if (B.nprde >= 0)
    c->nprde = B.nprde;
  begin_solve(c, B.tol, B.min_it, B.max_it, B.seq, B.iters0);
}

struct GridDims {
  int nx, ny, nz;     // cells per axis (dense box)
  int ox, oy, oz;     // integer cell index of the box origin
  uint32_t ncells;
  double radius;      // cell edge = neighbourhood radius
};

template <class T>
struct SceneView {
  using V4 = vec4_t<T>;
  int nf, nb;
  CubicSpline<T> K;
  T thr2;  // neighbour iff squared distance < thr2 (exactly the reference's test)
  T dt, rho0, omega, mu_f, mu_b, fade, t, hh001;  // hh001 = (0.01*h)*h
  T g0, g1, g2;
  T gc0, gc1, gc2;  // gravity_center
  int axial;        // the step's gravity is initialize_axial_acceleration
  // the fused setup (or source pass) also takes IISPH iteration 1, which
  // needs no neighbours when every pressure starts at 0
  int first_iter;
  GridDims grid;
  const uint32_t *fstart;
  const uint32_t *bstart;
  V4 *xm, *v, *a, *c, *G, *vd;
  V4 *xq;  // (x, y, z, m p / rho^2): the one gather of the pressure-acceleration pass
  // (x, y, z, m) and (a, 0) records of the pressure pass's gathers, in
  // groups of kXaGroup particles (their x-records, then their a-records), so
  // both records of a particle share one cache line: xa_x() / xa_a() give
  // the index of a record. One particle per group (x and a side by side,
  // 32 bytes for FP32) measured faster than 4 per group (128-byte groups,
  // whole 64-byte segments per pass): the pressure pass gains more from the
  // adjacent pair than the acceleration pass loses on half-segment writes.
  V4 *xa;
  static constexpr size_t kXaGroup = 1;
  __host__ __device__ static size_t xa_x(size_t j) {
    return (j / kXaGroup) * (2 * kXaGroup) + j % kXaGroup;
  }
  __host__ __device__ static size_t xa_a(size_t j) { return xa_x(j) + kXaGroup; }
  T *rho, *p, *pr, *s, *AA, *Ap, *sp;
  T *gc;  // gradient correction L (3x3 column-major per fluid particle)
  // pt16 (scenes with implicit viscosity 3; empty otherwise): vorticity, the
  // right-hand sides of its two systems, the target velocity gradient (3x3
  // column-major), pt16_density, pt16_same_density and the scalar PCG vectors
  V4 *om, *orhs, *vrhs;
  T *tvg, *prho, *psrho;
  T *scx, *scr, *scp, *scq, *scy;
  // list form of the pt16 products (fused steps): m_j W_ij of list entry e of
  // particle i at pcoef[((e / 4) * nf + i) * 4 + e % 4], written once per step
  T *pcoef;
  int32_t *id;
  uint32_t *fcell;
  V4 *bxV;
  int32_t *bid;
  uint32_t *bcell;
  Ctl *ctl;
  double *partials;  // one per block of the largest fluid launch
  // PCG vectors (3-vectors in real4); the direction p is interleaved with
  // the (x, m / rho) records the list product gathers beside it:
  // cxp[2 j] = (x_j, m_j / rho_j), cxp[2 j + 1] = p_j, so both halves of a
  // neighbour come from one cache line
  V4 *cx, *cr, *cxp, *cq, *cy, *cb;
};

// gravity.initialize_axial_acceleration (gravity.hpp:189-215,
// gravity.prtcl:26-33) at position x, in the reference's operation order
// without contraction: xc = x - c, g_n = g rcp0(|g|), xc_o = xc - (xc.g_n) g_n,
// a = ((-|g|) xc_o) rcp0(|xc_o|), rcp0(v) = 1/v if |v| > 1e-9 else 0
template <class T>
__device__ __forceinline__ vec4_t<T> axial_acceleration(const SceneView<T> &S, T x, T y, T z) {
#pragma clang fp contract(off)
  const T eps = static_cast<T>(1e-9);
  const T xc0 = x - S.gc0, xc1 = y - S.gc1, xc2 = z - S.gc2;
  const T ng = sqrt(sqdist_exact(S.g0, S.g1, S.g2));
  const T rg = fabs(ng) > eps ? T(1) / ng : T(0);
  const T n0 = S.g0 * rg, n1 = S.g1 * rg, n2 = S.g2 * rg;
  const T d = dot3_exact(xc0, xc1, xc2, n0, n1, n2);
  const T o0 = xc0 - d * n0, o1 = xc1 - d * n1, o2 = xc2 - d * n2;
  const T no = sqrt(sqdist_exact(o0, o1, o2));
  const T ro = fabs(no) > eps ? T(1) / no : T(0);
  return make4<vec4_t<T>>(((-ng) * o0) * ro, ((-ng) * o1) * ro, ((-ng) * o2) * ro, T(0));
}

// the step's gravity at position x: g, or the axial acceleration
template <class T>
__device__ __forceinline__ void step_gravity(const SceneView<T> &S, T x, T y, T z, T &a0, T &a1,
                                             T &a2) {
  if (S.axial) {
    const vec4_t<T> a = axial_acceleration(S, x, y, z);
    a0 = a.x;
    a1 = a.y;
    a2 = a.z;
  } else {
    a0 = S.g0;
    a1 = S.g1;
    a2 = S.g2;
  }
}

struct BlockRanges;

// One scene of a batch (prtcl_hip_step over several scenes): its view and
// the buffers the batched sort and setup use. A batched kernel maps its block
// to (scene, local block) through a block map of words (scene << 20 | block).
template <class T>
struct BatchEntry {
  SceneView<T> S;
  uint16_t *nidx;             // neighbour list of the scene (NbrList)
  uint32_t *ninfo;
  BlockRanges *nblk;
  vec4_t<T> *vtmp;            // integrated velocity written by the setup pass
  uint32_t *count, *start;    // fluid grid counts / starts
  uint32_t *rank, *perm;      // sort buffers (result in rank)
  uint32_t *scan_tmp;         // tile sums
  vec4_t<T> *x_dst, *v_dst;   // gather destinations
  uint32_t *id_dst, *cell_dst;
  unsigned nblocks;           // fluid blocks of kBlock particles
  unsigned ntiles;            // scan tiles of the grid
  // parameters of the scene's IISPH solve in this step
  double tol;
  int min_it, max_it, nprde;
  unsigned seq;
  // iteration 1 is taken by the setup and cannot end the solve: the solve
  // starts at 1 iteration without its convergence step
  int first_preset;
  // the step ends with the WKBB18 path (implicit viscosity), not with the
  // fused integration
  int viscous;
  // the setup takes three launches (explicit viscosity or axial gravity);
  // explicit_visc: v* from the explicit viscosity pass (else axial gravity)
  int split_setup;
  int explicit_visc;
  // the scene's WKBB18 solve in this step: FP64 CG partials, tolerance, cap
  // and sequence number; cg_fold: alpha and beta derived in the vector
  // kernels (at most kFusedCgBlocks blocks)
  double *cg_part;
  double cg_tol;
  int cg_max_k;
  unsigned cg_seq;
  int cg_fold;
  // pt16: strain-rate viscosity and the tolerance and cap of the vorticity
  // diffusion (0) and velocity reconstruction (1) systems
  double pt16_xi;
  double pt16_tol[2];
  int pt16_max_k[2];
};
// the largest CG launch whose vector kernels derive alpha and beta
// themselves: every block sums all nb partials of the kernel before it
constexpr unsigned kFusedCgBlocks = 1024;
constexpr uint32_t kBatchLocalBits = 20;
constexpr uint32_t kBatchLocalMask = (1u << kBatchLocalBits) - 1u;

__host__ __device__ inline unsigned grid_blocks(long n, int bs = kBlock) {
  return static_cast<unsigned>((n + bs - 1) / bs);
}

// Blocks are dispatched round-robin over the 8 XCDs; a kernel whose blocks
// gather from cell-ordered neighbours remaps its block index so that each
// XCD works on one contiguous range of particles and its L2 holds that
// range's neighbourhood (a bijection of [0, nblocks)).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblocks) {
  constexpr unsigned kXcd = 8;
  const unsigned per = nblocks / kXcd, rem = nblocks % kXcd;
  const unsigned x = b % kXcd, k = b / kXcd;
  // XCD x owns logical blocks [x*per + min(x, rem), ...) of length per + (x < rem)
  return x * per + (x < rem ? x : rem) + k;
}

// host-side error helper
struct HipError {
  hipError_t err;
  std::string where;
};
#define PRTCL_HIP_CHECK(expr)                                        \
  do {                                                               \
    hipError_t e_ = (expr);                                          \
    if (e_ != hipSuccess)                                            \
      throw ::prtcl_amd::HipError{e_, std::string(#expr)};           \
  } while (0)

// ----- grid / neighbour search (nsearch.hip) -----
template <class T>
struct GridOps {
  // Cell keys, counts, scan and a stable counting sort of one group. Returns
  // the permutation in slot order (new slot s holds old slot perm[s]); it is
  // written to `rank` (perm is scratch).
  static const uint32_t *sort_group(hipStream_t st, const GridDims &g, const vec4_t<T> *x,
                                    int n, uint32_t *cell, uint32_t *count, uint32_t *start,
                                    uint32_t *rank, uint32_t *perm, uint32_t *scan_tmp);
  // sort_group and the reorder of the carried state (x -> xd, v -> vd, id ->
  // id_d, cell -> rank) in one fewer pass; no permutation array results
  static void sort_state(hipStream_t st, const GridDims &g, const vec4_t<T> *x, int n,
                         uint32_t *cell, uint32_t *count, uint32_t *start, uint32_t *rank,
                         uint32_t *perm, uint32_t *scan_tmp, vec4_t<T> *xd, vec4_t<T> *vd,
                         const vec4_t<T> *vs, uint32_t *id_d, const uint32_t *id_s);
  static void gather4(hipStream_t st, void *dst, const void *src, const uint32_t *perm, int n);
  static void gather_elems(hipStream_t st, void *dst, const void *src,
                           const uint32_t *perm, int n, int elem_bytes);
  // the whole fluid sort (keys .. reorder of the carried state into the
  // entries' destinations) of every scene of a batch; cell_dst must not be
  // the perm buffer
  static void sort_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                         unsigned npblocks, const uint32_t *tmap, unsigned ntiles,
                         unsigned nscenes);
  // x, v (real4), id and cell of the carried state, reordered in one pass
  static void gather_state(hipStream_t st, vec4_t<T> *xd, const vec4_t<T> *xs, vec4_t<T> *vd,
                           const vec4_t<T> *vs, uint32_t *id_d, const uint32_t *id_s,
                           uint32_t *c_d, const uint32_t *c_s, const uint32_t *perm, int n);
};
// exclusive scan of n uint32 counts into start[0..n] (start[n] = total); the
// counts are cleared to zero for the next build
void scan_counts(hipStream_t st, uint32_t *count, uint32_t *start, uint32_t n,
                 uint32_t *tmp);
size_t scan_tmp_words(uint32_t n);
unsigned scan_tiles(uint32_t ncells);

}  // namespace prtcl_amd
