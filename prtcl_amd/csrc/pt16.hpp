// Launchers of the pt16 vorticity-preserving viscosity kernels (pt16.hip,
// Peer and Teschner 2016 as pt16.hpp / pt16.prtcl of the reference): the
// set-up (densities, velocity gradient, vorticity, target velocity
// gradient), the right-hand sides of the vorticity diffusion and velocity
// reconstruction systems, the spin term, and the scalar PCG of
// cg_openmp.hpp:32-215 over one vector component at a time.
#pragma once

#include "nbrlist.hpp"
#include "scene.hpp"

namespace prtcl_amd {

template <class T>
struct Pt16Ops {
  // With a neighbour list L (fused steps) the fluid sums walk the step's
  // list, the set-up writes the pair coefficients m_j W_ij for the step, and
  // the vorticity right-hand side and the CG products stream them; without,
  // every sum walks the grid in the reference's order.
  // pt16.setup (pt16.hpp:213-342); xi = strain_rate_viscosity
  static void setup(hipStream_t st, const SceneView<T> &S, const NbrList *L, T xi);
  // vorticity diffusion rhs: xi sum_f (m_j (w_i - w_j)) W_ij (pt16.hpp:378-418)
  static void omega_rhs(hipStream_t st, const SceneView<T> &S, const NbrList *L, T xi);
  // velocity reconstruction rhs: sum_f ((m_j 0.5) ((G_i + G_j) x_ij)) W_ij
  // (pt16.hpp:588-606)
  static void vrec_rhs(hipStream_t st, const SceneView<T> &S, const NbrList *L);
  // target velocity gradient += cross_product_matrix(0.5 w) (pt16.hpp:521-535)
  static void spin(hipStream_t st, const SceneView<T> &S);
  // which = 0: vorticity diffusion (system diagonal pt16_same_density, guess
  // and result in vorticity[dim]); which = 1: velocity reconstruction
  // (diagonal pt16_density, guess and result in velocity[dim])
  static void cg_setup(hipStream_t st, const SceneView<T> &S, const NbrList *L, int which,
                       int dim, double *part, double tol, int max_k, unsigned seq);
  // split: the alpha and beta reductions in one-block kernels of their own
  // (always so above kFusedCgBlocks blocks); the results do not depend on it
  static void cg_iteration(hipStream_t st, const SceneView<T> &S, const NbrList *L, int which,
                           double *part, bool split);
  static void cg_apply(hipStream_t st, const SceneView<T> &S, int which, int dim);
  // the same over every pt16 scene of a batch (list forms; map: block map of
  // the scenes' blocks); each table entry carries the scene's xi, CG
  // partials, the tolerance and cap of both systems, the sequence number of
  // the step's first solve (solve s = 0..5 runs under cg_seq + s) and the
  // fold choice. any_split: some scene takes alpha and beta from the
  // one-block kernels
  static void setup_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                          unsigned nblocks);
  static void omega_rhs_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                              unsigned nblocks);
  static void spin_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                         unsigned nblocks);
  static void vrec_rhs_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                             unsigned nblocks);
  static void cg_setup_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                             unsigned nblocks, unsigned nscenes, int which, int dim, int solve);
  static void cg_iteration_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                 unsigned nblocks, unsigned nscenes, int which, bool any_split);
  static void cg_apply_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                             unsigned nblocks, int which, int dim);
};

}  // namespace prtcl_amd
