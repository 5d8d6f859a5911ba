// Launchers of the WKBB18 implicit-viscosity kernels (wkbb18.hip). A null
// list selects the exact (reference operation order, grid walk) product.
#pragma once

#include "fused.hpp"

namespace prtcl_amd {

// block maps of the WKBB18 scenes of a batch: all of them, those without
// and those with the gradient correction (wkbb18 / wkbb18_gc)
struct BatchMaps {
  const uint32_t *all, *plain, *gc;
  unsigned n_all, n_plain, n_gc;
};

template <class T>
struct Wkbb18Ops {
  // sp_i = (I - dt D_i)^-1 (and, with a list, the boundary matrix B_i and the
  // (x, m / rho) gather array of the list product); gc: wkbb18_gc, every
  // gradient of particle i's terms multiplied by its gradient correction
  static void diagonal(hipStream_t st, const SceneView<T> &S, const NbrList *L, bool gc);
  // x = v, r = A x - b, y = P r, p = -y and the FP64 set-up scalars
  static void setup(hipStream_t st, const SceneView<T> &S, const NbrList *L, bool gc,
                    double *part, double tol, int max_k, unsigned seq);
  // diagonal() and setup() over the list L in one walk (the same results)
  static void diagonal_setup(hipStream_t st, const SceneView<T> &S, const NbrList &L, bool gc,
                             double *part, double tol, int max_k, unsigned seq);
  // one CG iteration in two parts, the product q = A p and the vector work
  // (alpha, x, r, y, beta, p); every kernel exits at once after convergence
  static void iteration_product(hipStream_t st, const SceneView<T> &S, const NbrList *L,
                                bool gc, double *part);
  // split: the alpha and beta reductions in one-block kernels of their own
  // (always so above kFusedCgBlocks blocks); the results do not depend on it
  static void iteration_vectors(hipStream_t st, const SceneView<T> &S, double *part,
                                bool split);
  // vd = (x - v) H, a = vd / dt
  static void apply(hipStream_t st, const SceneView<T> &S);
  // the same over every scene of a batch (fused steps, the list forms):
  // each table entry carries the scene's CG partials, tolerance, cap,
  // sequence number and fold choice; the wkbb18_gc scenes' blocks take the
  // gradient-corrected kernels. any_split: some scene takes its alpha and
  // beta from the one-block kernels
  static void correction_batch(hipStream_t st, const BatchEntry<T> *tab, const BatchMaps &m);
  static void diagonal_setup_batch(hipStream_t st, const BatchEntry<T> *tab, const BatchMaps &m,
                                   unsigned nscenes);
  static void iteration_batch(hipStream_t st, const BatchEntry<T> *tab, const BatchMaps &m,
                              unsigned nscenes, bool any_split);
  static void apply_batch(hipStream_t st, const BatchEntry<T> *tab, const BatchMaps &m);
  // correction.compute_gradient_correction: L_i = pinv(sum_f (m_j / rho_i)
  // gradW_ij (x_j - x_i)^T), over the grid (L null) or the step's list
  static void correction(hipStream_t st, const SceneView<T> &S, const NbrList *L);
  // correction.identity_gradient_correction
  static void identity_correction(hipStream_t st, const SceneView<T> &S);
};

}  // namespace prtcl_amd
