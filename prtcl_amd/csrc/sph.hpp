// Host-side launchers of the SPH procedure kernels (sph.hip), the fused step
// kernels (fused.hip) and the WKBB18 kernels (wkbb18.hip).
#pragma once

#include "scene.hpp"

namespace prtcl_amd {

template <class T>
struct SphOps {
  static void volume(hipStream_t st, const SceneView<T> &S);
  static void density(hipStream_t st, const SceneView<T> &S);
  static void gravity(hipStream_t st, const SceneView<T> &S);
  static void gravity_axial(hipStream_t st, const SceneView<T> &S);
  static void gravity_accumulate(hipStream_t st, const SceneView<T> &S);
  static void viscosity(hipStream_t st, const SceneView<T> &S);
  static void integrate_v(hipStream_t st, const SceneView<T> &S);
  static void integrate_x(hipStream_t st, const SceneView<T> &S);
  // the two integrations over every scene of a batch (map: block map of the
  // scenes' blocks); from_xa: a = the last pressure acceleration first
  static void integrate_v_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                unsigned nblocks, bool from_xa);
  static void integrate_x_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                unsigned nblocks, unsigned nscenes);
  static void iisph_setup(hipStream_t st, const SceneView<T> &S);
  static void iisph_accel(hipStream_t st, const SceneView<T> &S, bool solve);
  static void iisph_pressure(hipStream_t st, const SceneView<T> &S, bool solve);
  static void count_neighbors(hipStream_t st, const SceneView<T> &S, int qgroup, int ngroup,
                              uint32_t *cnt);
  static void write_neighbors(hipStream_t st, const SceneView<T> &S, int qgroup, int ngroup,
                              const uint32_t *offsets, uint32_t *out);
};

}  // namespace prtcl_amd
