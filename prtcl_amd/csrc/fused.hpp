// Launchers of the fused step kernels (fused.hip).
#pragma once

#include "scene.hpp"

namespace prtcl_amd {

// per-step fluid neighbour list, slot-interleaved: entry k of particle i at
// idx[k * n + i]; count[i] may exceed cap (the particle is then walked over
// the grid)
struct NbrList {
  uint32_t *idx;
  uint32_t *count;
  uint32_t cap;
};

template <class T>
struct FusedOps {
  // density, neighbour list, IISPH c / AA / s, integrated velocity into vtmp
  static void setup(hipStream_t st, const SceneView<T> &S, const NbrList &L,
                    bool explicit_viscosity, vec4_t<T> *vtmp);
  static void accel(hipStream_t st, const SceneView<T> &S, const NbrList &L);
  static void pressure(hipStream_t st, const SceneView<T> &S, const NbrList &L);
  static void integrate(hipStream_t st, const SceneView<T> &S);
  // the same passes over every scene of a batch (pmap: block map)
  static void setup_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                          unsigned npblocks);
  static void accel_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                          unsigned npblocks);
  static void pressure_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                             unsigned npblocks, unsigned nscenes);
  static void integrate_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                              unsigned npblocks, unsigned nscenes);
};

}  // namespace prtcl_amd
