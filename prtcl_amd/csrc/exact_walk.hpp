// Exact neighbour visit of the per-procedure kernels: the 27 cells of the
// dense cell box around a particle's cell, x-runs of three cells, and the
// reference predicate (FP32 squared distance < thr2; the particle itself
// included, as CopyNeighbors, grouped_uniform_grid.hpp:407-422).
#pragma once

#include "device_common.hpp"
#include "scene.hpp"

namespace prtcl_amd {
namespace {

template <class Fn>
__device__ __forceinline__ void runs3(const GridDims &g, uint32_t cellc,
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

// exact neighbour visit (reference predicate), fn(j, d0, d1, d2, r2, pj)
template <class T, class Fn>
__device__ __forceinline__ void nbrs_exact(const SceneView<T> &S, const vec4_t<T> &xi,
                                           uint32_t cellc, const uint32_t *start,
                                           const vec4_t<T> *__restrict__ pts, Fn &&fn) {
  runs3(S.grid, cellc, start, [&](uint32_t b, uint32_t e) {
    for (uint32_t j = b; j < e; ++j) {
      const vec4_t<T> pj = pts[j];
      const T d0 = xi.x - pj.x, d1 = xi.y - pj.y, d2 = xi.z - pj.z;
      const T r2 = sqdist_exact(d0, d1, d2);
      if (r2 < S.thr2)
        fn(j, d0, d1, d2, r2, pj);
    }
  });
}

}  // namespace
}  // namespace prtcl_amd
