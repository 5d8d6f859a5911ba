// Host-side owner of one device-resident scene: the HBM arena, the uniform
// grid, the procedure dispatch behind prtcl_hip_run() and the device-driven
// IISPH loop behind prtcl_hip_iisph_solve() / prtcl_hip_step().
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/prtcl_hip.h"
#include "scene.hpp"

namespace prtcl_amd {

struct HostPinned;

// Shared per-context resources a scene borrows.
struct CtxResources {
  hipStream_t stream = nullptr;
  // device staging for host transfers, grown on demand
  void *staging = nullptr;
  size_t staging_bytes = 0;
  // pinned host words for asynchronous flag read-back
  int32_t *pinned = nullptr;  // 64 words
  hipEvent_t ev[2] = {nullptr, nullptr};
  void *ensure_staging(size_t bytes);
};

class SceneBase {
 public:
  virtual ~SceneBase() = default;
  prtcl_hip_scene_desc desc{};
  uint64_t bytes = 0;  // device bytes held
  std::string err;
  // admission hook: returns false when `extra` more bytes exceed the budget
  std::function<bool(uint64_t)> budget_check;

  virtual void upload(int group, int field, const void *host, size_t count) = 0;
  virtual void download(int group, int field, void *host, size_t count) = 0;
  virtual void run(int proc) = 0;
  virtual double get_global(int g) = 0;
  virtual void set_global(int g, double v) = 0;
  virtual int iisph_solve(double tol, int min_it, int max_it, int *iters, double *rel,
                          bool wait) = 0;
  virtual int step(int k) = 0;
  virtual void copy_neighbors(int group, int ngroup, uint32_t *offsets, uint32_t *indices,
                              uint64_t *total) = 0;
  virtual void last_permutation(int group, uint32_t *perm) = 0;
  virtual void step_info(int *iisph_iters, double *rel, int *wkbb18_iters) = 0;
  virtual uint64_t grid_bytes_for_domain() const = 0;
};

SceneBase *make_scene(const prtcl_hip_scene_desc &desc, CtxResources *res);
uint64_t scene_particle_bytes(const prtcl_hip_scene_desc &desc);

}  // namespace prtcl_amd
