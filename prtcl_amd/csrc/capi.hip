// The extern "C" boundary (include/prtcl_hip.h): contexts, scenes, field
// transfer, procedure dispatch, fused IISPH solve / step and neighbour
// export. Every entry point catches all exceptions and returns a status code.
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

#include "checkpoint.hpp"
#include "scene_impl.hpp"
#include "serve.hpp"
#include "trace.hpp"
#include "vtk.hpp"

using prtcl_amd::CtxResources;
using prtcl_amd::HipError;
using prtcl_amd::SceneBase;

// Home of the batched tables of prtcl_hip_step_noalloc: a fixed array in
// device memory (one copy per GPU, held until the process ends), so the entry
// point allocates nothing. Every call on one GPU shares it and leaves the last
// call's tables in it; each call writes all of the tables it reads, so nothing
// in it must start at zero.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
// Per GPU: the device address of `scratch`, looked up once with the GPU
// current (the array's name in host code is only a stand-in for the symbol),
// and the event after the last kernel that read it, so that a call on
// another stream (another context) copies its tables in only after that.
struct ScratchHome {
  void *dev = nullptr;
  hipEvent_t released = nullptr;
};
std::mutex g_scratch_mu;
std::map<int, ScratchHome> g_scratch;

ScratchHome scratch_home(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(device);
  if (it != g_scratch.end())
    return it->second;
  ScratchHome h;
  PRTCL_HIP_CHECK(hipGetSymbolAddress(&h.dev, HIP_SYMBOL(scratch)));
  PRTCL_HIP_CHECK(hipEventCreateWithFlags(&h.released, hipEventDisableTiming));
  g_scratch[device] = h;
  return h;
}
}  // namespace

struct prtcl_hip_ctx {
  int device = 0;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  CtxResources res;
  // prtcl_hip_step_noalloc: this GPU's `scratch` (and its release event),
  // the page-locked host copy of the tables (set up with the context) and the
  // end of its last copy
  ScratchHome scratch;
  void *noalloc_host = nullptr;
  hipEvent_t noalloc_ev = nullptr;
  std::map<int32_t, std::unique_ptr<SceneBase>> scenes;
  int32_t next_id = 0;
  std::string last_error;
  prtcl_amd::Server server;
};

namespace {

int fail(prtcl_hip_ctx *ctx, int code, const std::string &msg) {
  if (ctx)
    ctx->last_error = msg;
  return code;
}

template <class Fn>
int guarded(prtcl_hip_ctx *ctx, Fn &&fn) {
  if (!ctx)
    return PRTCL_HIP_INVALID_ARGUMENT;
  try {
    if (hipSetDevice(ctx->device) != hipSuccess)
      return fail(ctx, PRTCL_HIP_ERROR, "hipSetDevice failed");
    return fn();
  } catch (const HipError &e) {
    return fail(ctx, PRTCL_HIP_ERROR,
                std::string("HIP error ") + hipGetErrorString(e.err) + " in " + e.where);
  } catch (const std::length_error &e) {
    return fail(ctx, PRTCL_HIP_ADMISSION_DENIED, e.what());
  } catch (const std::out_of_range &e) {
    return fail(ctx, PRTCL_HIP_NOT_FOUND, e.what());
  } catch (const std::invalid_argument &e) {
    return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, e.what());
  } catch (const std::logic_error &e) {
    return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, e.what());
  } catch (const std::bad_alloc &) {
    return fail(ctx, PRTCL_HIP_ADMISSION_DENIED, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(ctx, PRTCL_HIP_ERROR, e.what());
  } catch (...) {
    return fail(ctx, PRTCL_HIP_ERROR, "unknown error");
  }
}

SceneBase *find(prtcl_hip_ctx *ctx, int32_t id) {
  auto it = ctx->scenes.find(id);
  if (it == ctx->scenes.end())
    throw std::out_of_range("scene does not exist");
  return it->second.get();
}

struct Named {
  const char *name;
  int id;
};

const Named kFields[] = {
    {"position", PRTCL_HIP_F_POSITION},
    {"velocity", PRTCL_HIP_F_VELOCITY},
    {"acceleration", PRTCL_HIP_F_ACCELERATION},
    {"density", PRTCL_HIP_F_DENSITY},
    {"pressure", PRTCL_HIP_F_PRESSURE},
    {"mass", PRTCL_HIP_F_MASS},
    {"iisph_helper_c", PRTCL_HIP_F_IISPH_HELPER_C},
    {"iisph_source_term", PRTCL_HIP_F_IISPH_SOURCE_TERM},
    {"iisph_diagonal_element", PRTCL_HIP_F_IISPH_DIAGONAL_ELEMENT},
    {"iisph_right_hand_side", PRTCL_HIP_F_IISPH_RIGHT_HAND_SIDE},
    {"time_of_birth", PRTCL_HIP_F_TIME_OF_BIRTH},
    {"volume", PRTCL_HIP_F_VOLUME},
    {"wkbb18_velocity_delta", PRTCL_HIP_F_WKBB18_VELOCITY_DELTA},
    {"wkbb18_system_diagonal", PRTCL_HIP_F_WKBB18_SYSTEM_DIAGONAL},
    {"gradient_correction", PRTCL_HIP_F_GRADIENT_CORRECTION},
    {"pt16_normalizer", PRTCL_HIP_F_PT16_NORMALIZER},
    {"pt16_density", PRTCL_HIP_F_PT16_DENSITY},
    {"pt16_same_density", PRTCL_HIP_F_PT16_SAME_DENSITY},
    {"vorticity", PRTCL_HIP_F_VORTICITY},
    {"target_velocity_gradient", PRTCL_HIP_F_TARGET_VELOCITY_GRADIENT},
    {"pt16_vorticity_diffusion_diagonal", PRTCL_HIP_F_PT16_VORTICITY_DIFFUSION_DIAGONAL},
    {"pt16_vorticity_diffusion_rhs", PRTCL_HIP_F_PT16_VORTICITY_DIFFUSION_RHS},
    {"pt16_velocity_reconstruction_diagonal",
     PRTCL_HIP_F_PT16_VELOCITY_RECONSTRUCTION_DIAGONAL},
    {"pt16_velocity_reconstruction_rhs", PRTCL_HIP_F_PT16_VELOCITY_RECONSTRUCTION_RHS},
};

const Named kProcs[] = {
    {"neighborhood.update", PRTCL_HIP_P_NSEARCH_UPDATE},
    {"neighborhood.permute", PRTCL_HIP_P_NSEARCH_PERMUTE},
    {"aiast12.compute_volume", PRTCL_HIP_P_AIAST12_COMPUTE_VOLUME},
    {"density.compute_density", PRTCL_HIP_P_DENSITY_COMPUTE_DENSITY},
    {"gravity.initialize_acceleration", PRTCL_HIP_P_GRAVITY_INITIALIZE_ACCELERATION},
    {"viscosity.accumulate_acceleration", PRTCL_HIP_P_VISCOSITY_ACCUMULATE_ACCELERATION},
    {"symplectic_euler.integrate_velocity_with_hard_fade",
     PRTCL_HIP_P_SE_INTEGRATE_VELOCITY_WITH_HARD_FADE},
    {"symplectic_euler.integrate_position", PRTCL_HIP_P_SE_INTEGRATE_POSITION},
    {"iisph.setup", PRTCL_HIP_P_IISPH_SETUP},
    {"iisph.iteration_pressure_acceleration",
     PRTCL_HIP_P_IISPH_ITERATION_PRESSURE_ACCELERATION},
    {"iisph.iteration_pressure", PRTCL_HIP_P_IISPH_ITERATION_PRESSURE},
    {"wkbb18.compute_diagonal", PRTCL_HIP_P_WKBB18_COMPUTE_DIAGONAL},
    {"wkbb18.accumulate_acceleration", PRTCL_HIP_P_WKBB18_ACCUMULATE_ACCELERATION},
    {"correction.identity_gradient_correction",
     PRTCL_HIP_P_CORRECTION_IDENTITY_GRADIENT_CORRECTION},
    {"correction.compute_gradient_correction", PRTCL_HIP_P_CORRECTION_COMPUTE_GRADIENT_CORRECTION},
    {"wkbb18_gc.compute_diagonal", PRTCL_HIP_P_WKBB18_GC_COMPUTE_DIAGONAL},
    {"wkbb18_gc.accumulate_acceleration", PRTCL_HIP_P_WKBB18_GC_ACCUMULATE_ACCELERATION},
    {"pt16.setup", PRTCL_HIP_P_PT16_SETUP},
    {"pt16.solve_vorticity_diffusion", PRTCL_HIP_P_PT16_SOLVE_VORTICITY_DIFFUSION},
    {"pt16.vorticity_preservation", PRTCL_HIP_P_PT16_VORTICITY_PRESERVATION},
    {"pt16.solve_velocity_reconstruction", PRTCL_HIP_P_PT16_SOLVE_VELOCITY_RECONSTRUCTION},
    {"gravity.initialize_axial_acceleration", PRTCL_HIP_P_GRAVITY_INITIALIZE_AXIAL_ACCELERATION},
    {"gravity.accumulate_acceleration", PRTCL_HIP_P_GRAVITY_ACCUMULATE_ACCELERATION},
};

}  // namespace

const char *prtcl_amd::field_name(int id) {
  for (const auto &f : kFields)
    if (f.id == id)
      return f.name;
  throw std::out_of_range("field does not exist");
}

namespace {

bool valid_desc(const prtcl_hip_scene_desc *d) {
  return d && d->struct_size == sizeof(prtcl_hip_scene_desc) && d->n_fluid >= 0 &&
         d->n_boundary >= 0 && d->n_fluid < (int64_t(1) << 30) &&
         d->n_boundary < (int64_t(1) << 30) && (d->real_type == 0 || d->real_type == 1) &&
         d->smoothing_scale > 0 && d->time_step > 0 && d->iisph_max_iterations >= 0 &&
         d->implicit_viscosity >= 0 && d->implicit_viscosity <= 3;
}

}  // namespace

extern "C" {

void prtcl_hip_scene_desc_init(prtcl_hip_scene_desc *d) {
  if (!d)
    return;
  std::memset(d, 0, sizeof(*d));
  d->struct_size = sizeof(prtcl_hip_scene_desc);
  d->real_type = 0;
  d->smoothing_scale = 0.025;
  d->time_step = 0.001;
  d->fade_duration = 0.0;
  d->iisph_relaxation = 0.5;
  d->gravity[2] = -9.81;
  d->rest_density = 1000.0;
  d->iisph_tolerance = 1e-3;
  d->iisph_min_iterations = 2;
  d->iisph_max_iterations = 50;
  d->wkbb18_maximum_error = 0.05;
  d->wkbb18_maximum_iterations = 100;
  d->permute_every = 8;
  d->pt16_strain_rate_viscosity = 0.0;
  d->pt16_vorticity_maximum_error = 1e-3;
  d->pt16_velocity_maximum_error = 1e-3;
  d->pt16_vorticity_maximum_iterations = 500;
  d->pt16_velocity_maximum_iterations = 500;
}

int prtcl_hip_ctx_create(int device, uint64_t budget, prtcl_hip_ctx **out) {
  if (!out)
    return PRTCL_HIP_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return PRTCL_HIP_ERROR;
  auto *ctx = new (std::nothrow) prtcl_hip_ctx;
  if (!ctx)
    return PRTCL_HIP_ERROR;
  ctx->device = device;
  ctx->budget = budget;
  int rc = guarded(ctx, [&] {
    PRTCL_HIP_CHECK(hipStreamCreateWithFlags(&ctx->res.stream, hipStreamNonBlocking));
    PRTCL_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&ctx->res.pinned), 64 * sizeof(int32_t),
                                  hipHostMallocDefault));
    std::memset(ctx->res.pinned, 0, 64 * sizeof(int32_t));
    for (auto &e : ctx->res.ev)
      PRTCL_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->scratch = scratch_home(device);
    PRTCL_HIP_CHECK(hipHostMalloc(&ctx->noalloc_host, kNoAllocScratchBytes, hipHostMallocDefault));
    PRTCL_HIP_CHECK(hipEventCreateWithFlags(&ctx->noalloc_ev, hipEventDisableTiming));
    return int(PRTCL_HIP_OK);
  });
  if (rc != PRTCL_HIP_OK) {
    prtcl_hip_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return PRTCL_HIP_OK;
}

int prtcl_hip_ctx_destroy(prtcl_hip_ctx *ctx) {
  if (!ctx)
    return PRTCL_HIP_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  if (ctx->res.stream)
    (void)hipStreamSynchronize(ctx->res.stream);
  ctx->scenes.clear();
  ctx->res.release();
  if (ctx->res.staging)
    (void)hipFree(ctx->res.staging);
  if (ctx->res.pinned)
    (void)hipHostFree(ctx->res.pinned);
  for (auto &e : ctx->res.ev)
    if (e)
      (void)hipEventDestroy(e);
  if (ctx->noalloc_host)
    (void)hipHostFree(ctx->noalloc_host);
  if (ctx->noalloc_ev)
    (void)hipEventDestroy(ctx->noalloc_ev);
  if (ctx->res.stream)
    (void)hipStreamDestroy(ctx->res.stream);
  delete ctx;
  return PRTCL_HIP_OK;
}

const char *prtcl_hip_last_error(const prtcl_hip_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

uint64_t prtcl_hip_ctx_bytes_in_use(const prtcl_hip_ctx *ctx) {
  if (!ctx)
    return 0;
  uint64_t b = 0;
  for (const auto &kv : ctx->scenes)
    b += kv.second->bytes;
  return b;
}

int prtcl_hip_synchronize(prtcl_hip_ctx *ctx) {
  return guarded(ctx, [&] {
    PRTCL_HIP_CHECK(hipStreamSynchronize(ctx->res.stream));
    return int(PRTCL_HIP_OK);
  });
}

uint64_t prtcl_hip_scene_footprint(const prtcl_hip_scene_desc *desc) {
  if (!valid_desc(desc))
    return 0;
  return prtcl_amd::scene_particle_bytes(*desc);
}

int prtcl_hip_scene_create(prtcl_hip_ctx *ctx, const prtcl_hip_scene_desc *desc,
                           int32_t *scene_id) {
  return guarded(ctx, [&]() -> int {
    if (!valid_desc(desc) || !scene_id)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "invalid scene description");
    const uint64_t need = prtcl_amd::scene_particle_bytes(*desc);
    const uint64_t used = prtcl_hip_ctx_bytes_in_use(ctx);
    if (ctx->budget && used + need > ctx->budget)
      return fail(ctx, PRTCL_HIP_ADMISSION_DENIED, "scene exceeds the HBM budget");
    std::unique_ptr<SceneBase> sc(prtcl_amd::make_scene(*desc, &ctx->res));
    sc->budget_check = [ctx](uint64_t extra) {
      return ctx->budget == 0 || prtcl_hip_ctx_bytes_in_use(ctx) + extra <= ctx->budget;
    };
    const int32_t id = ctx->next_id++;
    ctx->scenes.emplace(id, std::move(sc));
    *scene_id = id;
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_scene_destroy(prtcl_hip_ctx *ctx, int32_t scene) {
  return guarded(ctx, [&]() -> int {
    auto it = ctx->scenes.find(scene);
    if (it == ctx->scenes.end())
      return fail(ctx, PRTCL_HIP_NOT_FOUND, "scene does not exist");
    PRTCL_HIP_CHECK(hipStreamSynchronize(ctx->res.stream));
    ctx->scenes.erase(it);
    ctx->server.forget(scene);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_upload_field(prtcl_hip_ctx *ctx, int32_t scene, int32_t group, int32_t field,
                           const void *host, size_t count) {
  return guarded(ctx, [&]() -> int {
    if (!host && count)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "null host pointer");
    find(ctx, scene)->upload(group, field, host, count);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_download_field(prtcl_hip_ctx *ctx, int32_t scene, int32_t group, int32_t field,
                             void *host, size_t count) {
  return guarded(ctx, [&]() -> int {
    if (!host && count)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "null host pointer");
    find(ctx, scene)->download(group, field, host, count);
    return PRTCL_HIP_OK;
  });
}

int32_t prtcl_hip_field_id(const char *name) {
  if (!name)
    return -1;
  for (const auto &f : kFields)
    if (std::strcmp(f.name, name) == 0)
      return f.id;
  return -1;
}

int32_t prtcl_hip_proc_id(const char *name) {
  if (!name)
    return -1;
  for (const auto &p : kProcs)
    if (std::strcmp(p.name, name) == 0)
      return p.id;
  return -1;
}

int prtcl_hip_get_global(prtcl_hip_ctx *ctx, int32_t scene, int32_t global, double *value) {
  return guarded(ctx, [&]() -> int {
    if (!value)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "null value pointer");
    *value = find(ctx, scene)->get_global(global);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_set_global(prtcl_hip_ctx *ctx, int32_t scene, int32_t global, double value) {
  return guarded(ctx, [&]() -> int {
    find(ctx, scene)->set_global(global, value);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_run(prtcl_hip_ctx *ctx, int32_t scene, int32_t proc) {
  return guarded(ctx, [&]() -> int {
    if (proc < 0 || proc >= PRTCL_HIP_P_COUNT)
      return fail(ctx, PRTCL_HIP_NOT_FOUND, "procedure does not exist");
    find(ctx, scene)->run(proc);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_iisph_solve(prtcl_hip_ctx *ctx, int32_t scene, double tol, int32_t min_it,
                          int32_t max_it, int32_t *iters, double *rel_err) {
  return guarded(ctx, [&]() -> int {
    if (min_it < 0 || max_it < 0)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "iteration limits");
    int it = 0;
    double rel = 0;
    int rc = find(ctx, scene)->iisph_solve(tol, min_it, max_it, &it, &rel, true);
    if (iters)
      *iters = it;
    if (rel_err)
      *rel_err = rel;
    return rc;
  });
}

int prtcl_hip_step(prtcl_hip_ctx *ctx, const int32_t *scenes, int32_t n_scenes,
                   int32_t k_steps) {
  return guarded(ctx, [&]() -> int {
    if (!scenes || n_scenes < 0 || k_steps < 0)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "invalid scene list");
    std::vector<SceneBase *> list;
    for (int i = 0; i < n_scenes; ++i)
      list.push_back(find(ctx, scenes[i]));
    return step_scenes(list, k_steps, &ctx->res);
  });
}

int prtcl_hip_step_noalloc(prtcl_hip_ctx *ctx, const int32_t *scenes, int32_t n_scenes,
                           int32_t k_steps) {
  return guarded(ctx, [&]() -> int {
    if (!scenes || n_scenes < 0 || k_steps < 0)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "invalid scene list");
    std::vector<SceneBase *> list;
    for (int i = 0; i < n_scenes; ++i)
      list.push_back(find(ctx, scenes[i]));
    prtcl_amd::BatchTables t;
    t.host = ctx->noalloc_host;
    t.dev = ctx->scratch.dev;
    t.bytes = kNoAllocScratchBytes;
    t.copied = ctx->noalloc_ev;
    t.released = ctx->scratch.released;
    try {
      return prtcl_amd::step_scenes_noalloc(list, k_steps, &ctx->res, t);
    } catch (const prtcl_amd::TablesTooLarge &e) {
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT,
                  "the batch's tables take " + std::to_string(e.bytes) +
                      " bytes; prtcl_hip_step_noalloc holds at most " +
                      std::to_string(kNoAllocScratchBytes));
    }
  });
}

int prtcl_hip_last_step_info(prtcl_hip_ctx *ctx, int32_t scene, int32_t *iisph_iters,
                             double *iisph_rel_err, int32_t *wkbb18_iters) {
  return guarded(ctx, [&]() -> int {
    find(ctx, scene)->step_info(iisph_iters, iisph_rel_err, wkbb18_iters);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_copy_neighbors(prtcl_hip_ctx *ctx, int32_t scene, int32_t group,
                             int32_t neighbor_group, uint32_t *offsets, uint32_t *indices,
                             uint64_t *total) {
  return guarded(ctx, [&]() -> int {
    find(ctx, scene)->copy_neighbors(group, neighbor_group, offsets, indices, total);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_last_permutation(prtcl_hip_ctx *ctx, int32_t scene, int32_t group,
                               uint32_t *perm) {
  return guarded(ctx, [&]() -> int {
    if (!perm)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "null permutation pointer");
    find(ctx, scene)->last_permutation(group, perm);
    return PRTCL_HIP_OK;
  });
}

namespace {
prtcl_amd::Server::Hooks hooks(prtcl_hip_ctx *ctx) {
  prtcl_amd::Server::Hooks h;
  h.find = [ctx](int32_t id) { return find(ctx, id); };
  h.scene_ids = [ctx]() {
    std::vector<int32_t> v;
    for (const auto &kv : ctx->scenes)
      v.push_back(kv.first);
    return v;
  };
  h.bytes_in_use = [ctx]() { return prtcl_hip_ctx_bytes_in_use(ctx); };
  h.budget = ctx->budget;
  return h;
}
}  // namespace

int prtcl_hip_scene_describe(prtcl_hip_ctx *ctx, int32_t scene, prtcl_hip_scene_desc *out) {
  return guarded(ctx, [&]() -> int {
    if (!out)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "out is NULL");
    *out = find(ctx, scene)->desc;
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_scene_save_archive(prtcl_hip_ctx *ctx, int32_t scene, const char *path,
                                 uint64_t *bytes_written) {
  return guarded(ctx, [&]() -> int {
    if (!path)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "path is NULL");
    const uint64_t n = prtcl_amd::save_scene_archive(*find(ctx, scene), ctx->res, path);
    if (bytes_written)
      *bytes_written = n;
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_scene_load_archive(prtcl_hip_ctx *ctx, const char *path,
                                 const prtcl_hip_scene_desc *settings, int32_t *scene_id) {
  return guarded(ctx, [&]() -> int {
    if (!path || !scene_id)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "path or scene_id is NULL");
    prtcl_hip_scene_desc d;
    if (settings) {
      if (settings->struct_size != sizeof(prtcl_hip_scene_desc))
        return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "invalid scene description");
      d = *settings;
    } else {
      prtcl_hip_scene_desc_init(&d);
    }
    auto parsed = prtcl_amd::parse_scene_archive(path, d);
    if (!valid_desc(&d))
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT,
                  "the archive's parameters give an invalid scene description");
    int32_t id = -1;
    const int rc = prtcl_hip_scene_create(ctx, &d, &id);
    if (rc != PRTCL_HIP_OK)
      return rc;
    try {
      prtcl_amd::load_scene_archive(*find(ctx, id), *parsed);
    } catch (...) {
      prtcl_hip_scene_destroy(ctx, id);
      throw;
    }
    *scene_id = id;
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_scene_save_vtk(prtcl_hip_ctx *ctx, int32_t scene, int32_t group, const char *path,
                             uint64_t *bytes_written) {
  return guarded(ctx, [&]() -> int {
    if (!path)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "path is NULL");
    const uint64_t n = prtcl_amd::save_scene_vtk(*find(ctx, scene), ctx->res, group, path);
    if (bytes_written)
      *bytes_written = n;
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_scene_trace(prtcl_hip_ctx *ctx, int32_t scene, int32_t group,
                          const prtcl_hip_camera *camera, double threshold, int32_t max_steps,
                          double *image, double *parameter) {
  return guarded(ctx, [&]() -> int {
    if (!camera || !image)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "camera or image is NULL");
    prtcl_amd::trace_scene(*find(ctx, scene), ctx->res, group, *camera, threshold, max_steps,
                           image, parameter);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_scene_evict(prtcl_hip_ctx *ctx, int32_t scene) {
  return guarded(ctx, [&]() -> int {
    find(ctx, scene)->evict();
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_scene_resident(prtcl_hip_ctx *ctx, int32_t scene, int32_t *resident) {
  return guarded(ctx, [&]() -> int {
    if (!resident)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "resident is NULL");
    *resident = find(ctx, scene)->resident() ? 1 : 0;
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_serve_make_room(prtcl_hip_ctx *ctx, uint64_t bytes) {
  return guarded(ctx, [&]() -> int {
    if (!ctx->server.make_room(hooks(ctx), bytes))
      return fail(ctx, PRTCL_HIP_ADMISSION_DENIED,
                  "not enough idle scenes to evict for the requested bytes");
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_serve_submit(prtcl_hip_ctx *ctx, int32_t scene, int32_t k_steps,
                           int64_t *ticket) {
  return guarded(ctx, [&]() -> int {
    if (!ticket || k_steps < 0)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "invalid request");
    (void)find(ctx, scene);
    *ticket = ctx->server.submit(scene, k_steps);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_serve_tick(prtcl_hip_ctx *ctx, int32_t *n_stepped) {
  return guarded(ctx, [&]() -> int {
    int status = PRTCL_HIP_OK;
    const int n = ctx->server.tick(hooks(ctx), &ctx->res, &status);
    if (n_stepped)
      *n_stepped = n;
    return status;
  });
}

int prtcl_hip_serve_poll(prtcl_hip_ctx *ctx, int64_t *tickets, int32_t cap, int32_t *n) {
  return guarded(ctx, [&]() -> int {
    if (!n || (cap > 0 && !tickets) || cap < 0)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "invalid poll buffer");
    *n = ctx->server.poll(tickets, cap);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_serve_pending(prtcl_hip_ctx *ctx, int64_t *requests, int64_t *steps) {
  return guarded(ctx, [&]() -> int {
    ctx->server.pending(requests, steps);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_set_schedule(prtcl_hip_ctx *ctx, int32_t scene, int32_t mask) {
  return guarded(ctx, [&]() -> int {
    if (mask < -1 || mask > 127)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "schedule mask must be -1 or 0..127");
    if (scene < 0) {
      for (auto &t : ctx->res.batch_tuner)
        t.force(mask);
      return PRTCL_HIP_OK;
    }
    find(ctx, scene)->set_schedule(mask);
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_get_schedule(prtcl_hip_ctx *ctx, int32_t scene, int32_t *mask) {
  return guarded(ctx, [&]() -> int {
    if (!mask)
      return fail(ctx, PRTCL_HIP_INVALID_ARGUMENT, "mask is NULL");
    *mask = scene < 0 ? ctx->res.batch_tuner[0].current().mask() : find(ctx, scene)->schedule();
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_profile_enable(prtcl_hip_ctx *ctx, int on) {
  return guarded(ctx, [&]() -> int {
    PRTCL_HIP_CHECK(hipStreamSynchronize(ctx->res.stream));
    ctx->res.prof.reset();
    ctx->res.prof.on = on != 0;
    return PRTCL_HIP_OK;
  });
}

int prtcl_hip_profile_read(prtcl_hip_ctx *ctx, int32_t stage, double *total_ms,
                           uint64_t *launches, const char **name) {
  return guarded(ctx, [&]() -> int {
    if (stage < 0 || stage >= prtcl_amd::kStageCount)
      return fail(ctx, PRTCL_HIP_NOT_FOUND, "stage does not exist");
    PRTCL_HIP_CHECK(hipStreamSynchronize(ctx->res.stream));
    ctx->res.prof.resolve();
    if (total_ms)
      *total_ms = ctx->res.prof.ms[stage];
    if (launches)
      *launches = ctx->res.prof.launches[stage];
    if (name)
      *name = prtcl_amd::stage_name(stage);
    return PRTCL_HIP_OK;
  });
}

}  // extern "C"
