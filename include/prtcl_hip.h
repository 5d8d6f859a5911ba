/*
 * prtcl_hip.h — C ABI of the MI355X-native SPH step hot path.
 *
 * The reference (siretty/prtcl) exposes this path as C++ classes: generated
 * schemes behind SchemeBase (src/prtcl/schemes/scheme_base.hpp:18-64), a
 * registry (scheme_base.hpp:66-114) and a Neighborhood
 * (src/prtcl/util/neighborhood.hpp:23-42) over a Model's uniform and varying
 * fields (src/prtcl/data/model.hpp:85-88, group.hpp:75-99). This header is the
 * flat boundary those classes are re-provided over on the GPU side: opaque
 * handles, plain pointers and sizes, int status codes, no exceptions and no
 * torch types. Each entry point names the reference interface it replaces.
 *
 * Threading: a context is used from one host thread at a time. All work of a
 * context runs on the context's own HIP stream.
 */
#ifndef PRTCL_HIP_H
#define PRTCL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
/* The reference throws string literals ("procedure does not exist",
 * scheme_base.hpp:37; "scheme is not registered", :91) or typed errors
 * (src/prtcl/errors/); here every failure is a status code, and
 * prtcl_hip_last_error() holds the message. */
enum prtcl_hip_status {
  PRTCL_HIP_OK = 0,
  PRTCL_HIP_INVALID_ARGUMENT = 1,
  PRTCL_HIP_ADMISSION_DENIED = 2, /* HBM budget exceeded */
  PRTCL_HIP_NOT_CONVERGED = 3,    /* warning: solver hit its iteration cap */
  PRTCL_HIP_NONFINITE = 4,        /* NaN/Inf detected in particle state */
  PRTCL_HIP_ERROR = 5,            /* HIP runtime error */
  PRTCL_HIP_NOT_FOUND = 6         /* unknown scene / field / procedure */
};

typedef struct prtcl_hip_ctx prtcl_hip_ctx;

/* ---- groups, fields, globals, procedures -------------------------------- */
enum prtcl_hip_group { PRTCL_HIP_BOUNDARY = 0, PRTCL_HIP_FLUID = 1 };

/* Varying fields, 1:1 with the names the reference schemes bind in Load()
 * (iisph.hpp:161-182, wkbb18.hpp:160-188, density/aiast12/symplectic_euler).
 * Host layout is the reference's AoS layout: f32/f64 [n][3] for vectors,
 * [n][3][3] column-major (Eigen) for the 3x3 system diagonal, [n] scalars. */
enum prtcl_hip_field {
  PRTCL_HIP_F_POSITION = 0,               /* "position" [3] (both groups) */
  PRTCL_HIP_F_VELOCITY = 1,               /* "velocity" [3] */
  PRTCL_HIP_F_ACCELERATION = 2,           /* "acceleration" [3] */
  PRTCL_HIP_F_DENSITY = 3,                /* "density" */
  PRTCL_HIP_F_PRESSURE = 4,               /* "pressure" */
  PRTCL_HIP_F_MASS = 5,                   /* "mass" */
  PRTCL_HIP_F_IISPH_HELPER_C = 6,         /* "iisph_helper_c" [3] */
  PRTCL_HIP_F_IISPH_SOURCE_TERM = 7,      /* "iisph_source_term" */
  PRTCL_HIP_F_IISPH_DIAGONAL_ELEMENT = 8, /* "iisph_diagonal_element" */
  PRTCL_HIP_F_IISPH_RIGHT_HAND_SIDE = 9,  /* "iisph_right_hand_side" */
  PRTCL_HIP_F_TIME_OF_BIRTH = 10,         /* "time_of_birth" */
  PRTCL_HIP_F_VOLUME = 11,                /* "volume" (boundary) */
  PRTCL_HIP_F_WKBB18_VELOCITY_DELTA = 12, /* "wkbb18_velocity_delta" [3] */
  PRTCL_HIP_F_WKBB18_SYSTEM_DIAGONAL = 13,/* "wkbb18_system_diagonal" [3][3] */
  PRTCL_HIP_F_GRADIENT_CORRECTION = 14,   /* "gradient_correction" [3][3] (fluid;
                                             correction.hpp, wkbb18_gc.hpp) */
  /* pt16.hpp:140-190 (fluid; scenes created with implicit_viscosity = 3) */
  PRTCL_HIP_F_PT16_NORMALIZER = 15,       /* "pt16_normalizer" (bound, unused) */
  PRTCL_HIP_F_PT16_DENSITY = 16,          /* "pt16_density" */
  PRTCL_HIP_F_PT16_SAME_DENSITY = 17,     /* "pt16_same_density" */
  PRTCL_HIP_F_VORTICITY = 18,             /* "vorticity" [3] */
  PRTCL_HIP_F_TARGET_VELOCITY_GRADIENT = 19, /* "target_velocity_gradient" [3][3] */
  PRTCL_HIP_F_PT16_VORTICITY_DIFFUSION_DIAGONAL = 20,    /* bound, unused */
  PRTCL_HIP_F_PT16_VORTICITY_DIFFUSION_RHS = 21,         /* [3] */
  PRTCL_HIP_F_PT16_VELOCITY_RECONSTRUCTION_DIAGONAL = 22,/* bound, unused */
  PRTCL_HIP_F_PT16_VELOCITY_RECONSTRUCTION_RHS = 23,     /* [3] */
  PRTCL_HIP_F_COUNT = 24
};

/* Global / uniform scalars read back or set (model.global and group
 * uniforms of the reference; iisph.hpp:136-142, wkbb18.hpp:135-158,
 * symplectic_euler max_speed). */
enum prtcl_hip_global {
  PRTCL_HIP_G_IISPH_APRDE = 0,       /* "iisph_aprde" (FP64 on device) */
  PRTCL_HIP_G_IISPH_NPRDE = 1,       /* "iisph_nprde" */
  PRTCL_HIP_G_MAXIMUM_SPEED = 2,     /* "maximum_speed" (running max) */
  PRTCL_HIP_G_WKBB18_ITERATIONS = 3, /* fluid uniform "wkbb18_iterations" */
  PRTCL_HIP_G_CURRENT_TIME = 4,      /* "current_time" */
  PRTCL_HIP_G_IISPH_ITERATIONS = 5,  /* iterations of the last IISPH solve */
  PRTCL_HIP_G_IISPH_REL_ERROR = 6,   /* aprde / nprde after the last solve */
  /* pt16 fluid uniforms (pt16.hpp:160-175) */
  PRTCL_HIP_G_PT16_MAXIMUM_ERROR = 7,      /* "pt16_maximum_error" */
  PRTCL_HIP_G_PT16_MAXIMUM_ITERATIONS = 8, /* "pt16_maximum_iterations" */
  PRTCL_HIP_G_PT16_ITERATIONS = 9,         /* "pt16_iterations" (last solve) */
  PRTCL_HIP_G_STRAIN_RATE_VISCOSITY = 10,  /* "strain_rate_viscosity" (xi) */
  PRTCL_HIP_G_COUNT = 11
};

/* One id per reference procedure (SchemeBase::RunProcedure names). */
enum prtcl_hip_proc {
  PRTCL_HIP_P_NSEARCH_UPDATE = 0,        /* Neighborhood::Update */
  PRTCL_HIP_P_NSEARCH_PERMUTE = 1,       /* Neighborhood::Permute */
  PRTCL_HIP_P_AIAST12_COMPUTE_VOLUME = 2,
  PRTCL_HIP_P_DENSITY_COMPUTE_DENSITY = 3,
  PRTCL_HIP_P_GRAVITY_INITIALIZE_ACCELERATION = 4,
  PRTCL_HIP_P_VISCOSITY_ACCUMULATE_ACCELERATION = 5,
  PRTCL_HIP_P_SE_INTEGRATE_VELOCITY_WITH_HARD_FADE = 6,
  PRTCL_HIP_P_SE_INTEGRATE_POSITION = 7,
  PRTCL_HIP_P_IISPH_SETUP = 8,
  PRTCL_HIP_P_IISPH_ITERATION_PRESSURE_ACCELERATION = 9,
  PRTCL_HIP_P_IISPH_ITERATION_PRESSURE = 10,
  PRTCL_HIP_P_WKBB18_COMPUTE_DIAGONAL = 11,
  PRTCL_HIP_P_WKBB18_ACCUMULATE_ACCELERATION = 12,
  /* correction.hpp:133-217 */
  PRTCL_HIP_P_CORRECTION_IDENTITY_GRADIENT_CORRECTION = 13,
  PRTCL_HIP_P_CORRECTION_COMPUTE_GRADIENT_CORRECTION = 14,
  /* wkbb18_gc.hpp: wkbb18 with the fluid particle's gradient_correction
   * applied to every kernel gradient of its terms */
  PRTCL_HIP_P_WKBB18_GC_COMPUTE_DIAGONAL = 15,
  PRTCL_HIP_P_WKBB18_GC_ACCUMULATE_ACCELERATION = 16,
  /* pt16.hpp:213-760 (scenes created with implicit_viscosity = 3) */
  PRTCL_HIP_P_PT16_SETUP = 17,
  PRTCL_HIP_P_PT16_SOLVE_VORTICITY_DIFFUSION = 18,
  PRTCL_HIP_P_PT16_VORTICITY_PRESERVATION = 19,
  PRTCL_HIP_P_PT16_SOLVE_VELOCITY_RECONSTRUCTION = 20,
  /* gravity.hpp:189-215 (gravity.prtcl:26-33): a = -|g| xc_o / |xc_o| with
   * xc_o the part of x - gravity_center orthogonal to g (0 on the axis) */
  PRTCL_HIP_P_GRAVITY_INITIALIZE_AXIAL_ACCELERATION = 21,
  /* gravity.hpp:157-186 (gravity.prtcl:21-25): a += g */
  PRTCL_HIP_P_GRAVITY_ACCUMULATE_ACCELERATION = 22,
  PRTCL_HIP_P_COUNT = 23
};

/* Scene description: counts, arithmetic type and the model's parameter
 * fields (global smoothing_scale/time_step/gravity/iisph_relaxation/
 * fade_duration, fluid rest_density/dynamic_viscosity/wkbb18_maximum_*,
 * boundary dynamic_viscosity; example.lua:122-186) and the caller-side loop
 * settings of example.lua:363-462. */
typedef struct prtcl_hip_scene_desc {
  uint32_t struct_size;   /* sizeof(prtcl_hip_scene_desc) */
  int32_t real_type;      /* 0 = f32 (the reference's T=f32), 1 = f64 */
  int64_t n_fluid;
  int64_t n_boundary;
  double radius;          /* Neighborhood::SetRadius; 0 selects 2h */
  double smoothing_scale;
  double time_step;
  double fade_duration;
  double iisph_relaxation;
  double gravity[3];
  double rest_density;
  double fluid_viscosity;
  double boundary_viscosity;
  double iisph_tolerance;      /* 1e-3 */
  int32_t iisph_min_iterations;/* 2 */
  int32_t iisph_max_iterations;/* 100 */
  double wkbb18_maximum_error;
  int32_t wkbb18_maximum_iterations;
  int32_t explicit_viscosity;  /* step runs viscosity.accumulate_acceleration */
  int32_t implicit_viscosity;  /* step runs 1: wkbb18; 2: wkbb18_gc, with
                                  correction.compute_gradient_correction after
                                  the density (prtcl-lua.bunny.lua:277-281);
                                  3: pt16 (example.lua:435-445) */
  int32_t permute_every;       /* host-visible permute period; 0 = never */
  int32_t step_mode;           /* prtcl_hip_step kernels: 0 = fused passes over a
                                  per-step neighbour list (default), 1 = one
                                  kernel per reference procedure */
  /* pt16 (implicit_viscosity = 3): fluid uniform strain_rate_viscosity, and
   * the pt16_maximum_error / _iterations the step sets before each solve
   * (prtcl-lua.bunny.lua:129-155, example.lua:436-443) */
  double pt16_strain_rate_viscosity;
  double pt16_vorticity_maximum_error;
  double pt16_velocity_maximum_error;
  int32_t pt16_vorticity_maximum_iterations;
  int32_t pt16_velocity_maximum_iterations;
  /* global gravity_center (gravity.prtcl:13; example.lua sets it to the
   * domain centre) and the step's gravity procedure: 0 runs
   * gravity.initialize_acceleration, 1 initialize_axial_acceleration (the
   * call example.lua:391 makes) */
  double gravity_center[3];
  int32_t axial_gravity;
} prtcl_hip_scene_desc;

/* Fills a descriptor with the reference script defaults. */
void prtcl_hip_scene_desc_init(prtcl_hip_scene_desc *desc);

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart (the reference has no device); owns the HIP
 * stream, the scene table and the HBM budget. budget 0 = no limit. */
int prtcl_hip_ctx_create(int device, uint64_t hbm_budget_bytes, prtcl_hip_ctx **out);
int prtcl_hip_ctx_destroy(prtcl_hip_ctx *ctx);
const char *prtcl_hip_last_error(const prtcl_hip_ctx *ctx);
/* bytes of device memory held by scenes of this context */
uint64_t prtcl_hip_ctx_bytes_in_use(const prtcl_hip_ctx *ctx);
int prtcl_hip_synchronize(prtcl_hip_ctx *ctx);

/* ---- scenes -------------------------------------------------------------
 * Replaces Model::AddGroup("f","fluid") / ("b","boundary") + the schemes'
 * Load(Model&) (iisph.hpp:134-185): allocates every field the schemes bind,
 * zero-filled like VaryingManager's new fields (varying_field.hpp:165-167).
 * Admission: fails with PRTCL_HIP_ADMISSION_DENIED when the scene's device
 * footprint would exceed the context budget. */
int prtcl_hip_scene_create(prtcl_hip_ctx *ctx, const prtcl_hip_scene_desc *desc,
                           int32_t *scene_id);
int prtcl_hip_scene_destroy(prtcl_hip_ctx *ctx, int32_t scene);
/* Device bytes a scene with this description would occupy. */
uint64_t prtcl_hip_scene_footprint(const prtcl_hip_scene_desc *desc);
/* The description a scene was created with (a loaded scene: as completed
 * from its archive). Replaces Group::GetItemCount (group.hpp:59) and reads of
 * the model's parameter fields. */
int prtcl_hip_scene_describe(prtcl_hip_ctx *ctx, int32_t scene, prtcl_hip_scene_desc *out);

/* ---- field transfer ----------------------------------------------------
 * Replaces reading/writing VaryingFieldSpan items (varying_field.hpp:25-58)
 * of the group's field. `host` holds count values of the scene's real type
 * (float or double) in the reference's host order and AoS layout.
 * count must equal n * components. */
int prtcl_hip_upload_field(prtcl_hip_ctx *ctx, int32_t scene, int32_t group,
                           int32_t field, const void *host, size_t count);
int prtcl_hip_download_field(prtcl_hip_ctx *ctx, int32_t scene, int32_t group,
                             int32_t field, void *host, size_t count);
/* Field id from the reference's field name ("position", ...); -1 if none. */
int32_t prtcl_hip_field_id(const char *name);

/* ---- globals ------------------------------------------------------------
 * Replaces model.global:get_field(name):get()/set() (example.lua:414,418). */
int prtcl_hip_get_global(prtcl_hip_ctx *ctx, int32_t scene, int32_t global, double *value);
int prtcl_hip_set_global(prtcl_hip_ctx *ctx, int32_t scene, int32_t global, double value);

/* ---- procedures ----------------------------------------------------------
 * Replaces SchemeBase::RunProcedure(name, nhood) (scheme_base.hpp:33-38) and
 * Neighborhood::Update/Permute (neighborhood.cpp:100-154). Unknown ids give
 * PRTCL_HIP_NOT_FOUND ("procedure does not exist"). Asynchronous: returns
 * after enqueueing, except where a global must be read back. */
int prtcl_hip_run(prtcl_hip_ctx *ctx, int32_t scene, int32_t proc);
/* Procedure id from "<scheme>.<procedure>" (e.g. "iisph.setup") or
 * "neighborhood.update"/"neighborhood.permute"; -1 if none. */
int32_t prtcl_hip_proc_id(const char *name);

/* ---- fused paths -------------------------------------------------------
 * The IISPH loop of example.lua:409-421 on the device: iterations run until
 * it >= min_it and aprde/nprde < tol, or max_it; aprde is reduced in FP64 and
 * the convergence decision is taken on the device. iters/rel_err may be NULL.
 * Returns PRTCL_HIP_NOT_CONVERGED (a warning) at the cap. */
int prtcl_hip_iisph_solve(prtcl_hip_ctx *ctx, int32_t scene, double tol, int32_t min_it,
                          int32_t max_it, int32_t *iters, double *rel_err);
/* k whole steps (example.lua:363-462 order) of each listed scene. Scenes are
 * advanced together: every kernel of the step sweeps all listed scenes, so a
 * batch of uneven scenes costs one launch per stage rather than one per
 * scene. Per-scene results are independent of which other scenes are listed.
 * Returns once the steps have finished on the device (as the reference's
 * procedures return when done): PRTCL_HIP_NONFINITE when a step of a listed
 * scene produced a NaN or infinite position or speed (the particle state is
 * left as computed), else PRTCL_HIP_NOT_CONVERGED when an IISPH solve stopped
 * at its iteration cap, else PRTCL_HIP_OK. */
int prtcl_hip_step(prtcl_hip_ctx *ctx, const int32_t *scenes, int32_t n_scenes,
                   int32_t k_steps);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). prtcl_hip_step's batched step of the listed scenes with
 * its per-step tables in a fixed 4 MiB device array of the library (one per
 * GPU, set up with the context) instead of context buffers grown on demand:
 * it calls no hipMalloc / hipFree, captures no graph and times nothing, and
 * its results equal prtcl_hip_step's bit for bit. Every listed scene must be
 * resident, take the batched step (the fused step, step mode 0) and have its
 * grid built (by a step or neighborhood.update); all FP32 or all FP64. Returns
 * PRTCL_HIP_INVALID_ARGUMENT otherwise, and when
 * the tables exceed 4 MiB (about 2,000 scenes of 10,000 particles). Calls on
 * one GPU share the array: a call copies its tables in only after the last
 * kernel of the call before it (any context) that read them, but calls that
 * overlap in time from different host threads race. */
int prtcl_hip_step_noalloc(prtcl_hip_ctx *ctx, const int32_t *scenes, int32_t n_scenes,
                           int32_t k_steps);
/* Per-scene statistics of the last step: IISPH iterations, relative density
 * error, WKBB18 iterations. */
int prtcl_hip_last_step_info(prtcl_hip_ctx *ctx, int32_t scene, int32_t *iisph_iters,
                             double *iisph_rel_err, int32_t *wkbb18_iters);

/* ---- neighbour export ---------------------------------------------------
 * Replaces Neighborhood::CopyNeighbors (neighborhood.cpp:156-167) for every
 * particle of `group`: the neighbours in `neighbor_group` as CSR over the
 * host index space. offsets has n+1 entries; with indices == NULL only the
 * offsets are written and *total receives the entry count. Rows are sorted
 * ascending. */
int prtcl_hip_copy_neighbors(prtcl_hip_ctx *ctx, int32_t scene, int32_t group,
                             int32_t neighbor_group, uint32_t *offsets,
                             uint32_t *indices, uint64_t *total);

/* Permutation applied by the last PRTCL_HIP_P_NSEARCH_PERMUTE for `group`:
 * new host index i holds the particle that had host index perm[i] (gather
 * semantics of VaryingManager::PermuteItems, varying_manager.hpp:86-102, as
 * pinned by group.test.cpp:110-120). perm has n entries. */
int prtcl_hip_last_permutation(prtcl_hip_ctx *ctx, int32_t scene, int32_t group,
                               uint32_t *perm);

/* ---- model archives: checkpoint and resume --------------------------------
 * Replaces Model::Save / Model::Load over NativeBinaryArchiveWriter /
 * NativeBinaryArchiveReader (model.cpp:54-75, util/archive.cpp:4-83), which
 * the Lua binding exposes as model:save_native_binary(path) and
 * prtcl.data.model.load_native_binary(path) (prtcl-lua/module_data.cpp:83-95).
 *
 * save writes the scene in that format, readable by the reference's
 * Model::Load and share/python/prtcl_archive.py: the global fields of the
 * schemes (smoothing_scale, time_step, fade_duration, iisph_relaxation,
 * gravity, current_time, maximum_speed, iisph_aprde, iisph_nprde), a group
 * "b" of type "boundary" and a group "f" of type "fluid" (tag "dynamic")
 * with every varying field the scene holds, in host order, and their
 * uniform fields (rest_density, dynamic_viscosity, the wkbb18 / pt16 solver
 * uniforms of a viscous scene). Three extra fields carry the device state
 * of a bit-identical continuation: global "device_clock" (f64),
 * "device_steps_since_permute" (s32) and "device_slot" (s32 per particle)
 * in each group. Works on an evicted scene too. bytes_written may be NULL.
 *
 * load creates a scene from an archive: particle counts, real type and the
 * model's parameters come from the archive, the loop settings (IISPH
 * tolerance and iterations, explicit viscosity, permute period, step mode,
 * radius, pt16 per-solve settings) from `settings` (NULL: the defaults of
 * prtcl_hip_scene_desc_init), the implicit viscosity scheme from the fluid
 * fields present (wkbb18_system_diagonal: wkbb18, with gradient_correction:
 * wkbb18_gc, pt16_same_density: pt16), else from `settings`. A scene holds
 * one group of type "fluid" and at most one of type "boundary"; varying
 * fields of other names are skipped, fields the scene holds but the archive
 * lacks stay zero (as new fields of the reference). Errors:
 * PRTCL_HIP_INVALID_ARGUMENT for a malformed or truncated archive or a field
 * of another type or item count, PRTCL_HIP_ADMISSION_DENIED over budget. */
int prtcl_hip_scene_save_archive(prtcl_hip_ctx *ctx, int32_t scene, const char *path,
                                 uint64_t *bytes_written);
int prtcl_hip_scene_load_archive(prtcl_hip_ctx *ctx, const char *path,
                                 const prtcl_hip_scene_desc *settings, int32_t *scene_id);

/* ---- VTK output -----------------------------------------------------------
 * Replaces SaveVTK(std::ostream&, Group const&) (util/save_vtk.cpp:57-129),
 * bound to Lua as group:save_vtk(path) (prtcl-lua/module_data.cpp:126-131):
 * writes one group of the scene (PRTCL_HIP_FLUID as "f fluid",
 * PRTCL_HIP_BOUNDARY as "b boundary") as the same ASCII legacy-VTK POLYDATA
 * text, byte for byte: POINTS from position, then POINT_DATA velocity and
 * acceleration (VECTORS), _index (the host index), mass, volume, density,
 * pressure, time_of_birth (SCALARS) for the fields the group holds, every
 * real as "%.6f". The numbers are formatted on the device. bytes_written
 * may be NULL. */
int prtcl_hip_scene_save_vtk(prtcl_hip_ctx *ctx, int32_t scene, int32_t group, const char *path,
                             uint64_t *bytes_written);

/* ---- sphere tracing --------------------------------------------------------
 * Replaces SphereTracer<PinholeCamera>::Trace(model) (util/sphere_tracer.hpp:
 * 46-268), bound to Lua as prtcl.util.sphere_tracer.new(camera):trace(model)
 * (prtcl-lua/module_util.cpp:162-175), for one group of a scene (the
 * tracer's "visible" group). The camera is the reference's PinholeCamera
 * (geometry/pinhole_camera.hpp:19-29). Writes width * height doubles to
 * image (and, if not NULL, the ray parameter of each pixel to parameter),
 * row iy at [iy * width + ix]: 0 where the march ran out of max_steps, 0.3
 * where the ray left the scene, else -dot(ray direction, surface normal).
 * threshold is SphereTracer::SetThreshold (0.5 by default), max_steps
 * SetMaxSteps (300). Double precision on the device. */
typedef struct prtcl_hip_camera {
  double origin[3];
  double principal[3];
  double up[3];
  double focal_length;
  int64_t width, height; /* sensor, pixels */
} prtcl_hip_camera;
int prtcl_hip_scene_trace(prtcl_hip_ctx *ctx, int32_t scene, int32_t group,
                          const prtcl_hip_camera *camera, double threshold, int32_t max_steps,
                          double *image, double *parameter);

/* ---- residency and serving ----------------------------------------------
 * No reference counterpart (the reference runs one scene per process from
 * its Lua loop). Device state of a scene can be evicted to host memory and is
 * brought back bit-identically (same slot order and buffers) by the next
 * operation on the scene. The request loop advances resident scenes by k
 * steps per request: prtcl_hip_serve_tick steps every scene with pending
 * work once, all of them in one batched step (continuous batching at step
 * granularity); a scene with work that does not fit the HBM budget is brought
 * back after idle scenes (no pending work, least recently stepped first) are
 * evicted, or waits. A request completes when its last step has finished on
 * the device; prtcl_hip_serve_poll returns completed tickets in completion
 * order. Each scene's results do not depend on the other requests. */
int prtcl_hip_scene_evict(prtcl_hip_ctx *ctx, int32_t scene);
int prtcl_hip_scene_resident(prtcl_hip_ctx *ctx, int32_t scene, int32_t *resident);
/* evicts idle scenes until `bytes` more fit the budget (admission of a new
 * scene); PRTCL_HIP_ADMISSION_DENIED if they cannot */
int prtcl_hip_serve_make_room(prtcl_hip_ctx *ctx, uint64_t bytes);
int prtcl_hip_serve_submit(prtcl_hip_ctx *ctx, int32_t scene, int32_t k_steps,
                           int64_t *ticket);
int prtcl_hip_serve_tick(prtcl_hip_ctx *ctx, int32_t *n_stepped);
int prtcl_hip_serve_poll(prtcl_hip_ctx *ctx, int64_t *tickets, int32_t cap, int32_t *n);
int prtcl_hip_serve_pending(prtcl_hip_ctx *ctx, int64_t *requests, int64_t *steps);

/* ---- schedule of the fused passes ---------------------------------------
 * No reference counterpart (tuning of this implementation; results do not
 * depend on it). The setup, pressure-acceleration and pressure passes of
 * prtcl_hip_step either gather the neighbours' fields from global memory or
 * stage them slab by slab in LDS; an autotuner times both on the first fused
 * steps of each scene (and of batched steps) and keeps the faster per pass.
 * mask: bit 0 setup, bit 1 pressure acceleration, bit 2 pressure in LDS;
 * bit 3 launches each IISPH iteration kernel by kernel instead of replaying
 * it from a captured hipGraph; bit 4 gives the alpha and beta reductions of
 * every CG iteration (WKBB18 and pt16 solves) one-block kernels of their own
 * (they are otherwise folded into the vector kernels of scenes of up to
 * 262,144 fluid particles); bit 5 walks the neighbour list twice for the
 * WKBB18 diagonal and the CG set-up of a fused step instead of once; bit 6
 * runs IISPH iteration 1 as its own pressure-acceleration and pressure
 * passes instead of folding it into the setup pass (with every pressure 0 at
 * the start of the solve it needs no neighbours); -1 restarts the tuner.
 * scene < 0 addresses the schedule of batched steps. */
int prtcl_hip_set_schedule(prtcl_hip_ctx *ctx, int32_t scene, int32_t mask);
int prtcl_hip_get_schedule(prtcl_hip_ctx *ctx, int32_t scene, int32_t *mask);

/* ---- stage timing -------------------------------------------------------
 * No reference counterpart (the reference logs host-side timestamps around
 * each RunProcedure, module_schemes.cpp:30-37). When enabled, hipEvents are
 * recorded on the context stream around each kernel stage of
 * prtcl_hip_step(); launches of solver iterations that exited at the
 * convergence flag are not counted. Enabling clears the accumulators. */
int prtcl_hip_profile_enable(prtcl_hip_ctx *ctx, int on);
int prtcl_hip_profile_read(prtcl_hip_ctx *ctx, int32_t stage, double *total_ms,
                           uint64_t *launches, const char **name);

#ifdef __cplusplus
}
#endif

#endif /* PRTCL_HIP_H */
