/*
 * ORACLE — test infrastructure only.
 *
 * C interface of the CPU/OpenMP restatement of prtcl's SPH step hot path
 * (uniform-grid neighbour search, density, gravity, explicit viscosity,
 * symplectic Euler, Akinci-2012 boundary volume, IISPH, WKBB18 + CG).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker or as the timed CPU baseline; the
 * product (prtcl_amd) never links or calls it.
 */
#ifndef PRTCL_ORACLE_CPU_REF_H
#define PRTCL_ORACLE_CPU_REF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene parameters: the model's global fields and group uniforms that the
 * reference schemes bind in Load() (iisph.hpp:136-182, wkbb18.hpp:135-188,
 * density/gravity/viscosity/symplectic_euler/aiast12 Load()), plus the
 * caller-side loop settings of share/scripts/prtcl-lua.example.lua:363-462. */
typedef struct oracle_params {
  double radius;               /* Neighborhood::SetRadius (example.lua:206: 2h) */
  float smoothing_scale;       /* global smoothing_scale (h) */
  float time_step;             /* global time_step */
  float fade_duration;         /* global fade_duration */
  float iisph_relaxation;      /* global iisph_relaxation (omega) */
  float gravity[3];            /* global gravity */
  float rest_density;          /* fluid uniform rest_density */
  float fluid_viscosity;       /* fluid uniform dynamic_viscosity */
  float boundary_viscosity;    /* boundary uniform dynamic_viscosity */
  float wkbb18_maximum_error;  /* fluid uniform wkbb18_maximum_error */
  int32_t wkbb18_maximum_iterations; /* fluid uniform */
  float iisph_tolerance;       /* example.lua:419 (0.001) */
  int32_t iisph_min_iterations;/* example.lua:419 (2) */
  int32_t iisph_max_iterations;/* example.lua:411 (50; 100 in the configs) */
  int32_t explicit_viscosity;  /* run viscosity.accumulate_acceleration */
  int32_t implicit_viscosity;  /* 1: run wkbb18; 2: correction after the
                                  density and wkbb18_gc (prtcl-lua.bunny.lua);
                                  3: pt16 (example.lua:435-445) */
  int32_t permute_every;       /* example.lua:381 (8); 0 disables */
  int32_t fp64_reductions;     /* 0: FP32 reductions as the reference */
  /* pt16 (implicit_viscosity = 3): fluid uniform strain_rate_viscosity and the
   * pt16_maximum_error / _iterations the step sets before each of its two
   * solves (prtcl-lua.bunny.lua:150-155, example.lua:436-443) */
  float strain_rate_viscosity;
  float pt16_vorticity_maximum_error;
  float pt16_velocity_maximum_error;
  int32_t pt16_vorticity_maximum_iterations;
  int32_t pt16_velocity_maximum_iterations;
  /* global gravity_center (gravity.prtcl:13) and the step's choice between
   * gravity.initialize_acceleration (0) and initialize_axial_acceleration (1,
   * example.lua:390-391) */
  float gravity_center[3];
  int32_t axial_gravity;
} oracle_params;

typedef struct oracle_step_info {
  int32_t iisph_iterations;
  float iisph_relative_error;   /* aprde / nprde after the last iteration */
  int32_t wkbb18_iterations;
  float maximum_speed;
} oracle_step_info;

void *oracle_scene_create(int64_t n_fluid, int64_t n_boundary);
void oracle_scene_destroy(void *scene);
void oracle_set_params(void *scene, const oracle_params *params);
void oracle_set_time(void *scene, double t);
double oracle_get_time(void *scene);

/* group: "fluid" or "boundary"; name: the reference field name
 * ("position", "velocity", ..., "id" for the carried original index).
 * count = number of scalar components. Returns 0 on success. */
int oracle_set_field(void *scene, const char *group, const char *name,
                     const void *data, int64_t count);
int oracle_get_field(void *scene, const char *group, const char *name,
                     void *out, int64_t count);

/* global reduction results: "iisph_aprde", "iisph_nprde", "maximum_speed",
 * "wkbb18_iterations", and the pt16 uniforms "pt16_maximum_error",
 * "pt16_maximum_iterations", "pt16_iterations", "strain_rate_viscosity" */
double oracle_get_global(void *scene, const char *name);
void oracle_set_global(void *scene, const char *name, double value);

void oracle_nsearch_update(void *scene);
void oracle_nsearch_permute(void *scene);

/* procedure: "<scheme>.<procedure>" as in SchemeBase::RunProcedure, e.g.
 * "density.compute_density", "iisph.iteration_pressure". Returns 0 on success,
 * -1 if the procedure does not exist. */
int oracle_run(void *scene, const char *procedure);

/* One step in the order of example.lua:363-462 (update / permute every
 * permute_every steps, density, gravity, [viscosity], integrate velocity,
 * IISPH setup + loop, integrate velocity, [WKBB18 + integrate velocity],
 * integrate position, clock advance). */
void oracle_step(void *scene, oracle_step_info *info);

/* Neighbour lists of (group g, all particles) restricted to neighbour group
 * ng, as CSR with each row sorted ascending. g/ng: 0 = boundary, 1 = fluid
 * (the reference's group index = rank of the group name, b < f). With
 * indices == NULL only offsets (n+1 entries) are written. */
int64_t oracle_neighbors_csr(void *scene, int g, int ng, int64_t *offsets,
                             int64_t *indices);
/* Same sets by O(N^2) brute force with the identical predicate. */
int64_t oracle_brute_neighbors_csr(void *scene, int g, int ng,
                                   int64_t *offsets, int64_t *indices);

/* Morton comparison of two int32 grid indices (morton_order.hpp:39-57):
 * returns -1, 0, 1. */
int oracle_morton_compare(const int32_t *lhs, const int32_t *rhs);

/* Cubic spline (cubic_spline_kernel.hpp, kernel_facade.hpp) in FP32. */
float oracle_kernel_w(const float *v, float h);
void oracle_kernel_grad(const float *v, float h, float *out);
/* 3x3 inverse as used by wkbb18 (column-major storage, like Eigen). */
void oracle_invert3(const float *m, float *out);
/* Moore-Penrose pseudo-inverse (invert_pm, eigen.hpp:266-276), column-major. */
void oracle_pinv3(const float *m, float *out);

/* Sorted order of the grid (group, index) pairs; returns count. */
int64_t oracle_sorted_order(void *scene, int32_t *group, int64_t *index);

int oracle_num_threads(void);

/* Sphere tracer (trace_ref.cpp): PinholeCamera fields (pinhole_camera.hpp:19-29). */
typedef struct oracle_camera {
  double origin[3];
  double principal[3];
  double up[3];
  double focal_length;
  int64_t width, height;
} oracle_camera;
/* 1/(4 pi) of the cubic spline kernel in 3 dimensions, as the reference rounds it. */
double oracle_spline_alpha3(void);
/* SphereTracer::Trace over the n positions x (double [n][3]) of the visible
 * particles: image and the ray parameter (may be NULL) per pixel, row iy at
 * [iy * width + ix]. Returns 0, or 1 on invalid arguments. */
int oracle_sphere_trace(const double *x, int64_t n, double smoothing_scale,
                        const oracle_camera *cam, double threshold, int64_t max_steps,
                        double *image, double *parameter);

#ifdef __cplusplus
}
#endif

#endif
