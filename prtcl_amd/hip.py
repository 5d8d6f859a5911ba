"""Python binding of the C ABI in include/prtcl_hip.h (ctypes).

This is plumbing for tests and the benchmark: every call goes through the
C-ABI library prtcl_amd/_lib/libprtcl_hip.so. Loading fails loudly when the
library has not been built; nothing here falls back to a CPU path.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import List, Dict, Optional, Sequence

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libprtcl_hip.so")

OK, INVALID_ARGUMENT, ADMISSION_DENIED, NOT_CONVERGED, NONFINITE, ERROR, NOT_FOUND = range(7)
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "ADMISSION_DENIED", 3: "NOT_CONVERGED",
                4: "NONFINITE", 5: "ERROR", 6: "NOT_FOUND"}
BOUNDARY, FLUID = 0, 1

FIELDS = {
    "position": 0, "velocity": 1, "acceleration": 2, "density": 3, "pressure": 4,
    "mass": 5, "iisph_helper_c": 6, "iisph_source_term": 7,
    "iisph_diagonal_element": 8, "iisph_right_hand_side": 9, "time_of_birth": 10,
    "volume": 11, "wkbb18_velocity_delta": 12, "wkbb18_system_diagonal": 13,
    "gradient_correction": 14, "pt16_normalizer": 15, "pt16_density": 16,
    "pt16_same_density": 17, "vorticity": 18, "target_velocity_gradient": 19,
    "pt16_vorticity_diffusion_diagonal": 20, "pt16_vorticity_diffusion_rhs": 21,
    "pt16_velocity_reconstruction_diagonal": 22, "pt16_velocity_reconstruction_rhs": 23,
}
FIELD_COMPONENTS = {0: 3, 1: 3, 2: 3, 3: 1, 4: 1, 5: 1, 6: 3, 7: 1, 8: 1, 9: 1, 10: 1,
                    11: 1, 12: 3, 13: 9, 14: 9, 15: 1, 16: 1, 17: 1, 18: 3, 19: 9, 20: 1,
                    21: 3, 22: 1, 23: 3}
GLOBALS = {"iisph_aprde": 0, "iisph_nprde": 1, "maximum_speed": 2, "wkbb18_iterations": 3,
           "current_time": 4, "iisph_iterations": 5, "iisph_rel_error": 6,
           "pt16_maximum_error": 7, "pt16_maximum_iterations": 8, "pt16_iterations": 9,
           "strain_rate_viscosity": 10}
PROCS = {
    "neighborhood.update": 0, "neighborhood.permute": 1, "aiast12.compute_volume": 2,
    "density.compute_density": 3, "gravity.initialize_acceleration": 4,
    "viscosity.accumulate_acceleration": 5,
    "symplectic_euler.integrate_velocity_with_hard_fade": 6,
    "symplectic_euler.integrate_position": 7, "iisph.setup": 8,
    "iisph.iteration_pressure_acceleration": 9, "iisph.iteration_pressure": 10,
    "wkbb18.compute_diagonal": 11, "wkbb18.accumulate_acceleration": 12,
    "correction.identity_gradient_correction": 13,
    "correction.compute_gradient_correction": 14,
    "wkbb18_gc.compute_diagonal": 15, "wkbb18_gc.accumulate_acceleration": 16,
    "pt16.setup": 17, "pt16.solve_vorticity_diffusion": 18, "pt16.vorticity_preservation": 19,
    "pt16.solve_velocity_reconstruction": 20,
    "gravity.initialize_axial_acceleration": 21,
    "gravity.accumulate_acceleration": 22,
}

# exported symbols declared in include/prtcl_hip.h
EXPORTS = [
    "prtcl_hip_scene_desc_init", "prtcl_hip_ctx_create", "prtcl_hip_ctx_destroy",
    "prtcl_hip_last_error", "prtcl_hip_ctx_bytes_in_use", "prtcl_hip_synchronize",
    "prtcl_hip_scene_create", "prtcl_hip_scene_destroy", "prtcl_hip_scene_footprint",
    "prtcl_hip_scene_describe", "prtcl_hip_scene_save_archive", "prtcl_hip_scene_load_archive",
    "prtcl_hip_scene_save_vtk", "prtcl_hip_scene_trace",
    "prtcl_hip_upload_field", "prtcl_hip_download_field", "prtcl_hip_field_id",
    "prtcl_hip_get_global", "prtcl_hip_set_global", "prtcl_hip_run", "prtcl_hip_proc_id",
    "prtcl_hip_iisph_solve", "prtcl_hip_step", "prtcl_hip_step_noalloc",
    "prtcl_hip_last_step_info",
    "prtcl_hip_set_schedule", "prtcl_hip_get_schedule",
    "prtcl_hip_scene_evict", "prtcl_hip_scene_resident", "prtcl_hip_serve_make_room",
    "prtcl_hip_serve_submit", "prtcl_hip_serve_tick", "prtcl_hip_serve_poll",
    "prtcl_hip_serve_pending",
    "prtcl_hip_copy_neighbors", "prtcl_hip_last_permutation", "prtcl_hip_profile_enable",
    "prtcl_hip_profile_read",
]


class Camera(C.Structure):
    """prtcl_hip_camera: the reference's PinholeCamera."""
    _fields_ = [("origin", C.c_double * 3), ("principal", C.c_double * 3),
                ("up", C.c_double * 3), ("focal_length", C.c_double),
                ("width", C.c_int64), ("height", C.c_int64)]


def camera(origin, principal, up, focal_length=1.0, width=800, height=600) -> Camera:
    c = Camera()
    for k in range(3):
        c.origin[k], c.principal[k], c.up[k] = origin[k], principal[k], up[k]
    c.focal_length, c.width, c.height = focal_length, width, height
    return c


class SceneDesc(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("real_type", C.c_int32),
        ("n_fluid", C.c_int64),
        ("n_boundary", C.c_int64),
        ("radius", C.c_double),
        ("smoothing_scale", C.c_double),
        ("time_step", C.c_double),
        ("fade_duration", C.c_double),
        ("iisph_relaxation", C.c_double),
        ("gravity", C.c_double * 3),
        ("rest_density", C.c_double),
        ("fluid_viscosity", C.c_double),
        ("boundary_viscosity", C.c_double),
        ("iisph_tolerance", C.c_double),
        ("iisph_min_iterations", C.c_int32),
        ("iisph_max_iterations", C.c_int32),
        ("wkbb18_maximum_error", C.c_double),
        ("wkbb18_maximum_iterations", C.c_int32),
        ("explicit_viscosity", C.c_int32),
        ("implicit_viscosity", C.c_int32),
        ("permute_every", C.c_int32),
        ("step_mode", C.c_int32),
        ("pt16_strain_rate_viscosity", C.c_double),
        ("pt16_vorticity_maximum_error", C.c_double),
        ("pt16_velocity_maximum_error", C.c_double),
        ("pt16_vorticity_maximum_iterations", C.c_int32),
        ("pt16_velocity_maximum_iterations", C.c_int32),
        ("gravity_center", C.c_double * 3),
        ("axial_gravity", C.c_int32),
    ]


class HipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib = None


def lib():
    """Load the C-ABI library; raise if it is missing (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "or python prtcl_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u32p = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_uint32)
        L.prtcl_hip_scene_desc_init.argtypes = [C.POINTER(SceneDesc)]
        L.prtcl_hip_ctx_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
        L.prtcl_hip_ctx_destroy.argtypes = [vp]
        L.prtcl_hip_last_error.argtypes = [vp]
        L.prtcl_hip_last_error.restype = C.c_char_p
        L.prtcl_hip_ctx_bytes_in_use.argtypes = [vp]
        L.prtcl_hip_ctx_bytes_in_use.restype = C.c_uint64
        L.prtcl_hip_synchronize.argtypes = [vp]
        L.prtcl_hip_scene_create.argtypes = [vp, C.POINTER(SceneDesc), C.POINTER(i32)]
        L.prtcl_hip_scene_destroy.argtypes = [vp, i32]
        L.prtcl_hip_scene_footprint.argtypes = [C.POINTER(SceneDesc)]
        L.prtcl_hip_scene_footprint.restype = C.c_uint64
        L.prtcl_hip_upload_field.argtypes = [vp, i32, i32, i32, vp, C.c_size_t]
        L.prtcl_hip_download_field.argtypes = [vp, i32, i32, i32, vp, C.c_size_t]
        L.prtcl_hip_field_id.argtypes = [C.c_char_p]
        L.prtcl_hip_proc_id.argtypes = [C.c_char_p]
        L.prtcl_hip_get_global.argtypes = [vp, i32, i32, C.POINTER(C.c_double)]
        L.prtcl_hip_set_global.argtypes = [vp, i32, i32, C.c_double]
        L.prtcl_hip_run.argtypes = [vp, i32, i32]
        L.prtcl_hip_iisph_solve.argtypes = [vp, i32, C.c_double, i32, i32, C.POINTER(i32),
                                            C.POINTER(C.c_double)]
        L.prtcl_hip_step.argtypes = [vp, C.POINTER(i32), i32, i32]
        L.prtcl_hip_step_noalloc.argtypes = [vp, C.POINTER(i32), i32, i32]
        L.prtcl_hip_set_schedule.argtypes = [vp, i32, i32]
        L.prtcl_hip_scene_describe.argtypes = [vp, i32, C.POINTER(SceneDesc)]
        L.prtcl_hip_scene_save_archive.argtypes = [vp, i32, C.c_char_p, C.POINTER(C.c_uint64)]
        L.prtcl_hip_scene_load_archive.argtypes = [vp, C.c_char_p, C.POINTER(SceneDesc),
                                                   C.POINTER(i32)]
        L.prtcl_hip_scene_save_vtk.argtypes = [vp, i32, i32, C.c_char_p, C.POINTER(C.c_uint64)]
        L.prtcl_hip_scene_trace.argtypes = [vp, i32, i32, C.POINTER(Camera), C.c_double, i32,
                                            vp, vp]
        L.prtcl_hip_scene_evict.argtypes = [vp, i32]
        L.prtcl_hip_scene_resident.argtypes = [vp, i32, C.POINTER(i32)]
        L.prtcl_hip_serve_make_room.argtypes = [vp, C.c_uint64]
        L.prtcl_hip_serve_submit.argtypes = [vp, i32, i32, C.POINTER(i64)]
        L.prtcl_hip_serve_tick.argtypes = [vp, C.POINTER(i32)]
        L.prtcl_hip_serve_poll.argtypes = [vp, C.POINTER(i64), i32, C.POINTER(i32)]
        L.prtcl_hip_serve_pending.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
        L.prtcl_hip_get_schedule.argtypes = [vp, i32, C.POINTER(i32)]
        L.prtcl_hip_last_step_info.argtypes = [vp, i32, C.POINTER(i32), C.POINTER(C.c_double),
                                               C.POINTER(i32)]
        L.prtcl_hip_copy_neighbors.argtypes = [vp, i32, i32, i32, u32p, u32p,
                                               C.POINTER(C.c_uint64)]
        L.prtcl_hip_last_permutation.argtypes = [vp, i32, i32, u32p]
        L.prtcl_hip_profile_enable.argtypes = [vp, C.c_int]
        L.prtcl_hip_profile_read.argtypes = [vp, i32, C.POINTER(C.c_double),
                                             C.POINTER(C.c_uint64), C.POINTER(C.c_char_p)]
        _lib = L
    return _lib


def desc_for(scene, real_type: int = 0, step_mode: int = 0) -> SceneDesc:
    """Scene description from a prtcl_amd.scenes.Scene."""
    d = SceneDesc()
    lib().prtcl_hip_scene_desc_init(C.byref(d))
    d.real_type = real_type
    d.step_mode = step_mode
    d.n_fluid = scene.n_fluid
    d.n_boundary = scene.n_boundary
    d.radius = scene.radius
    d.smoothing_scale = float(np.float32(scene.smoothing_scale))
    d.time_step = float(np.float32(scene.time_step))
    d.fade_duration = float(np.float32(scene.fade_duration))
    d.iisph_relaxation = float(np.float32(scene.iisph_relaxation))
    for k in range(3):
        d.gravity[k] = float(np.float32(scene.gravity[k]))
    d.rest_density = float(np.float32(scene.rest_density))
    d.fluid_viscosity = float(np.float32(scene.fluid_viscosity))
    d.boundary_viscosity = float(np.float32(scene.boundary_viscosity))
    d.iisph_tolerance = float(scene.iisph_tolerance)
    d.iisph_min_iterations = scene.iisph_min_iterations
    d.iisph_max_iterations = scene.iisph_max_iterations
    d.wkbb18_maximum_error = float(np.float32(scene.wkbb18_maximum_error))
    d.wkbb18_maximum_iterations = scene.wkbb18_maximum_iterations
    d.explicit_viscosity = int(scene.explicit_viscosity)
    d.implicit_viscosity = int(scene.implicit_viscosity)
    d.permute_every = scene.permute_every
    d.pt16_strain_rate_viscosity = float(np.float32(scene.pt16_strain_rate_viscosity))
    d.pt16_vorticity_maximum_error = float(np.float32(scene.pt16_vorticity_maximum_error))
    d.pt16_velocity_maximum_error = float(np.float32(scene.pt16_velocity_maximum_error))
    d.pt16_vorticity_maximum_iterations = int(scene.pt16_vorticity_maximum_iterations)
    d.pt16_velocity_maximum_iterations = int(scene.pt16_velocity_maximum_iterations)
    for k in range(3):
        d.gravity_center[k] = float(np.float32(scene.gravity_center[k]))
    d.axial_gravity = int(scene.axial_gravity)
    return d


class Context:
    def __init__(self, device: int = 0, budget: int = 0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.prtcl_hip_ctx_create(device, budget, C.byref(h))
        if rc != OK:
            raise HipError(rc, "prtcl_hip_ctx_create failed")
        self.h = h

    def check(self, rc, what=""):
        if rc != OK:
            raise HipError(rc, f"{what}: {self.L.prtcl_hip_last_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.prtcl_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bytes_in_use(self) -> int:
        return int(self.L.prtcl_hip_ctx_bytes_in_use(self.h))

    def synchronize(self):
        self.check(self.L.prtcl_hip_synchronize(self.h), "synchronize")

    def profile(self, on: bool):
        self.check(self.L.prtcl_hip_profile_enable(self.h, int(on)), "profile_enable")

    def profile_read(self) -> Dict[str, tuple]:
        """{stage name: (total_ms, launches)} of the stages timed since
        profile(True)."""
        out = {}
        k = 0
        while True:
            ms, n, name = C.c_double(), C.c_uint64(), C.c_char_p()
            rc = self.L.prtcl_hip_profile_read(self.h, k, C.byref(ms), C.byref(n), C.byref(name))
            if rc == NOT_FOUND:
                break
            self.check(rc, "profile_read")
            out[name.value.decode()] = (ms.value, n.value)
            k += 1
        return out

    # ---- residency and serving
    def make_room(self, nbytes: int):
        """Evict idle scenes until nbytes more fit the HBM budget."""
        self.check(self.L.prtcl_hip_serve_make_room(self.h, int(nbytes)), "make_room")

    def submit(self, scene: "HipScene", k: int) -> int:
        t = C.c_int64()
        self.check(self.L.prtcl_hip_serve_submit(self.h, scene.id, int(k), C.byref(t)),
                   "serve_submit")
        return t.value

    def tick(self) -> int:
        """One step of every scene with pending requests (one batched step)."""
        n = C.c_int32()
        rc = self.L.prtcl_hip_serve_tick(self.h, C.byref(n))
        if rc not in (OK, NOT_CONVERGED):
            self.check(rc, "serve_tick")
        return n.value

    def poll(self, cap: int = 4096) -> List[int]:
        buf = (C.c_int64 * cap)()
        n = C.c_int32()
        self.check(self.L.prtcl_hip_serve_poll(self.h, buf, cap, C.byref(n)), "serve_poll")
        return [buf[k] for k in range(n.value)]

    def pending(self):
        r, st = C.c_int64(), C.c_int64()
        self.check(self.L.prtcl_hip_serve_pending(self.h, C.byref(r), C.byref(st)), "pending")
        return r.value, st.value

    def set_batch_schedule(self, mask: int):
        """Force the LDS/global schedule of batched steps (-1: autotune)."""
        self.check(self.L.prtcl_hip_set_schedule(self.h, -1, mask), "set_schedule")

    def batch_schedule(self) -> int:
        m = C.c_int32()
        self.check(self.L.prtcl_hip_get_schedule(self.h, -1, C.byref(m)), "get_schedule")
        return m.value

    def step(self, scenes: Sequence["HipScene"], k: int = 1) -> int:
        ids = (C.c_int32 * len(scenes))(*[s.id for s in scenes])
        rc = self.L.prtcl_hip_step(self.h, ids, len(scenes), k)
        if rc not in (OK, NOT_CONVERGED):
            self.check(rc, "step")
        return rc


class HipScene:
    """One device-resident scene created from a prtcl_amd.scenes.Scene."""

    def __init__(self, ctx: Context, scene, real_type: int = 0, upload: bool = True,
                 step_mode: int = 0):
        self.ctx, self.L = ctx, ctx.L
        self.scene = scene
        self.real_type = real_type
        self.dtype = np.float64 if real_type == 1 else np.float32
        self.desc = desc_for(scene, real_type, step_mode)
        sid = C.c_int32()
        ctx.check(self.L.prtcl_hip_scene_create(ctx.h, C.byref(self.desc), C.byref(sid)),
                  "scene_create")
        self.id = sid.value
        self.nf, self.nb = scene.n_fluid, scene.n_boundary
        if upload:
            self.upload("fluid", "position", scene.fluid_x)
            self.upload("fluid", "velocity", scene.fluid_v)
            self.upload("fluid", "mass", scene.fluid_m)
            self.upload("fluid", "time_of_birth", scene.time_of_birth)
            self.upload("boundary", "position", scene.boundary_x)

    @classmethod
    def load_archive(cls, ctx: Context, path, settings=None, step_mode: int = 0):
        """A scene created from a model archive (prtcl_hip_scene_load_archive).
        settings: a prtcl_amd.scenes.Scene whose loop settings apply (IISPH
        tolerance and iterations, viscosity options, permute period), or None
        for the defaults; counts, real type and the model's parameters come
        from the archive."""
        self = cls.__new__(cls)
        self.ctx, self.L = ctx, ctx.L
        self.scene = settings
        if settings is not None:
            desc = desc_for(settings, 0, step_mode)
        else:
            desc = SceneDesc()
            self.L.prtcl_hip_scene_desc_init(C.byref(desc))
            desc.step_mode = step_mode
        sid = C.c_int32()
        ctx.check(self.L.prtcl_hip_scene_load_archive(ctx.h, os.fsencode(path), C.byref(desc),
                                                      C.byref(sid)), "scene_load_archive")
        self.id = sid.value
        d = SceneDesc()
        ctx.check(self.L.prtcl_hip_scene_describe(ctx.h, self.id, C.byref(d)), "scene_describe")
        self.desc = d
        self.real_type = d.real_type
        self.dtype = np.float64 if d.real_type == 1 else np.float32
        self.nf, self.nb = d.n_fluid, d.n_boundary
        return self

    def save_archive(self, path) -> int:
        """Writes the scene as a model archive (prtcl_hip_scene_save_archive);
        returns the bytes written."""
        n = C.c_uint64()
        self.ctx.check(self.L.prtcl_hip_scene_save_archive(self.ctx.h, self.id,
                                                            os.fsencode(path), C.byref(n)),
                       "scene_save_archive")
        return n.value

    def save_vtk(self, group, path) -> int:
        """Writes one group as legacy VTK text (prtcl_hip_scene_save_vtk);
        returns the bytes written."""
        n = C.c_uint64()
        self.ctx.check(self.L.prtcl_hip_scene_save_vtk(self.ctx.h, self.id, self._group(group),
                                                        os.fsencode(path), C.byref(n)),
                       "scene_save_vtk")
        return n.value

    def trace(self, cam: "Camera", group="fluid", threshold=0.5, max_steps=300):
        """SphereTracer::Trace of one group (prtcl_hip_scene_trace): (image,
        ray parameter), both (height, width) float64."""
        img = np.zeros((cam.height, cam.width), np.float64)
        par = np.zeros((cam.height, cam.width), np.float64)
        self.ctx.check(self.L.prtcl_hip_scene_trace(self.ctx.h, self.id, self._group(group),
                                                     C.byref(cam), float(threshold),
                                                     int(max_steps), img.ctypes.data,
                                                     par.ctypes.data), "scene_trace")
        return img, par

    def destroy(self):
        if self.id is not None:
            self.ctx.check(self.L.prtcl_hip_scene_destroy(self.ctx.h, self.id), "scene_destroy")
            self.id = None

    @staticmethod
    def _group(g):
        return FLUID if g in ("fluid", "f", FLUID) else BOUNDARY

    def upload(self, group, field, arr):
        g = self._group(group)
        fid = FIELDS[field]
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        self.ctx.check(self.L.prtcl_hip_upload_field(self.ctx.h, self.id, g, fid,
                                                      a.ctypes.data, a.size), f"upload {field}")

    def download(self, group, field):
        g = self._group(group)
        fid = FIELDS[field]
        n = self.nf if g == FLUID else self.nb
        nc = FIELD_COMPONENTS[fid]
        out = np.empty((n, nc) if nc > 1 else n, self.dtype)
        self.ctx.check(self.L.prtcl_hip_download_field(self.ctx.h, self.id, g, fid,
                                                        out.ctypes.data, out.size),
                       f"download {field}")
        return out

    def run(self, proc: str):
        pid = PROCS.get(proc, -1)
        rc = self.L.prtcl_hip_run(self.ctx.h, self.id, pid)
        self.ctx.check(rc, f"run {proc}")

    def get_global(self, name: str) -> float:
        v = C.c_double()
        self.ctx.check(self.L.prtcl_hip_get_global(self.ctx.h, self.id, GLOBALS[name],
                                                    C.byref(v)), f"get_global {name}")
        return v.value

    def set_global(self, name: str, value: float):
        self.ctx.check(self.L.prtcl_hip_set_global(self.ctx.h, self.id, GLOBALS[name],
                                                    float(value)), f"set_global {name}")

    def solve(self, tol=None, min_it=None, max_it=None):
        s = self.scene
        it, rel = C.c_int32(), C.c_double()
        rc = self.L.prtcl_hip_iisph_solve(
            self.ctx.h, self.id, s.iisph_tolerance if tol is None else tol,
            s.iisph_min_iterations if min_it is None else min_it,
            s.iisph_max_iterations if max_it is None else max_it, C.byref(it), C.byref(rel))
        if rc not in (OK, NOT_CONVERGED):
            self.ctx.check(rc, "iisph_solve")
        return it.value, rel.value, rc

    def step(self, k: int = 1) -> int:
        return self.ctx.step([self], k)

    def evict(self):
        """Move the device state to host memory (restored by the next use)."""
        self.ctx.check(self.L.prtcl_hip_scene_evict(self.ctx.h, self.id), "scene_evict")

    def resident(self) -> bool:
        r = C.c_int32()
        self.ctx.check(self.L.prtcl_hip_scene_resident(self.ctx.h, self.id, C.byref(r)),
                       "scene_resident")
        return bool(r.value)

    def set_schedule(self, mask: int):
        """Force the schedule of the fused passes: bit 0 setup, bit 1
        pressure acceleration, bit 2 pressure in LDS; bit 3 eager IISPH
        iterations; bit 4 separate one-block alpha/beta kernels in the CG
        solves; bit 5 separate WKBB18 diagonal and CG set-up passes; bit 6
        IISPH iteration 1 in its own passes (-1: autotune)."""
        self.ctx.check(self.L.prtcl_hip_set_schedule(self.ctx.h, self.id, mask), "set_schedule")

    def schedule(self) -> int:
        m = C.c_int32()
        self.ctx.check(self.L.prtcl_hip_get_schedule(self.ctx.h, self.id, C.byref(m)),
                       "get_schedule")
        return m.value

    def step_info(self) -> Dict[str, float]:
        it, rel, wk = C.c_int32(), C.c_double(), C.c_int32()
        self.ctx.check(self.L.prtcl_hip_last_step_info(self.ctx.h, self.id, C.byref(it),
                                                        C.byref(rel), C.byref(wk)), "step_info")
        return dict(iisph_iterations=it.value, iisph_relative_error=rel.value,
                    wkbb18_iterations=wk.value)

    def prepare(self):
        """Scene set-up as example.lua:342-344: grid update, boundary volume."""
        self.run("neighborhood.update")
        self.run("aiast12.compute_volume")

    def neighbors(self, g: int, ng: int):
        n = self.nf if g == FLUID else self.nb
        off = np.empty(n + 1, np.uint32)
        total = C.c_uint64()
        u32p = C.POINTER(C.c_uint32)
        self.ctx.check(self.L.prtcl_hip_copy_neighbors(
            self.ctx.h, self.id, g, ng, off.ctypes.data_as(u32p), None, C.byref(total)),
            "copy_neighbors")
        idx = np.empty(total.value, np.uint32)
        self.ctx.check(self.L.prtcl_hip_copy_neighbors(
            self.ctx.h, self.id, g, ng, off.ctypes.data_as(u32p), idx.ctypes.data_as(u32p),
            C.byref(total)), "copy_neighbors")
        return off.astype(np.int64), idx.astype(np.int64)

    def last_permutation(self, group):
        g = self._group(group)
        n = self.nf if g == FLUID else self.nb
        out = np.empty(n, np.uint32)
        self.ctx.check(self.L.prtcl_hip_last_permutation(
            self.ctx.h, self.id, g, out.ctypes.data_as(C.POINTER(C.c_uint32))),
            "last_permutation")
        return out
