"""ORACLE — test infrastructure only.

ctypes wrapper around oracle/_build/libprtcl_oracle.so, the CPU/OpenMP
restatement of the reference hot path. Imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libprtcl_oracle.so")


class Params(C.Structure):
    _fields_ = [
        ("radius", C.c_double),
        ("smoothing_scale", C.c_float),
        ("time_step", C.c_float),
        ("fade_duration", C.c_float),
        ("iisph_relaxation", C.c_float),
        ("gravity", C.c_float * 3),
        ("rest_density", C.c_float),
        ("fluid_viscosity", C.c_float),
        ("boundary_viscosity", C.c_float),
        ("wkbb18_maximum_error", C.c_float),
        ("wkbb18_maximum_iterations", C.c_int32),
        ("iisph_tolerance", C.c_float),
        ("iisph_min_iterations", C.c_int32),
        ("iisph_max_iterations", C.c_int32),
        ("explicit_viscosity", C.c_int32),
        ("implicit_viscosity", C.c_int32),
        ("permute_every", C.c_int32),
        ("fp64_reductions", C.c_int32),
        ("strain_rate_viscosity", C.c_float),
        ("pt16_vorticity_maximum_error", C.c_float),
        ("pt16_velocity_maximum_error", C.c_float),
        ("pt16_vorticity_maximum_iterations", C.c_int32),
        ("pt16_velocity_maximum_iterations", C.c_int32),
        ("gravity_center", C.c_float * 3),
        ("axial_gravity", C.c_int32),
    ]


class StepInfo(C.Structure):
    _fields_ = [
        ("iisph_iterations", C.c_int32),
        ("iisph_relative_error", C.c_float),
        ("wkbb18_iterations", C.c_int32),
        ("maximum_speed", C.c_float),
    ]


class Camera(C.Structure):
    """PinholeCamera (pinhole_camera.hpp:19-29)."""
    _fields_ = [("origin", C.c_double * 3), ("principal", C.c_double * 3),
                ("up", C.c_double * 3), ("focal_length", C.c_double),
                ("width", C.c_int64), ("height", C.c_int64)]


def camera(origin, principal, up, focal_length, width, height) -> Camera:
    c = Camera()
    for k in range(3):
        c.origin[k], c.principal[k], c.up[k] = origin[k], principal[k], up[k]
    c.focal_length, c.width, c.height = focal_length, width, height
    return c


def sphere_trace(x, smoothing_scale, cam: Camera, threshold=0.5, max_steps=300):
    """SphereTracer::Trace of the positions x (n, 3): (image, ray parameter),
    both (height, width) float64."""
    L = lib()
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    img = np.zeros((cam.height, cam.width), np.float64)
    par = np.zeros((cam.height, cam.width), np.float64)
    rc = L.oracle_sphere_trace(x.ctypes.data, len(x), float(smoothing_scale), C.byref(cam),
                               float(threshold), int(max_steps), img.ctypes.data,
                               par.ctypes.data)
    if rc != 0:
        raise ValueError("oracle_sphere_trace: invalid arguments")
    return img, par


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return LIB_PATH


_libs = {}


def lib(path: str = None):
    """The oracle library (default build), or the build at `path` (the
    invariant tests load builds with a planted mistake)."""
    path = LIB_PATH if path is None else path
    if path not in _libs:
        if path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(path)
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        L.oracle_scene_create.restype = vp
        L.oracle_scene_create.argtypes = [i64, i64]
        L.oracle_scene_destroy.argtypes = [vp]
        L.oracle_set_params.argtypes = [vp, C.POINTER(Params)]
        L.oracle_set_time.argtypes = [vp, C.c_double]
        L.oracle_get_time.argtypes = [vp]
        L.oracle_get_time.restype = C.c_double
        L.oracle_set_field.argtypes = [vp, C.c_char_p, C.c_char_p, vp, i64]
        L.oracle_get_field.argtypes = [vp, C.c_char_p, C.c_char_p, vp, i64]
        L.oracle_get_global.argtypes = [vp, C.c_char_p]
        L.oracle_get_global.restype = C.c_double
        L.oracle_set_global.argtypes = [vp, C.c_char_p, C.c_double]
        L.oracle_nsearch_update.argtypes = [vp]
        L.oracle_nsearch_permute.argtypes = [vp]
        L.oracle_run.argtypes = [vp, C.c_char_p]
        L.oracle_step.argtypes = [vp, C.POINTER(StepInfo)]
        for fn in (L.oracle_neighbors_csr, L.oracle_brute_neighbors_csr):
            fn.argtypes = [vp, C.c_int, C.c_int, vp, vp]
            fn.restype = i64
        L.oracle_morton_compare.argtypes = [vp, vp]
        L.oracle_kernel_w.argtypes = [vp, C.c_float]
        L.oracle_kernel_w.restype = C.c_float
        L.oracle_kernel_grad.argtypes = [vp, C.c_float, vp]
        L.oracle_invert3.argtypes = [vp, vp]
        L.oracle_pinv3.argtypes = [vp, vp]
        L.oracle_sorted_order.argtypes = [vp, vp, vp]
        L.oracle_sorted_order.restype = i64
        L.oracle_num_threads.restype = C.c_int
        L.oracle_spline_alpha3.restype = C.c_double
        L.oracle_sphere_trace.argtypes = [vp, i64, C.c_double, C.POINTER(Camera), C.c_double,
                                          i64, vp, vp]
        L.oracle_sphere_trace.restype = C.c_int
        _libs[path] = L
    return _libs[path]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


FLUID_FIELDS = {
    "position": 3, "velocity": 3, "acceleration": 3, "iisph_helper_c": 3,
    "wkbb18_velocity_delta": 3, "density": 1, "pressure": 1, "mass": 1,
    "iisph_source_term": 1, "iisph_diagonal_element": 1,
    "iisph_right_hand_side": 1, "time_of_birth": 1, "wkbb18_system_diagonal": 9,
    "gradient_correction": 9, "pt16_normalizer": 1, "pt16_density": 1,
    "pt16_same_density": 1, "vorticity": 3, "target_velocity_gradient": 9,
    "pt16_vorticity_diffusion_diagonal": 1, "pt16_vorticity_diffusion_rhs": 3,
    "pt16_velocity_reconstruction_diagonal": 1, "pt16_velocity_reconstruction_rhs": 3,
}
BOUNDARY_FIELDS = {"position": 3, "volume": 1}


class OracleScene:
    """One scene in the CPU restatement."""

    def __init__(self, scene, fp64_reductions: bool = False, permute_every=None,
                 lib_path: str = None):
        self.L = lib(lib_path)
        self.nf, self.nb = scene.n_fluid, scene.n_boundary
        self.h = self.L.oracle_scene_create(self.nf, self.nb)
        p = Params()
        p.radius = scene.radius
        p.smoothing_scale = scene.smoothing_scale
        p.time_step = scene.time_step
        p.fade_duration = scene.fade_duration
        p.iisph_relaxation = scene.iisph_relaxation
        for k in range(3):
            p.gravity[k] = scene.gravity[k]
        p.rest_density = scene.rest_density
        p.fluid_viscosity = scene.fluid_viscosity
        p.boundary_viscosity = scene.boundary_viscosity
        p.wkbb18_maximum_error = scene.wkbb18_maximum_error
        p.wkbb18_maximum_iterations = scene.wkbb18_maximum_iterations
        p.iisph_tolerance = scene.iisph_tolerance
        p.iisph_min_iterations = scene.iisph_min_iterations
        p.iisph_max_iterations = scene.iisph_max_iterations
        p.explicit_viscosity = int(scene.explicit_viscosity)
        p.implicit_viscosity = int(scene.implicit_viscosity)
        p.permute_every = scene.permute_every if permute_every is None else permute_every
        p.fp64_reductions = int(fp64_reductions)
        p.strain_rate_viscosity = scene.pt16_strain_rate_viscosity
        p.pt16_vorticity_maximum_error = scene.pt16_vorticity_maximum_error
        p.pt16_velocity_maximum_error = scene.pt16_velocity_maximum_error
        p.pt16_vorticity_maximum_iterations = scene.pt16_vorticity_maximum_iterations
        p.pt16_velocity_maximum_iterations = scene.pt16_velocity_maximum_iterations
        for k in range(3):
            p.gravity_center[k] = scene.gravity_center[k]
        p.axial_gravity = int(scene.axial_gravity)
        self.params = p
        self.L.oracle_set_params(self.h, C.byref(p))
        self.set("fluid", "position", scene.fluid_x)
        self.set("fluid", "velocity", scene.fluid_v)
        self.set("fluid", "mass", scene.fluid_m)
        self.set("fluid", "time_of_birth", scene.time_of_birth)
        self.set("boundary", "position", scene.boundary_x)

    def close(self):
        if self.h:
            self.L.oracle_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self.L.oracle_set_params(self.h, C.byref(self.params))

    def set(self, group, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.int64 if name == "id" else np.float32)
        rc = self.L.oracle_set_field(self.h, group.encode(), name.encode(), _ptr(a), a.size)
        if rc != 0:
            raise ValueError(f"oracle_set_field({group},{name}) -> {rc}")

    def get(self, group, name):
        n = self.nf if group == "fluid" else self.nb
        if name == "id":
            out = np.empty(n, np.int64)
        else:
            nc = (FLUID_FIELDS if group == "fluid" else BOUNDARY_FIELDS)[name]
            out = np.empty((n, nc) if nc > 1 else n, np.float32)
        rc = self.L.oracle_get_field(self.h, group.encode(), name.encode(), _ptr(out), out.size)
        if rc != 0:
            raise ValueError(f"oracle_get_field({group},{name}) -> {rc}")
        return out

    def get_global(self, name):
        return self.L.oracle_get_global(self.h, name.encode())

    def set_global(self, name, value):
        self.L.oracle_set_global(self.h, name.encode(), float(value))

    def update(self):
        self.L.oracle_nsearch_update(self.h)

    def permute(self):
        self.L.oracle_nsearch_permute(self.h)

    def run(self, proc):
        if self.L.oracle_run(self.h, proc.encode()) != 0:
            raise KeyError(f"procedure does not exist: {proc}")

    def step(self):
        info = StepInfo()
        self.L.oracle_step(self.h, C.byref(info))
        return dict(iisph_iterations=info.iisph_iterations,
                    iisph_relative_error=info.iisph_relative_error,
                    wkbb18_iterations=info.wkbb18_iterations,
                    maximum_speed=info.maximum_speed)

    def prepare(self):
        """Scene set-up as example.lua:342-344: grid update, boundary volume."""
        self.update()
        self.run("aiast12.compute_volume")

    def neighbors(self, g: int, ng: int, brute: bool = False):
        fn = self.L.oracle_brute_neighbors_csr if brute else self.L.oracle_neighbors_csr
        n = self.nf if g == 1 else self.nb
        off = np.empty(n + 1, np.int64)
        total = fn(self.h, g, ng, _ptr(off), None)
        idx = np.empty(total, np.int64)
        fn(self.h, g, ng, _ptr(off), _ptr(idx))
        return off, idx


def kernel_w(v, h):
    v = np.ascontiguousarray(v, np.float32)
    return lib().oracle_kernel_w(_ptr(v), float(h))


def kernel_grad(v, h):
    v = np.ascontiguousarray(v, np.float32)
    out = np.empty(3, np.float32)
    lib().oracle_kernel_grad(_ptr(v), float(h), _ptr(out))
    return out


def invert3(m):
    m = np.ascontiguousarray(m, np.float32).reshape(9)
    out = np.empty(9, np.float32)
    lib().oracle_invert3(_ptr(m), _ptr(out))
    return out


def pinv3(m):
    """Moore-Penrose pseudo-inverse of a column-major 3x3 (invert_pm)."""
    m = np.ascontiguousarray(m, np.float32).reshape(9)
    out = np.empty(9, np.float32)
    lib().oracle_pinv3(_ptr(m), _ptr(out))
    return out


def morton_compare(a, b):
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    return lib().oracle_morton_compare(_ptr(a), _ptr(b))
