"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function include/prtcl_hip.h declares, the name tables match the
reference's field and procedure names, and the synthetic scenes have the
documented sizes."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from prtcl_amd import hip, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "prtcl_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(prtcl_hip_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(hip.EXPORTS)


def test_name_tables():
    L = hip.lib()
    for name, fid in hip.FIELDS.items():
        assert L.prtcl_hip_field_id(name.encode()) == fid
    assert L.prtcl_hip_field_id(b"no_such_field") == -1
    for name, pid in hip.PROCS.items():
        assert L.prtcl_hip_proc_id(name.encode()) == pid
    assert L.prtcl_hip_proc_id(b"iisph.no_such_procedure") == -1


def test_desc_defaults_and_footprint():
    L = hip.lib()
    d = hip.SceneDesc()
    L.prtcl_hip_scene_desc_init(C.byref(d))
    assert d.struct_size == C.sizeof(hip.SceneDesc)
    assert d.iisph_min_iterations == 2 and d.permute_every == 8
    d.n_fluid, d.n_boundary = 1000, 500
    b1 = L.prtcl_hip_scene_footprint(C.byref(d))
    d.n_fluid = 2000
    b2 = L.prtcl_hip_scene_footprint(C.byref(d))
    assert 0 < b1 < b2
    d.struct_size = 3  # wrong struct size is rejected
    assert L.prtcl_hip_scene_footprint(C.byref(d)) == 0


def test_null_context_is_an_error_not_a_crash():
    L = hip.lib()
    assert L.prtcl_hip_run(None, 0, 0) == hip.INVALID_ARGUMENT
    assert L.prtcl_hip_synchronize(None) == hip.INVALID_ARGUMENT
    assert L.prtcl_hip_ctx_destroy(None) == hip.INVALID_ARGUMENT


def test_scene_sizes():
    s0 = scenes.config0()
    assert (s0.n_fluid, s0.n_boundary) == (8000, 16002)
    assert s0.fluid_m[0] == np.float32(np.float32(0.025) ** 3 * 1000.0) or \
        abs(float(s0.fluid_m[0]) - 0.015625) < 1e-9
    b = scenes.box_boundary(401, 201, 201, 0.01)
    assert len(b) == 400002
    b2 = scenes.box_boundary(101, 101, 151, 0.01)
    assert len(b2) == 80002
    # jitter bounded by 1e-4 and reproducible
    again = scenes.config0()
    assert np.array_equal(again.fluid_x, s0.fluid_x)
    lat = scenes.dambreak(20, 20, 20, (2.0, 1.0, 1.0), 0.025, seed=0, steps=1, name="x",
                          jitter=0.0).fluid_x
    assert np.max(np.abs(s0.fluid_x - lat)) <= 1.0001e-4


def test_batched_scene_generator():
    sc = scenes.config3(count=64)
    n = np.array([s.n_fluid for s in sc])
    assert len(sc) == 64
    assert n.min() >= 1000 and n.max() <= 400000
    assert abs(np.log(n).mean() - (np.log(2000) + np.log(200000)) / 2) < 0.5
    v = sc[0].fluid_v
    assert v.min() >= -0.5 and v.max() <= 0.5


def test_splitmix_stream_is_counter_based():
    base = scenes.stream_base(3, 7)
    a = scenes.uniform(base, 0, 10)
    b = scenes.uniform(base, 5, 5)
    assert np.array_equal(a[5:], b)
    assert np.all((a >= 0) & (a < 1))
