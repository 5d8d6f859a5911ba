"""The C++ host interface (prtcl_amd/compat): Model / Group / field spans /
Neighborhood / SchemeBase / SchemeRegistry semantics on the CPU, and the
example.lua step loop driven through the GPU schemes against prtcl_hip_step
on the GPU (bit-identical fields and iteration counts)."""

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "prtcl_amd", "_lib")


def build_check(tmp_path):
    for lib in ("libprtcl_hip.so", "libprtcl_compat.so"):
        if not os.path.exists(os.path.join(LIBDIR, lib)):
            pytest.fail(f"{lib} missing: build with __graft_entry__.build()")
    exe = str(tmp_path / "compat_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "compat",
                                                                     "compat_check.cpp"),
                           "-L" + LIBDIR, "-lprtcl_compat", "-lprtcl_hip",
                           "-Wl,-rpath," + LIBDIR, "-o", exe])
    return exe


def test_compat_model_and_registry(tmp_path):
    exe = build_check(tmp_path)
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")


def mixed_model():
    """the model compat_check.cpp's build_mixed() builds"""
    from oracle import archive as A
    i = np.arange(4)
    return {
        "globals": {"gravity": A.field("f32", [3], [0, 0, -9.81]),
                    "iisph_nprde": A.field("s32", [], [7]),
                    "smoothing_scale": A.field("f32", [], [0.025]),
                    "big": A.field("s64", [], [-5]),
                    "flag": A.field("b", [], [True]),
                    "z_f64": A.field("f64", [], [1.5])},
        "groups": {
            "f": {"type": "fluid", "tags": ["visible", "dynamic"], "items": 4,
                  "varying": {
                      "position": A.field("f32", [3], np.stack([i, 10 + i, 20 + i], 1)),
                      "wkbb18_system_diagonal": A.field("f32", [3, 3],
                                                        np.arange(36).reshape(4, 9)),
                      "id": A.field("s64", [], 100 + i),
                      "alive": A.field("b", [], i % 2 == 0),
                      "mass": A.field("f64", [], 0.5 * i)},
                  "uniform": {"rest_density": A.field("f32", [], [1000]),
                              "wkbb18_maximum_iterations": A.field("s32", [], [100])}},
            "b": {"type": "boundary", "tags": [], "items": 0,
                  "varying": {"position": A.field("f32", [3], np.zeros((0, 3)))}},
            "a_empty": {"type": "other", "tags": [], "items": 0},
        },
    }


def test_compat_archive_matches_restatement(tmp_path):
    """Model::Save / Load of the C++ interface: the reference's round-trip
    tests, and byte equality with the restatement in both directions."""
    from oracle import archive as A
    want = A.write_model(mixed_model())
    (tmp_path / "from_oracle.bin").write_bytes(want)
    exe = build_check(tmp_path)
    r = subprocess.run([exe, "archive", str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
    assert (tmp_path / "mixed.bin").read_bytes() == want
    assert (tmp_path / "resaved.bin").read_bytes() == want
    m = A.read_model(want)
    assert m["groups"]["f"]["varying"]["wkbb18_system_diagonal"]["data"][2, 5] == 23


@pytest.mark.gpu
@pytest.mark.parametrize("gravity", ["uniform", "axial"])
def test_compat_step_loop_matches_c_abi(tmp_path, gravity):
    from prtcl_amd import scenes
    sc = scenes.tiny(8, steps=6)
    path = str(tmp_path / "scene.bin")
    with open(path, "wb") as fh:
        np.array([sc.n_fluid, sc.n_boundary], np.int64).tofile(fh)
        for a in (sc.fluid_x, sc.fluid_v, sc.fluid_m, sc.time_of_birth, sc.boundary_x):
            np.ascontiguousarray(a, np.float32).tofile(fh)
        np.array([sc.smoothing_scale, sc.time_step, sc.rest_density, sc.iisph_relaxation,
                  *sc.gravity, sc.iisph_tolerance, sc.iisph_max_iterations,
                  sc.fade_duration], np.float64).tofile(fh)
    exe = build_check(tmp_path)
    args = [exe, "gpu", path, "6"] + (["axial"] if gravity == "axial" else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
