"""Checkpoint and resume of device-resident scenes through the C ABI
(prtcl_hip_scene_save_archive / prtcl_hip_scene_load_archive), checked with
the archive restatement (oracle/archive.py):

* every varying field in the archive equals the field's download, bit for
  bit, in host order, and the globals hold the scene's values;
* a scene loaded from an archive continues bit-identically to the scene
  that was saved (fused and per-procedure steps, FP32 and FP64, wkbb18 and
  pt16 scenes), after host-visible permutes;
* an archive written by the restatement (other group names, only the
  fields the reference's scripts set) loads into the same scene as the
  arrays uploaded directly;
* malformed archives are refused without leaving a scene behind.
"""

import numpy as np
import pytest

from oracle import archive as A
from prtcl_amd import scenes

pytestmark = pytest.mark.gpu


def _hip():
    from prtcl_amd import hip
    return hip


def _fields(s, group):
    hip = _hip()
    out = {}
    for name, fid in hip.FIELDS.items():
        try:
            out[name] = s.download(group, name)
        except hip.HipError:
            pass  # a field this scene does not hold
    return out


# the varying fields the schemes of a scene without viscosity bind
# (iisph.hpp:161-182, density, gravity, symplectic_euler, aiast12)
PLAIN_FLUID = {"position", "velocity", "acceleration", "density", "pressure", "mass",
               "iisph_helper_c", "iisph_source_term", "iisph_diagonal_element",
               "iisph_right_hand_side", "time_of_birth"}
PLAIN_BOUNDARY = {"position", "volume"}


def _same_state(a, b):
    for g in ("fluid", "boundary"):
        fa, fb = _fields(a, g), _fields(b, g)
        assert set(fa) == set(fb), g
        for k in fa:
            assert np.array_equal(fa[k], fb[k]), (g, k)
    assert a.step_info() == b.step_info()
    assert a.get_global("current_time") == b.get_global("current_time")


def _viscous_block(implicit_viscosity=1):
    s = 0.025
    bx = scenes.box_boundary(14, 11, 13, s)
    fx = scenes.fluid_block((7, 5, 6), (3, 3, 1), s, 1e-4, scenes.stream_base(31, 0))
    rng = np.random.default_rng(5)
    v = rng.uniform(-0.3, 0.3, fx.shape).astype(np.float32)
    return scenes._finish("viscous_block", s, fx, bx, 4, v=v,
                          implicit_viscosity=implicit_viscosity, fluid_viscosity=2.0,
                          boundary_viscosity=1.0)


def _gc_block():
    return _viscous_block(implicit_viscosity=2)


def _pt16_block():
    s = 0.025
    bx = scenes.box_boundary(15, 13, 13, s)
    fx = scenes.fluid_block((8, 6, 8), (3, 3, 1), s, 1e-4, scenes.stream_base(23, 0))
    rng = np.random.default_rng(3)
    v = rng.uniform(-0.3, 0.3, fx.shape).astype(np.float32)
    return scenes._finish("pt16_block", s, fx, bx, 4, v=v, implicit_viscosity=3,
                          pt16_vorticity_maximum_error=1e-4, pt16_velocity_maximum_error=1e-4)


def test_archive_fields_equal_downloads(hip_ctx, tmp_path):
    hip = _hip()
    sc = scenes.tiny(8)
    print("case: tiny8 dam break, 11 steps, save", flush=True)
    s = hip.HipScene(hip_ctx, sc)
    s.prepare()
    s.step(11)  # a host-visible permute after step 8
    path = tmp_path / "scene.bin"
    n = s.save_archive(path)
    buf = path.read_bytes()
    assert n == len(buf)
    m = A.read_model(buf)
    assert sorted(m["groups"]) == ["b", "f"]
    f, b = m["groups"]["f"], m["groups"]["b"]
    assert f["type"] == "fluid" and f["tags"] == ["dynamic"] and f["items"] == sc.n_fluid
    assert b["type"] == "boundary" and b["items"] == sc.n_boundary
    for group, g, bound in (("fluid", f, PLAIN_FLUID), ("boundary", b, PLAIN_BOUNDARY)):
        want = _fields(s, group)
        names = [k for k in g["varying"] if k != "device_slot"]
        assert set(names) == bound, group
        for k in names:
            fld = g["varying"][k]
            assert fld["ctype"] == "f32"
            assert fld["shape"] == {1: [], 3: [3], 9: [3, 3]}[fld["data"].shape[1]]
            assert np.array_equal(fld["data"].reshape(want[k].shape), want[k]), (group, k)
        slot = g["varying"]["device_slot"]["data"].reshape(-1)
        assert np.array_equal(np.sort(slot), np.arange(g["items"]))
    gl = m["globals"]
    assert gl["current_time"]["data"][0] == np.float32(s.get_global("current_time"))
    assert abs(gl["device_clock"]["data"][0] - s.get_global("current_time")) < 1e-6
    assert gl["smoothing_scale"]["data"][0] == np.float32(sc.smoothing_scale)
    assert np.array_equal(gl["gravity"]["data"], np.float32(sc.gravity))
    assert int(gl["iisph_nprde"]["data"][0]) == sc.n_fluid
    assert int(gl["device_steps_since_permute"]["data"][0]) == 3
    assert f["uniform"]["rest_density"]["data"][0] == np.float32(sc.rest_density)
    # saving an evicted scene gives the same bytes
    s.evict()
    s.save_archive(tmp_path / "evicted.bin")
    assert (tmp_path / "evicted.bin").read_bytes() == buf


@pytest.mark.parametrize("real_type,step_mode", [(0, 0), (0, 1), (1, 0)],
                         ids=["fp32-fused", "fp32-procedures", "fp64-fused"])
def test_resume_is_bit_identical(hip_ctx, tmp_path, real_type, step_mode):
    hip = _hip()
    sc = scenes.tiny(8)
    print(f"case: tiny8 resume, real_type {real_type}, step_mode {step_mode}", flush=True)
    a = hip.HipScene(hip_ctx, sc, real_type=real_type, step_mode=step_mode)
    a.prepare()
    a.step(11)
    path = tmp_path / "scene.bin"
    a.save_archive(path)
    b = hip.HipScene.load_archive(hip_ctx, path, settings=sc, step_mode=step_mode)
    assert b.real_type == real_type and (b.nf, b.nb) == (sc.n_fluid, sc.n_boundary)
    assert b.desc.implicit_viscosity == 0 and b.desc.step_mode == step_mode
    for _ in range(3):  # crosses the next permute (step 16)
        a.step(2)
        b.step(2)
        _same_state(a, b)
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("make", [_viscous_block, _gc_block, _pt16_block],
                         ids=["wkbb18", "wkbb18_gc", "pt16"])
def test_resume_viscous_scenes(hip_ctx, tmp_path, make):
    hip = _hip()
    sc = make()
    print(f"case: {sc.name} resume", flush=True)
    a = hip.HipScene(hip_ctx, sc)
    a.prepare()
    a.step(3)
    path = tmp_path / "scene.bin"
    a.save_archive(path)
    m = A.read_model(path.read_bytes())
    fv = m["groups"]["f"]["varying"]
    iv = sc.implicit_viscosity
    assert ("wkbb18_system_diagonal" in fv) == (iv in (1, 2))
    assert ("wkbb18_velocity_delta" in fv) == (iv in (1, 2))
    assert ("gradient_correction" in fv) == (iv == 2)
    assert ("pt16_same_density" in fv) == (iv == 3)
    # the scheme comes from the archive's fields, not from the settings
    plain = scenes.tiny(4)
    b = hip.HipScene.load_archive(hip_ctx, path, settings=plain)
    assert b.desc.implicit_viscosity == sc.implicit_viscosity
    a2 = hip.HipScene.load_archive(hip_ctx, path, settings=sc)
    a.step(2)
    a2.step(2)
    _same_state(a, a2)
    for s in (a, b, a2):
        s.destroy()


def test_archive_from_the_restatement_loads(hip_ctx, tmp_path):
    hip = _hip()
    sc = scenes.tiny(7, seed=11)
    print("case: tiny7 archive written by the restatement", flush=True)
    h = np.float32(sc.smoothing_scale)
    model = {
        "globals": {"smoothing_scale": A.field("f32", [], [h]),
                    "time_step": A.field("f32", [], [sc.time_step]),
                    "gravity": A.field("f32", [3], sc.gravity),
                    "iisph_relaxation": A.field("f32", [], [sc.iisph_relaxation]),
                    "fade_duration": A.field("f32", [], [sc.fade_duration]),
                    "gravity_center": A.field("f32", [3], [0, 0, 0])},
        "groups": {
            "water": {"type": "fluid", "tags": ["dynamic", "visible"], "items": sc.n_fluid,
                      "varying": {"position": A.field("f32", [3], sc.fluid_x),
                                  "velocity": A.field("f32", [3], sc.fluid_v),
                                  "mass": A.field("f32", [], sc.fluid_m),
                                  "time_of_birth": A.field("f32", [], sc.time_of_birth),
                                  "colour": A.field("s32", [], np.zeros(sc.n_fluid))},
                      "uniform": {"rest_density": A.field("f32", [], [sc.rest_density])}},
            "walls": {"type": "boundary", "tags": [], "items": sc.n_boundary,
                      "varying": {"position": A.field("f32", [3], sc.boundary_x)}},
        },
    }
    path = tmp_path / "model.bin"
    path.write_bytes(A.write_model(model))
    b = hip.HipScene.load_archive(hip_ctx, path, settings=sc)
    assert (b.nf, b.nb) == (sc.n_fluid, sc.n_boundary)
    assert np.float32(b.desc.smoothing_scale) == h
    a = hip.HipScene(hip_ctx, sc)
    a.step(5)
    b.step(5)
    _same_state(a, b)
    a.destroy()
    b.destroy()


def test_malformed_archives_are_refused(hip_ctx, tmp_path):
    hip = _hip()
    sc = scenes.tiny(5)
    print("case: malformed archives", flush=True)
    a = hip.HipScene(hip_ctx, sc)
    a.step(1)
    path = tmp_path / "scene.bin"
    a.save_archive(path)
    buf = path.read_bytes()
    before = hip_ctx.bytes_in_use()
    bad = []
    for cut in (0, 9, len(buf) // 3, len(buf) - 1):
        p = tmp_path / f"cut{cut}.bin"
        p.write_bytes(buf[:cut])
        bad.append(p)
    m = A.read_model(buf)
    v = m["groups"]["f"]["varying"]["velocity"]
    m["groups"]["f"]["varying"]["velocity"] = A.field("f64", [3], v["data"])
    p = tmp_path / "f64_velocity.bin"
    p.write_bytes(A.write_model(m))
    bad.append(p)
    m = A.read_model(buf)
    m["groups"]["g2"] = dict(m["groups"]["f"])
    p = tmp_path / "two_fluids.bin"
    p.write_bytes(A.write_model(m))
    bad.append(p)
    m = A.read_model(buf)
    m["groups"]["f"]["varying"]["device_slot"] = A.field(
        "s32", [], np.zeros(sc.n_fluid, np.int32))
    p = tmp_path / "bad_slots.bin"
    p.write_bytes(A.write_model(m))
    bad.append(p)
    bad.append(tmp_path / "does_not_exist.bin")
    for p in bad:
        with pytest.raises(hip.HipError):
            hip.HipScene.load_archive(hip_ctx, p, settings=sc)
        assert hip_ctx.bytes_in_use() == before, p.name
    a.destroy()
