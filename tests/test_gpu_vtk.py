"""VTK output of device scenes through the C ABI (prtcl_hip_scene_save_vtk):
the text formatted on the device equals, byte for byte, the restatement of
the reference's SaveVTK (oracle/vtk.py) applied to the downloaded fields,
for adversarial values in FP32 and FP64 scenes (ties at the sixth decimal,
carries, signed zeros, denormals, integers beyond 2^64, the largest finite
values, NaN and infinities) and for a stepped dam break."""

import numpy as np
import pytest

from oracle import vtk as V
from prtcl_amd import scenes
from vtk_values import adversarial

pytestmark = pytest.mark.gpu

FLUID = ("position", "velocity", "acceleration", "mass", "density", "pressure", "time_of_birth")
BOUNDARY = ("position", "volume")


def _expected(s, group, names):
    f = {k: s.download(group, k) for k in names}
    return V.write_vtk("f" if group == "fluid" else "b", group, f)


@pytest.mark.parametrize("real_type", [0, 1], ids=["fp32", "fp64"])
def test_vtk_of_adversarial_values(hip_ctx, tmp_path, real_type):
    from prtcl_amd import hip
    dt = np.float64 if real_type == 1 else np.float32
    nf, nb = 3000, 700
    print(f"case: adversarial values, real_type {real_type}", flush=True)
    sc = scenes.Scene(name="vtk", spacing=0.025,
                      fluid_x=np.zeros((nf, 3), np.float32), fluid_v=np.zeros((nf, 3), np.float32),
                      fluid_m=np.ones(nf, np.float32), time_of_birth=np.zeros(nf, np.float32),
                      boundary_x=np.zeros((nb, 3), np.float32), smoothing_scale=0.025)
    s = hip.HipScene(hip_ctx, sc, real_type=real_type)
    for k, name in enumerate(FLUID):
        c = 3 if name in ("position", "velocity", "acceleration") else 1
        a = adversarial(nf * c, dt, seed=100 + k, nonfinite=name != "position")
        s.upload("fluid", name, a.reshape(nf, c) if c == 3 else a)
    s.upload("boundary", "position", adversarial(nb * 3, dt, seed=200).reshape(nb, 3))
    s.upload("boundary", "volume", adversarial(nb, dt, seed=201, nonfinite=True))
    for group, names in (("fluid", FLUID), ("boundary", BOUNDARY)):
        path = tmp_path / f"{group}.vtk"
        n = s.save_vtk(group, path)
        got = path.read_bytes()
        assert n == len(got)
        want = _expected(s, group, names)
        if got != want:
            gl, wl = got.split(b"\n"), want.split(b"\n")
            bad = next(i for i, (x, y) in enumerate(zip(gl, wl)) if x != y)
            raise AssertionError(f"line {bad}: {gl[bad]!r} != {wl[bad]!r}")
    s.destroy()


def test_vtk_of_a_stepped_dam_break(hip_ctx, tmp_path):
    from prtcl_amd import hip
    sc = scenes.tiny(8)
    print("case: tiny8 dam break, 3 steps, vtk", flush=True)
    s = hip.HipScene(hip_ctx, sc)
    s.prepare()
    s.step(3)
    for group, names in (("fluid", FLUID), ("boundary", BOUNDARY)):
        path = tmp_path / f"{group}.vtk"
        s.save_vtk(group, path)
        assert path.read_bytes() == _expected(s, group, names), group
    # an unwritable path is an error, not a crash
    with pytest.raises(hip.HipError):
        s.save_vtk("fluid", tmp_path / "no_such_dir" / "x.vtk")
    s.destroy()
