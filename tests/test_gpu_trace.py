"""Sphere tracing of device scenes through the C ABI (prtcl_hip_scene_trace)
against the restatement of the reference's SphereTracer (oracle/trace_ref.cpp)
on the same positions: a stepped dam break (FP32, and FP64), the boundary
lattice, a step limit that stops rays early, and an empty group. Both sides
run the same double-precision expressions in the same order without
contraction, with the particles of a cell in host order, so the images and
ray parameters are bit-identical."""

import numpy as np
import pytest

from oracle import pyoracle as P
from prtcl_amd import scenes

pytestmark = pytest.mark.gpu


def _compare(s, group, cam_args, threshold=0.5, max_steps=300):
    from prtcl_amd import hip
    x = s.download(group, "position").astype(np.float64)
    h = float(np.float32(s.desc.smoothing_scale)) if s.real_type == 0 else s.desc.smoothing_scale
    want_img, want_par = P.sphere_trace(x, h, P.camera(*cam_args), threshold, max_steps)
    img, par = s.trace(hip.camera(*cam_args), group, threshold, max_steps)
    exact = float(np.mean((img == want_img) & (par == want_par)))
    print(f"  {group}: {np.count_nonzero(img != 0.3)} pixels on the surface, "
          f"{100 * exact:.2f}% bit-identical", flush=True)
    assert np.array_equal(img, want_img)
    assert np.array_equal(par, want_par)
    return img


CAM = ((0.1, -0.45, 0.12), (0, 1, 0.05), (0, 0, 1), 1.0, 64, 48)


@pytest.mark.parametrize("real_type", [0, 1], ids=["fp32", "fp64"])
def test_trace_of_a_dam_break(hip_ctx, real_type):
    from prtcl_amd import hip
    sc = scenes.tiny(6)
    print(f"case: tiny6 dam break, trace, real_type {real_type}", flush=True)
    s = hip.HipScene(hip_ctx, sc, real_type=real_type)
    s.prepare()
    s.step(3)
    img = _compare(s, "fluid", CAM)
    assert np.count_nonzero(img != 0.3) > 100  # the block is in view
    s.destroy()


def test_trace_of_the_boundary_and_step_limit(hip_ctx):
    from prtcl_amd import hip
    sc = scenes.tiny(5)
    print("case: tiny5 boundary lattice, trace; step limit", flush=True)
    s = hip.HipScene(hip_ctx, sc)
    cam = ((0.2, -0.4, 0.25), (0, 1, -0.3), (0, 0, 1), 1.0, 48, 40)
    _compare(s, "boundary", cam)
    img = _compare(s, "fluid", cam, threshold=0.95, max_steps=4)
    assert np.any(img == 0.0)  # rays stopped by the step limit
    s.destroy()


def test_trace_of_an_empty_group(hip_ctx):
    from prtcl_amd import hip
    sc = scenes.Scene(name="no_boundary", spacing=0.025,
                      fluid_x=np.zeros((4, 3), np.float32), fluid_v=np.zeros((4, 3), np.float32),
                      fluid_m=np.ones(4, np.float32), time_of_birth=np.zeros(4, np.float32),
                      boundary_x=np.zeros((0, 3), np.float32), smoothing_scale=0.025)
    print("case: empty boundary group, trace", flush=True)
    s = hip.HipScene(hip_ctx, sc)
    img, par = s.trace(hip.camera((0, -1, 0), (0, 1, 0), (0, 0, 1), 1.0, 16, 12), "boundary")
    assert np.all(img == 0.3)
    with pytest.raises(hip.HipError):
        s.trace(hip.camera((0, -1, 0), (0, 1, 0), (0, 0, 1), 1.0, 0, 12), "fluid")
    s.destroy()
