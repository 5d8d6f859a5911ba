"""GPU parity of the gradient correction (correction.hpp:133-217) and of
wkbb18_gc (wkbb18_gc.hpp), through the C ABI: the correction, the corrected
system diagonal and the PCG solution against the FP64 golden fixture
(numpy.linalg.pinv, dense solve), L = I reducing wkbb18_gc to wkbb18 bit for
bit, and whole steps with the correction and wkbb18_gc against the oracle in
both step modes."""

import numpy as np
import pytest

from test_gpu_wkbb18 import viscous_column
from test_oracle_correction import golden, rel, scene_from_golden

pytestmark = pytest.mark.gpu


def _golden_scene(hip_ctx, g):
    from prtcl_amd import hip
    sc = scene_from_golden(g, implicit_viscosity=2, fluid_viscosity=float(g["mu"]),
                           boundary_viscosity=float(g["mu"]), wkbb18_maximum_error=1e-4,
                           wkbb18_maximum_iterations=200)
    s = hip.HipScene(hip_ctx, sc)
    s.prepare()
    s.run("density.compute_density")
    s.upload("fluid", "velocity", g["vstar"].astype(np.float32))
    return s


def _mat(s, name, n):
    return s.download("fluid", name).reshape(n, 3, 3).transpose(0, 2, 1)


def test_correction_and_wkbb18_gc_vs_golden(hip_ctx):
    g = golden()
    n = len(g["xf"])
    print("case: golden tiny scene, correction + wkbb18_gc procedures", flush=True)
    s = _golden_scene(hip_ctx, g)
    s.run("correction.compute_gradient_correction")
    assert rel(_mat(s, "gradient_correction", n), g["L"]) < 1e-4
    s.run("wkbb18_gc.compute_diagonal")
    assert rel(_mat(s, "wkbb18_system_diagonal", n), g["sp"]) < 1e-4
    s.run("wkbb18_gc.accumulate_acceleration")
    x = g["vstar"] + s.download("fluid", "wkbb18_velocity_delta")
    assert rel(x, g["solution"]) < 2e-3
    assert rel(x, g["cg_solution"]) < 1e-4
    assert abs(s.get_global("wkbb18_iterations") - int(g["cg_iterations"])) <= 1


def test_identity_correction_reduces_to_wkbb18(hip_ctx):
    g = golden()
    n = len(g["xf"])
    print("case: golden tiny scene, L = I", flush=True)
    out = []
    for proc in ("wkbb18", "wkbb18_gc"):
        s = _golden_scene(hip_ctx, g)
        s.run("correction.identity_gradient_correction")
        assert np.array_equal(_mat(s, "gradient_correction", n), np.tile(np.eye(3), (n, 1, 1)))
        s.run(f"{proc}.compute_diagonal")
        s.run(f"{proc}.accumulate_acceleration")
        out.append((s.download("fluid", "wkbb18_system_diagonal"),
                    s.download("fluid", "wkbb18_velocity_delta"),
                    s.get_global("wkbb18_iterations")))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("step_mode", [1, 0], ids=["procedures", "fused"])
def test_gc_viscous_steps_vs_oracle(hip_ctx, oracle_lib, step_mode):
    """prtcl-lua.bunny.lua's order: correction after the density, wkbb18_gc
    as the implicit viscosity"""
    from prtcl_amd import hip
    sc = viscous_column()
    sc.implicit_viscosity = 2
    sc.permute_every = 0
    print(f"case: viscous column with wkbb18_gc, step_mode {step_mode}", flush=True)
    o = oracle_lib.OracleScene(sc, fp64_reductions=True)
    g = hip.HipScene(hip_ctx, sc, step_mode=step_mode)
    o.prepare()
    g.prepare()
    wo, wg = [], []
    for _ in range(5):
        info = o.step()
        g.step(1)
        wo.append(info["wkbb18_iterations"])
        wg.append(g.step_info()["wkbb18_iterations"])
    assert max(abs(a - b) for a, b in zip(wo, wg)) <= 1, (wo, wg)
    assert min(wo) >= 1
    n = sc.n_fluid
    Lg = g.download("fluid", "gradient_correction")
    Lo = o.get("fluid", "gradient_correction").reshape(n, 9)
    assert rel(Lg, Lo) < 5e-3
    for field, tol in (("position", 1e-5), ("velocity", 5e-3), ("density", 1e-3)):
        x, y = g.download("fluid", field), o.get("fluid", field)
        assert np.all(np.isfinite(x))
        err = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        assert err < tol, (field, err)
