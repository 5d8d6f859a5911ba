"""GPU parity of the WKBB18 implicit-viscosity path (wkbb18.hpp:194-425,
cg_openmp.hpp:32-236): the system diagonal inverse and the PCG solution
against the FP64 golden fixture, and whole viscous steps against the
oracle."""

import numpy as np
import pytest

from prtcl_amd import scenes
from test_oracle_golden import golden, rel, scene_from_golden

pytestmark = pytest.mark.gpu


def test_wkbb18_procedures_vs_golden(hip_ctx):
    from prtcl_amd import hip
    g = golden()
    sc = scene_from_golden(g, implicit_viscosity=True, fluid_viscosity=float(g["mu"]),
                           boundary_viscosity=float(g["mu"]), wkbb18_maximum_error=1e-4,
                           wkbb18_maximum_iterations=200)
    s = hip.HipScene(hip_ctx, sc)
    s.prepare()
    s.run("density.compute_density")
    s.upload("fluid", "velocity", g["vstar"].astype(np.float32))
    s.run("wkbb18.compute_diagonal")
    n = len(g["xf"])
    sp = s.download("fluid", "wkbb18_system_diagonal").reshape(n, 3, 3).transpose(0, 2, 1)
    assert rel(sp, g["wkbb18_sp"]) < 1e-4
    s.run("wkbb18.accumulate_acceleration")
    x = g["vstar"] + s.download("fluid", "wkbb18_velocity_delta")
    assert rel(x, g["wkbb18_solution"]) < 2e-3
    assert rel(x, g["wkbb18_cg_solution"]) < 1e-4
    assert abs(s.get_global("wkbb18_iterations") - int(g["wkbb18_cg_iterations"])) <= 1
    a = s.download("fluid", "acceleration")
    dt = np.float32(sc.time_step)
    assert rel(a, (x - g["vstar"]) / float(dt)) < 1e-3


def viscous_column(mu=50.0, nx=10, nz=20):
    s = 0.01
    bx = scenes.box_boundary(nx + 11, nx + 11, nz + 6, s)
    fx = scenes.fluid_block((nx, nx, nz), (5, 5, 1), s, 1e-4, scenes.stream_base(22, 0))
    return scenes._finish("column", s, fx, bx, 5, implicit_viscosity=True,
                          fluid_viscosity=mu, boundary_viscosity=mu,
                          wkbb18_maximum_error=1e-4, wkbb18_maximum_iterations=200)


@pytest.mark.parametrize("step_mode", [1, 0], ids=["procedures", "fused"])
def test_viscous_steps_vs_oracle(hip_ctx, oracle_lib, step_mode):
    from prtcl_amd import hip
    sc = viscous_column()
    sc.permute_every = 0
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
    for field, tol in (("position", 1e-5), ("velocity", 5e-3), ("density", 1e-3)):
        x, y = g.download("fluid", field), o.get("fluid", field)
        assert np.all(np.isfinite(x))
        err = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        assert err < tol, (field, err)


@pytest.mark.parametrize("implicit_viscosity", [1, 2], ids=["wkbb18", "wkbb18_gc"])
def test_cg_reductions_in_vector_kernels_equal_split(hip_ctx, implicit_viscosity):
    """The CG iterations whose vector kernels derive alpha and beta themselves
    (every block sums all partials, in the order of the one-block kernels)
    equal, bit for bit, the iterations with separate one-block alpha and beta
    kernels (schedule bit 4), and the one list walk for the diagonal and the
    CG set-up equals the two walks (schedule bit 5), over steps of the viscous
    column."""
    from prtcl_amd import hip
    sc = viscous_column()
    sc.implicit_viscosity = implicit_viscosity
    sc.permute_every = 0
    print(f"case: viscous column, implicit_viscosity {implicit_viscosity}, folded vs split",
          flush=True)
    out = []
    masks = (8, 8 | 16, 8 | 32, 8 | 16 | 32)
    for mask in masks:
        g = hip.HipScene(hip_ctx, sc)
        g.set_schedule(mask)
        g.prepare()
        iters = []
        for _ in range(4):
            g.step(1)
            iters.append(g.step_info()["wkbb18_iterations"])
        assert g.schedule() == mask
        out.append((iters, {f: g.download("fluid", f) for f in ("position", "velocity")}))
        g.destroy()
    ia, fa = out[0]
    assert min(ia) >= 1
    for mask, (ib, fb) in zip(masks[1:], out[1:]):
        assert ia == ib, mask
        for f in fa:
            np.testing.assert_array_equal(fa[f], fb[f], err_msg=f"{f}, mask {mask}")
