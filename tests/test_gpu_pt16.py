"""GPU parity of pt16 (pt16.hpp / pt16.prtcl) through the C ABI: the set-up
and both scalar CG solves against the FP64 golden fixture, steps against the
oracle in both step modes, and eviction. The solves run every component
(x, y, z), the vorticity system first: a solve over the last component is the
case a component select compiled without its address would fault on."""

import numpy as np
import pytest

from prtcl_amd import scenes
from test_oracle_pt16 import golden, mats, rel, scene_from_golden

pytestmark = pytest.mark.gpu


def _golden_setup(hip_ctx, g):
    from prtcl_amd import hip
    s = hip.HipScene(hip_ctx, scene_from_golden(g))
    s.prepare()
    s.run("density.compute_density")
    s.upload("fluid", "velocity", g["vstar"].astype(np.float32))
    s.run("pt16.setup")
    return s


def test_pt16_setup_vs_golden(hip_ctx):
    g = golden()
    print("case: golden tiny scene, pt16.setup", flush=True)
    s = _golden_setup(hip_ctx, g)
    assert rel(s.download("fluid", "pt16_same_density"), g["same"]) < 1e-5
    assert rel(s.download("fluid", "pt16_density"), g["prho"]) < 1e-5
    assert rel(s.download("fluid", "vorticity"), g["om"]) < 1e-4
    assert rel(mats(s.download("fluid", "target_velocity_gradient")), g["tvg"]) < 1e-4


def test_pt16_solves_vs_golden(hip_ctx):
    g = golden()
    print("case: golden tiny scene, pt16 solves", flush=True)
    s = _golden_setup(hip_ctx, g)
    s.set_global("pt16_maximum_error", 1e-4)
    s.set_global("pt16_maximum_iterations", 500)
    s.run("pt16.solve_vorticity_diffusion")
    assert rel(s.download("fluid", "pt16_vorticity_diffusion_rhs"), g["orhs"]) < 1e-4
    assert rel(s.download("fluid", "vorticity"), g["om_cg"]) < 2e-3
    assert abs(s.get_global("pt16_iterations") - int(g["om_iterations"][-1])) <= 1
    assert rel(mats(s.download("fluid", "target_velocity_gradient")), g["tvg_spin"]) < 2e-3
    s.run("pt16.solve_velocity_reconstruction")
    assert rel(s.download("fluid", "pt16_velocity_reconstruction_rhs"), g["vrhs"]) < 2e-3
    v = s.download("fluid", "velocity")
    assert rel(v, g["v_cg"]) < 2e-3
    assert rel(v, g["v_solution"]) < 5e-3
    assert abs(s.get_global("pt16_iterations") - int(g["v_iterations"][-1])) <= 1


def pt16_block(nx=8, ny=6, nz=8):
    s = 0.025
    bx = scenes.box_boundary(nx + 7, ny + 7, nz + 5, s)
    fx = scenes.fluid_block((nx, ny, nz), (3, 3, 1), s, 1e-4, scenes.stream_base(23, 0))
    rng = np.random.default_rng(3)
    v = rng.uniform(-0.3, 0.3, fx.shape).astype(np.float32)
    return scenes._finish("pt16_block", s, fx, bx, 4, v=v, implicit_viscosity=3,
                          pt16_vorticity_maximum_error=1e-4, pt16_velocity_maximum_error=1e-4)


@pytest.mark.parametrize("step_mode", [1, 0], ids=["procedures", "fused"])
def test_pt16_steps_vs_oracle(hip_ctx, oracle_lib, step_mode):
    from prtcl_amd import hip
    sc = pt16_block()
    sc.permute_every = 0
    print(f"case: pt16 block, step_mode {step_mode}", flush=True)
    o = oracle_lib.OracleScene(sc, fp64_reductions=True)
    g = hip.HipScene(hip_ctx, sc, step_mode=step_mode)
    o.prepare()
    g.prepare()
    for _ in range(4):
        o.step()
        g.step(1)
    for field, tol in (("position", 1e-5), ("velocity", 5e-3), ("density", 1e-3),
                       ("vorticity", 2e-2)):
        x, y = g.download("fluid", field), o.get("fluid", field)
        assert np.all(np.isfinite(x))
        err = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        assert err < tol, (field, err)


def test_pt16_scene_evict_restore(hip_ctx):
    from prtcl_amd import hip
    print("case: pt16 block, evict and restore", flush=True)
    out = []
    for evict in (False, True):
        s = hip.HipScene(hip_ctx, pt16_block())
        s.prepare()
        s.step(2)
        if evict:
            s.evict()
            assert not s.resident()
        s.step(2)
        out.append({f: s.download("fluid", f) for f in ("position", "velocity", "vorticity",
                                                       "target_velocity_gradient")})
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f


@pytest.mark.parametrize("step_mode", [1, 0], ids=["procedures", "fused"])
def test_pt16_cg_reductions_in_vector_kernels_equal_split(hip_ctx, step_mode):
    """The pt16 solves whose CG vector kernels derive alpha and beta
    themselves equal, bit for bit, the solves with separate one-block alpha
    and beta kernels (schedule bit 4)."""
    from prtcl_amd import hip
    sc = pt16_block()
    sc.permute_every = 0
    print(f"case: pt16 block, step_mode {step_mode}, folded vs split", flush=True)
    out = []
    for mask in (8, 8 | 16):
        g = hip.HipScene(hip_ctx, sc, step_mode=step_mode)
        g.set_schedule(mask)
        g.prepare()
        g.step(3)
        out.append((g.get_global("pt16_iterations"),
                    {f: g.download("fluid", f) for f in ("velocity", "vorticity")}))
        g.destroy()
    (ka, fa), (kb, fb) = out
    assert ka == kb and ka >= 1
    for f in fa:
        np.testing.assert_array_equal(fa[f], fb[f], err_msg=f)
