"""Edge cases of the step path on the device:

* free fall of isolated particles with no boundary group (no neighbours but
  themselves, so no pressure): after n steps of symplectic Euler,
  v_n = v_0 + n dt g and x_n = x_0 + n dt v_0 + dt^2 g n (n + 1) / 2 (the
  step order of example.lua:363-462 with t_birth = -1 s and no fade), in both
  step modes, and the same against the oracle;
* a scene whose fluid group is empty steps without error and only advances
  the clock;
* particles moved far outside the cell box planned at the first update
  (their cells are clamped to the box's border) keep exact neighbour sets:
  the device CSR equals the brute-force sets of the oracle.
"""

import numpy as np
import pytest

from prtcl_amd import scenes

pytestmark = pytest.mark.gpu


def _isolated(n_side=2, gap=0.5, seed=5):
    g = np.arange(n_side) * gap
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    x += 1.0
    rng = np.random.default_rng(seed)
    v = rng.uniform(-0.2, 0.2, x.shape).astype(np.float32)
    nf = len(x)
    return scenes.Scene(name="isolated", spacing=0.025, fluid_x=x, fluid_v=v,
                        fluid_m=np.full(nf, np.float32(0.025 ** 3 * 1000.0)),
                        time_of_birth=np.full(nf, -1.0, np.float32),
                        boundary_x=np.zeros((0, 3), np.float32), smoothing_scale=0.025,
                        permute_every=0)  # host order stays the input order


@pytest.mark.parametrize("step_mode", [0, 1], ids=["fused", "procedures"])
def test_free_fall_closed_form(hip_ctx, oracle_lib, step_mode):
    from prtcl_amd import hip
    sc = _isolated()
    n = 20
    print(f"case: free fall of 8 isolated particles, step_mode {step_mode}", flush=True)
    s = hip.HipScene(hip_ctx, sc, step_mode=step_mode)
    s.prepare()
    s.step(n)
    x, v = s.download("fluid", "position"), s.download("fluid", "velocity")
    dt = np.float64(np.float32(sc.time_step))
    g = np.array(sc.gravity, np.float64)
    x0, v0 = sc.fluid_x.astype(np.float64), sc.fluid_v.astype(np.float64)
    v_want = v0 + n * dt * g
    x_want = x0 + n * dt * v0 + dt * dt * g * n * (n + 1) / 2
    # FP32 rounding of 20 position updates near 1.5 m (ulp 1.2e-7); the
    # gravity displacement checked is 5e-4 m
    assert np.max(np.abs(v - v_want)) < 1e-5
    assert np.max(np.abs(x - x_want)) < 5e-6
    assert np.all(s.download("fluid", "pressure") == 0)
    assert s.step_info()["iisph_iterations"] >= 2
    o = oracle_lib.OracleScene(sc)
    o.prepare()
    for _ in range(n):
        o.step()
    assert np.max(np.abs(x - o.get("fluid", "position"))) < 5e-6
    np.testing.assert_array_equal(s.download("fluid", "density"), o.get("fluid", "density"))
    s.destroy()


def test_empty_fluid_group_steps(hip_ctx):
    from prtcl_amd import hip
    base = scenes.tiny(4)
    sc = scenes.Scene(name="no_fluid", spacing=base.spacing,
                      fluid_x=np.zeros((0, 3), np.float32), fluid_v=np.zeros((0, 3), np.float32),
                      fluid_m=np.zeros(0, np.float32), time_of_birth=np.zeros(0, np.float32),
                      boundary_x=base.boundary_x, smoothing_scale=base.smoothing_scale)
    print("case: empty fluid group, 3 steps", flush=True)
    for mode in (0, 1):
        s = hip.HipScene(hip_ctx, sc, step_mode=mode)
        s.prepare()
        s.step(3)
        assert abs(s.get_global("current_time") - 3 * np.float32(sc.time_step)) < 1e-9
        assert s.step_info()["iisph_iterations"] == 0
        vol = s.download("boundary", "volume")
        assert np.all(np.isfinite(vol)) and np.all(vol > 0)
        s.destroy()


def test_neighbors_outside_the_planned_box(hip_ctx, oracle_lib):
    from prtcl_amd import hip
    sc = scenes.tiny(4)
    print("case: particles moved outside the planned cell box, neighbour sets", flush=True)
    s = hip.HipScene(hip_ctx, sc)
    s.prepare()  # plans the cell box around the initial positions
    x = sc.fluid_x.copy()
    r = sc.radius
    x[0] = (5.0, 5.0, 5.0)
    x[1] = (5.0 + 0.6 * r, 5.0, 5.0)      # a neighbour of 0, far outside the box
    x[2] = (-4.0, 3.0, -2.0)              # alone, outside on the other side
    x[3] = (5.0, 5.0 + 3.0 * r, 5.0)      # outside, near 0 and 1 in cells, not in reach
    s.upload("fluid", "position", x)
    s.run("neighborhood.update")
    off, idx = s.neighbors(hip.FLUID, hip.FLUID)
    sc2 = scenes.tiny(4)
    sc2.fluid_x = x
    o = oracle_lib.OracleScene(sc2)
    o.prepare()
    o_off, o_idx = o.neighbors(1, 1, brute=True)
    for i in range(len(x)):
        got = sorted(idx[off[i]:off[i + 1]])
        want = sorted(o_idx[o_off[i]:o_off[i + 1]])
        assert got == want, i
    assert sorted(idx[off[0]:off[1]]) == [0, 1]
    assert sorted(idx[off[2]:off[3]]) == [2]
    s.destroy()
