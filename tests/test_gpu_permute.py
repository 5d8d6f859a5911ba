"""Host-visible permutes on the device (Neighborhood::Permute,
neighborhood.cpp:106-154; VaryingManager::PermuteItems, varying_manager.hpp:
86-102; gather semantics pinned by group.test.cpp:102-121).

* The procedure: after neighborhood.permute every varying field of both
  groups equals old[perm] bit for bit, with perm from
  prtcl_hip_last_permutation; the new host order is the grid's cell order
  (the library's dense box of cells, x fastest; the reference sorts cells in
  Morton order instead, which no caller can observe beyond "grid order");
  and the neighbour sets in the new index space equal the oracle's brute
  force on the permuted positions.
* Permutes inside steps (permute_every = 8 on both sides): composing the
  recorded permutations maps the device's host order back to the original
  indices, and the state after 10 steps matches the oracle's, which permutes
  in Morton order, in original-index space.
"""

import numpy as np
import pytest

from prtcl_amd import scenes

pytestmark = pytest.mark.gpu


def _fields(h, group):
    from prtcl_amd import hip
    out = {}
    for name, fid in hip.FIELDS.items():
        try:
            out[name] = h.download(group, name)
        except hip.HipError:
            pass  # a field this group does not hold
    return out


def test_permute_procedure_gathers_every_field(hip_ctx, oracle_lib):
    from prtcl_amd import hip
    sc = scenes.tiny(10, steps=3, implicit_viscosity=1, fluid_viscosity=5.0,
                     boundary_viscosity=5.0)
    sc.permute_every = 0
    h = hip.HipScene(hip_ctx, sc)
    h.prepare()
    h.step(3)
    h.run("neighborhood.update")  # the grid of the current positions (example.lua:380-383)
    before = {g: _fields(h, g) for g in ("fluid", "boundary")}
    assert len(before["fluid"]) >= 12 and len(before["boundary"]) >= 2
    h.run("neighborhood.permute")
    for g in ("fluid", "boundary"):
        perm = h.last_permutation(g).astype(np.int64)
        n = h.nf if g == "fluid" else h.nb
        assert np.array_equal(np.sort(perm), np.arange(n)), g
        after = _fields(h, g)
        for name, old in before[g].items():
            assert np.array_equal(after[name], old[perm]), (g, name)
    # the host order is the cell order: cells of consecutive particles never
    # go back in the box order (z, then y, then x)
    r = sc.radius
    for g in ("fluid", "boundary"):
        x = h.download(g, "position").astype(np.float64)
        c = np.floor(x / r).astype(np.int64)
        c -= c.min(axis=0)
        span = c.max(axis=0) + 1
        key = (c[:, 2] * span[1] + c[:, 1]) * span[0] + c[:, 0]
        assert np.all(np.diff(key) >= 0), g
    # neighbour sets in the new index space
    o = oracle_lib.OracleScene(sc, permute_every=0)
    o.set("fluid", "position", h.download("fluid", "position"))
    o.set("boundary", "position", h.download("boundary", "position"))
    o.update()
    for qg, ng in [(1, 1), (1, 0), (0, 1), (0, 0)]:
        a, b = h.neighbors(qg, ng), o.neighbors(qg, ng, brute=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (qg, ng)
    # a second permute right away changes nothing
    h.run("neighborhood.permute")
    assert np.array_equal(h.last_permutation("fluid"), np.arange(h.nf))
    h.destroy()


@pytest.mark.parametrize("step_mode", [0, 1], ids=["fused", "procedures"])
def test_steps_across_permutes_match_oracle(hip_ctx, oracle_lib, step_mode):
    from prtcl_amd import hip
    sc = scenes.config0()
    sc.permute_every = 8
    print(f"permute parity step_mode={step_mode}", flush=True)
    g = hip.HipScene(hip_ctx, sc, step_mode=step_mode)
    o = oracle_lib.OracleScene(sc, fp64_reductions=True, permute_every=8)
    g.prepare()
    o.prepare()
    orig = np.arange(sc.n_fluid)  # original index of each device host index
    permutes = 0
    for s in range(1, 11):
        o.step()
        g.step(1)
        if s % 8 == 0:
            perm = g.last_permutation("fluid").astype(np.int64)
            assert not np.array_equal(perm, np.arange(sc.n_fluid))
            orig = orig[perm]
            permutes += 1
    assert permutes == 1
    # the device's host order after the permute is its cell order, the
    # oracle's is its Morton order: compare in original-index space
    ids = o.get("fluid", "id")
    for f, tol in (("position", 1e-5), ("velocity", 1e-3), ("density", 5e-4),
                   ("pressure", 2e-2)):
        a = np.empty_like(g.download("fluid", f))
        a[orig] = g.download("fluid", f)
        b = np.empty_like(o.get("fluid", f))
        b[ids] = o.get("fluid", f)
        err = float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))
        print(f"  {f}: normwise relative error {err:.2e}", flush=True)
        assert err < tol, (f, err)
    g.destroy()
