"""Batched stepping: prtcl_hip_step over several scenes advances them with one
launch per stage for all of them. Each scene's result must be the same alone
and in the batch: the per-scene kernels, their inputs and the order of every
reduction are those of the single-scene step, so the comparison is exact
(bit-identical fields and identical IISPH iteration counts)."""

import numpy as np
import pytest

from prtcl_amd import scenes

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "density", "pressure")


def small_mix(count=5, seed=3):
    # config-3-like scenes (log-uniform N, uneven aspect ratios, random v) at
    # 2k..20k particles so the test runs in seconds
    return [scenes.random_scene(seed, i, 2000, 20000, 4) for i in range(count)]


def run(ctx, scs, k, together, **kw):
    from prtcl_amd import hip
    hs = [hip.HipScene(ctx, sc, **kw) for sc in scs]
    for h in hs:
        h.prepare()
    iters = [[] for _ in hs]
    for _ in range(k):
        if together:
            ctx.step(hs, 1)
        else:
            for h in hs:
                h.step(1)
        for j, h in enumerate(hs):
            iters[j].append(h.step_info()["iisph_iterations"])
    out = [{f: h.download("fluid", f) for f in FIELDS} for h in hs]
    for h in hs:
        h.destroy()
    return out, iters


def test_batch_equals_alone(hip_ctx):
    mix = small_mix()
    # 9 steps: the host-visible permutation (every 8 steps) happens inside
    alone, it_a = run(hip_ctx, mix, 9, together=False)
    batch, it_b = run(hip_ctx, mix, 9, together=True)
    assert it_a == it_b
    for j, (a, b) in enumerate(zip(alone, batch)):
        for f in FIELDS:
            assert np.all(np.isfinite(b[f])), (j, f)
            assert np.array_equal(a[f], b[f]), (j, f, float(np.max(np.abs(a[f] - b[f]))))


def run_info(ctx, scs, k, together, kinds=None, **kw):
    """Like run(), with per-scene keyword sets and the WKBB18 counts too."""
    from prtcl_amd import hip
    kinds = kinds or [dict() for _ in scs]
    hs = [hip.HipScene(ctx, sc, **kw, **kd) for sc, kd in zip(scs, kinds)]
    for h in hs:
        h.prepare()
    info = [[] for _ in hs]
    for _ in range(k):
        if together:
            ctx.step(hs, 1)
        else:
            for h in hs:
                h.step(1)
        for j, h in enumerate(hs):
            d = h.step_info()
            info[j].append((d["iisph_iterations"], d["wkbb18_iterations"]))
    out = [{f: h.download("fluid", f) for f in FIELDS} for h in hs]
    for h in hs:
        h.destroy()
    return out, info


def viscous_mix(seed=7):
    # config-4-like scenes: three with WKBB18 implicit viscosity, one with
    # WKBB18 and the gradient correction, one with pt16, one with explicit
    # viscosity, two with axial gravity (one of them also WKBB18), so every
    # setup form and every step end are in one batch
    mix = small_mix(8, seed=seed)
    mix[7].implicit_viscosity = 3
    mix[7].fluid_viscosity = mix[7].boundary_viscosity = 10.0
    for j, mu in ((1, 5.0), (3, 20.0), (4, 60.0)):
        mix[j].implicit_viscosity = 1
        mix[j].fluid_viscosity = mix[j].boundary_viscosity = mu
    mix[2].implicit_viscosity = 2
    mix[2].fluid_viscosity = mix[2].boundary_viscosity = 10.0
    mix[5].explicit_viscosity = True
    mix[5].fluid_viscosity = mix[5].boundary_viscosity = 0.05
    for j in (4, 6):
        mix[j].axial_gravity = True
        mix[j].gravity_center = tuple(float(c) for c in mix[j].fluid_x.mean(axis=0))
    return mix


@pytest.mark.parametrize("real_type", [0, 1], ids=["fp32", "fp64"])
def test_batch_with_viscous_scenes(hip_ctx, real_type):
    """Viscous scenes are stepped inside the batch: after the batched IISPH
    solve, the WKBB18 scenes' gradient correction (wkbb18_gc scenes),
    diagonal and CG set-up, PCG iterations (each scene's FP64 CG control and
    its own convergence), apply and integrations run over all of them at
    once, and the pt16 scenes' set-up, six scalar PCG solves and
    integrations over theirs. Scenes with explicit viscosity or axial gravity take
    their three setup launches over their own blocks. Every scene matches its
    result alone bit for bit, with the same IISPH and WKBB18 iteration counts
    (the last step crosses a host-visible permutation)."""
    mix = viscous_mix()
    alone, it_a = run_info(hip_ctx, mix, 9, together=False, real_type=real_type)
    batch, it_b = run_info(hip_ctx, mix, 9, together=True, real_type=real_type)
    assert it_a == it_b
    assert all(w > 0 for j in (1, 3, 4) for _, w in it_b[j])
    for j, (a, b) in enumerate(zip(alone, batch)):
        for f in FIELDS:
            assert np.all(np.isfinite(b[f])), (j, f)
            assert np.array_equal(a[f], b[f]), (j, f, float(np.max(np.abs(a[f] - b[f]))))


def test_batch_viscous_split_cg_equals_alone(hip_ctx):
    """The batched PCG with the alpha / beta reductions in one-block kernels
    per scene (schedule bit 4) against each scene alone with them folded into
    the vector kernels: the same fields."""
    mix = viscous_mix(seed=8)
    alone, it_a = run_info(hip_ctx, mix, 3, together=False)
    hip_ctx.set_batch_schedule(8 | 16)
    try:
        batch, it_b = run_info(hip_ctx, mix, 3, together=True)
    finally:
        hip_ctx.set_batch_schedule(-1)
    assert it_a == it_b
    for j, (a, b) in enumerate(zip(alone, batch)):
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), (j, f)


def test_batch_with_unbatchable_scenes(hip_ctx):
    """A list that mixes batched scenes (pt16 and gradient-corrected WKBB18
    among them) with a scene the batch does not take (per-procedure kernels)
    and an FP64 scene (a batch of its own): every scene still matches its
    result alone."""
    mix = small_mix(5, seed=5)
    mix[1].implicit_viscosity = 3
    mix[1].fluid_viscosity = mix[1].boundary_viscosity = 10.0
    mix[2].implicit_viscosity = 2
    mix[2].fluid_viscosity = mix[2].boundary_viscosity = 10.0
    kinds = [dict(), dict(), dict(), dict(real_type=1), dict(step_mode=1)]
    alone, it_a = run_info(hip_ctx, mix, 3, together=False, kinds=kinds)
    batch, it_b = run_info(hip_ctx, mix, 3, together=True, kinds=kinds)
    assert it_a == it_b
    for j, (a, b) in enumerate(zip(alone, batch)):
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), (j, f)


def test_batch_rejects_duplicate(hip_ctx):
    from prtcl_amd import hip
    sc = small_mix(1)[0]
    h = hip.HipScene(hip_ctx, sc)
    h.prepare()
    with pytest.raises(hip.HipError):
        hip_ctx.step([h, h], 1)
    h.destroy()


@pytest.mark.parametrize("real_type", [0, 1], ids=["fp32", "fp64"])
def test_schedules_bit_identical(hip_ctx, real_type):
    """The global-memory and the LDS-staged schedule of each fused pass visit
    the same neighbours in the same order: forced schedules give identical
    fields, alone and batched."""
    from prtcl_amd import hip
    mix = small_mix(3, seed=7)
    out = {}
    for mask in (0, 7, 5):
        hs = [hip.HipScene(hip_ctx, sc, real_type=real_type) for sc in mix]
        for h in hs:
            h.prepare()
            h.set_schedule(mask)
        for _ in range(3):
            for h in hs:
                h.step(1)
        hip_ctx.set_batch_schedule(mask)
        for _ in range(3):
            hip_ctx.step(hs, 1)
        assert all(h.schedule() == mask for h in hs)
        assert hip_ctx.batch_schedule() == mask
        out[mask] = [{f: h.download("fluid", f) for f in FIELDS} for h in hs]
        for h in hs:
            h.destroy()
    hip_ctx.set_batch_schedule(-1)
    for mask in (7, 5):
        for a, b in zip(out[0], out[mask]):
            for f in FIELDS:
                assert np.array_equal(a[f], b[f]), (mask, f)


def test_fp32_and_fp64_batch_graphs_stay_apart(hip_ctx):
    """Batched IISPH iterations replayed from captured graphs, FP32 scenes
    first, then FP64 scenes of the same mix: equal block counts and, for three
    scenes, equal 256-byte-aligned table offsets, so the launch arguments of
    the two batches are the same bytes. Each type must replay its own graphs:
    the FP64 batch equals the FP64 scenes stepped with eager launches (an FP32
    graph replayed on FP64 tables reads them as FP32 entries and dereferences
    garbage pointers)."""
    from prtcl_amd import hip
    mix = small_mix(3, seed=7)

    def go(real_type, mask):
        hs = [hip.HipScene(hip_ctx, sc, real_type=real_type) for sc in mix]
        for h in hs:
            h.prepare()
        hip_ctx.set_batch_schedule(mask)
        for _ in range(3):
            hip_ctx.step(hs, 1)
        out = [{f: h.download("fluid", f) for f in FIELDS} for h in hs]
        for h in hs:
            h.destroy()
        return out

    print("fp32 batch from graphs, then fp64 batch from graphs and eager", flush=True)
    try:
        go(0, 0)
        from_graphs = go(1, 0)
        eager = go(1, 8)
    finally:
        hip_ctx.set_batch_schedule(-1)
    for j, (a, b) in enumerate(zip(eager, from_graphs)):
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), (j, f)


@pytest.mark.parametrize("real_type", [0, 1], ids=["fp32", "fp64"])
def test_first_iteration_fold_bit_identical(hip_ctx, real_type):
    """IISPH iteration 1 folded into the setup pass (every pressure is 0 at
    the start of the solve, so Ap = 0 and the Jacobi update needs no
    neighbours) equals iteration 1 run as its own two passes (schedule bit 6),
    bit for bit in every field and in the iteration counts and density
    errors: batched scenes, a scene with explicit viscosity and one with
    axial gravity (the separate source pass takes the iteration there), and a
    scene whose solve stops after iteration 1 (min 1 iteration, loose
    tolerance)."""
    from prtcl_amd import hip
    mix = small_mix(3, seed=11)
    visc = scenes.random_scene(11, 3, 2000, 6000, 4)
    visc.explicit_viscosity = True
    visc.fluid_viscosity = visc.boundary_viscosity = 0.05
    axial = scenes.random_scene(11, 4, 2000, 6000, 4)
    axial.axial_gravity = True
    axial.gravity_center = tuple(float(c) for c in axial.fluid_x.mean(axis=0))
    one = scenes.random_scene(11, 5, 2000, 6000, 4)
    one.iisph_min_iterations = 1
    one.iisph_tolerance = 0.5
    scs = mix + [visc, axial, one]
    out = {}
    for mask in (8, 8 | 64):
        print(f"fold check real_type={real_type} mask={mask}", flush=True)
        hs = [hip.HipScene(hip_ctx, sc, real_type=real_type) for sc in scs]
        for h in hs:
            h.prepare()
            h.set_schedule(mask)
        hip_ctx.set_batch_schedule(mask)
        info = []
        try:
            for _ in range(4):
                hip_ctx.step(hs, 1)
                info.append([h.step_info() for h in hs])
        finally:
            hip_ctx.set_batch_schedule(-1)
        out[mask] = ([{f: h.download("fluid", f) for f in FIELDS + ("acceleration",)}
                      for h in hs], info)
        for h in hs:
            h.destroy()
    (a, ia), (b, ib) = out[8], out[8 | 64]
    assert ia == ib
    assert all(row[-1]["iisph_iterations"] == 1 for row in ia)  # the scene `one`
    for j, (x, y) in enumerate(zip(a, b)):
        for f in x:
            assert np.array_equal(x[f], y[f]), (j, f)


def test_graph_replay_equals_eager(hip_ctx):
    """The IISPH iteration replayed from captured hipGraphs (default) and
    launched kernel by kernel (schedule bit 3) give identical fields, alone
    and batched; several steps, so the cached graphs are replayed with the
    step-to-step buffer swaps and the per-step tables."""
    from prtcl_amd import hip
    mix = small_mix(3, seed=9)
    out = {}
    for mask in (0, 8):
        hs = [hip.HipScene(hip_ctx, sc) for sc in mix]
        for h in hs:
            h.prepare()
            h.set_schedule(mask)
        for _ in range(4):
            for h in hs:
                h.step(1)
        hip_ctx.set_batch_schedule(mask)
        for _ in range(4):
            hip_ctx.step(hs, 1)
        out[mask] = [{f: h.download("fluid", f) for f in FIELDS} for h in hs]
        for h in hs:
            h.destroy()
    hip_ctx.set_batch_schedule(-1)
    for a, b in zip(out[0], out[8]):
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), f
