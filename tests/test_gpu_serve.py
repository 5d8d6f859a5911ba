"""Residency and the served stream: eviction to host memory and back is
bit-identical, and every scene's result is the same alone and in the mix of a
served request stream (continuous batching, admission against a small HBM
budget that forces evictions)."""

import numpy as np
import pytest

from prtcl_amd import scenes

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "density", "pressure")


def mix(count=6, seed=11):
    out = [scenes.random_scene(seed, i, 2000, 12000, 4) for i in range(count)]
    if count > 2:
        out[2].implicit_viscosity = True
        out[2].fluid_viscosity = out[2].boundary_viscosity = 20.0
    return out


def state(h):
    return {f: h.download("fluid", f) for f in FIELDS}


def test_evict_restore_bit_identical(hip_ctx):
    from prtcl_amd import hip
    sc = mix(1)[0]
    a = hip.HipScene(hip_ctx, sc)
    b = hip.HipScene(hip_ctx, sc)
    for h in (a, b):
        h.prepare()
        h.step(3)
    used = hip_ctx.bytes_in_use()
    b.evict()
    assert not b.resident()
    assert hip_ctx.bytes_in_use() < used
    for h in (a, b):
        h.step(3)  # b comes back on first use
    assert b.resident()
    sa, sb = state(a), state(b)
    for f in FIELDS:
        assert np.array_equal(sa[f], sb[f]), f
    a.destroy()
    b.destroy()


def test_served_mix_equals_alone():
    """Requests for six scenes (one viscous, one FP64) arrive interleaved;
    the budget holds about half of them, so the loop evicts idle scenes and
    brings them back. Each scene's final state equals the same scene stepped
    alone by the total number of steps it was asked for."""
    from prtcl_amd import hip
    scs = mix()
    kinds = [dict(), dict(), dict(), dict(real_type=1), dict(), dict()]
    requests = [(0, 3), (1, 2), (2, 2), (0, 1), (3, 3), (4, 2), (5, 4), (1, 3), (2, 1),
                (4, 2), (0, 2), (5, 1)]
    steps = [0] * len(scs)
    for s, k in requests:
        steps[s] += k

    # alone: a context with no budget, one scene at a time
    ref = []
    ctx = hip.Context(0)
    for sc, kw, n in zip(scs, kinds, steps):
        h = hip.HipScene(ctx, sc, **kw)
        h.prepare()
        h.step(n)
        ref.append(state(h))
        h.destroy()
    ctx.close()

    # served: budget of about half the total footprint
    probe = hip.Context(0)
    hs = [hip.HipScene(probe, sc, **kw) for sc, kw in zip(scs, kinds)]
    for h in hs:
        h.prepare()
    total = probe.bytes_in_use()
    probe.close()
    ctx = hip.Context(0, budget=int(total * 0.55))
    live = {}

    def get(i):
        if i not in live:
            desc = hip.desc_for(scs[i], kinds[i].get("real_type", 0))
            need = int(ctx.L.prtcl_hip_scene_footprint(hip.C.byref(desc))) + total // len(scs)
            # admission waits (other requests progress) until idle scenes
            # can be evicted to make room
            while True:
                try:
                    ctx.make_room(need)
                    break
                except hip.HipError:
                    assert ctx.pending()[0] > 0
                    ctx.tick()
            h = hip.HipScene(ctx, scs[i], **kinds[i])
            h.prepare()
            live[i] = h
        return live[i]

    tickets = {}
    evictions_seen = False
    for n, (s, k) in enumerate(requests):
        tickets[ctx.submit(get(s), k)] = s
        if n % 3 == 2:
            ctx.tick()
            evictions_seen |= not all(h.resident() for h in live.values())
    while ctx.pending()[0] > 0:
        ctx.tick()
        evictions_seen |= not all(h.resident() for h in live.values())
    done = ctx.poll()
    assert sorted(done) == sorted(tickets)
    assert ctx.bytes_in_use() <= int(total * 0.55)
    assert evictions_seen
    for h in live.values():
        h.evict()
    for i, h in live.items():
        st = state(h)  # brought back alone
        h.evict()
        for f in FIELDS:
            assert np.array_equal(st[f], ref[i][f]), (i, f)
    ctx.close()


def test_admission_denied_without_idle_scenes():
    from prtcl_amd import hip
    sc = mix(1)[0]
    desc = hip.desc_for(sc)
    ctx = hip.Context(0)
    need = int(ctx.L.prtcl_hip_scene_footprint(hip.C.byref(desc)))
    ctx.close()
    ctx = hip.Context(0, budget=int(need * 1.5))
    a = hip.HipScene(ctx, sc)
    with pytest.raises(hip.HipError) as e:
        hip.HipScene(ctx, sc)
    assert e.value.status == hip.ADMISSION_DENIED
    ctx.submit(a, 2)  # a has pending work: it cannot be evicted
    with pytest.raises(hip.HipError):
        ctx.make_room(need)
    while ctx.pending()[0]:
        ctx.tick()
    ctx.make_room(need)  # now idle: evicted
    assert not a.resident()
    ctx.close()


def test_replay_small_trace():
    """The real-time driver on a short Poisson trace: every request
    completes and latencies are positive."""
    from prtcl_amd import hip, serving
    scs = mix(4, seed=13)
    ctx = hip.Context(0)
    ss = serving.SceneSet(ctx, lambda i: scs[i])
    arr = scenes.poisson_arrivals(4, 200.0, 0.2, len(scs))
    out = serving.replay(ctx, ss, arr, k=2)
    assert out["requests"] == len(arr)
    assert np.all(out["latencies"] > 0)
    ctx.close()
