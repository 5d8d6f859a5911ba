"""prtcl_hip_step_noalloc: the batched step of prtcl_hip_step with its tables in
the library's fixed device array instead of memory it allocates. Called
through the C ABI, its results must equal prtcl_hip_step's bit for bit (the
same kernels in the same order), on two calls in a row, the second with a
smaller input; and it must neither allocate device memory for the context
nor accept scenes it cannot step without allocating."""

import ctypes as C

import numpy as np
import pytest

from prtcl_amd import hip, scenes

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "density", "pressure")


def _noalloc(ctx, hs, k):
    L = ctx.L
    L.prtcl_hip_step_noalloc.restype = C.c_int
    L.prtcl_hip_step_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32]
    ids = (C.c_int32 * len(hs))(*[h.id for h in hs])
    return L.prtcl_hip_step_noalloc(ctx.h, ids, len(hs), k)


def _state(h):
    return {f: h.download("fluid", f) for f in FIELDS}


def _pair(ctx, sc):
    a, b = hip.HipScene(ctx, sc), hip.HipScene(ctx, sc)
    a.prepare()
    b.prepare()
    return a, b


def test_noalloc_step_equals_step():
    ctx = hip.Context(0)
    try:
        mix = [scenes.random_scene(13, i, 2000, 12000, 4) for i in range(3)]
        # one scene with WKBB18 implicit viscosity and one with pt16 (their
        # solves run batched too)
        mix[1].implicit_viscosity = 1
        mix[1].fluid_viscosity = mix[1].boundary_viscosity = 10.0
        mix[2].implicit_viscosity = 3
        mix[2].fluid_viscosity = mix[2].boundary_viscosity = 10.0
        pairs = [_pair(ctx, sc) for sc in mix]
        ref = [p[0] for p in pairs]
        new = [p[1] for p in pairs]
        ctx.synchronize()
        used = ctx.bytes_in_use()
        # call 1: three scenes, three steps
        rc_ref = ctx.step(ref, 3)
        rc_new = _noalloc(ctx, new, 3)
        assert rc_new in (hip.OK, hip.NOT_CONVERGED), ctx.L.prtcl_hip_last_error(ctx.h)
        assert rc_new == rc_ref
        assert ctx.bytes_in_use() == used
        for a, b in zip(ref, new):
            assert a.step_info() == b.step_info()
            sa, sb = _state(a), _state(b)
            for f in FIELDS:
                assert np.all(np.isfinite(sb[f])), f
                assert np.array_equal(sa[f], sb[f]), (f, float(np.max(np.abs(sa[f] - sb[f]))))
        # call 2: a smaller input (one scene, the smallest), two steps
        j = int(np.argmin([sc.n_fluid for sc in mix]))
        rc_ref = ctx.step([ref[j]], 2)
        rc_new = _noalloc(ctx, [new[j]], 2)
        assert rc_new == rc_ref
        assert ctx.bytes_in_use() == used
        assert ref[j].step_info() == new[j].step_info()
        sa, sb = _state(ref[j]), _state(new[j])
        for f in FIELDS:
            assert np.array_equal(sa[f], sb[f]), f
        for h in ref + new:
            h.destroy()
    finally:
        ctx.close()


def test_noalloc_step_refuses_what_would_allocate():
    ctx = hip.Context(0)
    try:
        sc = scenes.tiny(6, steps=2)
        fresh = hip.HipScene(ctx, sc)  # no grid yet: a step would allocate it
        assert _noalloc(ctx, [fresh], 1) == hip.INVALID_ARGUMENT
        # the per-procedure step is stepped scene by scene, not batched
        visc = hip.HipScene(ctx, scenes.tiny(6, steps=2), step_mode=1)
        visc.prepare()
        assert _noalloc(ctx, [visc], 1) == hip.INVALID_ARGUMENT
        ok = hip.HipScene(ctx, sc)
        ok.prepare()
        ok.evict()
        assert _noalloc(ctx, [ok], 1) == hip.INVALID_ARGUMENT
        assert _noalloc(ctx, [ok, ok], 1) == hip.INVALID_ARGUMENT
        ok.prepare()  # brings it back
        assert _noalloc(ctx, [ok], 1) in (hip.OK, hip.NOT_CONVERGED)
        for h in (fresh, visc, ok):
            h.destroy()
    finally:
        ctx.close()
