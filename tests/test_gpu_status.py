"""Status codes of prtcl_hip_step and the gravity procedures on the device.

* A NaN velocity uploaded into one particle makes the next step report
  PRTCL_HIP_NONFINITE (the integration finds the particle's position or speed
  non-finite), in both step modes and for a scene stepped in a batch; the
  other scenes of the batch are unaffected. Injecting a NaN cannot fault: the
  cell box is planned from the finite positions and a NaN position's cell key
  is clamped into the box.
* An unattainable IISPH tolerance with a small iteration cap makes the step
  report PRTCL_HIP_NOT_CONVERGED (a warning; the step is still taken).
* gravity.initialize_axial_acceleration / accumulate_acceleration against the
  oracle (whose axial form is pinned by its closed form on the CPU), and a
  whole step with axial gravity against the oracle's step.
"""

import numpy as np
import pytest

from prtcl_amd import scenes

pytestmark = pytest.mark.gpu


def _scene(**kw):
    sc = scenes.tiny(8, steps=3, **kw)
    sc.permute_every = 0
    return sc


@pytest.mark.parametrize("step_mode", [0, 1], ids=["fused", "procedures"])
def test_nan_velocity_reports_nonfinite(hip_ctx, step_mode):
    from prtcl_amd import hip
    print(f"nonfinite step_mode={step_mode}", flush=True)
    h = hip.HipScene(hip_ctx, _scene(), step_mode=step_mode)
    h.prepare()
    assert h.step(1) in (hip.OK, hip.NOT_CONVERGED)
    v = h.download("fluid", "velocity")
    v[17, 1] = np.nan
    h.upload("fluid", "velocity", v)
    with pytest.raises(hip.HipError) as e:
        h.step(1)
    assert e.value.status == hip.NONFINITE
    x = h.download("fluid", "position")
    assert not np.all(np.isfinite(x))
    h.destroy()


def test_nan_in_one_scene_of_a_batch(hip_ctx):
    from prtcl_amd import hip
    print("nonfinite batch", flush=True)
    a = hip.HipScene(hip_ctx, _scene())
    b = hip.HipScene(hip_ctx, _scene())
    c = hip.HipScene(hip_ctx, _scene())  # the reference for b, stepped alone
    for h in (a, b, c):
        h.prepare()
    v = a.download("fluid", "velocity")
    v[3, 2] = np.inf
    a.upload("fluid", "velocity", v)
    with pytest.raises(hip.HipError) as e:
        hip_ctx.step([a, b], 2)
    assert e.value.status == hip.NONFINITE
    assert c.step(2) in (hip.OK, hip.NOT_CONVERGED)
    for f in ("position", "velocity", "density"):
        assert np.array_equal(b.download("fluid", f), c.download("fluid", f)), f
    # the report is per call: a finite scene alone is OK again
    assert b.step(1) in (hip.OK, hip.NOT_CONVERGED)
    for h in (a, b, c):
        h.destroy()


@pytest.mark.parametrize("step_mode", [0, 1], ids=["fused", "procedures"])
def test_unattainable_tolerance_reports_not_converged(hip_ctx, step_mode):
    from prtcl_amd import hip
    print(f"not converged step_mode={step_mode}", flush=True)
    sc = _scene(iisph_tolerance=1e-12, iisph_max_iterations=3)
    h = hip.HipScene(hip_ctx, sc, step_mode=step_mode)
    h.prepare()
    assert h.step(1) == hip.NOT_CONVERGED
    assert h.step_info()["iisph_iterations"] == 3
    assert np.all(np.isfinite(h.download("fluid", "position")))
    h.destroy()


def test_gravity_procedures_match_oracle(hip_ctx, oracle_lib):
    from prtcl_amd import hip
    print("gravity procedures", flush=True)
    sc = _scene(gravity=(0.5, -1.0, -9.81), gravity_center=(0.05, 0.1, 0.0))
    h = hip.HipScene(hip_ctx, sc)
    o = oracle_lib.OracleScene(sc)
    h.prepare()
    o.prepare()
    for proc in ("gravity.initialize_axial_acceleration", "gravity.accumulate_acceleration",
                 "gravity.initialize_acceleration", "gravity.accumulate_acceleration"):
        h.run(proc)
        o.run(proc)
        a, b = h.download("fluid", "acceleration"), o.get("fluid", "acceleration")
        assert np.array_equal(a, b), (proc, float(np.max(np.abs(a - b))))
    h.destroy()


@pytest.mark.parametrize("step_mode", [0, 1], ids=["fused", "procedures"])
def test_axial_gravity_steps_match_oracle(hip_ctx, oracle_lib, step_mode):
    from prtcl_amd import hip
    print(f"axial steps step_mode={step_mode}", flush=True)
    sc = _scene(axial_gravity=True, gravity_center=(0.1, 0.1, 0.0))
    h = hip.HipScene(hip_ctx, sc, step_mode=step_mode)
    o = oracle_lib.OracleScene(sc, fp64_reductions=True)
    h.prepare()
    o.prepare()
    for _ in range(3):
        h.step(1)
        o.step()
    for f, tol in (("position", 1e-5), ("velocity", 1e-3), ("density", 5e-4)):
        x, y = h.download("fluid", f), o.get("fluid", f)
        err = float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-30))
        print(f"  {f}: normwise relative error {err:.3e}", flush=True)
        assert err < tol, (f, err)
    h.destroy()
