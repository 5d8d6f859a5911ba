"""Pins of the oracle's pt16 (pt16.hpp / pt16.prtcl): the set-up fields,
both scalar solves and the spin term against the golden fixture written by
the independent FP64 numpy restatement (tests/golden/make_golden.py: dense
systems, numpy.linalg.solve, the CG of cg_openmp.hpp), and the closed forms
of linear velocity fields on a lattice (the SPH velocity gradient of
v = A x is A inside the fluid, so the vorticity is the curl of A and a
rigid rotation has no strain)."""

import os

import numpy as np
import pytest

from prtcl_amd import scenes

GOLD = os.path.join(os.path.dirname(__file__), "golden", "tiny_pt16.npz")


def golden():
    return dict(np.load(GOLD, allow_pickle=False))


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def scene_from_golden(g, **kw):
    nf = len(g["xf"])
    return scenes.Scene(
        name="golden_pt16", spacing=float(g["h"]), fluid_x=g["xf"].astype(np.float32),
        fluid_v=g["v"].astype(np.float32), fluid_m=g["m"].astype(np.float32),
        time_of_birth=np.full(nf, -1.0, np.float32), boundary_x=g["xb"].astype(np.float32),
        smoothing_scale=float(g["h"]), time_step=float(g["dt"]), rest_density=float(g["rho0"]),
        permute_every=0, implicit_viscosity=3, pt16_strain_rate_viscosity=float(g["xi"]),
        pt16_vorticity_maximum_error=1e-4, pt16_velocity_maximum_error=1e-4, **kw)


def mats(a):
    return np.asarray(a, np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)


@pytest.fixture(scope="module")
def run(oracle_lib):
    g = golden()
    o = oracle_lib.OracleScene(scene_from_golden(g), fp64_reductions=True)
    o.prepare()
    o.run("density.compute_density")
    o.set("fluid", "velocity", g["vstar"].astype(np.float32))
    res = {}
    o.run("pt16.setup")
    for f in ("pt16_same_density", "pt16_density", "vorticity", "target_velocity_gradient"):
        res[f] = o.get("fluid", f)
    o.set_global("pt16_maximum_error", 1e-4)
    o.set_global("pt16_maximum_iterations", 500)
    o.run("pt16.solve_vorticity_diffusion")
    res["orhs"] = o.get("fluid", "pt16_vorticity_diffusion_rhs")
    res["om"] = o.get("fluid", "vorticity")
    res["tvg_spin"] = o.get("fluid", "target_velocity_gradient")
    res["om_it"] = o.get_global("pt16_iterations")
    o.run("pt16.solve_velocity_reconstruction")
    res["vrhs"] = o.get("fluid", "pt16_velocity_reconstruction_rhs")
    res["v"] = o.get("fluid", "velocity")
    res["v_it"] = o.get_global("pt16_iterations")
    return g, res


def test_pt16_setup(run):
    g, r = run
    assert rel(r["pt16_same_density"], g["same"]) < 1e-5
    assert rel(r["pt16_density"], g["prho"]) < 1e-5
    assert rel(r["vorticity"], g["om"]) < 1e-4
    assert rel(mats(r["target_velocity_gradient"]), g["tvg"]) < 1e-4


def test_pt16_vorticity_diffusion(run):
    g, r = run
    assert rel(r["orhs"], g["orhs"]) < 1e-4
    # the system's rows sum to zero (a graph Laplacian): the CG iterate from
    # the guess is the pinned quantity, with the reference's stopping rules
    assert rel(r["om"], g["om_cg"]) < 2e-3
    assert abs(r["om_it"] - int(g["om_iterations"][-1])) <= 1
    assert rel(mats(r["tvg_spin"]), g["tvg_spin"]) < 2e-3


def test_pt16_velocity_reconstruction(run):
    g, r = run
    assert rel(r["vrhs"], g["vrhs"]) < 2e-3
    assert rel(r["v"], g["v_cg"]) < 2e-3
    assert rel(r["v"], g["v_solution"]) < 5e-3
    assert abs(r["v_it"] - int(g["v_iterations"][-1])) <= 1


def _lattice_with_velocity(oracle_lib, A):
    sc = scenes.dambreak(9, 9, 9, (0.5, 0.5, 0.5), 0.025, 0, 1, "lattice", jitter=0.0,
                         implicit_viscosity=3)
    x = sc.fluid_x.astype(np.float64)
    c = x.mean(0)
    sc.fluid_v = ((x - c) @ np.asarray(A, np.float64).T).astype(np.float32)
    o = oracle_lib.OracleScene(sc)
    o.prepare()
    o.run("density.compute_density")
    o.run("pt16.setup")
    lo, hi = x.min(0) + 2.1 * 0.025, x.max(0) - 2.1 * 0.025
    inner = np.all((x > lo) & (x < hi), axis=1)
    assert inner.sum() >= 8
    return o, inner


def test_pt16_vorticity_of_a_linear_field(oracle_lib):
    A = np.array([[0.1, -0.7, 0.3], [0.5, -0.2, 0.4], [-0.6, 0.2, 0.05]])
    o, inner = _lattice_with_velocity(oracle_lib, A)
    om = o.get("fluid", "vorticity")[inner]
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert np.abs(om - curl).max() < 0.05 * np.abs(curl).max()
    # a transposed gradient would give -curl
    assert np.abs(om + curl).max() > 0.5 * np.abs(curl).max()


def test_pt16_rigid_rotation_has_no_strain(oracle_lib):
    W = np.array([0.3, -0.5, 0.8])
    A = np.array([[0, -W[2], W[1]], [W[2], 0, -W[0]], [-W[1], W[0], 0]])  # v = W x (x - c)
    o, inner = _lattice_with_velocity(oracle_lib, A)
    om = o.get("fluid", "vorticity")[inner]
    assert np.abs(om - 2 * W).max() < 0.05 * np.abs(2 * W).max()
    G = mats(o.get("fluid", "target_velocity_gradient"))[inner]
    assert np.abs(G).max() < 0.02 * np.abs(W).max()
