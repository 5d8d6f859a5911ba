"""Pins of the oracle's IISPH, explicit-viscosity, boundary-volume and WKBB18
functions by what the mathematics fixes (CPU only).

Each check is a property of the reference's formulas, not a second copy of
them, chosen so that a plausible slip (a dropped term, a wrong factor, sign or
index) breaks it:

* iisph (iisph.hpp:197-516): with the pressure field e_i (P at particle i, 0
  elsewhere), one pressure-acceleration pass and one pressure pass give
  Ap_i = P AA_i, because AA_i is the diagonal of the linear map p -> Ap. This
  ties the helper c, AA and the boundary factors 0.7 * 2 of c and of the
  pressure acceleration together. Ap is linear in p. Without a boundary, the
  pressure accelerations conserve momentum (sum m_i a_i = 0), and for a
  uniform velocity the source term is s = rho0 - rho exactly.
* viscosity.accumulate_acceleration (viscosity.hpp:171-240), without a
  boundary: a uniform translation gives a = 0 exactly, a rigid rotation gives
  a = 0 up to rounding, and random velocities conserve momentum.
* aiast12.compute_volume (aiast12.hpp:125-175): on an infinite plane lattice
  of spacing h the volume is 1 / (W(0) + 4 W(h) + 4 W(sqrt(2) h)) =
  h^3 / (alpha (8 + 4 (2 - sqrt 2)^3)).
* wkbb18 (wkbb18.hpp:194-425), without a boundary: the system maps a uniform
  field and a rigid rotation to themselves (so the solve returns the velocity
  unchanged: velocity delta 0), and each diagonal block D_i is symmetric (so
  is (I - dt D_i)^-1).

The oracle source carries numbered hooks for planted mistakes
(-DORACLE_MUTANT=k); every check is shown to fail on the builds whose mistake
it must catch.
"""

import concurrent.futures as cf
import math
import os
import subprocess

import numpy as np
import pytest

from prtcl_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# planted mistakes (hook numbers in oracle/cpu_ref.cpp)
MUTANTS = {
    1: "c: boundary factor 0.7 instead of 0.7 * 2",
    2: "AA: the -m_i/rho_i^2 |gradW|^2 term added instead of subtracted",
    3: "pressure acceleration: boundary term p_i instead of 2 p_i",
    4: "Ap: boundary term with the wrong sign",
    5: "viscosity: m_j / rho_i instead of m_j / rho_j",
    6: "viscosity: v_i instead of v_i - v_j",
    7: "pressure acceleration: p_j / rho_i^2 instead of p_j / rho_j^2",
    8: "source term: v_i instead of v_i - v_j",
    9: "volume: self term dropped",
    10: "wkbb18 system: w_i instead of w_i - w_j",
    11: "wkbb18 diagonal: x_i instead of x_ij in the outer product",
}


@pytest.fixture(scope="session")
def mutant_libs(tmp_path_factory):
    out = tmp_path_factory.mktemp("oracle_mutants")

    def one(k):
        path = str(out / f"libprtcl_oracle_m{k}.so")
        subprocess.check_call(["g++", "-std=c++17", "-O1", "-fopenmp", "-fPIC",
                               "-ffp-contract=off", f"-DORACLE_MUTANT={k}", "-shared", "-o",
                               path, os.path.join(ORACLE, "cpu_ref.cpp"),
                               os.path.join(ORACLE, "trace_ref.cpp")])
        return k, path

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return dict(ex.map(one, MUTANTS))


# ---------------------------------------------------------------- scenes
def _walled(seed=21):
    """A jittered 8^3 dam break in a closed tank (boundary neighbours)."""
    sc = scenes.tiny(8, seed=seed, steps=1)
    sc.permute_every = 0
    return sc


def _cloud(seed=4, n=(7, 6, 6), spacing=0.025):
    """A jittered fluid block with no boundary group."""
    x = scenes.fluid_block(n, (0, 0, 0), spacing, 0.2 * spacing, scenes.stream_base(seed, 0))
    nf = len(x)
    m = scenes._mass(spacing, 1000.0)
    rng = np.random.default_rng(seed)
    return scenes.Scene(name="cloud", spacing=spacing, fluid_x=x,
                        fluid_v=np.zeros((nf, 3), np.float32),
                        fluid_m=(m * rng.uniform(0.8, 1.2, nf)).astype(np.float32),
                        time_of_birth=np.full(nf, -1.0, np.float32),
                        boundary_x=np.zeros((0, 3), np.float32), smoothing_scale=spacing,
                        permute_every=0, fluid_viscosity=50.0, boundary_viscosity=50.0,
                        implicit_viscosity=1)


def _oracle(pyo, sc, path):
    o = pyo.OracleScene(sc, lib_path=path)
    o.prepare()
    o.run("density.compute_density")
    return o


def _rotation(x, omega=(0.3, -1.1, 0.7)):
    c = x.mean(axis=0)
    return np.cross(np.asarray(omega, np.float64), x - c).astype(np.float32)


# ---------------------------------------------------------------- checks
def check_iisph_diagonal(pyo, path):
    sc = _walled()
    o = _oracle(pyo, sc, path)
    rng = np.random.default_rng(1)
    o.set("fluid", "velocity", rng.uniform(-0.3, 0.3, (sc.n_fluid, 3)))
    o.run("iisph.setup")
    AA = o.get("fluid", "iisph_diagonal_element").astype(np.float64)
    # particles with the most boundary neighbours, and a few interior ones
    off, _ = o.neighbors(1, 0)
    nbnd = np.diff(off)
    pick = list(np.argsort(-nbnd)[:12]) + list(np.where(nbnd == 0)[0][:6])
    P = 1000.0
    for i in pick:
        p = np.zeros(sc.n_fluid, np.float32)
        p[i] = P
        o.set("fluid", "pressure", p)
        o.run("iisph.iteration_pressure_acceleration")
        o.run("iisph.iteration_pressure")
        Ap = float(o.get("fluid", "iisph_right_hand_side")[i])
        assert abs(Ap - P * AA[i]) <= 1e-4 * abs(P * AA[i]), (i, int(nbnd[i]), Ap, P * AA[i])


def check_iisph_linear(pyo, path):
    sc = _walled(seed=22)
    o = _oracle(pyo, sc, path)
    o.run("iisph.setup")
    rng = np.random.default_rng(2)

    def ap(p):
        o.set("fluid", "pressure", p)
        o.run("iisph.iteration_pressure_acceleration")
        o.run("iisph.iteration_pressure")
        return o.get("fluid", "iisph_right_hand_side").astype(np.float64)

    p = rng.uniform(0, 500, sc.n_fluid).astype(np.float32)
    q = rng.uniform(0, 500, sc.n_fluid).astype(np.float32)
    lhs = ap((2 * p + 3 * q).astype(np.float32))
    rhs = 2 * ap(p) + 3 * ap(q)
    assert np.linalg.norm(lhs - rhs) <= 1e-5 * np.linalg.norm(rhs)


def check_pressure_momentum(pyo, path):
    sc = _cloud()
    o = _oracle(pyo, sc, path)
    rng = np.random.default_rng(3)
    rho = (1000.0 * rng.uniform(0.9, 1.1, sc.n_fluid)).astype(np.float32)
    o.set("fluid", "density", rho)
    o.set("fluid", "pressure", rng.uniform(0, 800, sc.n_fluid).astype(np.float32))
    o.run("iisph.iteration_pressure_acceleration")
    a = o.get("fluid", "acceleration").astype(np.float64)
    m = sc.fluid_m.astype(np.float64)[:, None]
    total = np.abs((m * a).sum(axis=0))
    scale = np.abs(m * a).sum(axis=0)
    assert np.all(total <= 1e-5 * scale), (total, scale)


def check_source_uniform_velocity(pyo, path):
    sc = _cloud(seed=5)
    o = _oracle(pyo, sc, path)
    o.set("fluid", "velocity", np.tile(np.float32([0.4, -0.25, 0.1]), (sc.n_fluid, 1)))
    o.run("iisph.setup")
    rho = o.get("fluid", "density")
    s = o.get("fluid", "iisph_source_term")
    assert np.array_equal(s, np.float32(1000.0) - rho)


def check_viscosity_translation_rotation(pyo, path):
    sc = _cloud(seed=6)
    o = _oracle(pyo, sc, path)
    zero = np.zeros((sc.n_fluid, 3), np.float32)
    o.set("fluid", "velocity", np.tile(np.float32([0.7, 0.2, -0.4]), (sc.n_fluid, 1)))
    o.set("fluid", "acceleration", zero)
    o.run("viscosity.accumulate_acceleration")
    assert np.all(o.get("fluid", "acceleration") == 0.0)
    rng = np.random.default_rng(6)
    o.set("fluid", "velocity", rng.uniform(-0.5, 0.5, (sc.n_fluid, 3)).astype(np.float32))
    o.set("fluid", "acceleration", zero)
    o.run("viscosity.accumulate_acceleration")
    scale = float(np.abs(o.get("fluid", "acceleration")).max())
    o.set("fluid", "velocity", _rotation(sc.fluid_x.astype(np.float64)))
    o.set("fluid", "acceleration", zero)
    o.run("viscosity.accumulate_acceleration")
    assert float(np.abs(o.get("fluid", "acceleration")).max()) <= 1e-4 * scale


def check_viscosity_momentum(pyo, path):
    sc = _cloud(seed=7)
    o = _oracle(pyo, sc, path)
    rng = np.random.default_rng(7)
    o.set("fluid", "density", (1000.0 * rng.uniform(0.9, 1.1, sc.n_fluid)).astype(np.float32))
    o.set("fluid", "velocity", rng.uniform(-0.5, 0.5, (sc.n_fluid, 3)).astype(np.float32))
    o.set("fluid", "acceleration", np.zeros((sc.n_fluid, 3), np.float32))
    o.run("viscosity.accumulate_acceleration")
    a = o.get("fluid", "acceleration").astype(np.float64)
    m = sc.fluid_m.astype(np.float64)[:, None]
    total = np.abs((m * a).sum(axis=0))
    scale = np.abs(m * a).sum(axis=0)
    assert np.all(total <= 1e-5 * scale), (total, scale)


def check_volume_plane(pyo, path):
    h = 0.025
    n = 21
    g = np.arange(n) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    bx = np.stack([X.ravel(), Y.ravel(), np.zeros(n * n)], 1).astype(np.float32)
    fx = np.float32([[0.5 * n * h, 0.5 * n * h, 1.0]])
    sc = scenes.Scene(name="plane", spacing=h, fluid_x=fx, fluid_v=np.zeros_like(fx),
                      fluid_m=np.float32([h ** 3 * 1000.0]),
                      time_of_birth=np.float32([-1.0]), boundary_x=bx, smoothing_scale=h)
    o = pyo.OracleScene(sc, lib_path=path)
    o.prepare()
    V = o.get("boundary", "volume").reshape(n, n)
    alpha = pyo.lib(path).oracle_spline_alpha3()
    closed = h ** 3 / (alpha * (8.0 + 4.0 * (2.0 - math.sqrt(2.0)) ** 3))
    inner = V[3:-3, 3:-3]
    np.testing.assert_allclose(inner, closed, rtol=1e-5)


def check_wkbb18_rigid(pyo, path):
    sc = _cloud(seed=8)
    o = _oracle(pyo, sc, path)
    for v in (np.tile(np.float32([0.3, -0.6, 0.2]), (sc.n_fluid, 1)),
              _rotation(sc.fluid_x.astype(np.float64))):
        o.set("fluid", "velocity", v)
        o.run("wkbb18.compute_diagonal")
        o.run("wkbb18.accumulate_acceleration")
        vd = o.get("fluid", "wkbb18_velocity_delta")
        assert float(np.abs(vd).max()) <= 1e-6 * float(np.abs(v).max()), float(np.abs(vd).max())


def check_wkbb18_diagonal_symmetric(pyo, path):
    sc = _cloud(seed=9)
    o = _oracle(pyo, sc, path)
    o.run("wkbb18.compute_diagonal")
    sp = o.get("fluid", "wkbb18_system_diagonal").reshape(-1, 3, 3).astype(np.float64)
    asym = np.abs(sp - sp.transpose(0, 2, 1)).max(axis=(1, 2))
    size = np.abs(sp - np.eye(3)).max(axis=(1, 2))
    assert np.all(asym <= 1e-4 * size + 1e-7), float((asym / (size + 1e-30)).max())


CHECKS = {
    "iisph_diagonal": (check_iisph_diagonal, [1, 2, 3, 4]),
    "iisph_linear": (check_iisph_linear, []),
    "pressure_momentum": (check_pressure_momentum, [7]),
    "source_uniform_velocity": (check_source_uniform_velocity, [8]),
    "viscosity_translation_rotation": (check_viscosity_translation_rotation, [6]),
    "viscosity_momentum": (check_viscosity_momentum, [5]),
    "volume_plane": (check_volume_plane, [9]),
    "wkbb18_rigid": (check_wkbb18_rigid, [10]),
    "wkbb18_diagonal_symmetric": (check_wkbb18_diagonal_symmetric, [11]),
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_invariant_holds(oracle_lib, name):
    CHECKS[name][0](oracle_lib, None)


@pytest.mark.parametrize("name,k", [(n, k) for n in sorted(CHECKS) for k in CHECKS[n][1]])
def test_invariant_catches_planted_mistake(oracle_lib, mutant_libs, name, k):
    with pytest.raises(AssertionError):
        CHECKS[name][0](oracle_lib, mutant_libs[k])


def test_every_planted_mistake_is_caught():
    caught = {k for _, ks in CHECKS.values() for k in ks}
    assert caught == set(MUTANTS)
