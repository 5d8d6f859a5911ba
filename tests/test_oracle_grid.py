"""Pins of the oracle's grid and neighbour search
(grouped_uniform_grid.hpp, morton_order.hpp, neighborhood.cpp): Morton order
against bit interleaving, neighbour sets against O(N^2) brute force with the
same predicate, the permutation's gather semantics (group.test.cpp:110-120),
and closed-form step results."""

import numpy as np
import pytest

from prtcl_amd import scenes


def unsign(v):
    v = int(v)
    smax = 2 ** 31 - 1
    return (v + smax) if v < 0 else v + smax


def morton_key(g):
    """Z-order key: bits interleaved from the most significant level down,
    dimension 0 first within a level."""
    u = [unsign(x) for x in g]
    key = 0
    for b in range(31, -1, -1):
        for d in range(3):
            key = (key << 1) | ((u[d] >> b) & 1)
    return key


def test_morton_compare_matches_interleaving(oracle_lib):
    rng = np.random.default_rng(3)
    for _ in range(3000):
        a = rng.integers(-40, 40, 3)
        b = rng.integers(-40, 40, 3) if rng.random() < 0.7 else a.copy()
        want = (morton_key(a) > morton_key(b)) - (morton_key(a) < morton_key(b))
        assert oracle_lib.morton_compare(a, b) == want, (a, b)


def _brute_equal(o):
    for (g, ng) in [(1, 1), (1, 0), (0, 0), (0, 1)]:
        a = o.neighbors(g, ng)
        b = o.neighbors(g, ng, brute=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (g, ng)


def test_neighbors_config0_brute_force(oracle_lib):
    o = oracle_lib.OracleScene(scenes.config0())
    o.prepare()
    _brute_equal(o)
    off, _ = o.neighbors(1, 1)
    assert np.all(np.diff(off) >= 1)  # self is always a neighbour


def test_neighbors_negative_coordinates_and_ties(oracle_lib):
    # lattice straddling the origin, unjittered: pairs at exactly r = 2 s
    sc = scenes.dambreak(6, 6, 6, (0.2, 0.2, 0.2), 0.025, seed=1, steps=1, name="t",
                         jitter=0.0)
    sc.fluid_x = (sc.fluid_x - np.float32(0.1)).astype(np.float32)
    sc.boundary_x = (sc.boundary_x - np.float32(0.1)).astype(np.float32)
    o = oracle_lib.OracleScene(sc)
    o.prepare()
    _brute_equal(o)


def test_neighbors_random_cloud(oracle_lib):
    rng = np.random.default_rng(9)
    sc = scenes.tiny(4)
    n = 1500
    sc.fluid_x = rng.normal(0, 0.05, (n, 3)).astype(np.float32)
    sc.fluid_v = np.zeros((n, 3), np.float32)
    sc.fluid_m = np.full(n, sc.fluid_m[0], np.float32)
    sc.time_of_birth = np.full(n, -1.0, np.float32)
    o = oracle_lib.OracleScene(sc)
    o.prepare()
    _brute_equal(o)


def test_permute_gathers_every_field(oracle_lib):
    sc = scenes.config0()
    o = oracle_lib.OracleScene(sc)
    o.prepare()
    o.run("density.compute_density")
    before = {f: o.get("fluid", f) for f in ("position", "density", "mass")}
    bpos = o.get("boundary", "position")
    o.permute()
    ids = o.get("fluid", "id")
    assert sorted(ids.tolist()) == list(range(sc.n_fluid))
    assert not np.array_equal(ids, np.arange(sc.n_fluid))
    for f, old in before.items():
        assert np.array_equal(o.get("fluid", f), old[ids])  # new[i] = old[perm[i]]
    bids = o.get("boundary", "id")
    assert np.array_equal(o.get("boundary", "position"), bpos[bids])
    # after the permute each group is stored in cell order: the Morton keys
    # of the particles' cells are non-decreasing in index order
    r = sc.radius
    for grp_name in ("fluid", "boundary"):
        x = o.get(grp_name, "position").astype(np.float64)
        cells = np.floor(x / r).astype(np.int64)
        keys = [morton_key(c) for c in cells]
        assert all(keys[k] <= keys[k + 1] for k in range(len(keys) - 1)), grp_name


def test_gather_semantics_known_answer():
    # group.test.cpp:110-120: values 1..5, perm {3,4,1,2,0} -> 4,5,2,3,1
    vals = np.array([1, 2, 3, 4, 5])
    assert vals[np.array([3, 4, 1, 2, 0])].tolist() == [4, 5, 2, 3, 1]


def test_free_fall_closed_form(oracle_lib):
    """A lone particle far from everything: no neighbours but itself, AA = 0
    so p = 0, and symplectic Euler gives v_n = n dt g,
    x_n = x_0 + dt^2 g n (n + 1) / 2."""
    sc = scenes.tiny(2)
    sc.fluid_x = np.array([[5.0, 5.0, 5.0]], np.float32)
    sc.fluid_v = np.zeros((1, 3), np.float32)
    sc.fluid_m = sc.fluid_m[:1]
    sc.time_of_birth = np.full(1, -1.0, np.float32)
    o = oracle_lib.OracleScene(sc)
    o.prepare()
    n = 25
    for _ in range(n):
        info = o.step()
        assert info["iisph_iterations"] == 2  # min iterations, rel = 0
    dt, g = np.float64(np.float32(sc.time_step)), -9.81
    v = o.get("fluid", "velocity")[0]
    x = o.get("fluid", "position")[0]
    assert v[2] == pytest.approx(n * dt * g, rel=1e-5)
    assert x[2] == pytest.approx(5.0 + dt * dt * g * n * (n + 1) / 2, abs=2e-6)
    assert o.get("fluid", "pressure")[0] == 0.0


def test_lattice_interior_density(oracle_lib):
    """Interior of a perfect lattice with h = spacing: rho = rho0 h^3 sum W
    over the lattice (the Shepard sum), computed here independently."""
    sc = scenes.dambreak(11, 11, 11, (0.5, 0.5, 0.5), 0.025, seed=0, steps=1, name="lat",
                         jitter=0.0)
    o = oracle_lib.OracleScene(sc)
    o.prepare()
    o.run("density.compute_density")
    rho = o.get("fluid", "density")
    x = sc.fluid_x.astype(np.float64)
    centre = np.argmin(((x - x.mean(0)) ** 2).sum(1))
    h = float(np.float32(0.025))
    k = np.arange(-2, 3)
    I, J, K = np.meshgrid(k, k, k, indexing="ij")
    r = np.sqrt(I ** 2 + J ** 2 + K ** 2).ravel()
    q = r  # spacing = h
    wq = np.where(q < 2, (2 - q) ** 3, 0) - np.where(q < 1, 4 * (1 - q) ** 3, 0)
    shepard = (wq / (4 * np.pi)).sum()
    assert rho[centre] == pytest.approx(1000.0 * shepard, rel=1e-5)


def test_empty_fluid_group(oracle_lib):
    sc = scenes.tiny(2)
    sc.fluid_x = np.zeros((0, 3), np.float32)
    sc.fluid_v = np.zeros((0, 3), np.float32)
    sc.fluid_m = np.zeros(0, np.float32)
    sc.time_of_birth = np.zeros(0, np.float32)
    o = oracle_lib.OracleScene(sc)
    o.prepare()
    info = o.step()
    assert info["iisph_iterations"] == 0  # nprde < 1: the loop is skipped
