"""Pins of the oracle's cubic spline kernel (cubic_spline_kernel.hpp:43-102,
kernel_facade.hpp:204-229) against closed forms and invariants that a
dropped term, wrong power, wrong sign or wrong normalisation would break."""

import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


H = np.float32(0.025)
ALPHA = 1.0 / (4.0 * math.pi)


def test_w_known_values(O):
    # q = 0: alpha/h^3 * (2^3 - 4) ; q = 1: alpha/h^3 * 1 ; q >= 2: 0
    h = float(H)
    w0 = O.kernel_w([0, 0, 0], H)
    assert w0 == pytest.approx(ALPHA / h ** 3 * 4.0, rel=1e-6)
    w1 = O.kernel_w([h, 0, 0], H)
    assert w1 == pytest.approx(ALPHA / h ** 3 * 1.0, rel=1e-6)
    assert O.kernel_w([2 * h, 0, 0], H) == 0.0
    assert O.kernel_w([0, 3 * h, 0], H) == 0.0
    # q = 1.5: (0.5)^3 = 0.125
    assert O.kernel_w([0, 0, 1.5 * h], H) == pytest.approx(ALPHA / h ** 3 * 0.125, rel=1e-5)
    # q = 0.5: 1.5^3 - 4 * 0.5^3 = 3.375 - 0.5
    assert O.kernel_w([0.5 * h, 0, 0], H) == pytest.approx(ALPHA / h ** 3 * 2.875, rel=1e-5)


def test_w_normalised(O):
    """integral of W over R^3 = 1 (midpoint quadrature on a fine lattice)"""
    h = float(H)
    n = 48
    step = 4 * h / n
    g = (np.arange(n) + 0.5) * step - 2 * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    # the kernel is radial: evaluate through the oracle on a radial table
    r = np.sqrt((pts.astype(np.float64) ** 2).sum(1))
    table_r = np.linspace(0, 2 * h, 2001)
    table_w = np.array([O.kernel_w([x, 0, 0], H) for x in table_r])
    w = np.interp(r, table_r, table_w, right=0.0)
    total = w.sum() * step ** 3
    assert total == pytest.approx(1.0, abs=3e-3)


def test_gradient_matches_derivative_and_is_antisymmetric(O):
    h = float(H)
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.uniform(-1.8 * h, 1.8 * h, 3).astype(np.float32)
        r = float(np.linalg.norm(v.astype(np.float64)))
        if r > 1.95 * h or r < 0.05 * h:
            continue
        g = O.kernel_grad(v, H)
        gm = O.kernel_grad(-v, H)
        np.testing.assert_allclose(gm, -g, rtol=1e-6, atol=1e-3)
        # central finite difference of W along each axis
        eps = 1e-3 * h
        fd = []
        for k in range(3):
            e = np.zeros(3)
            e[k] = eps
            fd.append((O.kernel_w((v + e).astype(np.float32), H) -
                       O.kernel_w((v - e).astype(np.float32), H)) / (2 * eps))
        np.testing.assert_allclose(g, fd, rtol=2e-2, atol=2e-2 * np.abs(g).max())
        # gradient points towards the centre (W decreasing in r)
        assert float(np.dot(g, v)) < 0


def test_gradient_cutoff(O):
    # |x| below sqrt(FLT_EPSILON) = 3.4527e-4 (absolute) gives a zero gradient
    cut = math.sqrt(np.finfo(np.float32).eps)
    assert np.all(O.kernel_grad([0.9 * cut, 0, 0], H) == 0)
    assert np.any(O.kernel_grad([1.1 * cut, 0, 0], H) != 0)
    assert np.all(O.kernel_grad([0, 0, 0], H) == 0)


def test_invert3_matches_numpy(O):
    rng = np.random.default_rng(1)
    for _ in range(20):
        m = (np.eye(3) + 0.3 * rng.standard_normal((3, 3))).astype(np.float32)
        inv = O.invert3(m.T.ravel())          # column-major storage
        ref = np.linalg.inv(m.astype(np.float64))
        np.testing.assert_allclose(inv.reshape(3, 3).T, ref, rtol=1e-4, atol=1e-5)
