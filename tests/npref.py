"""Independent FP64 numpy restatement of the hot path's mathematics, written
from the scheme definitions (share/schemes/{density,iisph,wkbb18,aiast12,
viscosity}.prtcl) rather than from the generated C++: brute-force neighbour
sets, pair arrays and vectorised sums. Used to pin the oracle and to write
the golden fixtures under tests/golden/. Only for small inputs (O(N^2))."""

from __future__ import annotations

import math

import numpy as np

ALPHA = 1.0 / (4.0 * math.pi)


def w(r, h):
    q = r / h
    out = np.where(q < 2, (2 - q) ** 3, 0.0) - np.where(q < 1, 4 * (1 - q) ** 3, 0.0)
    return ALPHA / h ** 3 * out


def dw(r, h):
    q = r / h
    out = np.where(q < 1, 12 * (1 - q) ** 2, 0.0) - np.where(q < 2, 3 * (2 - q) ** 2, 0.0)
    return ALPHA / h ** 4 * out


def grad(d, h):
    """gradient of W at x_ij = d (pairs x 3); zero below sqrt(FLT_EPSILON)"""
    r = np.linalg.norm(d, axis=1)
    f = np.where(r < math.sqrt(np.finfo(np.float32).eps), 0.0,
                 dw(r, h) / np.where(r > 0, r, 1.0))
    return f[:, None] * d


def pairs(xa, xb, radius, same=False):
    """(i, j) with |xa_i - xb_j|^2 < radius^2 using the FP32 predicate of the
    reference (float differences, float squared norm, promoted to double)."""
    xa32, xb32 = xa.astype(np.float32), xb.astype(np.float32)
    d = xa32[:, None, :] - xb32[None, :, :]
    r2 = (d[..., 0] * d[..., 0] + (d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]))
    ii, jj = np.nonzero(r2.astype(np.float64) < radius * radius)
    return ii, jj


class Tiny:
    """Precomputed neighbour pairs and kernel values of one small scene."""

    def __init__(self, xf, xb, h, radius, m, rho0, dt):
        self.xf, self.xb = np.asarray(xf, np.float64), np.asarray(xb, np.float64)
        self.h, self.radius, self.rho0, self.dt = float(h), float(radius), float(rho0), float(dt)
        self.m = np.asarray(m, np.float64)
        nf = len(self.xf)
        self.nf = nf
        self.fi, self.fj = pairs(xf, xf, radius)
        self.bi, self.bj = pairs(xf, xb, radius)
        self.ci, self.cj = pairs(xb, xb, radius)
        self.df = self.xf[self.fi] - self.xf[self.fj]
        self.db = self.xf[self.bi] - self.xb[self.bj]
        self.gf = grad(self.df, self.h)
        self.gb = grad(self.db, self.h)

    def volume(self):
        d = self.xb[self.ci] - self.xb[self.cj]
        s = np.zeros(len(self.xb))
        np.add.at(s, self.ci, w(np.linalg.norm(d, axis=1), self.h))
        return 1.0 / s

    def density(self, V):
        rho = np.zeros(self.nf)
        np.add.at(rho, self.fi, self.m[self.fj] * w(np.linalg.norm(self.df, axis=1), self.h))
        np.add.at(rho, self.bi, V[self.bj] * self.rho0 * w(np.linalg.norm(self.db, axis=1),
                                                           self.h))
        return rho

    def _sum3(self, idx, vals):
        out = np.zeros((self.nf, 3))
        for k in range(3):
            np.add.at(out[:, k], idx, vals[:, k])
        return out

    def _sum1(self, idx, vals):
        out = np.zeros(self.nf)
        np.add.at(out, idx, vals)
        return out

    def setup(self, V, rho, v):
        """c, s, AA of iisph.setup (iisph.prtcl:44-134)"""
        m, rho0, dt = self.m, self.rho0, self.dt
        c = -self._sum3(self.fi, (m[self.fj] / rho[self.fi] ** 2)[:, None] * self.gf)
        c -= self._sum3(self.bi, (1.4 * rho0 * V[self.bj] / rho[self.bi] ** 2)[:, None] * self.gb)
        dv = v[self.fi] - v[self.fj]
        s = self._sum1(self.fi, m[self.fj] * (dv * self.gf).sum(1))
        s += self._sum1(self.bi, rho0 * V[self.bj] * (v[self.bi] * self.gb).sum(1))
        s = (rho0 - rho) - dt * (rho0 / rho) * s
        AA = self._sum1(self.fi, m[self.fj] * (c[self.fi] * self.gf).sum(1))
        AA -= self._sum1(self.fi, m[self.fj] * (m[self.fi] / rho[self.fi] ** 2) *
                         (self.gf * self.gf).sum(1))
        AA += self._sum1(self.bi, rho0 * V[self.bj] * (c[self.bi] * self.gb).sum(1))
        AA *= dt * dt * (rho0 / rho)
        return c, s, AA

    def pressure_acceleration(self, V, rho, p):
        m, rho0 = self.m, self.rho0
        pr = p / rho ** 2
        a = -self._sum3(self.fi, (m[self.fj] * (pr[self.fi] + pr[self.fj]))[:, None] * self.gf)
        a -= self._sum3(self.bi, (0.7 * rho0 * V[self.bj] * 2 * pr[self.bi])[:, None] * self.gb)
        return a

    def pressure(self, V, rho, a, s, AA, p, omega):
        m, rho0, dt = self.m, self.rho0, self.dt
        da = a[self.fi] - a[self.fj]
        Ap = self._sum1(self.fi, m[self.fj] * (da * self.gf).sum(1))
        Ap += self._sum1(self.bi, rho0 * V[self.bj] * (a[self.bi] * self.gb).sum(1))
        Ap *= dt * dt * rho0 / rho
        rcp = np.where(np.abs(AA) > 1e-9, 1.0 / np.where(AA != 0, AA, 1.0), 0.0)
        pn = np.maximum(0.0, p + omega * (s - Ap) * rcp) * (np.abs(AA) > 1e-9)
        aprde = (((Ap - s) / rho0) * (pn > 1e-9)).sum()
        return Ap, pn, aprde

    # ---- pt16.prtcl (Peer & Teschner 2016): setup and its two scalar systems
    def pt16_setup(self, V, rho, v, xi):
        """pt16.prtcl:48-109: same-group and smoothed densities, velocity
        gradient, vorticity and target velocity gradient (without spin)"""
        m, rho0, h = self.m, self.rho0, self.h
        wf = w(np.linalg.norm(self.df, axis=1), h)
        wb = w(np.linalg.norm(self.db, axis=1), h)
        same = self._sum1(self.fi, m[self.fj] * wf)
        prho = (self._sum1(self.fi, m[self.fj] / rho0 * wf) + self._sum1(self.bi, V[self.bj] * wb)) * rho0
        vg = np.zeros((self.nf, 3, 3))
        np.add.at(vg, self.fi, m[self.fj][:, None, None] *
                  np.einsum("ta,tb->tab", v[self.fj] - v[self.fi], self.gf))
        vg /= rho[:, None, None]
        div = np.trace(vg, axis1=1, axis2=2)
        R = (vg - vg.transpose(0, 2, 1)) / 2
        Vm = (div / 3)[:, None, None] * np.eye(3)
        S = (vg + vg.transpose(0, 2, 1)) / 2 - Vm
        om = 2 * np.stack([R[:, 2, 1], R[:, 0, 2], R[:, 1, 0]], 1)
        step = ((rho - rho0 >= 0) | (-div >= 0)).astype(float)
        tvg = Vm * step[:, None, None] + xi * S
        return same, prho, om, tvg

    def pt16_matrix(self, diag):
        """(A x)_i = diag_i x_i - sum_f m_j W_ij x_j over fluid pairs (self included)"""
        A = np.diag(np.asarray(diag, np.float64))
        wf = w(np.linalg.norm(self.df, axis=1), self.h)
        np.add.at(A, (self.fi, self.fj), -self.m[self.fj] * wf)
        return A

    def pt16_omega_rhs(self, om, xi):
        wf = w(np.linalg.norm(self.df, axis=1), self.h)
        return xi * self._sum3(self.fi, (self.m[self.fj] * wf)[:, None] * (om[self.fi] - om[self.fj]))

    def pt16_vrec_rhs(self, tvg):
        wf = w(np.linalg.norm(self.df, axis=1), self.h)
        M = tvg[self.fi] + tvg[self.fj]
        return self._sum3(self.fi, (0.5 * self.m[self.fj] * wf)[:, None] *
                          np.einsum("tab,tb->ta", M, self.df))

    @staticmethod
    def cross_matrix(a):
        z = np.zeros(len(a))
        return np.stack([np.stack([z, -a[:, 2], a[:, 1]], 1), np.stack([a[:, 2], z, -a[:, 0]], 1),
                         np.stack([-a[:, 1], a[:, 0], z], 1)], 1)

    # ---- correction.prtcl:21-38: L_i = pinv(sum_f (m_j / rho_i) gradW_ij (x_j - x_i)^T)
    def gradient_correction(self, rho, rcond=3 * np.finfo(np.float32).eps):
        n, m = self.nf, self.m
        Lf = np.zeros((n, 3, 3))
        for t in range(len(self.fi)):
            i, j = self.fi[t], self.fj[t]
            Lf[i] += (m[j] / rho[i]) * np.outer(self.gf[t], -self.df[t])
        return np.stack([np.linalg.pinv(Lf[i], rcond=rcond) for i in range(n)])

    # ---- WKBB18 (wkbb18.prtcl:49-173) as a dense 3N x 3N system; with L,
    # wkbb18_gc (wkbb18_gc.prtcl): every gradient of particle i's terms is L_i gradW
    def wkbb18_matrix(self, V, rho, mu_f, mu_b, L=None):
        """A with (A x)_i = x_i - dt (sum_f k_ij ((x_i - x_j).x_ij) gradW_ij +
        sum_b k_ib (x_i.x_ib) gradW_ib)"""
        n, h, dt, rho0, m = self.nf, self.h, self.dt, self.rho0, self.m
        A = np.eye(3 * n)
        gf = self.gf if L is None else np.einsum("tab,tb->ta", L[self.fi], self.gf)
        gb = self.gb if L is None else np.einsum("tab,tb->ta", L[self.bi], self.gb)
        kf = 10 * (mu_f / rho[self.fi]) * (m[self.fj] / rho[self.fj]) / (
            (self.df ** 2).sum(1) + 0.01 * h * h)
        for t in range(len(self.fi)):
            i, j = self.fi[t], self.fj[t]
            B = kf[t] * np.outer(gf[t], self.df[t])
            A[3 * i:3 * i + 3, 3 * i:3 * i + 3] -= dt * B
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] += dt * B
        kb = 10 * (mu_b / rho[self.bi]) * (rho0 * V[self.bj] / rho[self.bi]) / (
            (self.db ** 2).sum(1) + 0.01 * h * h)
        for t in range(len(self.bi)):
            i = self.bi[t]
            A[3 * i:3 * i + 3, 3 * i:3 * i + 3] -= dt * kb[t] * np.outer(gb[t], self.db[t])
        return A


def cg_reference(A, b, x0, M, tol, max_k):
    """Preconditioned CG with the conventions of cg_openmp.hpp:32-215
    (r = Ax - b, p0 = -y0, absolute breaks, stop at |r|^2 < max(tol^2 |b|^2,
    FLT_MIN), at most max_k + 1 iterations)."""
    x = x0.copy()
    r = A @ x - b
    b_nsq = b @ b
    thr = max(tol * tol * b_nsq, np.finfo(np.float32).tiny)
    y = M @ r
    p = -y
    r_nsq = r @ r
    ry = r @ y
    k = 0
    while True:
        if not (k < 1 or (k <= max_k and not (r_nsq < thr))):
            break
        prev = ry
        q = A @ p
        pq = p @ q
        if abs(pq) < tol:
            break
        alpha = prev / pq
        x = x + alpha * p
        r = r + alpha * q
        r_nsq = r @ r
        y = M @ r
        ry = r @ y
        if abs(prev) < tol:
            break
        beta = ry / prev
        p = -y + beta * p
        k += 1
    return x, k
