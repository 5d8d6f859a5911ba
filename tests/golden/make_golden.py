"""Writes the golden fixtures of tests/golden/ from the independent FP64
numpy restatement in tests/npref.py (not from the oracle, not from the HIP
path). Run from the repository root:  python tests/golden/make_golden.py"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import npref  # noqa: E402


def tiny_scene():
    """5x5x5 jittered fluid block one spacing above a 9x9 floor patch and
    beside a 9x6 wall patch (single-layer boundary), spacing 0.025."""
    s = np.float32(0.025)
    rng = np.random.default_rng(1234)
    i, j, k = np.meshgrid(np.arange(5), np.arange(5), np.arange(5), indexing="ij")
    xf = np.stack([i.ravel() + 2, j.ravel() + 2, k.ravel() + 1], 1) * float(s)
    xf = (xf + rng.uniform(-2e-3, 2e-3, xf.shape)).astype(np.float32)
    a, b = np.meshgrid(np.arange(9), np.arange(9), indexing="ij")
    floor = np.stack([a.ravel(), b.ravel(), np.zeros(81)], 1) * float(s)
    a, c = np.meshgrid(np.arange(9), np.arange(1, 7), indexing="ij")
    wall = np.stack([a.ravel(), np.zeros(a.size), c.ravel()], 1) * float(s)
    xb = np.concatenate([floor, wall]).astype(np.float32)
    v = rng.uniform(-0.3, 0.3, xf.shape).astype(np.float32)
    return s, xf, xb, v


def main():
    s, xf, xb, v = tiny_scene()
    h = float(s)
    radius = 2.0 * h
    rho0, dt, omega = 1000.0, 5e-4, 0.5
    m = np.full(len(xf), np.float32(h * h * h * rho0), np.float64)
    T = npref.Tiny(xf, xb, h, radius, m, rho0, dt)
    V = T.volume()
    rho = T.density(V)
    # gravity + integrate velocity (hard fade open): v* = v + dt g
    g = np.array([0.0, 0.0, -9.81])
    vs = v.astype(np.float64) + dt * g
    c, src, AA = T.setup(V, rho, vs)
    p = np.zeros(len(xf))
    out = dict(xf=xf, xb=xb, v=v, m=m.astype(np.float32), h=h, radius=radius, rho0=rho0,
               dt=dt, omega=omega, V=V, rho=rho, vstar=vs, c=c, s=src, AA=AA)
    for it in range(3):
        a = T.pressure_acceleration(V, rho, p)
        Ap, p, aprde = T.pressure(V, rho, a, src, AA, p, omega)
        out[f"a{it}"], out[f"Ap{it}"], out[f"p{it}"], out[f"aprde{it}"] = a, Ap, p, aprde
    # WKBB18 with mu = 50 on the same state
    mu = 50.0
    A = T.wkbb18_matrix(V, rho, mu, mu)
    n = len(xf)
    D = np.zeros((n, 3, 3))
    for i in range(n):
        D[i] = np.linalg.inv(A[3 * i:3 * i + 3, 3 * i:3 * i + 3])
    M = np.zeros_like(A)
    for i in range(n):
        M[3 * i:3 * i + 3, 3 * i:3 * i + 3] = D[i]
    b = vs.ravel()
    xsol = np.linalg.solve(A, b)
    xcg, kcg = npref.cg_reference(A, b, b.copy(), M, 1e-4, 200)
    out.update(mu=mu, wkbb18_A=A, wkbb18_sp=D, wkbb18_solution=xsol.reshape(n, 3),
               wkbb18_cg_solution=xcg.reshape(n, 3), wkbb18_cg_iterations=kcg)
    np.savez_compressed(os.path.join(HERE, "tiny_iisph_wkbb18.npz"), **out)
    print("wrote tiny_iisph_wkbb18.npz", n, "fluid", len(xb), "boundary, cg iters", kcg)

    # gradient correction (numpy.linalg.pinv) and WKBB18 with it (wkbb18_gc)
    L = T.gradient_correction(rho)
    Ag = T.wkbb18_matrix(V, rho, mu, mu, L)
    Dg = np.stack([np.linalg.inv(Ag[3 * i:3 * i + 3, 3 * i:3 * i + 3]) for i in range(n)])
    Mg = np.zeros_like(Ag)
    for i in range(n):
        Mg[3 * i:3 * i + 3, 3 * i:3 * i + 3] = Dg[i]
    xg = np.linalg.solve(Ag, b)
    xgcg, kgcg = npref.cg_reference(Ag, b, b.copy(), Mg, 1e-4, 200)
    gc = dict(xf=xf, xb=xb, v=v, m=m.astype(np.float32), h=h, radius=radius, rho0=rho0,
              dt=dt, mu=mu, V=V, rho=rho, vstar=vs, L=L, sp=Dg, solution=xg.reshape(n, 3),
              cg_solution=xgcg.reshape(n, 3), cg_iterations=kgcg)
    np.savez_compressed(os.path.join(HERE, "tiny_wkbb18_gc.npz"), **gc)
    print("wrote tiny_wkbb18_gc.npz, cg iters", kgcg)

    # pt16 on the same state (velocity v*, strain_rate_viscosity 0.9,
    # tolerance 1e-4): setup, vorticity diffusion (dense solve and the CG of
    # cg_openmp.hpp per component), spin, velocity reconstruction
    xi = 0.9
    same, prho, om, tvg = T.pt16_setup(V, rho, vs, xi)
    pre = np.diag(1.0 / (same - m * npref.w(0.0, h)))
    Ao, Av = T.pt16_matrix(same), T.pt16_matrix(prho)
    orhs = T.pt16_omega_rhs(om, xi)
    om_sol, om_cg, om_k = np.zeros((n, 3)), np.zeros((n, 3)), []
    for d in range(3):
        om_sol[:, d] = np.linalg.solve(Ao, orhs[:, d])
        xcg_d, k_d = npref.cg_reference(Ao, orhs[:, d], om[:, d].copy(), pre, 1e-4, 500)
        om_cg[:, d] = xcg_d
        om_k.append(k_d)
    tvg2 = tvg + T.cross_matrix(0.5 * om_cg)
    vrhs = T.pt16_vrec_rhs(tvg2)
    v_sol, v_cg, v_k = np.zeros((n, 3)), np.zeros((n, 3)), []
    for d in range(3):
        v_sol[:, d] = np.linalg.solve(Av, vrhs[:, d])
        xcg_d, k_d = npref.cg_reference(Av, vrhs[:, d], vs[:, d].copy(), pre, 1e-4, 500)
        v_cg[:, d] = xcg_d
        v_k.append(k_d)
    pt = dict(xf=xf, xb=xb, v=v, m=m.astype(np.float32), h=h, radius=radius, rho0=rho0, dt=dt,
              xi=xi, V=V, rho=rho, vstar=vs, same=same, prho=prho, om=om, tvg=tvg,
              orhs=orhs, om_solution=om_sol, om_cg=om_cg, om_iterations=np.array(om_k),
              tvg_spin=tvg2, vrhs=vrhs, v_solution=v_sol, v_cg=v_cg, v_iterations=np.array(v_k))
    np.savez_compressed(os.path.join(HERE, "tiny_pt16.npz"), **pt)
    print("wrote tiny_pt16.npz, cg iters", om_k, v_k)


if __name__ == "__main__":
    main()
