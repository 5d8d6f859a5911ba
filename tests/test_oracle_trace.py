"""Pins of the sphere-tracer restatement (oracle/trace_ref.cpp) on the CPU.

For one particle the implicit function of the tracer is radial, so its
surface is the sphere |x - P| = r0 with W(r0) = threshold W(0) (cubic spline,
h = 2 smoothing_scale, solved here by bisection on the closed form). Every
ray that meets that sphere must stop on it, at the parameter of the
textbook ray-sphere intersection, with the pixel -d.n of its outward normal;
rays that miss it leave the scene (0.3). The orientation of the pinhole
camera (pinhole_camera.hpp:35-62) is pinned by a particle off the axis: with
principal +y and up +z, +x appears at larger column and +z at smaller row.
An empty scene is all 0.3, and running out of steps gives 0.
"""

import numpy as np

from oracle import pyoracle as P


def spline(q):
    return max(2 - q, 0) ** 3 - 4 * max(1 - q, 0) ** 3


def surface_radius(h, threshold):
    half = 2 * h  # grid_diameter / 2 = 4 h / 2
    lo, hi = 0.0, 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if spline(mid) > threshold * spline(0):
            lo = mid
        else:
            hi = mid
    return lo * half


def directions(cam):
    O, p, up = (np.array(a, float) for a in (cam.origin, cam.principal, cam.up))
    v = up / np.linalg.norm(up)
    p = p / np.linalg.norm(p)
    hh = np.cross(v, p)
    hh /= np.linalg.norm(hh)
    out = np.zeros((cam.height, cam.width, 3))
    for iy in range(cam.height):
        for ix in range(cam.width):
            px = (-(cam.width - 1) / 2 + ix) / cam.width
            py = (-(cam.height - 1) / 2 + iy) / cam.width
            d = O - (O + px * hh + py * v - cam.focal_length * p)
            out[iy, ix] = d / np.linalg.norm(d)
    return O, out


def test_single_particle_is_a_sphere():
    h, thr = 0.025, 0.5
    Pp = np.array([0.01, 0.02, -0.015])
    cam = P.camera((0, -0.3, 0), (0, 1, 0), (0, 0, 1), 1.0, 41, 31)
    img, par = P.sphere_trace(Pp[None], h, cam, thr, 300)
    r0 = surface_radius(h, thr)
    O, D = directions(cam)
    hits = 0
    for iy in range(cam.height):
        for ix in range(cam.width):
            d = D[iy, ix]
            oc = O - Pp
            b = d @ oc
            disc = b * b - (oc @ oc - r0 * r0)
            if disc > 1e-4:
                t = -b - np.sqrt(disc)
                n = (O + t * d - Pp) / r0
                assert abs(par[iy, ix] - t) < 2e-5, (ix, iy, par[iy, ix], t)
                assert abs(img[iy, ix] - (-(d @ n))) < 2e-3
                hits += 1
            elif disc < -1e-4:
                assert img[iy, ix] == 0.3, (ix, iy)
    assert hits > 20


def test_camera_orientation():
    h = 0.025
    cam = P.camera((0, -1, 0), (0, 1, 0), (0, 0, 1), 1.0, 21, 21)
    img_x, _ = P.sphere_trace(np.array([[0.2, 0, 0]]), h, cam)
    img_z, _ = P.sphere_trace(np.array([[0, 0, 0.2]]), h, cam)
    iy, ix = np.argwhere(img_x != 0.3).mean(0)
    assert ix > 12 and abs(iy - 10) < 1
    iy, ix = np.argwhere(img_z != 0.3).mean(0)
    assert iy < 8 and abs(ix - 10) < 1


def test_empty_scene_and_step_limit():
    cam = P.camera((0, -1, 0), (0, 1, 0), (0, 0, 1), 1.0, 8, 6)
    img, par = P.sphere_trace(np.zeros((0, 3)), 0.025, cam)
    assert np.all(img == 0.3) and np.all(par > 4 * 0.025 * 10000)
    img, par = P.sphere_trace(np.array([[0.0, 0, 0]]), 0.025, cam, 0.5, 2)
    assert np.all(img == 0.0) and np.all(par == 4 * 0.025 * 10000)
