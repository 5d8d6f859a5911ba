"""Synthetic dam-break scenes for the SPH step hot path.

The scenes are lattices, generated identically for the HIP path and for any
CPU checker: fluid blocks on a cubic lattice with seeded jitter inside closed
single-layer boundary boxes, as the dam-break scripts of the reference build
them (share/scripts/prtcl-lua.example.lua:106-110 groups "f"/"b",
:122-186 parameters, :311-316 mass = h^3 rho0, :206 radius = 2h).

Random numbers come from a counter-based SplitMix64 stream, so every value is
a pure function of (seed, stream, counter) and can be regenerated anywhere:
``u = (splitmix64(base + k) >> 11) * 2**-53`` with
``base = splitmix64(seed * 2**32 + stream)``.
"""

from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(x):
    """SplitMix64 finaliser over numpy uint64 arrays (or Python ints)."""
    scalar = not isinstance(x, np.ndarray)
    z = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return int(z) if scalar else z


def stream_base(seed: int, stream: int) -> int:
    return splitmix64(((seed << 32) + stream) & _M64)


def uniform(base: int, start: int, count: int) -> np.ndarray:
    """``count`` doubles in [0, 1) from counters start..start+count-1."""
    k = np.arange(start, start + count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = splitmix64(np.uint64(base) + k)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclasses.dataclass
class Scene:
    """One scene: parameters (the reference's global fields and group
    uniforms) plus the initial particle data in AoS float32 layout."""

    name: str
    spacing: float
    fluid_x: np.ndarray            # (nf, 3) float32
    fluid_v: np.ndarray            # (nf, 3) float32
    fluid_m: np.ndarray            # (nf,) float32
    time_of_birth: np.ndarray      # (nf,) float32
    boundary_x: np.ndarray         # (nb, 3) float32
    smoothing_scale: float
    time_step: float = 5e-4
    rest_density: float = 1000.0
    gravity: tuple = (0.0, 0.0, -9.81)
    iisph_relaxation: float = 0.5
    iisph_tolerance: float = 1e-3
    iisph_min_iterations: int = 2
    iisph_max_iterations: int = 100
    fade_duration: float = 0.0
    explicit_viscosity: bool = False
    fluid_viscosity: float = 0.0
    boundary_viscosity: float = 0.0
    implicit_viscosity: int = 0  # 1: wkbb18; 2: correction + wkbb18_gc; 3: pt16
    # pt16 settings of prtcl-lua.bunny.lua:150-155
    pt16_strain_rate_viscosity: float = 0.9
    pt16_vorticity_maximum_error: float = 0.1
    pt16_vorticity_maximum_iterations: int = 500
    pt16_velocity_maximum_error: float = 0.1
    pt16_velocity_maximum_iterations: int = 500
    wkbb18_maximum_error: float = 1e-4
    wkbb18_maximum_iterations: int = 200
    permute_every: int = 8
    # gravity.initialize_axial_acceleration instead of initialize_acceleration
    # (example.lua:390-391) about the axis through gravity_center along g
    axial_gravity: bool = False
    gravity_center: tuple = (0.0, 0.0, 0.0)
    steps: int = 100

    @property
    def n_fluid(self) -> int:
        return int(self.fluid_x.shape[0])

    @property
    def n_boundary(self) -> int:
        return int(self.boundary_x.shape[0])

    @property
    def radius(self) -> float:
        # Neighborhood radius = 2h in double (example.lua:206)
        return 2.0 * float(np.float32(self.smoothing_scale))


def box_boundary(nx: int, ny: int, nz: int, spacing: float) -> np.ndarray:
    """Single-layer closed box: every lattice point (i, j, k) * spacing with
    0 <= i < nx, ... that lies on the box surface.
    Count = nx*ny*nz - (nx-2)*(ny-2)*(nz-2)."""
    # lexicographic (i, j, k) order, built slab by slab: the two end slabs
    # are full faces, every slab between them is the same (j, k) ring
    j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    face = np.stack([j.ravel(), k.ravel()], axis=1)
    on_ring = (face[:, 0] == 0) | (face[:, 0] == ny - 1) | (face[:, 1] == 0) | (face[:, 1] == nz - 1)
    ring = face[on_ring]
    blocks = [np.column_stack([np.zeros(len(face), np.int64), face])]
    if nx > 2:
        ii = np.repeat(np.arange(1, nx - 1), len(ring))
        blocks.append(np.column_stack([ii, np.tile(ring, (nx - 2, 1))]))
    if nx > 1:
        blocks.append(np.column_stack([np.full(len(face), nx - 1, np.int64), face]))
    pts = np.concatenate(blocks).astype(np.float64) * spacing
    return pts.astype(np.float32)


def fluid_block(n: tuple, offset: tuple, spacing: float, jitter: float,
                base: Optional[int]) -> np.ndarray:
    """Cubic lattice block of n = (nx, ny, nz) particles at
    (offset_d + i_d) * spacing plus per-component jitter (2u - 1) * jitter."""
    nx, ny, nz = n
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    lat = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    lat += np.asarray(offset, dtype=np.float64)
    pts = lat * spacing
    if jitter and base is not None:
        u = uniform(base, 0, pts.size).reshape(pts.shape)
        pts = pts + (2.0 * u - 1.0) * jitter
    return pts.astype(np.float32)


def _mass(h: float, rho0: float) -> np.float32:
    # example.lua:311-316: h * h * h * rho0 in Lua doubles, stored as float
    hd = float(np.float32(h))
    return np.float32(hd * hd * hd * float(np.float32(rho0)))


def _finish(name, spacing, fx, bx, steps, v=None, **kw) -> Scene:
    nf = fx.shape[0]
    h = float(np.float32(spacing))
    rho0 = kw.pop("rest_density", 1000.0)
    return Scene(
        name=name, spacing=spacing, fluid_x=fx,
        fluid_v=np.zeros((nf, 3), np.float32) if v is None else v.astype(np.float32),
        fluid_m=np.full(nf, _mass(h, rho0), np.float32),
        # time_of_birth = -1 s with fade_duration = 0: forces act from step 1
        time_of_birth=np.full(nf, -1.0, np.float32),
        boundary_x=bx, smoothing_scale=h, rest_density=rho0, steps=steps, **kw)


def dambreak(nfx: int, nfy: int, nfz: int, box: tuple, spacing: float, seed: int,
             steps: int, name: str, jitter: float = 1e-4, **kw) -> Scene:
    """Fluid block in the corner of a closed box, one spacing inside the
    walls (offset (1,1,1) lattice units)."""
    nbx, nby, nbz = (int(round(L / spacing)) + 1 for L in box)
    bx = box_boundary(nbx, nby, nbz, spacing)
    fx = fluid_block((nfx, nfy, nfz), (1, 1, 1), spacing, jitter, stream_base(seed, 0))
    return _finish(name, spacing, fx, bx, steps, **kw)


def config0(steps: int = 200) -> Scene:
    """20^3 fluid at 0.025 m in a 2x1x1 m box (16,002 boundary particles),
    explicit viscosity 0.01, 200 steps."""
    return dambreak(20, 20, 20, (2.0, 1.0, 1.0), 0.025, seed=0, steps=steps,
                    name="dambreak_8k", explicit_viscosity=True,
                    fluid_viscosity=0.01, boundary_viscosity=0.01)


def config1(steps: int = 100) -> Scene:
    """100^3 = 1e6 fluid at 0.01 m in a closed 4x2x2 m box (400,002 boundary
    particles), 100 steps."""
    return dambreak(100, 100, 100, (4.0, 2.0, 2.0), 0.01, seed=1, steps=steps,
                    name="dambreak_1m")


def config2(steps: int = 100) -> Scene:
    """50x50x100 viscous column (mu = 50 Pa s) centred on the floor of a closed
    1x1x1.5 m tank (80,002 boundary particles); WKBB18 PCG tol 1e-4, <= 200."""
    s = 0.01
    nbx, nby, nbz = 101, 101, 151
    bx = box_boundary(nbx, nby, nbz, s)
    off = ((100 - 49) // 2, (100 - 49) // 2, 1)
    fx = fluid_block((50, 50, 100), off, s, 1e-4, stream_base(2, 0))
    return _finish("viscous_column_250k", s, fx, bx, steps, implicit_viscosity=True,
                   fluid_viscosity=50.0, boundary_viscosity=50.0,
                   wkbb18_maximum_error=1e-4, wkbb18_maximum_iterations=200)


def random_scene(seed: int, index: int, n_lo: int, n_hi: int, steps: int,
                 viscous_fraction: float = 0.0, mu_range=(1.0, 100.0),
                 spacing: float = 0.01) -> Scene:
    """One scene of the batched / served workloads: N log-uniform in
    [n_lo, n_hi], block aspect ratios U[0.5, 2] per axis, tank twice the block
    extent, initial velocities U[-0.5, 0.5] m/s per component."""
    base = stream_base(seed, index)
    u = uniform(base, 0, 8)
    n_target = math.exp(math.log(n_lo) + u[0] * math.log(n_hi / n_lo))
    r = 0.5 + 1.5 * u[1:4]
    scale = (n_target / float(np.prod(r))) ** (1.0 / 3.0)
    n = tuple(max(1, int(round(rd * scale))) for rd in r)
    tank = tuple(max(2 * nd, nd + 2) for nd in n)  # lattice cells per axis
    bx = box_boundary(tank[0] + 1, tank[1] + 1, tank[2] + 1, spacing)
    fbase = stream_base(seed, 100000 + index)
    fx = fluid_block(n, (1, 1, 1), spacing, 1e-4, fbase)
    nf = fx.shape[0]
    vbase = stream_base(seed, 200000 + index)
    v = (uniform(vbase, 0, 3 * nf).reshape(nf, 3) - 0.5).astype(np.float32)
    kw = {}
    if viscous_fraction > 0 and u[4] < viscous_fraction:
        mu = mu_range[0] + (mu_range[1] - mu_range[0]) * u[5]
        kw.update(implicit_viscosity=True, fluid_viscosity=float(mu),
                  boundary_viscosity=float(mu))
    return _finish(f"scene{seed}_{index}", spacing, fx, bx, steps, v=v, **kw)


def config3(steps: int = 100, count: int = 64) -> List[Scene]:
    """64 independent dam-break scenes, N_f log-uniform in [2k, 200k]."""
    return [random_scene(3, i, 2000, 200000, steps) for i in range(count)]


def config4_scenes(count: int = 512) -> List[Scene]:
    """Resident scenes of the served stream: N_f log-uniform [5k, 500k],
    30% with implicit viscosity mu ~ U[1, 100]."""
    return [random_scene(4, i, 5000, 500000, 10, viscous_fraction=0.3)
            for i in range(count)]


def poisson_arrivals(seed: int, rate: float, duration: float, n_scenes: int):
    """(arrival_time, scene_index) pairs of a Poisson stream."""
    base = stream_base(seed, 999)
    out = []
    t, k = 0.0, 0
    while True:
        u = uniform(base, 2 * k, 2)
        k += 1
        t += -math.log(1.0 - u[0]) / rate
        if t >= duration:
            break
        out.append((t, int(u[1] * n_scenes)))
    return out


def tiny(nf_side: int = 6, seed: int = 7, **kw) -> Scene:
    """A small dam break (nf_side^3 fluid) for fast checks."""
    s = 0.025
    L = (nf_side + 3) * s
    return dambreak(nf_side, nf_side, nf_side, (2 * L, L, L), s, seed=seed,
                    steps=kw.pop("steps", 5), name=f"tiny{nf_side}", **kw)
