"""Synthetic inputs and ``.points`` files for the explicit MPM step.

``lattice_block`` places ``k`` points per cell per dimension at offsets ``(j+1/2)/k * h``
inside an axis-aligned box, with ``V = h^D / k^D`` and ``m = rho * V`` (the layout of
mpm_pygen/Grid3d.py:20-54). ``read_points``/``write_points`` handle the reference's
``.points`` text format (default_points.cpp:104-193): a count line, then per point
``m v x.. x_t.. active`` separated by single spaces.

Named workloads (sizes from the benchmark configurations):

* ``block_drop_2d``     -- 20x20 grid on 1x1 m, 0.4 m block at 2x2 points/cell, 0.3 m up.
* ``two_block_3d``      -- 256^3 grid on a 1 m cube, two 0.3 m cubes at 2^3 points/cell,
                           centres 0.4 m apart in x, velocities +-1 m/s (about 7.2M points).
* ``column_3d``         -- 256x256x128 grid on 2x2x1 m, a column at 2^3 points/cell with
                           positions jittered by +-0.25 cell from a seed; ``wide=True`` gives
                           the 1.0x1.0x0.953125 m column of 15,990,784 points.
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional, Sequence

import numpy as np


@dataclasses.dataclass
class Body:
    name: str
    m: np.ndarray          # (n,)
    v: np.ndarray          # (n,) volume
    x: np.ndarray          # (n, dim)
    xt: np.ndarray         # (n, dim) velocity
    active: np.ndarray     # (n,) int32
    E: Optional[float] = 1e5
    nu: Optional[float] = 0.3
    boundary: int = 1       # 0 none, 1 CartesianBox, 2 CartesianSmoothBox, 3 CartesianFrictionalBox
    bc_mu: float = 0.0      # CartesianFrictionalBox mu_f
    # Sand_SachithLocal instead of IsotropicLinearElasticity (E, nu as above):
    # (grain diameter, mu_s, grains_rho, rho_critical, mu_2, I_0) and USE_DENSITY_FOR_PRESSURE
    sand: Optional[tuple] = None
    sand_dp: bool = False

    @property
    def n(self) -> int:
        return int(self.m.shape[0])


@dataclasses.dataclass
class Contact:
    body1: int
    body2: int
    mu: float


@dataclasses.dataclass
class Workload:
    name: str
    dim: int
    L: Sequence[float]
    N: Sequence[int]
    bodies: List[Body]
    contacts: List[Contact]
    dt: float
    steps: int
    gravity: Sequence[float]
    cpdi_spec: int = 1
    contact_spec: int = 1

    @property
    def n_points(self) -> int:
        return sum(b.n for b in self.bodies)


def lattice_block(dim, h, lo, hi, k, rho, velocity=None, jitter=0.0, seed=0):
    """Points at ``k`` per cell per dim on the lattice of cells of size ``h`` inside [lo, hi).

    ``jitter`` moves every coordinate by U(-jitter, +jitter) * h (seeded)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    h = np.asarray(h, np.float64)
    axes = []
    for d in range(dim):
        c0 = int(np.floor(lo[d] / h[d] + 1e-9))
        c1 = int(np.ceil(hi[d] / h[d] - 1e-9))
        cells = np.arange(c0, c1)
        offs = (np.arange(k) + 0.5) / k
        coords = ((cells[:, None] + offs[None, :]) * h[d]).reshape(-1)
        coords = coords[(coords >= lo[d]) & (coords < hi[d])]
        axes.append(coords)
    grids = np.meshgrid(*axes, indexing="ij")
    # dim 0 fastest in the file order, as the reference's generators write them
    x = np.stack([g.transpose(*reversed(range(dim))).reshape(-1) for g in grids], axis=1)
    n = x.shape[0]
    vol = float(np.prod(h)) / k ** dim
    if jitter:
        rng = np.random.default_rng(seed)
        x = x + rng.uniform(-jitter, jitter, size=x.shape) * h[None, :dim]
    m = np.full(n, rho * vol)
    v = np.full(n, vol)
    xt = np.zeros((n, dim))
    if velocity is not None:
        xt[:] = np.asarray(velocity, np.float64)[None, :dim]
    return m, v, x, xt


def block_drop_2d(points_per_cell=2, steps=2000):
    L, N = (1.0, 1.0), (20, 20)
    h = np.array([L[0] / N[0], L[1] / N[1]])
    m, v, x, xt = lattice_block(2, h, (0.3, 0.3), (0.7, 0.7), points_per_cell, 1000.0)
    body = Body("block", m, v, x, xt, np.ones(len(m), np.int32), 1e5, 0.3)
    return Workload("block_drop_2d_20x20", 2, L, N, [body], [], 1e-4, steps, (0.0, -9.81))


def two_block_3d(n_cells=256, gap_cells=None, steps=1000, E=1e6, mu=0.4, gravity=(0, 0, -9.81)):
    """Two 0.3 m cubes, centres 0.4 m apart along x (or ``gap_cells`` cells apart at the
    faces when given), velocities +-1 m/s, ContactHuang between them."""
    L = (1.0, 1.0, 1.0)
    N = (n_cells,) * 3
    h = np.array([1.0 / n_cells] * 3)
    half = 0.15
    if gap_cells is None:
        c1, c2 = 0.3, 0.7
    else:
        gap = gap_cells * h[0]
        c1 = 0.5 - gap / 2 - half
        c2 = 0.5 + gap / 2 + half
    bodies = []
    for name, cx, vx in (("block", c1, 1.0), ("block2", c2, -1.0)):
        lo = (cx - half, 0.5 - half, 0.5 - half)
        hi = (cx + half, 0.5 + half, 0.5 + half)
        m, v, x, xt = lattice_block(3, h, lo, hi, 2, 1000.0, velocity=(vx, 0, 0))
        bodies.append(Body(name, m, v, x, xt, np.ones(len(m), np.int32), E, 0.3))
    return Workload(f"two_block_3d_{n_cells}^3", 3, L, N, bodies, [Contact(0, 1, mu)], 2e-5, steps,
                    gravity)


def sliding_block(dim=2, boundary=3, mu=0.3, wall_dim=None, upper=False, speed=1.0, steps=500,
                  E=1e6, rho=1000.0, g=9.81):
    """An elastic block resting on one face of the domain box and sliding along it.

    The face is ``wall_dim`` (default the last dimension) at 0, or at L when ``upper``;
    gravity ``g`` points into it. The block moves at ``speed`` along the tangent
    dimension(s): the other dimension in 2D, the diagonal of the two tangent dimensions in
    3D. 2D: 40x40 grid on 1x1 m, a 0.2x0.1 m block (128 points); 3D: 32^3 grid on a 1 m
    cube, a 0.25x0.25x0.125 m block (2048 points); 2 points per cell per dimension,
    dt = 1e-4.
    ``boundary`` and ``mu`` are the body's boundary type and friction coefficient."""
    if wall_dim is None:
        wall_dim = dim - 1
    L = (1.0,) * dim
    N = (40, 40) if dim == 2 else (32, 32, 32)
    h = np.array([L[d] / N[d] for d in range(dim)])
    tang = [d for d in range(dim) if d != wall_dim]
    width, depth = (0.2, 0.1) if dim == 2 else (0.25, 0.125)
    lo, hi = [0.0] * dim, [0.0] * dim
    for d in tang:
        lo[d], hi[d] = 0.3, 0.3 + width
    lo[wall_dim], hi[wall_dim] = (L[wall_dim] - depth, L[wall_dim]) if upper else (0.0, depth)
    vel = np.zeros(dim)
    vel[tang] = speed / np.sqrt(len(tang))
    m, v, x, xt = lattice_block(dim, h, lo, hi, 2, rho, velocity=vel)
    body = Body("slider", m, v, x, xt, np.ones(len(m), np.int32), E, 0.3, boundary, mu)
    grav = [0.0] * dim
    grav[wall_dim] = g if upper else -g
    face = f"{'upper' if upper else 'lower'}{wall_dim}"
    return Workload(f"sliding_block_{dim}d_{face}_bc{boundary}_mu{mu}", dim, L, N, [body], [], 1e-4, steps,
                    tuple(grav))


def sand_props(mu_s=0.577, grains_d=0.01, grains_rho=1650.0, rho_critical=1000.0, mu_2=None, I_0=1.0):
    """Sand_SachithLocal's material properties after E, nu, in the order the C-ABI takes them.
    With six properties the reference sets mu_2 = mu_s and I_0 = 1 (sand_sachithlocal.cpp:66-79)."""
    return (grains_d, mu_s, grains_rho, rho_critical, mu_s if mu_2 is None else mu_2, I_0)


def column_collapse_example(dim, steps=100):
    """The shipped columncollapse2D / columncollapse3D examples: their point files with the
    configuration's grid (1 x 0.5 m at 40 x 20; 1 x 1 x 0.5 m at 20 x 20 x 10), dt = 1e-4,
    Sand_SachithLocal {1e7, 0.3, 0.01, 0.577, 1650, 1000} and CartesianBox walls."""
    name = f"columncollapse{dim}D"
    m, v, x, xt, a = read_points(os.path.join(example_dir(), name, f"test{dim}D.points"), dim)
    body = Body("block", m, v, x, xt, a, 1e7, 0.3, 1, 0.0, sand_props())
    L, N = ((1.0, 0.5), (40, 20)) if dim == 2 else ((1.0, 1.0, 0.5), (20, 20, 10))
    g = (0.0, -9.81) if dim == 2 else (0.0, 0.0, -9.81)
    return Workload(name, dim, L, N, [body], [], 1e-4, steps, g)


def column_3d(wide=True, steps=1000, seed=0xC0FFEE + 2, grid=(256, 256, 128), domain=(2.0, 2.0, 1.0),
              size=None):
    """Column collapse at h = 7.8125 mm: 2^3 points per cell jittered by +-0.25 cell (seeded),
    E = 1e7 Pa, nu = 0.3, rho = 1500, dt = 1e-5, gravity -9.81 in z, CartesianBox walls.

    ``wide`` selects the 1.0x1.0x0.953125 m column (15,990,784 points on the 256x256x128
    grid over 2x2x1 m); otherwise the 0.5x0.5x0.9 m column (about 3.77M points).
    ``size`` overrides the column dimensions (m); the column stands centred on the floor."""
    N = tuple(grid)
    L = tuple(domain)
    h = np.array([L[i] / N[i] for i in range(3)])
    if size is None:
        size = (1.0, 1.0, 0.953125) if wide else (0.5, 0.5, 0.9)
    wx, wy, hz = size
    lo = (L[0] / 2 - wx / 2, L[1] / 2 - wy / 2, 0.0)
    hi = (L[0] / 2 + wx / 2, L[1] / 2 + wy / 2, hz)
    m, v, x, xt = lattice_block(3, h, lo, hi, 2, 1500.0, jitter=0.25, seed=seed)
    body = Body("column", m, v, x, xt, np.ones(len(m), np.int32), 1e7, 0.3)
    name = f"column_3d_{len(m)}pts_{N[0]}x{N[1]}x{N[2]}"
    return Workload(name, 3, L, N, [body], [], 1e-5, steps, (0.0, 0.0, -9.81))


def column_3d_sample(steps=4, seed=0xC0FFEE + 2):
    """A 1/16-footprint sample of the 16M column: same h, material and dt; a
    0.25x0.25x0.953125 m column on a 64x64x128 grid over 0.5x0.5x1 m (999,424 points)."""
    return column_3d(steps=steps, seed=seed, grid=(64, 64, 128), domain=(0.5, 0.5, 1.0),
                     size=(0.25, 0.25, 0.953125))


def read_points(path, job_dim, grid_dim=None):
    """Read a ``.points`` file like DefaultPoints::readFromFile (default_points.cpp:104-193).

    Columns: ``m v x[DIM] x_t[DIM] active`` split on single spaces; a line with
    ``GRID_DIM`` columns per vector is also accepted. Returns (m, v, x, xt, active) with
    ``x``/``xt`` of width ``job_dim``."""
    grid_dim = job_dim if grid_dim is None else grid_dim
    with open(path) as f:
        n = int(f.readline())
        m = np.zeros(n)
        v = np.zeros(n)
        x = np.zeros((n, job_dim))
        xt = np.zeros((n, job_dim))
        act = np.zeros(n, np.int32)
        for i in range(n):
            line = f.readline().rstrip("\n")
            tok = line.split(" ")
            if len(tok) == 3 + 2 * job_dim:
                d = job_dim
            elif len(tok) == 3 + 2 * grid_dim:
                d = grid_dim
            else:
                raise ValueError(f"unable to read line {i + 2} of {path}")
            m[i] = float(tok[0])
            v[i] = float(tok[1])
            x[i, :d] = [float(t) for t in tok[2:2 + d]]
            xt[i, :d] = [float(t) for t in tok[2 + d:2 + 2 * d]]
            act[i] = int(float(tok[2 + 2 * d]))
    return m, v, x, xt, act


def write_points(path, m, v, x, xt, active):
    """Write a ``.points`` file (17 significant digits, single-space separated)."""
    n = len(m)
    with open(path, "w") as f:
        f.write(f"{n}\n")
        for i in range(n):
            cols = [m[i], v[i], *x[i], *xt[i]]
            f.write(" ".join(f"{c:.17g}" for c in cols) + f" {int(active[i])}\n")


def example_dir():
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                        "golden", "examples")


def wave_speed(E, nu, rho):
    """P-wave speed sqrt((K + 4G/3) / rho) of IsotropicLinearElasticity."""
    G = E / (2 * (1 + nu))
    K = E / (3 * (1 - 2 * nu))
    return float(np.sqrt((K + 4 * G / 3) / rho))


def ensemble_2d(n_sims=512, seed=0xC0FFEE + 3, steps=2000, rho=1000.0):
    """Batched 2D ensemble: independent block drops drawn from `seed`. Grid N x N on a
    1 m square with N uniform in 16..256; points log-uniform in 1e3..2e5 at 2x2 per cell,
    clamped to half the domain's cells; E log-uniform in 1e4..1e7 Pa, nu uniform in
    0.2..0.4, drop height (block bottom) uniform in 0.1..0.5 m; dt = 0.2 h / c."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_sims):
        N = int(rng.integers(16, 257))
        h = 1.0 / N
        npts = float(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        npts = min(npts, 4 * N * N * 0.5)
        side = max(2, int(round(np.sqrt(npts / 4))))  # cells per side (2x2 points per cell)
        side = min(side, N - 4)
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        nu = float(rng.uniform(0.2, 0.4))
        y0 = float(rng.uniform(0.1, 0.5))
        c0 = (N - side) // 2
        j0 = int(round(y0 / h))
        j0 = max(2, min(j0, N - side - 2))
        lo = (c0 * h, j0 * h)
        hi = ((c0 + side) * h, (j0 + side) * h)
        m, v, x, xt = lattice_block(2, np.array([h, h]), lo, hi, 2, rho)
        body = Body("block", m, v, x, xt, np.ones(len(m), np.int32), E, nu, True)
        dt = 0.2 * h / wave_speed(E, nu, rho)
        out.append(Workload(f"ens{i}_N{N}_{len(m)}pts", 2, (1.0, 1.0), (N, N), [body], [], dt, steps, (0.0, -9.81)))
    return out


def serve_requests(n, seed=0xC0FFEE + 4, rho=1000.0):
    """Served 3D stream: block drops or two-block collisions drawn from `seed`. Grid N^3
    on a 1 m cube with N uniform in 32..128, 2^3 points per cell, 1e4..2e6 points, E
    log-uniform in 1e5..1e8 Pa, nu = 0.3, steps uniform in 500..5000, dt = 0.2 h / c.
    Returns (workload, steps) pairs."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        N = int(rng.integers(32, 129))
        h = 1.0 / N
        E = float(np.exp(rng.uniform(np.log(1e5), np.log(1e8))))
        nu = 0.3
        steps = int(rng.integers(500, 5001))
        collide = bool(rng.integers(0, 2))
        dt = 0.2 * h / wave_speed(E, nu, rho)
        hv = np.array([h, h, h])
        if not collide:
            s_max = min(N // 2 - 2, 62)
            s = int(rng.integers(11, s_max + 1))  # cells per side: 8 s^3 points
            c0 = (N - s) // 2
            k0 = int(rng.integers(2, max(3, N - s - 2)))
            lo = (c0 * h, c0 * h, k0 * h)
            hi = ((c0 + s) * h, (c0 + s) * h, (k0 + s) * h)
            m, v, x, xt = lattice_block(3, hv, lo, hi, 2, rho)
            bodies = [Body("block", m, v, x, xt, np.ones(len(m), np.int32), E, nu, True)]
            contacts = []
        else:
            s_max = min((N - 6) // 2 - 1, 49)
            s = int(rng.integers(9, max(10, s_max + 1)))
            c0 = (N - 2 * s - 2) // 2
            cy = (N - s) // 2
            bodies = []
            for b, (i0, vx) in enumerate(((c0, 1.0), (c0 + s + 2, -1.0))):
                lo = (i0 * h, cy * h, cy * h)
                hi = ((i0 + s) * h, (cy + s) * h, (cy + s) * h)
                m, v, x, xt = lattice_block(3, hv, lo, hi, 2, rho, velocity=(vx, 0, 0))
                bodies.append(Body(f"block{b}", m, v, x, xt, np.ones(len(m), np.int32), E, nu, True))
            contacts = [Contact(0, 1, 0.4)]
        w = Workload(f"req{i}_N{N}_{'collide' if collide else 'drop'}", 3, (1.0, 1.0, 1.0), (N, N, N), bodies,
                     contacts, dt, steps, (0.0, 0.0, -9.81))
        out.append((w, steps))
    return out
