"""Shared helpers for the parity tests: build the same workload in the CPU oracle and on the
GPU, and compare their states.

Tolerances (FP64 plain form). The GPU sums node contributions in a different order than
the reference's COO loops (separable weights, LDS and global atomics), so results agree
to rounding, growing slowly with the step count:
  positions   |dx| <= XTOL_CELLS * h              (XTOL_CELLS = 1e-9)
  velocities  |dv| <= VTOL * (max|v| + |g| t + 1e-12)      (VTOL = 1e-8)
  stresses    |dT| <= TTOL * (max|T| + E t max|v| / h)   (TTOL = 1e-7); the second term is
              the elastic stress the run can build up, so free-fall runs whose exact
              stress is zero are compared against that scale rather than their noise
  volumes     |dV| <= 1e-12 * V
Connectivity (element of every GIMP corner and the CPDI flag) must be identical.
"""
from __future__ import annotations

import os

import numpy as np

from mpm_3d_amd import workloads as W

XTOL_CELLS = 1e-9
VTOL = 1e-8
TTOL = 1e-7
VOLTOL = 1e-12


def oracle_of(w, native=False):
    from oracle.pyoracle import OracleSim
    s = OracleSim(w.dim, w.L, w.N, w.cpdi_spec, w.contact_spec, native=native)
    for bi, b in enumerate(w.bodies):
        s.add_body(b.m, b.v, b.x, b.xt, b.active, E=b.E, nu=b.nu, boundary=b.boundary, bc_mu=b.bc_mu)
        if getattr(b, "sand", None) is not None:
            s.set_sand(bi, b.sand, b.sand_dp)
    for c in w.contacts:
        s.add_contact(c.body1, c.body2, c.mu)
    s.init(g=w.gravity)
    return s


def gpu_of(w, **kw):
    from mpm_3d_amd.gpu import GpuSim
    return GpuSim(w, **kw)


def example_workload(name, dt=None, steps=10):
    """The reference's shipped example inputs with their configuration values."""
    ex = W.example_dir()
    spec = {
        "blockdrop2D": (2, ["test2D.points"], [], 1e-3),
        "collision2D": (2, ["test2D.points", "test2D_2.points"], [W.Contact(0, 1, 0.4)], 1e-3),
        "minimal2D": (2, ["test2D.points", "test2D_2.points"], [W.Contact(0, 1, 0.4)], 1e-3),
        "blockdrop3D": (3, ["test3D.points"], [], 1e-3),
        "collision3D": (3, ["test3D.points", "test3D_2.points"], [W.Contact(0, 1, 0.4)], 1e-3),
    }[name]
    dim, files, contacts, dt0 = spec
    bodies = []
    for i, f in enumerate(files):
        m, v, x, xt, a = W.read_points(os.path.join(ex, name, f), dim)
        bodies.append(W.Body(f"block{i}", m, v, x, xt, a, 1e5, 0.3, True))
    L = (1.0,) * dim
    N = (20,) * dim
    g = (0.0, -9.81) if dim == 2 else (0.0, 0.0, -9.81)
    return W.Workload(name, dim, L, N, bodies, contacts, dt or dt0, steps, g)


def compare_states(w, ora, gpu, t, label=""):
    """Assert GPU state == oracle state within the stated tolerances; return max errors."""
    h = min(w.L[d] / w.N[d] for d in range(w.dim))
    gmag = float(np.linalg.norm(w.gravity))
    errs = {}
    for bi in range(len(w.bodies)):
        o = ora.points(bi)
        g = gpu.points(bi)
        dim = w.dim
        np.testing.assert_array_equal(g["active"], o["active"], err_msg=f"{label} active flags body {bi}")
        ex = np.max(np.abs(g["x"] - o["x"][:, :dim])) if len(g["x"]) else 0.0
        vscale = float(np.max(np.abs(o["xt"]))) + gmag * t + 1e-12
        ev = np.max(np.abs(g["xt"] - o["xt"][:, :dim])) if len(g["x"]) else 0.0
        T9o = o["T"]
        from mpm_3d_amd.gpu import stress6_to9
        T9g = stress6_to9(g["T6"])
        E = w.bodies[bi].E or 0.0
        tscale = float(np.max(np.abs(T9o))) + E * t * vscale / h + 1e-300
        eT = np.max(np.abs(T9g - T9o)) if len(g["x"]) else 0.0
        eV = np.max(np.abs(g["v"] - o["v"]) / np.abs(o["v"])) if len(g["x"]) else 0.0
        errs[bi] = dict(x_cells=ex / h, v_rel=ev / vscale, T_rel=eT / tscale, V_rel=eV)
        print(f"{label} body {bi}: |dx|/h={ex / h:.3e} |dv|/v={ev / vscale:.3e} "
              f"|dT|/T={eT / tscale:.3e} |dV|/V={eV:.3e}")
        assert ex <= XTOL_CELLS * h, f"{label} body {bi}: position error {ex / h:.3e} cells"
        assert ev <= VTOL * vscale, f"{label} body {bi}: velocity error {ev / vscale:.3e}"
        assert eT <= TTOL * tscale, f"{label} body {bi}: stress error {eT / tscale:.3e}"
        assert eV <= VOLTOL, f"{label} body {bi}: volume error {eV:.3e}"
    return errs
