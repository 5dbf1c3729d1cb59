"""Precision form (FP32 particle state and constitutive update, FP64 nodal accumulation,
FP64 grid update and contact normals) against the FP64 oracle and the FP64 GPU form.

Bounds, in the units the FP32 state resolves:
  positions   |dx| <= 1e-5 h (the form's stated target; FP32 offsets resolve 6e-8 h and the
              FP32 velocity error, ~sqrt(steps) * 6e-8 relative, integrates into x)
  velocities  |dv| <= 1e-5 (max|v| + |g| t)
  stresses    |dT| <= 1e-4 (max|T| + E t max|v| / h)   (FP32 Jaumann increments)
  volumes     |dV| <= 1e-5 V   (V is FP32 and updated through expf every step)
  energy      |dE| <= 1e-4 of the kinetic + elastic energy the FP64 run reaches (the total
              energy itself is nearly conserved, so its change is no scale)
"""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_PREC_FP32, MPM_PREC_FP64, stress6_to9
from tests._harness import example_workload, gpu_of, oracle_of

pytestmark = pytest.mark.gpu


def _errors(w, ref, test, t):
    """max errors of `test` against `ref` (dicts with x, xt, T9, v), in the bound units"""
    h = min(w.L[d] / w.N[d] for d in range(w.dim))
    gmag = float(np.linalg.norm(w.gravity))
    dim = w.dim
    vscale = float(np.max(np.abs(ref["xt"]))) + gmag * t + 1e-12
    E = max(b.E or 0.0 for b in w.bodies)
    tscale = float(np.max(np.abs(ref["T9"]))) + E * t * vscale / h + 1e-300
    return dict(
        x_cells=float(np.max(np.abs(test["x"][:, :dim] - ref["x"][:, :dim]))) / h,
        v_rel=float(np.max(np.abs(test["xt"][:, :dim] - ref["xt"][:, :dim]))) / vscale,
        T_rel=float(np.max(np.abs(test["T9"] - ref["T9"]))) / tscale,
        V_rel=float(np.max(np.abs(test["v"] - ref["v"]) / np.abs(ref["v"]))),
    )


def _gpu_state(g, gb):
    p = g.points(gb)
    return {"x": p["x"], "xt": p["xt"], "T9": stress6_to9(p["T6"]), "v": p["v"], "active": p["active"]}


def _ora_state(o, b):
    p = o.points(b)
    return {"x": p["x"], "xt": p["xt"], "T9": p["T"], "v": p["v"], "active": p["active"]}


@pytest.mark.parametrize("name,steps", [("blockdrop2D", 200), ("collision2D", 200), ("collision3D", 50)])
def test_precision_form_vs_oracle(name, steps):
    w = example_workload(name)
    ora = oracle_of(w)
    g = gpu_of(w, precision=MPM_PREC_FP32)
    ora.step(w.dt, steps)
    g.step(w.dt, steps)
    for b in range(len(w.bodies)):
        ro, rg = _ora_state(ora, b), _gpu_state(g, b)
        np.testing.assert_array_equal(rg["active"], ro["active"])
        e = _errors(w, ro, rg, steps * w.dt)
        print(f"{name} body {b}: {e}")
        assert e["x_cells"] <= 1e-5 and e["v_rel"] <= 1e-5 and e["T_rel"] <= 1e-4 and e["V_rel"] <= 1e-5, e


def test_precision_form_contact_3d():
    w = W.two_block_3d(n_cells=32, gap_cells=1, steps=60, E=1e6)
    for b in w.bodies:
        b.x = b.x + np.random.default_rng(5).uniform(-0.1, 0.1, b.x.shape) / 32
    ora = oracle_of(w)
    g = gpu_of(w, precision=MPM_PREC_FP32)
    ora.step(w.dt * 10, 60)
    g.step(w.dt * 10, 60)
    for b in range(2):
        e = _errors(w, _ora_state(ora, b), _gpu_state(g, b), 60 * w.dt * 10)
        print(f"contact3d body {b}: {e}")
        assert e["x_cells"] <= 1e-5 and e["v_rel"] <= 1e-5 and e["T_rel"] <= 1e-4 and e["V_rel"] <= 1e-5, e


def energy(w, st, body):
    """(kinetic, gravitational, elastic = 1/2 V T : C^-1 : T isotropic) of one body"""
    b = w.bodies[body]
    dim = w.dim
    g = np.asarray(w.gravity, float)
    kin = 0.5 * float(np.sum(b.m * np.sum(st["xt"][:, :dim] ** 2, axis=1)))
    grav = -float(np.sum(b.m * (st["x"][:, :dim] @ g[:dim])))
    G = b.E / (2 * (1 + b.nu))
    K = b.E / (3 * (1 - 2 * b.nu))
    lam = K - 2 * G / 3
    T = st["T9"].reshape(-1, 3, 3)
    TT = np.einsum("pij,pij->p", T, T)
    tr = np.trace(T, axis1=1, axis2=2)
    w_el = TT / (4 * G) - lam / (4 * G * (2 * G + 3 * lam)) * tr ** 2
    ela = float(np.sum(st["v"] * w_el))
    return kin, grav, ela


@pytest.mark.parametrize("size", ["sample", "16M"])
def test_precision_column_1000_steps(size):
    """The column collapse at h = 7.8125 mm, 1000 steps: FP32 form vs FP64 form per field
    and in total energy. The FP64 form matches the oracle to 1e-9 h (test_gpu_parity)."""
    w = W.column_3d_sample() if size == "sample" else W.column_3d(wide=True)
    a = gpu_of(w, precision=MPM_PREC_FP64)
    a.step(w.dt, 1000)
    ref = _gpu_state(a, 0)
    ka, ga, la = energy(w, ref, 0)
    a.close()
    b = gpu_of(w, precision=MPM_PREC_FP32)
    b.step(w.dt, 1000)
    st = _gpu_state(b, 0)
    kb, gb, lb = energy(w, st, 0)
    b.close()
    bad = np.nonzero(st["active"] != ref["active"])[0]
    if len(bad):
        h = w.L[0] / w.N[0]
        x0 = w.bodies[0].x
        for i in bad[:12]:
            print(f"active differs at {i}: fp64 {ref['active'][i]} x {ref['x'][i] / h} v {ref['xt'][i]}; "
                  f"fp32 {st['active'][i]} x {st['x'][i] / h} v {st['xt'][i]}; x0 {x0[i] / h}")
        print(f"{len(bad)} differ; fp64 inactive {int((ref['active'] == 0).sum())}, fp32 inactive "
              f"{int((st['active'] == 0).sum())}")
    np.testing.assert_array_equal(st["active"], ref["active"])
    e = _errors(w, ref, st, 1000 * w.dt)
    e["energy_rel"] = abs((kb + gb + lb) - (ka + ga + la)) / (ka + la)
    print(f"column {size} x1000: {e}  (FP64 kinetic {ka:.6e} elastic {la:.6e} grav {ga:.6e}; "
          f"FP32 kinetic {kb:.6e} elastic {lb:.6e} grav {gb:.6e})")
    assert e["x_cells"] <= 1e-5, e
    assert e["v_rel"] <= 1e-5 and e["T_rel"] <= 1e-4 and e["V_rel"] <= 1e-5, e
    assert e["energy_rel"] <= 1e-4, e


def test_precision_form_points_on_node_lines_stay_finite():
    """CPDI-off points (within a corner offset of the floor) sitting exactly on node lines
    of the fixed-point grid: the FP32 form takes the one-sided gradient of the element the
    point is in, so nothing turns NaN and the points stay active. Exactly on the lines the
    FP64 form reproduces the reference's 0/0, so the comparison is with FP64 run 1e-12 h to
    the side of the lines. That input is ill-conditioned in the reference's model itself:
    the CPDI-off gradient jumps at node lines, and points drifting at ~1e-10 cells per step
    cross a line at a step that an offset of 1e-12 h already shifts, so the bounds here are
    positions 1e-5 h (the form's target) and velocities 1e-3 of their scale."""
    h = 1.0 / 20
    xs = np.array([i * h for i in range(4, 16)])
    x = np.stack([np.repeat(xs, 3), np.tile([0.001, 0.002, 0.03], len(xs))], axis=1)
    n = len(x)
    V = np.full(n, (h / 2) ** 2)

    def wl(xx):
        body = W.Body("row", 1000.0 * V, V, xx, np.zeros((n, 2)), np.ones(n, np.int32), 1e5, 0.3, True)
        return W.Workload("nodeline", 2, (1.0, 1.0), (20, 20), [body], [], 1e-4, 50, (0.0, -9.81))

    on, off = wl(x), wl(x + np.array([1e-12 * h, 0.0]))
    ref = gpu_of(off, precision=MPM_PREC_FP64)
    f_on = gpu_of(on, precision=MPM_PREC_FP32)
    for g in (ref, f_on):
        g.step(1e-4, 50)
    sr, son = _gpu_state(ref, 0), _gpu_state(f_on, 0)
    assert np.all(np.isfinite(son["xt"])) and np.all(son["active"] == 1)
    e = _errors(on, sr, son, 50e-4)
    print(f"node lines: {e}")
    assert e["x_cells"] <= 1e-5 and e["v_rel"] <= 1e-3


def test_precision_fused_matches_split():
    """The fused G2P+P2G pass of the precision form gives the split passes' state to FP32
    rounding (it scatters with the volume updated in the same pass)."""
    w = example_workload("collision3D")
    a = gpu_of(w, precision=MPM_PREC_FP32, fused=True)
    b = gpu_of(w, precision=MPM_PREC_FP32, fused=False)
    a.step(w.dt, 40)
    b.step(w.dt, 40)
    for bi in range(2):
        e = _errors(w, _gpu_state(b, bi), _gpu_state(a, bi), 40 * w.dt)
        print(f"fused vs split body {bi}: {e}")
        assert e["x_cells"] <= 1e-6 and e["v_rel"] <= 1e-6 and e["T_rel"] <= 1e-6 and e["V_rel"] <= 1e-6
