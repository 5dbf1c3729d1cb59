"""Size-independent properties of the two-block collision at its full size, through the C-ABI.

Two 0.3 m cubes on the 256^3 grid of the 1 m cube (7.2M points at 2^3 per cell, two
bodies with their own grid fields, ContactHuang with mu = 0.4 between them), one cell
apart at the faces so that contact engages within the run, with zero gravity and the
walls far away. The second cube moves at -0.5 m/s instead of -1 so the total momentum
is not zero by symmetry.

What the step conserves (explicit_usl.cpp:48-87 with contact_huang.cpp:104-197):
  * total momentum: P2G gives sum_i p_i = sum_p m_p v_p (the weights sum to one), the
    contact forces are equal and opposite (f1 + f2 = 0 at every node both bodies reach),
    the internal forces sum to zero over the nodes, and the FLIP update returns
    sum_p m_p dv_p = sum_i m_i dv_i = dt sum_i f_i = 0;
  * the centre of mass moves with it: sum_p m_p dx_p = sum_i m_i u_i = dt sum_i p_i, so
    x_cm(k) = x_cm(0) + k dt P / M.
A dropped or doubled contact force, a particle block taken twice, a wrong body's grid
fields or a lost corner weight breaks one of them by a visible amount. Contact must also
act: the first cube loses more than 1% of its momentum.

Tolerances: FP64 form 1e-11 of M * 1 m/s (rounding of 7.2M-term sums over 200 steps) and
1e-9 h for the centre of mass; precision form 1e-5 and 1e-6 h (FP32 point velocities)."""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_PREC_FP32, MPM_PREC_FP64
from tests._harness import gpu_of

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,ptol,xtol", [(MPM_PREC_FP64, 1e-11, 1e-9), (MPM_PREC_FP32, 1e-5, 1e-6)])
def test_full_collision_conserves_momentum(prec, ptol, xtol):
    w = W.two_block_3d(n_cells=256, gap_cells=1, steps=200, gravity=(0.0, 0.0, 0.0))
    w.bodies[1].xt[:, 0] = -0.5
    n = [b.n for b in w.bodies]
    assert sum(n) > 7_000_000
    h = 1.0 / 256
    steps = 200
    g = gpu_of(w, precision=prec)
    m = [np.asarray(b.m, np.float64) for b in w.bodies]
    if prec == MPM_PREC_FP32:
        m = [mm.astype(np.float32).astype(np.float64) for mm in m]
    M = sum(float(mm.sum()) for mm in m)
    st0 = [g.points(b) for b in range(2)]
    P0 = sum((m[b][:, None] * st0[b]["xt"]).sum(0) for b in range(2))
    X0 = sum((m[b][:, None] * st0[b]["x"]).sum(0) for b in range(2)) / M
    P1_0 = (m[0][:, None] * st0[0]["xt"]).sum(0)
    g.step(w.dt, steps)
    st = [g.points(b) for b in range(2)]
    g.close()
    for b in range(2):
        assert (st[b]["active"] != 0).all()
        assert np.isfinite(st[b]["x"]).all() and np.isfinite(st[b]["xt"]).all()
    P = sum((m[b][:, None] * st[b]["xt"]).sum(0) for b in range(2))
    X = sum((m[b][:, None] * st[b]["x"]).sum(0) for b in range(2)) / M
    P1 = (m[0][:, None] * st[0]["xt"]).sum(0)
    dP = np.abs(P - P0).max() / M
    dX = np.abs(X - (X0 + steps * w.dt * P0 / M)).max() / h
    lost = 1.0 - P1[0] / P1_0[0]
    print(f"collision 256^3 ({sum(n)} points): |dP|/M {dP:.2e} m/s, centre of mass {dX:.2e} h, "
          f"body 1 lost {lost:.3%} of its momentum")
    assert dP <= ptol
    assert dX <= xtol
    assert lost > 0.01
