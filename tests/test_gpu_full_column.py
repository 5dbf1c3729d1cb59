"""Size-independent properties of the P2G map on the full 16M-point column, through the C-ABI.

The nodal mass of a step is sum_p S_ip m_p. The GIMP/CPDI weights sum to one over the nodes
and reproduce linear functions (sum_i S_ip x_i = x_p: the corners are symmetric about the
point, and CPDI-off points of the floor layer use the linear basis at the point), so over
the grid
    sum_i m_i = sum_p m_p   and   sum_i m_i x_i = sum_p m_p x_p
for the positions the step started from (the oracle holds both to 1e-13:
tests/test_oracle_pins.py::test_p2g_mass_first_moment_reproduces_positions). Checked at
step 40 of the 15,990,784-point column on 256x256x128 (after sorts at steps 1, 17 and 33,
with ~31k particle blocks taken from the block counter): a particle block dropped or taken
twice, a lane that scatters another lane's point, a node index or patch origin off by one
or a corner weight lost would move the mass by whole point masses or the moment by M h.

Tolerances: 1e-12 relative (FP64 form: rounding of the 16M-term sums); precision form 1e-6
(FP32 weights; the point masses are the FP32 values the form stores).

A rigid translation closes the whole step at the same size: a 15.86M-point block on the same
grid, 3 or more cells from every wall, with zero gravity and zero stress, moving at a uniform
v0. Every node it touches gets p_i / m_i = v0, no force, and the gradients of the map sum to
zero, so after k steps x = x0 + k dt v0, v = v0 and T = 0 (the oracle holds this to 5e-14 h,
1e-15 m/s and 3e-10 Pa on a small block). Tolerances: FP64 form 1e-9 h, 1e-12 m/s and 1e-3 Pa;
precision form 1e-5 h (its target), 1e-6 m/s and 1 Pa (FP32 gradients sum to zero only to
about 1e-7 / h, a strain rate of 1e-5 / s at v0 = 1 m/s).
"""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_OPT_NODE_RECORD, MPM_PREC_FP32, MPM_PREC_FP64
from tests._harness import gpu_of

pytestmark = pytest.mark.gpu


def _node_x(N, h):
    """Node positions in the reference's node order (dimension 0 fastest)."""
    nx, ny = N[0] + 1, N[1] + 1
    idx = np.arange(nx * ny * (N[2] + 1))
    return np.stack([idx % nx, (idx // nx) % ny, idx // (nx * ny)], 1) * np.asarray(h)


@pytest.mark.parametrize("prec,rel", [(MPM_PREC_FP64, 1e-12), (MPM_PREC_FP32, 1e-6)])
def test_full_column_mass_and_first_moment(prec, rel):
    w = W.column_3d(wide=True)
    h = [w.L[d] / w.N[d] for d in range(3)]
    g = gpu_of(w, precision=prec)
    g.step(w.dt, 39)
    pt = g.points(0)
    g.set_option(MPM_OPT_NODE_RECORD, 1)
    g.step(w.dt, 1)
    nd = g.nodes(0)
    g.close()
    m = np.asarray(w.bodies[0].m, np.float64)
    if prec == MPM_PREC_FP32:
        m = m.astype(np.float32).astype(np.float64)
    act = pt["active"] != 0
    assert act.all(), "every point of the column stays in the domain"
    M = float(m.sum())
    mass = float(nd["m"].sum())
    print(f"mass: nodes {mass:.15e}, points {M:.15e}, rel {(mass - M) / M:.2e}")
    assert abs(mass - M) <= rel * M
    xn = _node_x(w.N, h)
    got = nd["m"] @ xn
    want = m @ pt["x"]
    err = np.abs(got - want) / np.abs(want)
    print(f"first moment rel err {err}")
    assert np.all(err <= rel), (got, want)


@pytest.mark.parametrize("prec,tol_x,tol_v,tol_T", [(MPM_PREC_FP64, 1e-9, 1e-12, 1e-3),
                                                   (MPM_PREC_FP32, 1e-5, 1e-6, 1.0)])
def test_full_size_rigid_translation(prec, tol_x, tol_v, tol_T):
    N, L = (256, 256, 128), (2.0, 2.0, 1.0)
    h = np.array([L[d] / N[d] for d in range(3)])
    v0 = np.array([1.0, 0.5, -0.25])
    m, vol, x, xt = W.lattice_block(3, h, (0.5, 0.5, 3 * h[2]), (1.5, 1.5, 124 * h[2]), 2, 1500.0, velocity=v0,
                                    jitter=0.25, seed=7)
    body = W.Body("block", m, vol, x, xt, np.ones(len(m), np.int32), 1e7, 0.3)
    w = W.Workload(f"translate_{len(m)}pts", 3, L, N, [body], [], 1e-5, 40, (0.0, 0.0, 0.0))
    g = gpu_of(w, precision=prec)
    g.step(w.dt, 40)
    pt = g.points(0)
    g.close()
    assert np.all(pt["active"] != 0)
    ex = float(np.max(np.abs(pt["x"] - (x + 40 * w.dt * v0)))) / h[0]
    ev = float(np.max(np.abs(pt["xt"] - v0)))
    eT = float(np.max(np.abs(pt["T6"])))
    print(f"{len(m)} points: x {ex:.2e} h, v {ev:.2e} m/s, T {eT:.2e} Pa")
    assert ex <= tol_x and ev <= tol_v and eT <= tol_T
