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
