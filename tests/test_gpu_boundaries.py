"""GPU parity against the CPU oracle for the three box boundaries, through the C-ABI:
CartesianBox (every component on a face node), CartesianSmoothBox (the face-normal
component, cartesian_smooth_box.cpp:70-84) and CartesianFrictionalBox (normal component
plus Coulomb friction against the closing impulse, cartesian_frictional_box.cpp:98-151).

Blocks slide along the floor, the ceiling and a side wall in 2D and 3D, some of their
points moving into the face and some away from it, so the face has closing and separating
nodes; one case sits in a corner (nodes on two faces), one puts two frictional blocks in
ContactHuang contact on the floor (the friction takes its normal impulse from the force
before contact). Tolerances are those of tests/_harness.py.
"""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_PREC_FP32
from tests._harness import compare_states, gpu_of, oracle_of

pytestmark = pytest.mark.gpu

FACES = [(2, 1, False), (2, 1, True), (2, 0, False), (3, 2, False), (3, 2, True), (3, 0, True)]


def _slider(dim, wd, upper, bc, mu, corner=False):
    w = W.sliding_block(dim, bc, mu, wall_dim=wd, upper=upper)
    b = w.bodies[0]
    t0 = [d for d in range(dim) if d != wd][0]
    if corner:
        b.x[:, t0] -= 0.3
    into = 1.0 if upper else -1.0
    low = b.x[:, t0] < np.median(b.x[:, t0])
    b.xt[:, wd] = np.where(low, -0.5 * into, 0.05 * into)
    return w


def _check(w, steps, label, **kw):
    ora = oracle_of(w)
    gpu = gpu_of(w, **kw)
    ora.step(w.dt, steps)
    gpu.step(w.dt, steps)
    return compare_states(w, ora, gpu, steps * w.dt, label=label)


@pytest.mark.parametrize("bc,mu", [(1, 0.0), (2, 0.0), (3, 0.3), (3, 1e12)])
@pytest.mark.parametrize("dim,wd,upper", FACES)
def test_box_boundaries_match_oracle(dim, wd, upper, bc, mu):
    steps = 200 if dim == 2 else 100
    w = _slider(dim, wd, upper, bc, mu)
    _check(w, steps, f"{w.name}x{steps}")


@pytest.mark.parametrize("bc,mu", [(2, 0.0), (3, 0.3)])
@pytest.mark.parametrize("dim", [2, 3])
def test_box_boundaries_in_a_corner(dim, bc, mu):
    w = _slider(dim, dim - 1, False, bc, mu, corner=True)
    _check(w, 150, f"corner {w.name}")


def test_frictional_box_under_contact():
    """Two blocks on a frictional floor slide into each other under ContactHuang."""
    a = W.sliding_block(2, 3, 0.3, speed=1.0)
    b = W.sliding_block(2, 3, 0.3, speed=-1.0)
    b.bodies[0].x[:, 0] += 0.2 + 1.0 / 40
    a.bodies.append(b.bodies[0])
    a.contacts.append(W.Contact(0, 1, 0.4))
    a.name = "two sliders"
    ora = oracle_of(a)
    gpu = gpu_of(a)
    ora.step(a.dt, 300)
    gpu.step(a.dt, 300)
    # they met and bounced: each block's mean velocity along the floor has turned round
    assert ora.points(0)["xt"][:, 0].mean() < 0 < ora.points(1)["xt"][:, 0].mean()
    compare_states(a, ora, gpu, 300 * a.dt, label="two sliders")


def test_frictional_box_precision_form():
    """The precision form on the frictional floor, at the bounds of test_gpu_precision."""
    from tests.test_gpu_precision import _errors, _gpu_state, _ora_state
    w = _slider(3, 2, False, 3, 0.3)
    ora = oracle_of(w)
    g = gpu_of(w, precision=MPM_PREC_FP32)
    ora.step(w.dt, 100)
    g.step(w.dt, 100)
    ro, rg = _ora_state(ora, 0), _gpu_state(g, 0)
    np.testing.assert_array_equal(rg["active"], ro["active"])
    e = _errors(w, ro, rg, 100 * w.dt)
    print(f"frictional fp32: {e}")
    assert e["x_cells"] <= 1e-5 and e["v_rel"] <= 1e-5 and e["T_rel"] <= 1e-4 and e["V_rel"] <= 1e-5, e
