"""mpm_gpu_connectivity_noalloc (no device allocation; its kernel writes into the library's
fixed device array) against mpm_gpu_connectivity (per-call device buffers), through the
C-ABI: identical corner elements and codes, bit for bit, on two calls in a row where the
second has fewer points, in both forms and after steps (so the slots are in sorted order
and some points have left the domain); bodies above the array's size are refused."""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_PREC_FP32, MPM_PREC_FP64, GpuSim, MpmGpuError
from tests._harness import example_workload

pytestmark = pytest.mark.gpu

MAX_POINTS = 524288  # MPM_NOALLOC_MAX_POINTS


def _both(sim, gb):
    e0, c0 = sim.connectivity(gb)
    e1, c1 = sim.connectivity(gb, noalloc=True)
    np.testing.assert_array_equal(c1, c0)
    np.testing.assert_array_equal(e1, e0)
    return len(c0), np.bincount(c0, minlength=3)


def _escaping_block():
    """A 3D block of 216,000 points moving fast towards the +x wall of a 60^3 grid, so that
    some points leave the domain (code 2) and the wall layer takes the CPDI-off branch."""
    h = np.full(3, 1.0 / 60)
    m, v, x, xt = W.lattice_block(3, h, (0.55, 0.2, 0.2), (0.95, 0.6, 0.6), 2, 1000.0, velocity=(40.0, 0, 0),
                                  jitter=0.2, seed=3)
    body = W.Body("block", m, v, x, xt, np.ones(len(m), np.int32), 1e6, 0.3, False)
    return W.Workload("escaping_block", 3, (1.0,) * 3, (60,) * 3, [body], [], 1e-4, 40, (0.0, 0.0, -9.81))


@pytest.mark.parametrize("prec", [MPM_PREC_FP64, MPM_PREC_FP32])
def test_noalloc_matches_connectivity(prec):
    big = _escaping_block()
    small = example_workload("collision3D")
    assert big.n_points <= MAX_POINTS
    a = GpuSim(big, precision=prec)
    a.step(big.dt, 40)
    n1, counts = _both(a, 0)
    print(f"call 1: {n1} points, codes {counts.tolist()}")
    assert counts[2] > 0 and counts[1] > 0
    b = GpuSim(small, precision=prec)
    b.step(small.dt, 7)
    n2, _ = _both(b, 1)
    n3, _ = _both(b, 0)
    assert n2 < n1 and n3 < n1
    # and the first handle again, after the smaller calls
    _both(a, 0)
    a.close()
    b.close()


def test_noalloc_refuses_above_its_size():
    h = np.full(3, 1.0 / 64)
    m, v, x, xt = W.lattice_block(3, h, (11 / 64, 11 / 64, 11 / 64), (53 / 64, 53 / 64, 53 / 64), 2, 1000.0)
    body = W.Body("block", m, v, x, xt, np.ones(len(m), np.int32), 1e6, 0.3, True)
    w = W.Workload("big_block", 3, (1.0,) * 3, (64,) * 3, [body], [], 1e-5, 1, (0.0, 0.0, -9.81))
    assert w.n_points > MAX_POINTS
    sim = GpuSim(w)
    with pytest.raises(MpmGpuError, match="at most 524288"):
        sim.connectivity(0, noalloc=True)
    e0, c0 = sim.connectivity(0)  # the allocating entry point still serves it
    assert (c0 == 1).all()
    sim.close()
