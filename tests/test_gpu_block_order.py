"""Block order of the particle passes on the 16M-point column, through the C-ABI.

With many particle blocks per workgroup the 3D P2G passes and the FP64 G2P take their
blocks from a device counter; with a large enough workgroup count (MPM_OPT_P2G_GRID /
MPM_OPT_PASS_GRID) they keep the fixed stride. A block missed or taken twice would change
the state by whole point masses; the two orders may differ only by the order of the FP64
atomic sums. The counter is cleared by the last workgroup of each launch, which must also
hold under graph replay (MPM_OPT_GRAPH, several sort periods).

Tolerances: positions 1e-9 h (FP64 form) and 1e-6 h (precision form: an FP64 sum that rounds
the other way can flip the last bit of an FP32 grid output).
"""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_OPT_GRAPH, MPM_OPT_P2G_GRID, MPM_OPT_PASS_GRID, MPM_PREC_FP32, MPM_PREC_FP64
from tests._harness import gpu_of

pytestmark = pytest.mark.gpu

STEPS = 40


def _x(g):
    return g.points(0)["x"]


@pytest.mark.parametrize("prec,tol", [(MPM_PREC_FP64, 1e-9), (MPM_PREC_FP32, 1e-6)])
def test_counter_order_matches_fixed_stride_and_graph_replay(prec, tol):
    w = W.column_3d(wide=True)
    h = w.L[0] / w.N[0]
    a = gpu_of(w, precision=prec)  # ~31k blocks, 1024 workgroups: the counter
    a.step(w.dt, STEPS)
    xa = _x(a)
    a.close()
    b = gpu_of(w, precision=prec)  # 4096 workgroups: fewer than 8 blocks each, fixed stride
    b.set_option(MPM_OPT_PASS_GRID, 4096)
    b.set_option(MPM_OPT_P2G_GRID, 4096)
    b.step(w.dt, STEPS)
    xb = _x(b)
    b.close()
    c = gpu_of(w, precision=prec)  # the counter, under graph replay
    c.set_option(MPM_OPT_GRAPH, 1)
    c.step(w.dt, STEPS)
    xc = _x(c)
    c.close()
    for name, x in (("fixed stride", xb), ("graph replay", xc)):
        err = float(np.max(np.abs(x - xa))) / h
        print(f"{name}: max |dx| / h = {err:.3e}")
        assert err <= tol, name
