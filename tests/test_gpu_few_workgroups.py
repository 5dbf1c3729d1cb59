"""Particle passes with fewer persistent workgroups than particle blocks, through the C-ABI.

A workgroup that runs several particle blocks takes the paths a single block never does:
the G2P pass loads the next block's node patch by LDS DMA while it runs the current block
and lays it out in LDS at the next one (zeros outside the domain; the precision form moves
four nodes per access), and the P2G pass keeps its accumulator patch across blocks, cleared
by its flush. The small workloads here have fewer blocks than the default 1024 workgroups,
so MPM_OPT_PASS_GRID (both particle passes) is lowered: 1 workgroup takes its blocks from
the device counter, 3 keep the fixed stride on these sizes. The workloads put window slots
over the edge of the grid into the patches (a block on the floor, a column collapsing
against the walls) and run two bodies with contact normals.

Checks: the FP64 form against the oracle at the parity bounds of tests/_harness.py, and both
forms against the same library at its default workgroup count, where every block stages its
patch from global memory: positions to 1e-9 h (FP64) and 1e-6 h (precision form; only the
order of the FP64 atomic sums differs, which can flip the last bit of an FP32 grid output).
"""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_OPT_PASS_GRID, MPM_PREC_FP32, MPM_PREC_FP64
from tests._harness import compare_states, example_workload, gpu_of, oracle_of

pytestmark = pytest.mark.gpu


def _contact_engaged():
    w = W.two_block_3d(n_cells=32, gap_cells=1, steps=60, E=1e6)
    for b in w.bodies:
        b.x = b.x + np.random.default_rng(5).uniform(-0.1, 0.1, b.x.shape) / 32
    w.dt = w.dt * 10
    return w


CASES = {
    "slider3d": (lambda: W.sliding_block(3, 3, 0.3), 100),
    "collapse2d": (lambda: W.column_collapse_example(2), 100),
    "collision3d": (lambda: example_workload("collision3D"), 50),
    "contact3d": (_contact_engaged, 60),
}


@pytest.mark.parametrize("wgs", [1, 3])
@pytest.mark.parametrize("case", list(CASES))
def test_few_workgroups(case, wgs):
    make, steps = CASES[case]
    w = make()
    h = min(w.L[d] / w.N[d] for d in range(w.dim))
    print(f"{case}: {sum(len(b.m) for b in w.bodies)} points, {wgs} workgroup(s)", flush=True)
    ora = oracle_of(w)
    ora.step(w.dt, steps)
    for prec, tol in ((MPM_PREC_FP64, 1e-9), (MPM_PREC_FP32, 1e-6)):
        few = gpu_of(w, precision=prec)
        few.set_option(MPM_OPT_PASS_GRID, wgs)
        few.step(w.dt, steps)
        ref = gpu_of(w, precision=prec)
        ref.step(w.dt, steps)
        if prec == MPM_PREC_FP64:
            compare_states(w, ora, few, steps * w.dt, label=f"{case} {wgs} wg")
        for bi in range(len(w.bodies)):
            pf, pr = few.points(bi), ref.points(bi)
            np.testing.assert_array_equal(pf["active"], pr["active"])
            err = float(np.max(np.abs(pf["x"] - pr["x"]))) / h if len(pf["x"]) else 0.0
            print(f"{case} body {bi} prec {prec}: |dx|/h against the default workgroup count {err:.3e}")
            assert err <= tol
        few.close()
        ref.close()
