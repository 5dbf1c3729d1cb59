"""The P2G lane order when points crowd their dual cells, against the oracle through the C-ABI.

The P2G pass places the point of in-block dual cell (x, y, z) on lane x + 4z + 16y of its
wavefront (2D: y + 8x); further points of an already claimed dual cell, and points that
have left the domain, fill the lanes no dual cell claimed. At 4 points per cell per
dimension every dual cell holds 4^D points, so nearly every wavefront has collisions; with
duplicated positions whole wavefronts share a few dual cells. Both forms, tolerances of
tests/_harness.py (FP64) and of tests/test_gpu_precision.py (precision form).
"""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_PREC_FP32
from tests._harness import compare_states, gpu_of, oracle_of
from tests.test_gpu_precision import _errors, _gpu_state, _ora_state

pytestmark = pytest.mark.gpu


def _crowded(dim, k, duplicate=1, seed=11):
    n = 12
    h = np.array([1.0 / n] * dim)
    lo = (3.0 / n,) * dim
    hi = (7.0 / n,) * dim  # 4 cells per dimension
    m, v, x, xt = W.lattice_block(dim, h, lo, hi, k, 1000.0, jitter=0.3 / k, seed=seed)
    if duplicate > 1:  # the same positions several times over
        m, v, x, xt = (np.concatenate([a] * duplicate) for a in (m / duplicate, v / duplicate, x, xt))
    rng = np.random.default_rng(seed)
    xt = rng.normal(0, 0.3, size=xt.shape)
    body = W.Body("crowd", m, v, x, xt, np.ones(len(m), np.int32), 1e6, 0.3, True)
    g = (0.0, -9.81) if dim == 2 else (0.0, 0.0, -9.81)
    return W.Workload(f"crowd{dim}d_k{k}_x{duplicate}", dim, (1.0,) * dim, (n,) * dim, [body], [], 2e-5, 40, g)


@pytest.mark.parametrize("dim,k,dup", [(3, 4, 1), (3, 2, 8), (2, 4, 1), (2, 2, 6)])
def test_crowded_dual_cells_fp64(dim, k, dup):
    w = _crowded(dim, k, dup)
    ora = oracle_of(w)
    gpu = gpu_of(w)
    ora.step(w.dt, 40)
    gpu.step(w.dt, 40)
    compare_states(w, ora, gpu, 40 * w.dt, label=w.name)


@pytest.mark.parametrize("dim,k,dup", [(3, 4, 1), (3, 2, 8)])
def test_crowded_dual_cells_precision_form(dim, k, dup):
    w = _crowded(dim, k, dup)
    ora = oracle_of(w)
    g = gpu_of(w, precision=MPM_PREC_FP32)
    ora.step(w.dt, 40)
    g.step(w.dt, 40)
    ro, rg = _ora_state(ora, 0), _gpu_state(g, 0)
    np.testing.assert_array_equal(rg["active"], ro["active"])
    e = _errors(w, ro, rg, 40 * w.dt)
    print(w.name, e)
    assert e["x_cells"] <= 1e-5 and e["v_rel"] <= 1e-5 and e["T_rel"] <= 1e-4 and e["V_rel"] <= 1e-5, e
