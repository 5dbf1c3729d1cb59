"""Points outside the particle passes' LDS path, against the oracle through the C-ABI.

The split passes hand such points to side kernels in some forms (P2G of the FP64 form, G2P
of the precision form) and process them inline in the others; these cases make them common:
* CPDI off (cpdi_spec = 0): every point takes the per-tuple path;
* extents above one cell on every third point: the per-tuple path for those;
* a block moving 2.4 cells in one call with no re-sort (sort interval 0 sorts once per
  call): windows leave their particle block's LDS patch and take the global path.
Tolerances of tests/_harness.py (FP64) and tests/test_gpu_precision.py (precision form).
"""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_PREC_FP32
from tests._harness import compare_states, gpu_of, oracle_of
from tests.test_gpu_precision import _errors, _gpu_state, _ora_state

pytestmark = pytest.mark.gpu


def _block(dim, velocity=None, cpdi=1, big_every=0, steps=60, seed=21):
    n = 24
    h = np.array([1.0 / n] * dim)
    m, v, x, xt = W.lattice_block(dim, h, (0.2,) * dim, (0.45,) * dim, 2, 1000.0, velocity=velocity, jitter=0.2,
                                  seed=seed)
    if big_every:
        v[::big_every] *= 100.0  # extent 0.5 * cbrt(V) (2D: sqrt) above one cell
    body = W.Body("blk", m, v, x, xt, np.ones(len(m), np.int32), 1e6, 0.3, True)
    g = (0.0, -9.81) if dim == 2 else (0.0, 0.0, -9.81)
    w = W.Workload(f"side{dim}d", dim, (1.0,) * dim, (n,) * dim, [body], [], 1e-4, steps, g)
    w.cpdi_spec = cpdi
    return w


CASES = {
    "cpdi_off": dict(cpdi=0),
    "big_extents": dict(big_every=3),
    "drift_unsorted": dict(velocity=(10.0, 4.0, 2.0), steps=100),
}


def _run(w, **kw):
    ora = oracle_of(w)
    gpu = gpu_of(w, **kw)
    ora.step(w.dt, w.steps)
    gpu.step(w.dt, w.steps)
    return ora, gpu


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("case", list(CASES))
def test_side_paths_fp64(case, dim):
    w = _block(dim, **CASES[case])
    ora, gpu = _run(w, sort_interval=0)
    compare_states(w, ora, gpu, w.steps * w.dt, label=f"{case} {dim}D")


@pytest.mark.parametrize("case", list(CASES))
def test_side_paths_precision_form(case):
    w = _block(3, **CASES[case])
    ora, g = _run(w, sort_interval=0, precision=MPM_PREC_FP32)
    ro, rg = _ora_state(ora, 0), _gpu_state(g, 0)
    np.testing.assert_array_equal(rg["active"], ro["active"])
    e = _errors(w, ro, rg, w.steps * w.dt)
    print(case, e)
    assert e["x_cells"] <= 1e-5 and e["v_rel"] <= 1e-5 and e["T_rel"] <= 1e-4 and e["V_rel"] <= 1e-5, e
