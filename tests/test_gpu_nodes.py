"""Node fields and contact fields of the last step, through the C-ABI (mpm_gpu_download_nodes,
mpm_gpu_download_contact with MPM_OPT_NODE_RECORD), against the oracle's nodes and contacts.

These are what DefaultNodes::writeFrame (default_nodes.cpp:68-82) and ContactHuang::writeFrame
(contact_huang.cpp:202-206) serialize: mass, momentum after moveGrid, force after contact and
boundary, velocity and its increment, and the contact normal and force. Nodes no point
reaches are 0 (normal: NaN) in both. Tolerances: mass 1e-12 relative; momentum, force,
velocity 1e-9 of their largest magnitude (summation order only); FP32 form (FP32 point
mass, velocity and weights, FP64 sums): mass 1e-6, the rest 1e-5 of their scale.
"""
import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_OPT_NODE_RECORD, MPM_PREC_FP32, MpmGpuError
from tests._harness import example_workload, gpu_of, oracle_of

pytestmark = pytest.mark.gpu


def _close(a, b, rel, what):
    scale = float(np.max(np.abs(b))) + 1e-300
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rel * scale, f"{what}: max error {err:.3e} vs scale {scale:.3e}"
    return err / scale


def _compare_nodes(w, ora, gpu, rel=1e-9, rel_v=1e-9, rel_m=1e-12):
    dim = w.dim
    for b in range(len(w.bodies)):
        o = ora.nodes(b)
        g = gpu.nodes(b)
        assert g["m"].shape == o["m"].shape
        _close(g["m"], o["m"], rel_m, f"body {b} mass")
        np.testing.assert_array_equal(g["m"] > 0, o["m"] > 0)
        for k in ("mxt", "f"):
            _close(g[k], o[k][:, :dim], rel, f"body {b} {k}")
        for k in ("xt", "dxt"):
            _close(g[k], o[k][:, :dim], rel_v, f"body {b} {k}")
        untouched = o["m"] == 0
        assert np.all(g["xt"][untouched] == 0) and np.all(g["f"][untouched] == 0)


@pytest.mark.parametrize("name,steps", [("blockdrop2D", 30), ("collision2D", 60), ("collision3D", 20)])
def test_nodes_match_oracle(name, steps):
    w = example_workload(name)
    ora = oracle_of(w)
    gpu = gpu_of(w)
    gpu.set_option(MPM_OPT_NODE_RECORD, 1)
    ora.step(w.dt, steps)
    gpu.step(w.dt, steps)
    _compare_nodes(w, ora, gpu)


def test_contact_fields_match_oracle():
    """Two jittered 3D blocks closing at 2 m/s: contact engages; normal and force per node."""
    w = W.two_block_3d(n_cells=32, gap_cells=1, steps=30, E=1e6)
    for b in w.bodies:
        b.x = b.x + np.random.default_rng(5).uniform(-0.1, 0.1, b.x.shape) / 32
    ora = oracle_of(w)
    gpu = gpu_of(w)
    gpu.set_option(MPM_OPT_NODE_RECORD, 1)
    ora.step(w.dt * 10, 30)
    gpu.step(w.dt * 10, 30)
    on, of = ora.contact(0)
    gn, gf = gpu.contact(0)
    on, of = on[:, :3], of[:, :3]
    engaged = np.any(of != 0, axis=1)
    assert engaged.sum() > 20
    np.testing.assert_array_equal(np.any(gf != 0, axis=1), engaged)
    _close(gf, of, 1e-8, "contact force")
    _close(gn[engaged], on[engaged], 1e-8, "contact normal at engaged nodes")
    # nodes no point of body 1 reaches: 0/0 in both
    m1 = ora.nodes(0)["m"]
    far = m1 == 0
    no_grad = far & np.all(np.isnan(on), axis=1)
    assert no_grad.sum() > 100 and np.all(np.isnan(gn[no_grad]))
    _compare_nodes(w, ora, gpu, rel=1e-8, rel_v=1e-8)


def test_nodes_precision_form():
    w = example_workload("collision3D")
    ora = oracle_of(w)
    gpu = gpu_of(w, precision=MPM_PREC_FP32)
    gpu.set_option(MPM_OPT_NODE_RECORD, 1)
    ora.step(w.dt, 20)
    gpu.step(w.dt, 20)
    _compare_nodes(w, ora, gpu, rel=1e-5, rel_v=1e-5, rel_m=1e-6)


def test_nodes_without_record_and_before_steps():
    w = example_workload("collision2D")
    gpu = gpu_of(w)
    with pytest.raises(MpmGpuError):
        gpu.nodes(0, record=False)  # no step yet
    gpu.step(w.dt, 3)
    with pytest.raises(MpmGpuError):
        gpu.nodes(0)  # m, mxt, f were not recorded
    with pytest.raises(MpmGpuError):
        gpu.contact(0)
    g = gpu.nodes(0, record=False)
    ora = oracle_of(w)
    ora.step(w.dt, 3)
    _close(g["xt"], ora.nodes(0)["xt"][:, :2], 1e-9, "velocity without record")


def test_nodes_in_a_batch():
    """Two sims of different grids in one handle: each body's nodes in its own sim's numbering."""
    a = example_workload("collision2D")   # 20 x 20 grid, two bodies
    b = W.sliding_block(2, 3, 0.3)        # 40 x 40 grid, frictional floor
    gpu = gpu_of([a, b])
    gpu.set_option(MPM_OPT_NODE_RECORD, 1)
    gpu.step(1e-4, 20)
    for w, gb0 in ((a, 0), (b, 2)):
        ora = oracle_of(w)
        ora.step(1e-4, 20)
        for bi in range(len(w.bodies)):
            o = ora.nodes(bi)
            g = gpu.nodes(gb0 + bi)
            _close(g["m"], o["m"], 1e-12, "batched mass")
            _close(g["mxt"], o["mxt"][:, :2], 1e-9, "batched momentum")
            _close(g["xt"], o["xt"][:, :2], 1e-9, "batched velocity")
