"""C-ABI argument and state errors on a live handle, through the C-ABI.

Every rejected call returns its MPM_ERR_* code, with a message (mpm_gpu_last_error) naming
the problem, and changes nothing. After a volley of rejected calls between its steps, a
handle reaches the same state as one that never saw them; the two differ only by the order
of FP64 atomic sums (positions to 1e-12 h, where a call that had changed the grid, a body or
the options would move them by far more). The descriptor errors that restate the
reference's keep its text ("Unrecognized Argument for use_cpdi", "ERROR: Unknown SPEC in
contact"). All of these are host-side checks made before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from mpm_3d_amd import gpu as G
from mpm_3d_amd import workloads as W
from tests._harness import gpu_of

pytestmark = pytest.mark.gpu

ARG, STATE = -1, -3
DP = C.POINTER(C.c_double)


def _expect(lib, rc, code, text):
    msg = lib.mpm_gpu_last_error().decode()
    assert rc == code, (rc, msg)
    assert text in msg, msg


def _workload():
    return W.two_block_3d(n_cells=16, gap_cells=1, steps=10)


def _desc(w, keep, **grid):
    d = G._sim_desc(w, keep)
    for k, v in grid.items():
        if k == "N":
            for i, n in enumerate(v):
                d.grid.N[i] = n
        else:
            setattr(d.grid, k, v)
    return d


def _volley(g, w):
    lib, h = g.lib, g.h
    n = w.bodies[0].n
    buf = np.zeros(3 * n)
    p = buf.ctypes.data_as(DP)
    act = np.ones(n, np.int32)
    pa = act.ctypes.data_as(C.POINTER(C.c_int32))
    _expect(lib, lib.mpm_gpu_upload(h, 99, p, p, None, p, p, p, pa), ARG, "body 99 out of range")
    _expect(lib, lib.mpm_gpu_upload(h, 0, None, p, None, p, p, p, pa), ARG, "are required")
    _expect(lib, lib.mpm_gpu_step(h, 0.0, 1), ARG, "dt must be positive and finite")
    _expect(lib, lib.mpm_gpu_step(h, float("nan"), 1), ARG, "dt must be positive and finite")
    _expect(lib, lib.mpm_gpu_step(h, w.dt, -1), ARG, "nsteps must be >= 0")
    dts = np.full(2, w.dt)
    _expect(lib, lib.mpm_gpu_step_batch(h, dts.ctypes.data_as(DP), 2, 1), ARG, "need 1 time steps, got 2")
    _expect(lib, lib.mpm_gpu_set_option(h, 99, 1), ARG, "unknown option 99")
    _expect(lib, lib.mpm_gpu_set_option(h, G.MPM_OPT_GRID_WGS, 0), ARG, "workgroup count out of range")
    _expect(lib, lib.mpm_gpu_set_option(h, G.MPM_OPT_SORT_INTERVAL, -1), ARG, "sort interval must be >= 0")
    v = C.c_int()
    _expect(lib, lib.mpm_gpu_get_option(h, 99, C.byref(v)), ARG, "unknown option 99")
    _expect(lib, lib.mpm_gpu_set_sort_interval(h, -3), ARG, "interval must be >= 0")
    cnt, ms = C.c_int64(), C.c_double()
    _expect(lib, lib.mpm_gpu_get_timing(h, 7, C.byref(cnt), C.byref(ms)), ARG, "kernel class out of range")
    _expect(lib, lib.mpm_gpu_download(h, 5, p, None, None, None, None), ARG, "body 5 out of range")
    _expect(lib, lib.mpm_gpu_download_contact(h, 3, None, None), ARG, "contact 3 out of range")
    _expect(lib, lib.mpm_gpu_connectivity(h, 0, None, None), ARG, "corner_elem and code are required")
    _expect(lib, lib.mpm_gpu_clear_sim(h, 4), ARG, "sim 4 out of range")
    # slot reconfiguration: every check runs before the slot is touched
    keep = []
    _expect(lib, lib.mpm_gpu_load_sim(h, 1, C.byref(_desc(w, keep))), ARG, "sim 1 out of range")
    _expect(lib, lib.mpm_gpu_load_sim(h, 0, C.byref(_desc(w, keep, N=(32, 16, 16)))), ARG,
            "need L > 0 and 1 <= N <= 16 in dimension 0")
    _expect(lib, lib.mpm_gpu_load_sim(h, 0, C.byref(_desc(w, keep, cpdi_spec=2))), ARG,
            "Unrecognized Argument for use_cpdi: 2")
    _expect(lib, lib.mpm_gpu_load_sim(h, 0, C.byref(_desc(w, keep, contact_spec=7))), ARG,
            "ERROR: Unknown SPEC in contact: 7")
    d = _desc(w, keep)
    d.bodies[1].n_points = w.bodies[1].n + 1
    _expect(lib, lib.mpm_gpu_load_sim(h, 0, C.byref(d)), ARG, f"body 1: at most {w.bodies[1].n} points")
    d = _desc(w, keep)
    d.bodies[0].has_boundary = 9
    _expect(lib, lib.mpm_gpu_load_sim(h, 0, C.byref(d)), ARG, "body 0: has_boundary must be 0..3, got 9")
    d = _desc(w, keep)
    d.contacts[0].body2 = d.contacts[0].body1
    _expect(lib, lib.mpm_gpu_load_sim(h, 0, C.byref(d)), ARG, "bodies must be two distinct sim-local indices")


def test_rejected_calls_change_nothing():
    w = _workload()
    h0 = w.L[0] / w.N[0]
    a = gpu_of(w)
    a.step(w.dt, 5)
    _volley(a, w)
    a.step(w.dt, 5)
    _volley(a, w)
    xa = [a.points(b)["x"] for b in range(2)]
    a.close()
    b = gpu_of(w)
    b.step(w.dt, 5)
    b.step(w.dt, 5)
    xb = [b.points(k)["x"] for k in range(2)]
    b.close()
    for k in range(2):
        err = float(np.max(np.abs(xa[k] - xb[k]))) / h0
        print(f"body {k}: max |dx| / h = {err:.2e}")
        assert err <= 1e-12


def test_state_errors():
    """Steps need every body of a reconfigured slot uploaded; node fields need a step since
    the last upload, and mass, momentum and force need the node record on during it."""
    w = _workload()
    g = gpu_of(w)
    lib, h = g.lib, g.h
    keep = []
    assert lib.mpm_gpu_load_sim(h, 0, C.byref(_desc(w, keep))) == 0
    _expect(lib, lib.mpm_gpu_step(h, w.dt, 1), STATE, "body 0 has not been uploaded")
    g.upload(0, w.bodies[0])
    _expect(lib, lib.mpm_gpu_step(h, w.dt, 1), STATE, "body 1 has not been uploaded")
    g.upload(1, w.bodies[1])
    nn = 17 ** 3
    m = np.zeros(nn)
    _expect(lib, lib.mpm_gpu_download_nodes(h, 0, m.ctypes.data_as(DP), None, None, None, None), STATE,
            "no step since the last upload or load")
    g.step(w.dt, 2)
    _expect(lib, lib.mpm_gpu_download_nodes(h, 0, m.ctypes.data_as(DP), None, None, None, None), STATE,
            "need MPM_OPT_NODE_RECORD on during the last step call")
    g.set_option(G.MPM_OPT_NODE_RECORD, 1)
    g.step(w.dt, 1)
    assert lib.mpm_gpu_download_nodes(h, 0, m.ctypes.data_as(DP), None, None, None, None) == 0
    assert abs(m.sum() - w.bodies[0].m.sum()) <= 1e-12 * w.bodies[0].m.sum()
    g.close()



def test_keep_L_off_refuses_the_velocity_gradient():
    """MPM_OPT_KEEP_L 0: no velocity-gradient plane; steps give the same state, the
    displacement still downloads, and asking for L returns MPM_ERR_STATE."""
    w = _workload()
    a = gpu_of(w)
    b = gpu_of(w)
    b.set_option(G.MPM_OPT_KEEP_L, 0)
    assert b.get_option(G.MPM_OPT_KEEP_L) == 0
    a.step(w.dt, 12)
    b.step(w.dt, 12)
    h = 1.0 / 16
    for gb in range(2):
        assert np.max(np.abs(a.points(gb)["x"] - b.points(gb)["x"])) <= 1e-12 * h
    n = w.bodies[0].n
    u = np.zeros(3 * n)
    L = np.zeros(9 * n)
    lib = b.lib
    assert lib.mpm_gpu_download_aux(b.h, 0, u.ctypes.data_as(DP), None) == 0
    ua = a.aux(0)["u"]
    assert np.max(np.abs(u.reshape(3, n).T - ua)) <= 1e-12 * h
    _expect(lib, lib.mpm_gpu_download_aux(b.h, 0, None, L.ctypes.data_as(DP)), STATE, "MPM_OPT_KEEP_L is 0")
    # turned back on, the next call writes L again
    b.set_option(G.MPM_OPT_KEEP_L, 1)
    b.step(w.dt, 1)
    a.step(w.dt, 1)
    assert np.max(np.abs(b.aux(0)["L"] - a.aux(0)["L"])) <= 1e-6 * max(1.0, np.max(np.abs(a.aux(0)["L"])))
    a.close()
    b.close()
