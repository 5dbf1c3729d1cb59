"""The served stream against the oracle: requests drawn from the served-stream generator
(block drops and two-block ContactHuang collisions on N^3 grids, 2^3 points per cell) pass
through a slot pool with fewer slots than requests, so that slots are reconfigured
(``mpm_gpu_load_sim``), refilled with fewer points than their capacity (``mpm_gpu_upload``
into the free slots of a layout that still holds the other slot's points) and freed
(``mpm_gpu_clear_sim``) while the other slot keeps stepping with its own dt. Every served
result is compared with a fresh oracle run of the same request, which is what the
reference does per request: a new Job (config.cpp:176-213, job.hpp:76-105) stepped by
ExplicitUSL::step (explicit_usl.cpp:48-87).

Bounds: the FP64 pool at the parity bounds of tests/_harness.py (positions 1e-9 h,
velocities 1e-8, stresses 1e-7 of their scales, volumes 1e-12 relative); the precision
pool at the precision-form bounds of tests/test_gpu_precision.py (1e-5 h, 1e-5, 1e-4,
1e-5). Step counts are cut to 32..96 so the oracle (serial, one thread per request)
finishes in seconds; sizes are the generator's smallest (N <= 40, 1e4..3e4 points)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from mpm_3d_amd import serve as S
from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_PREC_FP32, MPM_PREC_FP64, stress6_to9
from tests._harness import compare_states, oracle_of

pytestmark = pytest.mark.gpu


def _requests():
    """Six of the generator's smallest requests, three drops and three collisions, with
    step counts that differ so that slots free up at different times."""
    pool = [(w, s) for w, s in W.serve_requests(200, seed=31) if w.N[0] <= 40 and w.n_points <= 30000]
    drops = [r for r in pool if len(r[0].bodies) == 1][:3]
    coll = [r for r in pool if len(r[0].bodies) == 2][:3]
    reqs = [r for pair in zip(drops, coll) for r in pair]
    assert len(reqs) == 6
    steps = [32, 96, 64, 48, 80, 40]
    return [(w, k) for (w, _), k in zip(reqs, steps)]


def _oracle_result(req):
    w, steps = req
    ora = oracle_of(w)
    ora.step(w.dt, steps)
    out = [ora.points(b) for b in range(len(w.bodies))]
    ora.close()
    return out


class _Served:
    """The served result of one request, read like a GpuSim by compare_states."""

    def __init__(self, bodies):
        self.bodies = bodies

    def points(self, b):
        return self.bodies[b]


class _Ora:
    def __init__(self, bodies):
        self.bodies = bodies

    def points(self, b):
        return self.bodies[b]


@pytest.fixture(scope="module")
def oracle_results():
    reqs = _requests()
    # independent oracle handles, one serial run per request; ctypes releases the GIL
    with ThreadPoolExecutor(max_workers=6) as ex:
        return reqs, list(ex.map(_oracle_result, reqs))


def _serve(reqs, precision):
    caps = (max(w.bodies[0].n for w, _ in reqs), max(w.bodies[-1].n for w, _ in reqs))
    pool = S.SlotPool(2, dim=3, N_cap=40, body_caps=caps, precision=precision)
    log = []
    out = S.run_stream(pool, reqs, [0.0] * len(reqs), chunk_max=16, keep_results=True, log=log.append)
    pool.close()
    return out, log


def test_served_fp64_matches_oracle(oracle_results):
    reqs, ora = oracle_results
    out, log = _serve(reqs, MPM_PREC_FP64)
    # two slots for six requests: four admissions reuse a slot while the other slot steps
    assert out["chunks"] > len(reqs)
    starts = sorted(out["start"])
    assert starts[2] > 0.0
    for r, (w, steps) in enumerate(reqs):
        compare_states(w, _Ora(ora[r]), _Served(out["results"][r]), steps * w.dt, label=f"served {w.name}")


def test_served_precision_matches_oracle(oracle_results):
    reqs, ora = oracle_results
    out, _ = _serve(reqs, MPM_PREC_FP32)
    for r, (w, steps) in enumerate(reqs):
        h = w.L[0] / w.N[0]
        t = steps * w.dt
        for b in range(len(w.bodies)):
            o, g = ora[r][b], out["results"][r][b]
            np.testing.assert_array_equal(g["active"], o["active"])
            vscale = float(np.max(np.abs(o["xt"]))) + 9.81 * t + 1e-12
            tscale = float(np.max(np.abs(o["T"]))) + w.bodies[b].E * t * vscale / h + 1e-300
            ex = float(np.max(np.abs(g["x"] - o["x"]))) / h
            ev = float(np.max(np.abs(g["xt"] - o["xt"]))) / vscale
            eT = float(np.max(np.abs(stress6_to9(g["T6"]) - o["T"]))) / tscale
            eV = float(np.max(np.abs(g["v"] - o["v"]) / np.abs(o["v"])))
            print(f"served fp32 {w.name} body {b}: |dx|/h {ex:.2e} |dv| {ev:.2e} |dT| {eT:.2e} |dV| {eV:.2e}")
            assert ex <= 1e-5 and ev <= 1e-5 and eT <= 1e-4 and eV <= 1e-5


def test_served_slot_classes_match_oracle(oracle_results):
    """Two slot classes: the 27,000-point drops fit only the large slot, the others take the
    small slot first; every result against the oracle (FP64 pool, parity bounds)."""
    reqs, ora = oracle_results
    classes = [S.SlotClass(1, 40, (27000, 13824)), S.SlotClass(1, 39, (17576, 13824))]
    pool = S.SlotPool(classes=classes, dim=3, precision=MPM_PREC_FP64)
    assert pool.capacity(1) < pool.capacity(0)
    out = S.run_stream(pool, reqs, [0.0] * len(reqs), chunk_max=16, keep_results=True)
    pool.close()
    assert out["status"] == ["done"] * len(reqs)
    for r, (w, steps) in enumerate(reqs):
        compare_states(w, _Ora(ora[r]), _Served(out["results"][r]), steps * w.dt, label=f"classes {w.name}")


def test_reupload_next_to_a_stepping_sim():
    """A two-simulation batch steps; simulation 0 is then reloaded with fewer points than its
    capacity (its points go into the free slots of the layout simulation 1 still sorts)
    and both step on. Both forms: simulation 0 against the oracle of its new state, and
    simulation 1 against the oracle of its whole run, which never saw the re-upload."""
    reqs = _requests()
    w_big, w_small, w_other = reqs[2][0], reqs[0][0], reqs[3][0]  # two drops; a collision beside them
    assert w_small.bodies[0].n < w_big.bodies[0].n
    k1, k2 = 40, 56
    ora_other = oracle_of(w_other)
    ora_other.step(w_other.dt, k1 + k2)
    ora_small = oracle_of(w_small)
    ora_small.step(w_small.dt, k2)
    for prec in (MPM_PREC_FP64, MPM_PREC_FP32):
        caps = (max(w_big.bodies[0].n, w_other.bodies[0].n), max(w_big.bodies[-1].n, w_other.bodies[-1].n))
        pool = S.SlotPool(2, dim=3, N_cap=40, body_caps=caps, precision=prec)
        pool.load(0, 0, w_big, k1)
        pool.load(1, 1, w_other, k1 + k2)
        pool.sim.step_batch(pool.dt, k1)
        pool.free(0)
        pool.load(0, 2, w_small, k2)
        pool.sim.step_batch(pool.dt, k2)
        res_small, res_other = pool.result(0), pool.result(1)
        pool.close()
        if prec == MPM_PREC_FP64:
            compare_states(w_small, ora_small, _Served(res_small), k2 * w_small.dt, label="reloaded sim 0")
            compare_states(w_other, ora_other, _Served(res_other), (k1 + k2) * w_other.dt, label="sim 1")
        else:
            for w, ora, res, t in ((w_small, ora_small, res_small, k2 * w_small.dt),
                                   (w_other, ora_other, res_other, (k1 + k2) * w_other.dt)):
                h = w.L[0] / w.N[0]
                for b in range(len(w.bodies)):
                    o, g = ora.points(b), res[b]
                    vscale = float(np.max(np.abs(o["xt"]))) + 9.81 * t + 1e-12
                    ex = float(np.max(np.abs(g["x"] - o["x"]))) / h
                    ev = float(np.max(np.abs(g["xt"] - o["xt"]))) / vscale
                    print(f"reupload fp32 {w.name} body {b}: |dx|/h {ex:.2e} |dv| {ev:.2e}")
                    assert ex <= 1e-5 and ev <= 1e-5
