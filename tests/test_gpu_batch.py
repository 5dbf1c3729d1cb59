"""Batched and served forms: many simulations in one handle (per-simulation dt, HIP-graph
replay of sort periods, launch-shape autotuning) and a slot pool that admits, steps and
frees simulations while others run. Each simulation's result must match the same
simulation run alone on its own handle:
  FP64 form: positions to 1e-9 h, velocities to 1e-8 of their scale (the parity bounds;
  only atomic summation order differs); the oracle agrees on a sample.
  precision form: positions to 1e-6 h, velocities to 1e-6 of their scale (FP32
  rounding of the same arithmetic in a different summation order)."""
import numpy as np
import pytest

from mpm_3d_amd import serve as S
from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import (MPM_OPT_AUTOTUNE, MPM_OPT_GRAPH, MPM_OPT_GRAPH_CAPTURES, MPM_OPT_GRID_WGS,
                            MPM_OPT_P2G_GRID, MPM_OPT_PASS_GRID, MPM_OPT_SORT_INTERVAL, MPM_PREC_FP32, MPM_PREC_FP64,
                            GpuSim)
from tests._harness import oracle_of

pytestmark = pytest.mark.gpu


def _err(w, a, b):
    h = min(w.L[d] / w.N[d] for d in range(w.dim))
    vs = np.max(np.abs(a["xt"])) + 9.81 * w.steps * w.dt + 1e-12
    return np.max(np.abs(a["x"] - b["x"])) / h, np.max(np.abs(a["xt"] - b["xt"])) / vs


def _small_ensemble(n=12, steps=120):
    ws = W.ensemble_2d(n_sims=n, seed=11, steps=steps)
    return [w for w in ws if w.n_points < 40000][:8]


@pytest.mark.parametrize("prec,xt,vt", [(MPM_PREC_FP64, 1e-9, 1e-8), (MPM_PREC_FP32, 1e-6, 1e-6)])
def test_batch_matches_alone(prec, xt, vt):
    ws = _small_ensemble()
    steps = 120
    batch = GpuSim(ws, precision=prec)
    batch.step_batch([w.dt for w in ws], steps)
    for i, w in enumerate(ws):
        alone = GpuSim(w, precision=prec)
        alone.step(w.dt, steps)
        ex, ev = _err(w, alone.points(0), batch.points(i))
        print(f"{w.name}: batch vs alone |dx|/h {ex:.2e} |dv|/v {ev:.2e}")
        assert ex <= xt and ev <= vt
        alone.close()
    if prec == MPM_PREC_FP64:  # and the oracle, on the smallest
        w = min(ws, key=lambda w: w.n_points)
        i = ws.index(w)
        ora = oracle_of(w)
        ora.step(w.dt, steps)
        o = ora.points(0)
        ex, ev = _err(w, {"x": o["x"][:, :2], "xt": o["xt"][:, :2]}, batch.points(i))
        assert ex <= 1e-9 and ev <= 1e-8


def test_graph_replay_matches_eager():
    """Sort periods replayed from captured graphs, several calls in a row with eager work
    and synchronisation between them, give the eager result."""
    ws = _small_ensemble()
    a = GpuSim(ws, precision=MPM_PREC_FP64, sort_interval=8)
    b = GpuSim(ws, precision=MPM_PREC_FP64, sort_interval=8)
    b.set_option(MPM_OPT_GRAPH, 1)
    dts = [w.dt for w in ws]
    for chunk in (40, 1, 33, 17, 40):
        a.step_batch(dts, chunk)
        b.step_batch(dts, chunk)
        b.synchronize()
        for i, w in enumerate(ws):
            ex, ev = _err(w, a.points(i), b.points(i))
            assert ex <= 1e-9 and ev <= 1e-8, (chunk, w.name, ex, ev)
    assert b.get_option(MPM_OPT_GRAPH_CAPTURES) >= 1


def test_autotune_keeps_results():
    ws = _small_ensemble()
    a = GpuSim(ws, precision=MPM_PREC_FP64)
    b = GpuSim(ws, precision=MPM_PREC_FP64)
    b.set_option(MPM_OPT_AUTOTUNE, 1)
    dts = [w.dt for w in ws]
    a.step_batch(dts, 60)
    b.step_batch(dts, 60)
    assert b.get_option(MPM_OPT_AUTOTUNE) == 2
    assert b.get_option(MPM_OPT_PASS_GRID) in (512, 1024, 2048, 4096, 8192)
    assert b.get_option(MPM_OPT_P2G_GRID) in (768, 1024, 1536, 2048, 4096)
    assert b.get_option(MPM_OPT_GRID_WGS) in (512, 2048, 8192)
    for i, w in enumerate(ws):
        ex, ev = _err(w, a.points(i), b.points(i))
        assert ex <= 1e-9 and ev <= 1e-8


def test_autotune_sort_interval_keeps_results():
    """A call long enough to tune the sort interval too (8, 16 or 32 steps per binning), with
    graph replay after it: the same state as a handle that never tuned, to rounding."""
    ws = _small_ensemble()
    a = GpuSim(ws, precision=MPM_PREC_FP64)
    b = GpuSim(ws, precision=MPM_PREC_FP64)
    b.set_option(MPM_OPT_AUTOTUNE, 1)
    b.set_option(MPM_OPT_GRAPH, 1)
    dts = [w.dt for w in ws]
    a.step_batch(dts, 150)
    b.step_batch(dts, 150)
    assert b.get_option(MPM_OPT_AUTOTUNE) == 2
    assert b.get_option(MPM_OPT_SORT_INTERVAL) in (8, 16, 32)
    b.step_batch(dts, 70)
    a.step_batch(dts, 70)
    for i, w in enumerate(ws):
        ex, ev = _err(w, a.points(i), b.points(i))
        assert ex <= 1e-9 and ev <= 1e-8


@pytest.mark.parametrize("overlap", [False, True])
def test_served_results_match_alone(overlap):
    """Requests admitted into a 3-slot pool as slots free up (one slot reused twice), with
    others running: every result equals the request run alone; with the host work inline,
    and overlapped with the stepping (chunks stepped on a worker thread)."""
    reqs = [(w, s // 20) for w, s in W.serve_requests(40, seed=21) if w.n_points < 60000][:5]
    assert len(reqs) == 5
    pool = S.SlotPool(3, dim=3, N_cap=128, body_caps=(60000, 60000), precision=MPM_PREC_FP32)
    out = S.run_stream(pool, reqs, [0.0] * len(reqs), chunk_max=48, keep_results=True, overlap=overlap)
    pool.close()
    assert out["overlap"] == overlap and out["status"] == ["done"] * len(reqs)
    for r, (w, steps) in enumerate(reqs):
        alone = S.run_alone(w, steps)
        for b in range(len(w.bodies)):
            h = w.L[0] / w.N[0]
            ex = np.max(np.abs(alone[b]["x"] - out["results"][r][b]["x"])) / h
            vs = np.max(np.abs(alone[b]["xt"])) + 9.81 * steps * w.dt + 1e-12
            ev = np.max(np.abs(alone[b]["xt"] - out["results"][r][b]["xt"])) / vs
            print(f"{w.name} body {b}: served vs alone |dx|/h {ex:.2e} |dv|/v {ev:.2e}")
            assert ex <= 1e-6 and ev <= 1e-6
