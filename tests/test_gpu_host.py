"""The shipped example configurations run unchanged through the C++ host (mpm_v3_hip:
Configurator, DefaultDriver loop, HipExplicitUSL over the C-ABI) and end in the oracle's
state within the tolerances of tests/_harness.py."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from tests._harness import XTOL_CELLS, VTOL, example_workload, oracle_of

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mpm_3d_amd", "lib", "mpm_v3_hip")


def _run_example(name, tmp_path, extra):
    src = os.path.join(W.example_dir(), name)
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), tmp_path)
    os.makedirs(tmp_path / "output", exist_ok=True)
    os.makedirs(tmp_path / "save", exist_ok=True)
    cfg = [f for f in os.listdir(src) if f.endswith(".cfg")][0]
    print(f"running {name} {extra}", flush=True)
    r = subprocess.run([EXE, "-c", cfg, "--write-points", "final_", *extra], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def _compare(w, ora, tmp_path, steps):
    h = min(w.L[d] / w.N[d] for d in range(w.dim))
    for bi, b in enumerate(w.bodies):
        m, v, x, xt, act = W.read_points(str(tmp_path / f"final_{b.name}.points"), w.dim)
        o = ora.points(bi)
        np.testing.assert_array_equal(act, o["active"])
        ex = np.max(np.abs(x - o["x"][:, :w.dim])) / h
        vscale = np.max(np.abs(o["xt"])) + 9.81 * steps * w.dt
        ev = np.max(np.abs(xt - o["xt"][:, :w.dim])) / vscale
        print(f"{w.name} body {bi}: |dx|/h {ex:.3e} |dv|/v {ev:.3e}")
        assert ex <= XTOL_CELLS and ev <= VTOL


@pytest.mark.parametrize("name", ["collision2D", "collision3D"])
def test_cli_steps_match_oracle(name, tmp_path):
    w = example_workload(name)
    for i, b in enumerate(w.bodies):
        b.name = ["block", "block2"][i]
    steps = 120
    _run_example(name, tmp_path, ["--steps", str(steps)])
    ora = oracle_of(w)
    ora.step(w.dt, steps)
    _compare(w, ora, tmp_path, steps)


def test_driver_run_with_frames(tmp_path):
    """DefaultDriver::run to the configured stop time (2 s), writing frames at 60 fps."""
    w = example_workload("blockdrop2D")
    w.bodies[0].name = "block"
    r = _run_example("blockdrop2D", tmp_path, [])
    # steps of `while (t <= stop_time) { step; t += dt; }` with the reference's FP clock
    t, steps = 0.0, 0
    while t <= 2.0:
        steps += 1
        t += w.dt
    ora = oracle_of(w)
    ora.step(w.dt, steps)
    _compare(w, ora, tmp_path, steps)
    frames = sorted(os.listdir(tmp_path / "output"))
    pframes = [f for f in frames if f.startswith("fpd.0.block.")]
    nframes = [f for f in frames if f.startswith("fnd.0.block.")]
    assert len(pframes) >= 120 and len(nframes) == len(pframes) == len(frames) // 2
    assert "Simulation Complete" in r.stdout
