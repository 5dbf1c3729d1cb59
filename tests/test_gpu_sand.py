"""Sand_SachithLocal on the GPU against the oracle, through the C-ABI and the C++ host: the
shipped columncollapse2D / columncollapse3D examples (Sand {1e7, 0.3, 0.01, 0.577, 1650,
1000} in a CartesianBox), a sand with the 8-property form (mu_2 > mu_s, I_0) and
USE_DENSITY_FOR_PRESSURE, the precision form, and the host CLI with frames.

Tolerances: those of tests/_harness.py for the state; the plastic strain gammap and its rate
gammap_dot to 1e-7 of their largest value (FP64 form).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_PREC_FP32
from tests._harness import compare_states, gpu_of, oracle_of

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mpm_3d_amd", "lib", "mpm_v3_hip")


def _plastic_close(ora, gpu, b, rel):
    og, ogd = ora.plastic(b)
    gg, ggd = gpu.plastic(b)
    for name, o, g in (("gammap", og, gg), ("gammap_dot", ogd, ggd)):
        scale = float(np.max(np.abs(o))) + 1e-300
        err = float(np.max(np.abs(g - o)))
        print(f"{name}: max {scale:.3e}, error {err / scale:.3e}")
        assert scale > 0 and err <= rel * scale, name


@pytest.mark.parametrize("dim,steps", [(2, 300), (3, 100)])
def test_column_collapse_examples(dim, steps):
    w = W.column_collapse_example(dim)
    ora = oracle_of(w)
    gpu = gpu_of(w)
    ora.step(w.dt, steps)
    gpu.step(w.dt, steps)
    compare_states(w, ora, gpu, steps * w.dt, label=f"columncollapse{dim}D x{steps}")
    _plastic_close(ora, gpu, 0, 1e-7)


def test_sand_rate_dependent_with_density_pressure():
    """The 8-property form (mu_2 = 0.9 > mu_s, I_0 = 0.3) with USE_DENSITY_FOR_PRESSURE."""
    w = W.column_collapse_example(2)
    b = w.bodies[0]
    b.sand = W.sand_props(mu_2=0.9, I_0=0.3)
    b.sand_dp = True
    ora = oracle_of(w)
    gpu = gpu_of(w)
    ora.step(w.dt, 200)
    gpu.step(w.dt, 200)
    compare_states(w, ora, gpu, 200 * w.dt, label="sand mu(I) + density pressure")
    _plastic_close(ora, gpu, 0, 1e-7)


def test_sand_precision_form():
    """The precision form (FP32 state) on the 2D column: positions to the form's 1e-5 h target
    and stresses, volumes to the bounds of test_gpu_precision; velocities to 1e-4 of their
    scale (not 1e-5). Every point of the shipped column starts at exactly rho_critical, and
    the model zeroes the stress of a point whose density falls below it: with FP32 volumes
    (6e-8 relative) a few points near that threshold take the other branch for a step, and
    their neighbours' velocities differ at 6e-5."""
    from tests.test_gpu_precision import _errors, _gpu_state, _ora_state
    w = W.column_collapse_example(2)
    ora = oracle_of(w)
    g = gpu_of(w, precision=MPM_PREC_FP32)
    ora.step(w.dt, 200)
    g.step(w.dt, 200)
    ro, rg = _ora_state(ora, 0), _gpu_state(g, 0)
    np.testing.assert_array_equal(rg["active"], ro["active"])
    e = _errors(w, ro, rg, 200 * w.dt)
    print(f"sand fp32: {e}")
    assert e["x_cells"] <= 1e-5 and e["v_rel"] <= 1e-4 and e["T_rel"] <= 1e-4 and e["V_rel"] <= 1e-5, e


def test_sand_precision_form_3d():
    """The precision form on the shipped 3D column (100 steps): positions to 1e-5 h; velocities,
    stresses and volumes to the bounds of the 2D case above (same rho_critical threshold)."""
    from tests.test_gpu_precision import _errors, _gpu_state, _ora_state
    w = W.column_collapse_example(3)
    ora = oracle_of(w)
    g = gpu_of(w, precision=MPM_PREC_FP32)
    ora.step(w.dt, 100)
    g.step(w.dt, 100)
    ro, rg = _ora_state(ora, 0), _gpu_state(g, 0)
    np.testing.assert_array_equal(rg["active"], ro["active"])
    e = _errors(w, ro, rg, 100 * w.dt)
    print(f"sand fp32 3D: {e}")
    assert e["x_cells"] <= 1e-5 and e["v_rel"] <= 1e-4 and e["T_rel"] <= 1e-4 and e["V_rel"] <= 1e-5, e


def test_host_runs_the_shipped_column_example(tmp_path):
    """mpm_v3_hip -c test2D.cfg of examples/columncollapse2D, unchanged but for the stop time
    (0.02 s): the final points match the oracle, and the DefaultVTK frames carry the sand's
    gammap and gammap_dot after the node fields."""
    src = os.path.join(W.example_dir(), "columncollapse2D")
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), tmp_path)
    os.makedirs(tmp_path / "output", exist_ok=True)
    os.makedirs(tmp_path / "save", exist_ok=True)
    cfg = (tmp_path / "test2D.cfg").read_text()
    assert "properties = {1} #stop time [s]" in cfg
    (tmp_path / "run.cfg").write_text(cfg.replace("properties = {1} #stop time [s]", "properties = {0.02} #stop time [s]"))
    r = subprocess.run([EXE, "-c", "run.cfg", "--write-points", "final_"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Material properties (E = 1e+07, nu = 0.3" in r.stdout and "grains_rho = 1650" in r.stdout
    w = W.column_collapse_example(2)
    t, steps = 0.0, 0
    while t <= 0.02:
        steps += 1
        t += w.dt
    ora = oracle_of(w)
    ora.step(w.dt, steps)
    m, v, x, xt, act = W.read_points(str(tmp_path / "final_block.points"), 2)
    o = ora.points(0)
    h = 1.0 / 40
    assert np.max(np.abs(x - o["x"][:, :2])) <= 1e-9 * h
    frames = sorted(f for f in os.listdir(tmp_path / "output") if f.startswith("fpd."))
    last = open(tmp_path / "output" / frames[-1]).read()
    i_dv = last.index("VECTORS diff_velocity") if "VECTORS diff_velocity" in last else -1
    assert i_dv == -1  # point files carry no node fields
    assert "SCALARS gammap double 1" in last and "SCALARS gammap_dot double 1" in last
    nodes = open(tmp_path / "output" / frames[-1].replace("fpd.", "fnd.")).read()
    assert nodes.index("SCALARS active") < nodes.index("VECTORS bc_nodal_mask")


@pytest.mark.parametrize("cfg", ["test3D.cfg", "test3D_threadpool.cfg"])
def test_host_runs_the_shipped_parallel_example(tmp_path, cfg):
    """examples/parallel (the 3D sand column under ParallelExplicitUSL / ThreadPoolExplicitUSL,
    which run the same step on the GPU) to a stop time of 0.005 s, against the oracle."""
    src = os.path.join(W.example_dir(), "parallel")
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), tmp_path)
    os.makedirs(tmp_path / "output", exist_ok=True)
    os.makedirs(tmp_path / "save", exist_ok=True)
    text = (tmp_path / cfg).read_text()
    assert "properties = {1} #stop time [s]" in text
    (tmp_path / "run.cfg").write_text(text.replace("properties = {1} #stop time [s]", "properties = {0.005} #stop time [s]"))
    r = subprocess.run([EXE, "-c", "run.cfg", "--write-points", "final_"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    w = W.column_collapse_example(3)
    w.bodies[0].x = W.read_points(os.path.join(src, "test3D.points"), 3)[2]
    t, steps = 0.0, 0
    while t <= 0.005:
        steps += 1
        t += w.dt
    ora = oracle_of(w)
    ora.step(w.dt, steps)
    x = W.read_points(str(tmp_path / "final_block.points"), 3)[2]
    assert np.max(np.abs(x - ora.points(0)["x"])) <= 1e-9 / 20
