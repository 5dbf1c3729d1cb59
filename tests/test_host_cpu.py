"""The C++ host (mpm_v3_hip / libmpm_host.so) without a GPU: configuration grammar,
.points reading and the reference's error conventions (config.cpp, parser.hpp,
registry.hpp, isolin.cpp:26-32, contact_huang.cpp:25-33). `--check` configures and
initialises a job and stops before any GPU work."""
import os
import subprocess

import pytest

from mpm_3d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mpm_3d_amd", "lib", "mpm_v3_hip")


def run(args, cwd=None):
    return subprocess.run([EXE, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


@pytest.mark.parametrize("name,counts", [("blockdrop2D", [64]), ("collision2D", [64, 64]), ("minimal2D", [64, 64]),
                                         ("blockdrop3D", [512]), ("collision3D", [512, 512]),
                                         ("columncollapse2D", [1024]), ("columncollapse3D", [2048]),
                                         ("parallel", [2048])])
def test_examples_configure(name, counts):
    d = os.path.join(W.example_dir(), name)
    for cfg in sorted(f for f in os.listdir(d) if f.endswith(".cfg")):
        r = run(["-c", os.path.join(d, cfg), "--check"])
        assert r.returncode == 0, r.stderr
        got = [int(l.split(":")[1].split()[0]) for l in r.stdout.splitlines() if l.startswith("body ")]
        assert got == counts, cfg


BASE = """# comment line
job   # header comment
{{
    dt = 2e-4
    TYPE = 2
}}
serializer
{{
    class = "NullSerializer"
}}
driver
{{
    class = 'DefaultDriver'
    properties = {{0.01}}
}}
solver
{{
    class = "HipExplicitUSL"
    int-properties = {{1, 0, 1}}
}}
grid
{{
    class = "CartesianLinear"
    properties = {{1.0, 1.0}}
    int-properties = {{10,10}}
}}
body
{{
    class = "DefaultBody"
    name = "a"
    point-file = "p.points"
    point-class = "DefaultPoints"
    node-class = "DefaultNodes"
    material-class = "{mat}"
    material-props = {{{props}}}
    material-props = {{{props}}}
    boundary-class = "{bc}"{bc_props}
    colour = "blue"
}}
"""


def _write(tmp_path, mat="IsotropicLinearElasticity", props="1e5, 0.3", bc="CartesianBox", bc_props=None):
    (tmp_path / "p.points").write_text("2\n1 0.01 0.5 0.5 0 0 1\n1 0.01 0.25 0.5 0.1 0 1\n")
    cfg = tmp_path / "t.cfg"
    extra = "" if bc_props is None else f"\n    boundary-props = {{{bc_props}}}"
    cfg.write_text(BASE.format(mat=mat, props=props, bc=bc, bc_props=extra))
    return str(cfg)


def test_grammar_comments_quotes_duplicates_unknown_keys(tmp_path):
    r = run(["-c", _write(tmp_path), "--check"], cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert "precision = 1" in r.stdout and "contact_spec = 0" in r.stdout
    assert "E = 100000, nu = 0.3" in r.stdout
    assert "Already have seen key \"material-props\"" in r.stderr
    assert "Unknown key \"colour\"" in r.stderr
    assert "body 0 a: 2 points" in r.stdout


def test_unknown_class_raises_invalid_argument(tmp_path):
    r = run(["-c", _write(tmp_path, mat="NoSuchMaterial"), "--check"], cwd=str(tmp_path))
    assert r.returncode != 0
    assert "No class named 'NoSuchMaterial'" in r.stderr


def test_bad_material_properties_exit_zero_with_message(tmp_path):
    r = run(["-c", _write(tmp_path, props="1e5"), "--check"], cwd=str(tmp_path))
    assert r.returncode == 0
    assert "Need at least 2 properties defined ({E, nu})" in r.stderr
    assert "Job Initialized" not in r.stdout


def test_missing_config_file():
    r = run(["-c", "/nonexistent/x.cfg", "--check"])
    assert r.returncode == 0 and "Unable to open file" in r.stdout


@pytest.mark.parametrize("bc,props", [("CartesianSmoothBox", None), ("CartesianFrictionalBox", "0.35")])
def test_box_boundaries_configure(tmp_path, bc, props):
    """cartesian_smooth_box.cpp / cartesian_frictional_box.cpp are registered; the
    frictional box reads mu_f from boundary-props[0] (cartesian_frictional_box.cpp:26-37)."""
    r = run(["-c", _write(tmp_path, bc=bc, bc_props=props), "--check"], cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert "Job Initialized" in r.stdout
    if props:
        assert "Boundary properties ( mu_f = 0.35)." in r.stdout


def test_frictional_box_without_mu_exits_zero_with_message(tmp_path):
    r = run(["-c", _write(tmp_path, bc="CartesianFrictionalBox"), "--check"], cwd=str(tmp_path))
    assert r.returncode == 0
    assert "Need at least 1 property defined (mu_f)." in r.stderr
    assert "Job Initialized" not in r.stdout


def test_sand_material_properties(tmp_path):
    """Sand_SachithLocal::init: six properties (mu_2 = mu_s, I_0 = 1), eight, and fewer than
    six exit with EXIT_FAILURE and the reference's message (sand_sachithlocal.cpp:58-117)."""
    r = run(["-c", _write(tmp_path, mat="Sand_SachithLocal", props="1e7, 0.3, 0.01, 0.577, 1650, 1000"), "--check"],
            cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert ("Material properties (E = 1e+07, nu = 0.3, G = 3.84615e+06, K = 8.33333e+06, grain diameter = 0.01, "
            "mu_s = 0.577, grains_rho = 1650, rho_critical = 1000).") in r.stdout
    r = run(["-c", _write(tmp_path, mat="Sand_SachithLocal", props="1e7, 0.3, 0.01, 0.577, 1650, 1000, 0.9, 0.3"),
             "--check"], cwd=str(tmp_path))
    assert r.returncode == 0 and "mu_2 = 0.9, i_0 = 0.3)." in r.stdout
    r = run(["-c", _write(tmp_path, mat="Sand_SachithLocal", props="1e7, 0.3, 0.01, 0.577, 1650"), "--check"],
            cwd=str(tmp_path))
    assert r.returncode == 1
    assert "Need at least 6 properties defined (E, nu, grain diameter, mu_s, grains_rho, rho_critical)." in r.stderr
