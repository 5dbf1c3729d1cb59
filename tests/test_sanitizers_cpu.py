"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only; sanitizers are
for host code here, never GPU code).

* The oracle (oracle/mpm_oracle.c) through oracle/ora_selfcheck.c: ContactHuang IMPLICIT and
  EXPLICIT, CPDI off on a frictional floor, Sand_SachithLocal on a smooth box, a block leaving
  the domain, serial and threaded.
* The C++ host layer (config grammar, registry, plugins, .points reader, HipExplicitUSL up to
  its GPU calls) as mpm_v3_hip --check on every shipped example configuration.
Any memory error or undefined behaviour aborts the program (-fno-sanitize-recover=all); the
reference's exit(0) error convention leaks by design, so leak detection is off."""
import glob
import os
import shutil
import subprocess

import pytest

from mpm_3d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")


def _build(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and ("libasan" in r.stderr or "libubsan" in r.stderr):
        pytest.skip("sanitizer runtimes are not installed")
    assert r.returncode == 0, r.stderr


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "ora_selfcheck")
    _build(["gcc", "-std=c11", *SAN, "-o", exe, os.path.join(ROOT, "oracle", "ora_selfcheck.c"),
            os.path.join(ROOT, "oracle", "mpm_oracle.c"), "-lm", "-pthread"])
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all cases ran" in r.stdout


def test_host_check_under_sanitizers(tmp_path):
    lib = os.path.join(ROOT, "mpm_3d_amd", "lib")
    if not os.path.exists(os.path.join(lib, "libmpm_gpu.so")):
        pytest.skip("libmpm_gpu.so is not built")
    host = os.path.join(ROOT, "mpm_3d_amd", "host")
    srcs = sorted(glob.glob(os.path.join(host, "*.cpp")))
    exe = str(tmp_path / "mpm_v3_hip_san")
    _build(["g++", "-std=c++17", *SAN, f"-I{os.path.join(ROOT, 'include')}", f"-I{host}", "-o", exe, *srcs,
            f"-L{lib}", "-lmpm_gpu", f"-Wl,-rpath,{lib}"])
    ran = 0
    for cfg in sorted(glob.glob(os.path.join(W.example_dir(), "*", "*.cfg"))):
        r = subprocess.run([exe, "-c", cfg, "--check"], capture_output=True, text=True, env=ENV, timeout=120,
                           cwd=os.path.dirname(cfg))
        assert r.returncode == 0, (cfg, r.stderr[-4000:])
        assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, (cfg, r.stderr[-4000:])
        ran += 1
    assert ran >= 8
    shutil.rmtree(tmp_path, ignore_errors=True)
