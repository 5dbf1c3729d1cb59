"""The C-ABI library without a GPU: it loads, exports every function include/*.h declares,
and its host-only entry points (footprint, argument checks) behave. No compute calls."""
import ctypes as C
import glob
import os
import re

import pytest

from mpm_3d_amd import gpu as G
from mpm_3d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(mpm_gpu_\w+)\s*\(", src, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = G.load_library()
    names = _declared()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert names <= set(G.SYMBOLS), f"binding does not wire {sorted(names - set(G.SYMBOLS))}"


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(G, "_lib", None)
    with pytest.raises(G.MpmGpuError):
        G.load_library(str(tmp_path / "libmpm_gpu.so"))


def _sim_desc(w):
    grid = G.GridDesc(w.dim, (C.c_double * 3)(*(list(w.L) + [0.0] * (3 - w.dim))),
                      (C.c_int * 3)(*(list(w.N) + [0] * (3 - w.dim))), w.cpdi_spec, w.contact_spec,
                      (C.c_double * 3)(*(list(w.gravity) + [0.0] * (3 - w.dim))))
    bodies = (G.BodyDesc * len(w.bodies))(*[G.BodyDesc(b.n, 1, b.E, b.nu, 1) for b in w.bodies])
    contacts = (G.ContactDesc * max(1, len(w.contacts)))(*[G.ContactDesc(c.body1, c.body2, c.mu)
                                                          for c in w.contacts])
    return G.SimDesc(grid, len(w.bodies), bodies, len(w.contacts), contacts), (bodies, contacts)


def test_footprint_scales_with_points_and_grid():
    lib = G.load_library()
    small = W.two_block_3d(n_cells=16, gap_cells=1)
    big = W.two_block_3d(n_cells=32, gap_cells=1)
    ds, keep1 = _sim_desc(small)
    db, keep2 = _sim_desc(big)
    fs = lib.mpm_gpu_footprint(C.byref(ds), 1, G.MPM_PREC_FP64)
    fb = lib.mpm_gpu_footprint(C.byref(db), 1, G.MPM_PREC_FP64)
    assert fs > 0 and fb > fs
    # the FP64 particle state alone: x3 v3 T6 V m extent per point, double-buffered for the sort
    assert fb >= 2 * 15 * 8 * big.n_points


def test_footprint_rejects_bad_grids():
    lib = G.load_library()
    w = W.two_block_3d(n_cells=16, gap_cells=1)
    d, keep = _sim_desc(w)
    d.grid.dim = 4
    assert lib.mpm_gpu_footprint(C.byref(d), 1, G.MPM_PREC_FP64) < 0
    d.grid.dim = 3
    d.grid.N[1] = 0
    assert lib.mpm_gpu_footprint(C.byref(d), 1, G.MPM_PREC_FP64) < 0


def test_null_handle_is_an_argument_error():
    lib = G.load_library()
    assert lib.mpm_gpu_step(None, 1e-4, 1) == -1
    assert b"NULL" in lib.mpm_gpu_last_error() or b"null" in lib.mpm_gpu_last_error()
    assert lib.mpm_gpu_destroy(None) in (0, -1)


def test_body_descriptor_layout_and_boundary_check():
    """The ctypes BodyDesc has the C struct's stride (a wrong stride would read the second
    body's fields at the wrong offset), and has_boundary outside 0..3 is rejected with a
    message naming the body."""
    lib = G.load_library()
    assert C.sizeof(G.BodyDesc) == 104
    w = W.two_block_3d(n_cells=16, gap_cells=1)
    d, keep = _sim_desc(w)
    keep[0][1].has_boundary = 3
    keep[0][1].bc_mu = 0.3
    assert lib.mpm_gpu_footprint(C.byref(d), 1, G.MPM_PREC_FP64) > 0
    keep[0][1].has_boundary = 4
    assert lib.mpm_gpu_footprint(C.byref(d), 1, G.MPM_PREC_FP64) < 0
    assert b"body 1: has_boundary must be 0..3" in lib.mpm_gpu_last_error()


def test_create_rejects_bad_forms_before_allocating():
    """Descriptor and form errors are reported by create before any HIP call: an unknown
    precision form, a grid too tall for the precision form's packed cell (10 bits in z), and
    cpdi_spec outside {0, 1} with the reference's message."""
    lib = G.load_library()
    w = W.two_block_3d(n_cells=16, gap_cells=1)
    d, keep = _sim_desc(w)
    h = C.c_void_p()
    assert lib.mpm_gpu_create_batch(C.byref(d), 1, 7, C.byref(h)) == -1
    assert b"unknown precision form 7" in lib.mpm_gpu_last_error()
    tall = W.Workload("tall", 3, (1.0, 1.0, 8.0), (16, 16, 1100), w.bodies[:1], [], w.dt, 1, (0.0, 0.0, -9.81))
    d2, keep2 = _sim_desc(tall)
    assert lib.mpm_gpu_create_batch(C.byref(d2), 1, G.MPM_PREC_FP32, C.byref(h)) == -1
    assert b"precision form: N[2] = 1100 exceeds 1022 cells" in lib.mpm_gpu_last_error()
    d.grid.cpdi_spec = 3
    assert lib.mpm_gpu_create_batch(C.byref(d), 1, G.MPM_PREC_FP64, C.byref(h)) == -1
    assert b"Unrecognized Argument for use_cpdi: 3" in lib.mpm_gpu_last_error()
    assert not h.value
