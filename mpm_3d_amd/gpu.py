"""Python binding of libmpm_gpu.so (the C-ABI in include/mpm_gpu.h) over ctypes.

``GpuSim`` is the host-side mirror of the reference's solver surface for this path: it
takes a grid description (CartesianLinear {L, N}, TYPE, ExplicitUSL {cpdi_spec,
contact_spec}), bodies (points + IsotropicLinearElasticity {E, nu} + CartesianBox) and
ContactHuang contacts, keeps their state resident on the GPU and advances it with
``step``. There is no CPU fallback: if the HIP library is missing or no GPU is visible,
construction raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmpm_gpu.so")

MPM_PREC_FP64 = 0
MPM_PREC_FP32 = 1
MPM_MAX_BODIES_PER_SIM = 4

SYMBOLS = (
    "mpm_gpu_create", "mpm_gpu_create_batch", "mpm_gpu_upload", "mpm_gpu_step", "mpm_gpu_download",
    "mpm_gpu_download_aux", "mpm_gpu_connectivity", "mpm_gpu_connectivity_noalloc", "mpm_gpu_set_sort_interval",
    "mpm_gpu_set_timing", "mpm_gpu_get_timing", "mpm_gpu_stats", "mpm_gpu_synchronize", "mpm_gpu_destroy",
    "mpm_gpu_last_error", "mpm_gpu_footprint", "mpm_gpu_host_weights", "mpm_gpu_copy_probe",
    "mpm_gpu_set_option", "mpm_gpu_get_option", "mpm_gpu_step_batch", "mpm_gpu_load_sim", "mpm_gpu_clear_sim",
    "mpm_gpu_download_nodes", "mpm_gpu_download_contact", "mpm_gpu_download_plastic",
)
MPM_OPT_SORT_INTERVAL = 1
MPM_OPT_FUSED_PASS = 2
MPM_OPT_GRAPH = 3
MPM_OPT_AUTOTUNE = 4
MPM_OPT_PASS_GRID = 5
MPM_OPT_GRID_WGS = 6
MPM_OPT_GRAPH_CAPTURES = 7
MPM_OPT_P2G_GRID = 8
MPM_OPT_NODE_RECORD = 9
MPM_OPT_KEEP_L = 10
MPM_OPT_CELL_SORT = 11


class MpmGpuError(RuntimeError):
    pass


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("L", C.c_double * 3), ("N", C.c_int * 3), ("cpdi_spec", C.c_int),
                ("contact_spec", C.c_int), ("gravity", C.c_double * 3)]


class BodyDesc(C.Structure):
    _fields_ = [("n_points", C.c_int64), ("has_material", C.c_int), ("E", C.c_double), ("nu", C.c_double),
                ("has_boundary", C.c_int), ("bc_mu", C.c_double), ("sand", C.c_double * 6),
                ("sand_density_pressure", C.c_int)]


class ContactDesc(C.Structure):
    _fields_ = [("body1", C.c_int), ("body2", C.c_int), ("mu", C.c_double)]


class SimDesc(C.Structure):
    _fields_ = [("grid", GridDesc), ("nbody", C.c_int), ("bodies", C.POINTER(BodyDesc)), ("ncontact", C.c_int),
                ("contacts", C.POINTER(ContactDesc))]


_lib = None
_dp = C.POINTER(C.c_double)


def load_library(path: str = LIB_PATH):
    """Load libmpm_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MpmGpuError(f"{path} is missing: build it with `python -m mpm_3d_amd.build`")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.mpm_gpu_create.argtypes = [C.POINTER(GridDesc), C.POINTER(BodyDesc), C.c_int, C.POINTER(ContactDesc), C.c_int,
                                   C.c_int, C.POINTER(vp)]
    lib.mpm_gpu_create_batch.argtypes = [C.POINTER(SimDesc), C.c_int, C.c_int, C.POINTER(vp)]
    lib.mpm_gpu_upload.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_int32)]
    lib.mpm_gpu_step.argtypes = [vp, C.c_double, C.c_int]
    lib.mpm_gpu_download.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(C.c_int32)]
    lib.mpm_gpu_download_aux.argtypes = [vp, C.c_int, _dp, _dp]
    lib.mpm_gpu_download_nodes.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, _dp]
    lib.mpm_gpu_download_contact.argtypes = [vp, C.c_int, _dp, _dp]
    lib.mpm_gpu_download_plastic.argtypes = [vp, C.c_int, _dp, _dp]
    lib.mpm_gpu_connectivity.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_uint8)]
    lib.mpm_gpu_connectivity_noalloc.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_uint8)]
    lib.mpm_gpu_set_sort_interval.argtypes = [vp, C.c_int]
    lib.mpm_gpu_set_timing.argtypes = [vp, C.c_int]
    lib.mpm_gpu_set_option.argtypes = [vp, C.c_int, C.c_int]
    lib.mpm_gpu_get_option.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    lib.mpm_gpu_step_batch.argtypes = [vp, _dp, C.c_int, C.c_int]
    lib.mpm_gpu_load_sim.argtypes = [vp, C.c_int, C.POINTER(SimDesc)]
    lib.mpm_gpu_clear_sim.argtypes = [vp, C.c_int]
    lib.mpm_gpu_get_timing.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    lib.mpm_gpu_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.mpm_gpu_synchronize.argtypes = [vp]
    lib.mpm_gpu_destroy.argtypes = [vp]
    lib.mpm_gpu_last_error.restype = C.c_char_p
    lib.mpm_gpu_last_error.argtypes = []
    lib.mpm_gpu_footprint.restype = C.c_int64
    lib.mpm_gpu_footprint.argtypes = [C.POINTER(SimDesc), C.c_int, C.c_int]
    lib.mpm_gpu_copy_probe.argtypes = [C.c_int64, C.c_int, _dp]
    lib.mpm_gpu_host_weights.restype = C.c_int
    lib.mpm_gpu_host_weights.argtypes = [C.POINTER(GridDesc), _dp, C.c_double, C.POINTER(C.c_int32), _dp, _dp,
                                         C.c_int]
    for name in SYMBOLS:
        getattr(lib, name)
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise MpmGpuError(f"mpm_gpu error {rc}: {_lib.mpm_gpu_last_error().decode(errors='replace')}")


def extent_of(volume, dim):
    """GIMP half-width, computed on the host with the C library (default_points.cpp:49-57)."""
    volume = np.asarray(volume, np.float64)
    return 0.5 * (np.sqrt(volume) if dim == 2 else np.cbrt(volume))


def _grid_desc(dim, L, N, cpdi_spec, contact_spec, gravity):
    g = GridDesc()
    g.dim = dim
    for d in range(3):
        g.L[d] = float(L[d]) if d < dim else 0.0
        g.N[d] = int(N[d]) if d < dim else 0
        g.gravity[d] = float(gravity[d]) if d < len(gravity) else 0.0
    g.cpdi_spec = cpdi_spec
    g.contact_spec = contact_spec
    return g


def _sim_desc(w, keep):
    """SimDesc of a workload; `keep` holds the ctypes arrays it points to."""
    bds = (BodyDesc * len(w.bodies))()
    for bi, b in enumerate(w.bodies):
        bds[bi].n_points = b.n
        bds[bi].has_material = int(b.E is not None)
        bds[bi].E = float(b.E or 0.0)
        bds[bi].nu = float(b.nu or 0.0)
        bds[bi].has_boundary = int(b.boundary)
        bds[bi].bc_mu = float(getattr(b, "bc_mu", 0.0))
        sand = getattr(b, "sand", None)
        if sand is not None:
            bds[bi].has_material = 2
            for k in range(6):
                bds[bi].sand[k] = float(sand[k])
            bds[bi].sand_density_pressure = int(bool(getattr(b, "sand_dp", False)))
    cds = (ContactDesc * max(1, len(w.contacts)))()
    for ci, c in enumerate(w.contacts):
        cds[ci].body1, cds[ci].body2, cds[ci].mu = c.body1, c.body2, c.mu
    keep += [bds, cds]
    d = SimDesc()
    d.grid = _grid_desc(w.dim, w.L, w.N, w.cpdi_spec, w.contact_spec, w.gravity)
    d.nbody = len(w.bodies)
    d.bodies = bds
    d.ncontact = len(w.contacts)
    d.contacts = cds
    return d


class GpuSim:
    """One handle on the GPU holding one or more simulations (``workloads``)."""

    def __init__(self, workloads, precision=MPM_PREC_FP64, sort_interval=16, fused=False, upload=True):
        if not isinstance(workloads, (list, tuple)):
            workloads = [workloads]
        self.lib = load_library()
        self.workloads = list(workloads)
        self.dim = self.workloads[0].dim
        keep = []
        sims = (SimDesc * len(self.workloads))()
        self.body_index = []  # (sim, local) per global body
        self.body0 = []       # first global body of each sim
        for si, w in enumerate(self.workloads):
            self.body0.append(len(self.body_index))
            sims[si] = _sim_desc(w, keep)
            for bi in range(len(w.bodies)):
                self.body_index.append((si, bi))
        self._keep = keep
        self._sims = sims
        self.current = [self.workloads[si].bodies[bi] for si, bi in self.body_index]
        self.grid_N = [tuple(w.N) for w in self.workloads]  # current elements per sim
        h = C.c_void_p()
        _check(self.lib.mpm_gpu_create_batch(sims, len(self.workloads), precision, C.byref(h)))
        self.h = h
        _check(self.lib.mpm_gpu_set_sort_interval(self.h, sort_interval))
        _check(self.lib.mpm_gpu_set_option(self.h, MPM_OPT_FUSED_PASS, int(bool(fused))))
        if upload:
            for gb, (si, bi) in enumerate(self.body_index):
                self.upload(gb, self.workloads[si].bodies[bi])
        else:  # capacity only: every slot starts empty
            for si in range(len(self.workloads)):
                self.clear_sim(si)

    # -- state transfer -------------------------------------------------------------
    def _body(self, gb):
        return self.current[gb]

    def load_sim(self, si, w, preps=None):
        """Reconfigure slot si for workload w (within the slot's creation sizes) and upload it.
        preps: upload_arrays() of w's bodies, made beforehand (host work done off the GPU's
        critical path)."""
        keep = []
        d = _sim_desc(w, keep)
        _check(self.lib.mpm_gpu_load_sim(self.h, si, C.byref(d)))
        b0 = self.body0[si]
        self.grid_N[si] = tuple(w.N)
        for bi, b in enumerate(w.bodies):
            self.current[b0 + bi] = b
            self.upload(b0 + bi, b, prep=None if preps is None else preps[bi])

    def clear_sim(self, si):
        _check(self.lib.mpm_gpu_clear_sim(self.h, si))

    def step_batch(self, dts, nsteps):
        d = np.ascontiguousarray(dts, dtype=np.float64)
        _check(self.lib.mpm_gpu_step_batch(self.h, d.ctypes.data_as(_dp), len(d), int(nsteps)))

    def set_option(self, key, value):
        _check(self.lib.mpm_gpu_set_option(self.h, key, int(value)))

    def get_option(self, key):
        v = C.c_int()
        _check(self.lib.mpm_gpu_get_option(self.h, key, C.byref(v)))
        return v.value

    @staticmethod
    def upload_arrays(body, dim):
        """The host arrays mpm_gpu_upload takes for `body` (component planes, extents):
        numpy work only, so it can run while the GPU steps other simulations."""
        x = np.ascontiguousarray(np.asarray(body.x, np.float64)[:, :dim].T)
        v = np.ascontiguousarray(np.asarray(body.xt, np.float64)[:, :dim].T)
        V = np.ascontiguousarray(body.v, np.float64)
        m = np.ascontiguousarray(body.m, np.float64)
        ext = np.ascontiguousarray(extent_of(V, dim))
        act = np.ascontiguousarray(body.active, np.int32)
        return x, v, V, m, ext, act

    def upload(self, gb, body, T=None, prep=None):
        n = body.n
        x, v, V, m, ext, act = prep if prep is not None else self.upload_arrays(body, self.dim)
        Tp = None
        if T is not None:
            Tp = np.ascontiguousarray(np.asarray(T, np.float64).reshape(6, n))
        _check(self.lib.mpm_gpu_upload(self.h, gb, x.ctypes.data_as(_dp), v.ctypes.data_as(_dp),
                                       Tp.ctypes.data_as(_dp) if Tp is not None else None,
                                       V.ctypes.data_as(_dp), m.ctypes.data_as(_dp), ext.ctypes.data_as(_dp),
                                       act.ctypes.data_as(C.POINTER(C.c_int32))))

    def step(self, dt, nsteps=1):
        _check(self.lib.mpm_gpu_step(self.h, float(dt), int(nsteps)))

    def points_planes(self, gb):
        """State of global body gb in original order, as downloaded: x, xt (dim, n) and
        T (6, n) component planes; v (n,); active (n,)."""
        n = self._body(gb).n
        dim = self.dim
        x = np.empty((dim, n))
        v = np.empty((dim, n))
        T = np.empty((6, n))
        V = np.empty(n)
        a = np.empty(n, np.int32)
        _check(self.lib.mpm_gpu_download(self.h, gb, x.ctypes.data_as(_dp), v.ctypes.data_as(_dp),
                                         T.ctypes.data_as(_dp), V.ctypes.data_as(_dp),
                                         a.ctypes.data_as(C.POINTER(C.c_int32))))
        return {"x": x, "xt": v, "T6": T, "v": V, "active": a}

    @staticmethod
    def rows(planes):
        """points_planes() output in the points() layout (one row per point)."""
        return {"x": planes["x"].T.copy(), "xt": planes["xt"].T.copy(), "T6": planes["T6"].T.copy(),
                "v": planes["v"], "active": planes["active"]}

    def points(self, gb):
        """State of global body gb in original order: x, xt (n, dim); T (n, 6); v (n,); active."""
        return self.rows(self.points_planes(gb))

    def aux(self, gb):
        n = self._body(gb).n
        u = np.zeros((self.dim, n))
        L = np.zeros((9, n))
        _check(self.lib.mpm_gpu_download_aux(self.h, gb, u.ctypes.data_as(_dp), L.ctypes.data_as(_dp)))
        return {"u": u.T.copy(), "L": L.T.copy()}

    def plastic(self, gb):
        """Sand_SachithLocal gammap, gammap_dot of global body gb (original order)."""
        n = self._body(gb).n
        gp, gpd = np.zeros(n), np.zeros(n)
        _check(self.lib.mpm_gpu_download_plastic(self.h, gb, gp.ctypes.data_as(_dp), gpd.ctypes.data_as(_dp)))
        return gp, gpd

    def _node_count(self, si):
        nn = 1
        for d in range(self.dim):
            nn *= self.grid_N[si][d] + 1
        return nn

    def nodes(self, gb, record=True):
        """Node fields of global body gb after the last step, reference node order: m (nn,),
        mxt, f, xt, dxt (nn, dim); m, mxt, f need MPM_OPT_NODE_RECORD during that step
        (record=False skips them)."""
        nn = self._node_count(self.body_index[gb][0])
        dim = self.dim
        out = {"xt": np.zeros((dim, nn)), "dxt": np.zeros((dim, nn))}
        if record:
            out.update(m=np.zeros(nn), mxt=np.zeros((dim, nn)), f=np.zeros((dim, nn)))
        ptr = lambda k: out[k].ctypes.data_as(_dp) if k in out else None
        _check(self.lib.mpm_gpu_download_nodes(self.h, gb, ptr("m"), ptr("mxt"), ptr("f"), ptr("xt"), ptr("dxt")))
        return {k: (v if v.ndim == 1 else v.T.copy()) for k, v in out.items()}

    def contact(self, c, si=0):
        """(normal, force) of global contact c (of sim si), each (nn, dim)."""
        nn = self._node_count(si)
        n = np.zeros((self.dim, nn))
        f = np.zeros((self.dim, nn))
        _check(self.lib.mpm_gpu_download_contact(self.h, c, n.ctypes.data_as(_dp), f.ctypes.data_as(_dp)))
        return n.T.copy(), f.T.copy()

    def connectivity(self, gb, noalloc=False):
        """Corner elements (n, 2^dim) and codes (n,) of global body gb, original order; with
        `noalloc` through mpm_gpu_connectivity_noalloc (no device allocation)."""
        n = self._body(gb).n
        nc = 2 ** self.dim
        elem = np.zeros(n * nc, np.int32)
        code = np.zeros(n, np.uint8)
        fn = self.lib.mpm_gpu_connectivity_noalloc if noalloc else self.lib.mpm_gpu_connectivity
        _check(fn(self.h, gb, elem.ctypes.data_as(C.POINTER(C.c_int32)), code.ctypes.data_as(C.POINTER(C.c_uint8))))
        return elem.reshape(n, nc), code

    # -- instrumentation --------------------------------------------------------------
    def set_timing(self, on=True):
        _check(self.lib.mpm_gpu_set_timing(self.h, int(on)))

    def timing(self, cls):
        n = C.c_int64()
        ms = C.c_double()
        _check(self.lib.mpm_gpu_get_timing(self.h, cls, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def stats(self):
        an = C.c_int64()
        ap = C.c_int64()
        _check(self.lib.mpm_gpu_stats(self.h, C.byref(an), C.byref(ap)))
        return an.value, ap.value

    def synchronize(self):
        _check(self.lib.mpm_gpu_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpm_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def host_weights(dim, L, N, x, extent, cpdi_spec=1):
    """The kernels' shape-function tuples for one point, evaluated on the host (no GPU).

    Returns (node_ijk (k, dim), S (k,), G (k, dim)), or None when x is outside the domain."""
    lib = load_library()
    g = _grid_desc(dim, L, N, cpdi_spec, 1, (0.0, 0.0, 0.0))
    xa = np.zeros(3)
    xa[:dim] = x
    cap = 64
    nid = np.zeros(3 * cap, np.int32)
    S = np.zeros(cap)
    G = np.zeros(3 * cap)
    k = lib.mpm_gpu_host_weights(C.byref(g), xa.ctypes.data_as(_dp), float(extent),
                                 nid.ctypes.data_as(C.POINTER(C.c_int32)), S.ctypes.data_as(_dp),
                                 G.ctypes.data_as(_dp), cap)
    if k == -100:
        return None
    if k < 0:
        _check(k)
    k = min(k, cap)
    return nid.reshape(cap, 3)[:k, :dim], S[:k], G.reshape(cap, 3)[:k, :dim]


def copy_probe(nbytes=1 << 30, reps=20):
    """HBM GB/s of a plain device copy (read + write) on the current GPU."""
    lib = load_library()
    g = C.c_double()
    _check(lib.mpm_gpu_copy_probe(int(nbytes), int(reps), C.byref(g)))
    return g.value


def stress6_to9(T6):
    """(n, 6) symmetric xx, yy, zz, xy, xz, yz -> (n, 9) row-major."""
    T6 = np.asarray(T6)
    xx, yy, zz, xy, xz, yz = (T6[:, i] for i in range(6))
    return np.stack([xx, xy, xz, xy, yy, yz, xz, yz, zz], axis=1)
