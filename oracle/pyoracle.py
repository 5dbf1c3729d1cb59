"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of ``bench.py`` import
this module. It wraps ``oracle/_build/libmpm_oracle.so`` (the serial FP64 restatement of
mpm_3d's ExplicitUSL, see ``mpm_oracle.c``) and never touches the GPU library.

Arrays cross the boundary in the oracle's native layout: 3 doubles per point or node
(AoS, third component 0 in 2D) and row-major 3x3 tensors (9 doubles).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


def build() -> None:
    """Compile the oracle libraries (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(native: bool = False):
    name = "libmpm_oracle_native.so" if native else "libmpm_oracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, "_build", name)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    lib.ora_create.restype = C.c_void_p
    lib.ora_create.argtypes = [C.c_int, _dp, C.POINTER(C.c_int), C.c_int, C.c_int]
    lib.ora_add_body.restype = C.c_int
    lib.ora_add_body.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _ip, C.c_int,
                                 C.c_double, C.c_double, C.c_int, C.c_double]
    lib.ora_add_contact.restype = C.c_int
    lib.ora_add_contact.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
    lib.ora_init.restype = C.c_int
    lib.ora_init.argtypes = [C.c_void_p]
    lib.ora_set_gravity.argtypes = [C.c_void_p, _dp]
    lib.ora_step.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int]
    lib.ora_get_points.argtypes = [C.c_void_p, C.c_int] + [_dp] * 8 + [_ip]
    lib.ora_set_points.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _ip]
    lib.ora_get_nodes.argtypes = [C.c_void_p, C.c_int] + [_dp] * 6
    lib.ora_get_contact.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.ora_node_count.restype = C.c_int64
    lib.ora_node_count.argtypes = [C.c_void_p]
    lib.ora_element_count.restype = C.c_int64
    lib.ora_element_count.argtypes = [C.c_void_p]
    lib.ora_get_map.restype = C.c_int64
    lib.ora_get_map.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp, C.c_int64]
    lib.ora_connectivity.argtypes = [C.c_void_p, C.c_int, _ip, _u8p]
    lib.ora_which_element.restype = C.c_int
    lib.ora_which_element.argtypes = [C.c_void_p, _dp]
    lib.ora_in_domain.restype = C.c_int
    lib.ora_in_domain.argtypes = [C.c_void_p, _dp]
    lib.ora_basis_value.restype = C.c_int
    lib.ora_basis_value.argtypes = [C.c_void_p, _dp, _ip, _dp]
    lib.ora_basis_gradient.restype = C.c_int
    lib.ora_basis_gradient.argtypes = [C.c_void_p, _dp, _ip, _dp]
    lib.ora_isolin_update.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
    lib.ora_set_sand.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int]
    lib.ora_get_plastic.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.ora_sand_update.argtypes = [C.c_double, C.c_double, _dp, C.c_int, C.c_double, C.c_double, C.c_double,
                                    _dp, _dp]
    lib.ora_sand_update.restype = C.c_double
    lib.ora_destroy.argtypes = [C.c_void_p]
    _LIBS[name] = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _pad3(a, n):
    a = np.asarray(a, dtype=np.float64).reshape(n, -1)
    out = np.zeros((n, 3))
    out[:, : a.shape[1]] = a
    return out


class OracleSim:
    """One simulation in the CPU oracle: grid + bodies + contacts, FP64 serial steps."""

    def __init__(self, dim, L, N, cpdi_spec=1, contact_spec=1, native=False):
        self.lib = _lib(native)
        self.dim = dim
        Lc = (C.c_double * 3)(*(list(L) + [0.0] * (3 - len(L))))
        Nc = (C.c_int * 3)(*(list(N) + [0] * (3 - len(N))))
        self.h = self.lib.ora_create(dim, Lc, Nc, cpdi_spec, contact_spec)
        if not self.h:
            raise ValueError("oracle rejected grid description")
        self.sizes = []
        self._keep = []

    def add_body(self, m, v, x, xt, active=None, E=None, nu=None, boundary=True, bc_mu=0.0):
        """boundary: 0/False none, 1/True CartesianBox, 2 CartesianSmoothBox,
        3 CartesianFrictionalBox with mu_f = bc_mu"""
        m = np.ascontiguousarray(m, dtype=np.float64)
        n = m.shape[0]
        v = np.ascontiguousarray(v, dtype=np.float64)
        x3 = np.ascontiguousarray(_pad3(x, n))
        xt3 = np.ascontiguousarray(_pad3(xt, n))
        act = np.ones(n, np.int32) if active is None else np.ascontiguousarray(active, dtype=np.int32)
        has_mat = E is not None
        b = self.lib.ora_add_body(self.h, n, _d(m), _d(v), _d(x3), _d(xt3), _i(act), int(has_mat),
                                  float(E or 0.0), float(nu or 0.0), int(boundary), float(bc_mu))
        self.sizes.append(n)
        return b

    def set_sand(self, body, props6, use_density_pressure=False):
        """Sand_SachithLocal for body (added with E, nu; before init): props6 = (grain
        diameter, mu_s, grains_rho, rho_critical, mu_2, I_0)"""
        p = np.ascontiguousarray(props6, dtype=np.float64)
        if self.lib.ora_set_sand(self.h, body, _d(p), int(bool(use_density_pressure))) != 0:
            raise ValueError("oracle rejected the sand material")

    def plastic(self, body):
        n = self.sizes[body]
        gp, gpd = np.zeros(n), np.zeros(n)
        self.lib.ora_get_plastic(self.h, body, _d(gp), _d(gpd))
        return gp, gpd

    def add_contact(self, b1, b2, mu):
        return self.lib.ora_add_contact(self.h, b1, b2, mu)

    def init(self, g=None):
        if self.lib.ora_init(self.h) != 0:
            raise ValueError("oracle init failed")
        if g is None:
            g = [0.0, -9.81, 0.0] if self.dim == 2 else [0.0, 0.0, -9.81]
        gc = (C.c_double * 3)(*(list(g) + [0.0] * (3 - len(g))))
        self.lib.ora_set_gravity(self.h, gc)

    def step(self, dt, nsteps=1, nthreads=1):
        self.lib.ora_step(self.h, dt, nsteps, nthreads)

    def points(self, body):
        n = self.sizes[body]
        out = {k: np.zeros((n, 3)) for k in ("x", "u", "xt", "mxt")}
        out["T"] = np.zeros((n, 9))
        out["L"] = np.zeros((n, 9))
        out["v"] = np.zeros(n)
        out["extent"] = np.zeros(n)
        out["active"] = np.zeros(n, np.int32)
        self.lib.ora_get_points(self.h, body, _d(out["x"]), _d(out["u"]), _d(out["xt"]),
                                _d(out["mxt"]), _d(out["T"]), _d(out["L"]), _d(out["v"]),
                                _d(out["extent"]), _i(out["active"]))
        return out

    def set_points(self, body, x=None, xt=None, T=None, v=None, active=None):
        n = self.sizes[body]
        args = []
        for a, w in ((x, 3), (xt, 3), (T, 9), (v, 1)):
            if a is None:
                args.append(None)
            else:
                arr = np.ascontiguousarray(_pad3(a, n) if w == 3 else np.asarray(a, np.float64).reshape(n, -1))
                self._keep.append(arr)
                args.append(_d(arr))
        act = None
        if active is not None:
            act = np.ascontiguousarray(active, dtype=np.int32)
            self._keep.append(act)
            act = _i(act)
        self.lib.ora_set_points(self.h, body, *args, act)

    @property
    def node_count(self):
        return self.lib.ora_node_count(self.h)

    @property
    def element_count(self):
        return self.lib.ora_element_count(self.h)

    def nodes(self, body):
        nn = self.node_count
        out = {"m": np.zeros(nn)}
        for k in ("mxt", "f", "xt", "u", "dxt"):
            out[k] = np.zeros((nn, 3))
        self.lib.ora_get_nodes(self.h, body, _d(out["m"]), _d(out["mxt"]), _d(out["f"]),
                               _d(out["xt"]), _d(out["u"]), _d(out["dxt"]))
        return out

    def contact(self, c):
        nn = self.node_count
        nrm = np.zeros((nn, 3))
        frc = np.zeros((nn, 3))
        self.lib.ora_get_contact(self.h, c, _d(nrm), _d(frc))
        return nrm, frc

    def get_map(self, body, which):
        cnt = self.lib.ora_get_map(self.h, body, which, None, None, None, 0)
        w = 1 if which == 0 else 3
        node = np.zeros(cnt, np.int32)
        point = np.zeros(cnt, np.int32)
        val = np.zeros(cnt * w)
        self.lib.ora_get_map(self.h, body, which, _i(node), _i(point), _d(val), cnt)
        return node, point, val.reshape(cnt, w) if w == 3 else val

    def connectivity(self, body):
        n = self.sizes[body]
        nc = 2 ** self.dim
        elem = np.zeros(n * nc, np.int32)
        code = np.zeros(n, np.uint8)
        self.lib.ora_connectivity(self.h, body, _i(elem), code.ctypes.data_as(_u8p))
        return elem.reshape(n, nc), code

    def which_element(self, x):
        xa = np.ascontiguousarray(_pad3(x, 1)[0])
        return self.lib.ora_which_element(self.h, _d(xa))

    def in_domain(self, x):
        xa = np.ascontiguousarray(_pad3(x, 1)[0])
        return bool(self.lib.ora_in_domain(self.h, _d(xa)))

    def basis_value(self, x):
        xa = np.ascontiguousarray(_pad3(x, 1)[0])
        nid = np.zeros(8, np.int32)
        val = np.zeros(8)
        k = self.lib.ora_basis_value(self.h, _d(xa), _i(nid), _d(val))
        return nid[:k], val[:k]

    def basis_gradient(self, x):
        xa = np.ascontiguousarray(_pad3(x, 1)[0])
        nid = np.zeros(8, np.int32)
        g = np.zeros(24)
        k = self.lib.ora_basis_gradient(self.h, _d(xa), _i(nid), _d(g))
        return nid[:k], g.reshape(8, 3)[:k]

    def close(self):
        if self.h:
            self.lib.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def isolin_update(E, nu, dt, L9, T9):
    """One IsotropicLinearElasticity update of a single 3x3 stress (isolin.cpp:50-78)."""
    lib = _lib()
    L = np.ascontiguousarray(np.asarray(L9, np.float64).reshape(9))
    T = np.ascontiguousarray(np.asarray(T9, np.float64).reshape(9).copy())
    lib.ora_isolin_update(E, nu, dt, _d(L), _d(T))
    return T


def sand_update(E, nu, props6, dt, m, v, L9, T9, use_density_pressure=False):
    """Sand_SachithLocal::calculateStress on one point; returns (T9 after, nup_tau)."""
    lib = _lib()
    p = np.ascontiguousarray(props6, dtype=np.float64)
    L = np.ascontiguousarray(L9, dtype=np.float64).reshape(9)
    T = np.ascontiguousarray(T9, dtype=np.float64).reshape(9).copy()
    nup = lib.ora_sand_update(E, nu, _d(p), int(bool(use_density_pressure)), dt, m, v, _d(L), _d(T))
    return T, nup
