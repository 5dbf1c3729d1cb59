"""The kernels' shape-function code, evaluated on the host through mpm_gpu_host_weights
(the same __host__ __device__ functions the HIP kernels inline), against the oracle's COO
tuples (default_points.cpp:196-330) summed per node.

Covers GIMP/CPDI-on interior points (the separable 3^D stencil), points whose corners
reach the walls (CPDI off, per-tuple path), corners exactly on node coordinates and at
x = L, extents above one cell (per-tuple path), and the linear basis (cpdi_spec 0).
No GPU is needed.
"""
import numpy as np
import pytest

from mpm_3d_amd.gpu import host_weights
from oracle.pyoracle import OracleSim


def _oracle_tuples(dim, L, N, x, vol, cpdi):
    s = OracleSim(dim, L, N, cpdi_spec=cpdi)
    s.add_body([1.0], [vol], np.asarray(x, np.float64).reshape(1, dim), np.zeros((1, dim)), [1], E=1e5, nu=0.3)
    s.init()
    node, _, S = s.get_map(0, 0)
    gnode, _, G = s.get_map(0, 1)
    assert np.array_equal(node, gnode)
    active = s.points(0)["active"][0]
    return node, S, G[:, :dim], active


def _per_node(keys, vals):
    out = {}
    for k, v in zip(keys, vals):
        out[k] = out.get(k, 0.0) + v
    return out


def _compare(dim, L, N, x, vol, cpdi):
    ext = 0.5 * (np.sqrt(vol) if dim == 2 else np.cbrt(vol))
    node, S, G, active = _oracle_tuples(dim, L, N, x, vol, cpdi)
    hw = host_weights(dim, L, N, x, ext, cpdi)
    if not active:
        assert hw is None
        return
    assert hw is not None
    ijk, Sg, Gg = hw
    # node ids in the reference numbering (cartesian.cpp:128-140)
    strides = [1, N[0] + 1, (N[0] + 1) * (N[1] + 1) if dim == 3 else 0]
    inside = np.all((ijk >= 0) & (ijk <= np.array(N[:dim])), axis=1)
    # slots outside the grid (padding nodes -1 and N+1) must carry exact zeros
    assert np.all(Sg[~inside] == 0.0) and np.all(Gg[~inside] == 0.0)
    ids = (ijk[inside] * np.array(strides[:dim])).sum(axis=1)
    # the slots the kernels multiply are exactly the nodes of the reference's tuples (a slot no
    # corner's element covers is skipped, so 0 * NaN cannot reach a node the reference spares)
    assert set(ids.tolist()) == set(np.asarray(node).tolist()), (x, sorted(set(ids.tolist()) ^ set(np.asarray(node).tolist())))
    so = _per_node(node.tolist(), S)
    sg = _per_node(ids.tolist(), Sg[inside])
    h = min(L[d] / N[d] for d in range(dim))
    keys = set(so) | set(sg)
    for k in keys:
        assert abs(so.get(k, 0.0) - sg.get(k, 0.0)) <= 1e-14, (x, k, so.get(k), sg.get(k))
    for d in range(dim):
        go = _per_node(node.tolist(), G[:, d])
        gg = _per_node(ids.tolist(), Gg[inside][:, d])
        for k in set(go) | set(gg):
            assert abs(go.get(k, 0.0) - gg.get(k, 0.0)) <= 1e-12 / h, (x, d, k, go.get(k), gg.get(k))


@pytest.mark.parametrize("dim", [2, 3])
def test_weights_random_interior_and_walls(dim):
    rng = np.random.default_rng(11 + dim)
    L = (1.0, 0.8, 0.6)[:dim]
    N = (10, 8, 6)[:dim]
    h = np.array([L[d] / N[d] for d in range(dim)])
    cell_vol = float(np.prod(h))
    for trial in range(300):
        x = rng.uniform(0, 1, dim) * np.array(L)
        if trial % 5 == 0:  # near a wall: CPDI off
            d = trial % dim
            x[d] = rng.uniform(0, 0.1 * h[d]) if trial % 2 else L[d] - rng.uniform(0, 0.1 * h[d])
        vol = cell_vol / rng.choice([1, 2 ** dim, 3 ** dim, 4 ** dim])
        _compare(dim, L, N, x, vol, 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_weights_edge_cases(dim):
    L = (1.0, 1.0, 1.0)[:dim]
    N = (20, 20, 20)[:dim]
    h = 1.0 / 20
    vol = (h / 2) ** dim
    ext = 0.5 * (np.sqrt(vol) if dim == 2 else np.cbrt(vol))
    a = 0.5 * ext
    cases = []
    base = np.full(dim, 0.5 + 0.3 * h)
    cases.append(base.copy())
    for d in range(dim):
        on_node = base.copy()
        on_node[d] = 10 * h  # point on a node coordinate
        cases.append(on_node)
        corner_on_node = base.copy()
        corner_on_node[d] = 10 * h + a  # lower corner on a node
        cases.append(corner_on_node)
        corner_at_L = base.copy()
        corner_at_L[d] = L[d] - a  # upper corner exactly at x = L
        cases.append(corner_at_L)
        at_L = base.copy()
        at_L[d] = L[d]  # point at x = L: CPDI off, no element
        cases.append(at_L)
        outside = base.copy()
        outside[d] = -1e-9
        cases.append(outside)
    for x in cases:
        _compare(dim, L, N, x, vol, 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_weights_wide_extent_and_linear_basis(dim):
    rng = np.random.default_rng(5)
    L = (1.0,) * dim
    N = (12,) * dim
    h = 1.0 / 12
    for _ in range(60):
        x = rng.uniform(0.2, 0.8, dim)
        _compare(dim, L, N, x, (2.5 * h) ** dim, 1)  # extent 1.25 h: corners span 3 elements
        _compare(dim, L, N, x, (h / 2) ** dim, 0)    # cpdi_spec = 0: linear basis at the point
