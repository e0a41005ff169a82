"""Pins the oracle's topology and DSS layer (unmask / global_sum / mask / dot)
against a brute-force assembly built from physical node coordinates: copies of
one global node are the element-local nodes at the same (x, y)."""
import numpy as np
import pytest

from conftest import SMALL_MESH, structured


def global_numbering(o, boundary):
    X, Y = o.coords()
    key = np.round(X * 1e9).astype(np.int64) * 10**12 + np.round(Y * 1e9).astype(np.int64)
    uniq, gid = np.unique(key, return_inverse=True)
    groups = [[] for _ in range(uniq.size)]
    for k in range(X.size):  # ascending element order
        groups[gid[k]].append(k)
    bnd = np.array([boundary(X[g[0]], Y[g[0]]) for g in groups])
    return groups, bnd


def on_square(lo, hi):
    return lambda x, y: min(abs(x - lo), abs(x - hi), abs(y - lo), abs(y - hi)) < 1e-12


CASES = [
    ("small_mesh_p3", lambda o: o.Mesh.load(3, SMALL_MESH), on_square(-1, 1)),
    ("small_mesh_p6", lambda o: o.Mesh.load(6, SMALL_MESH), on_square(-1, 1)),
    ("jittered_p4", lambda o: o.Mesh.from_arrays(4, *structured(6, (0, 1, 0, 1), 0.15, 2)), on_square(0, 1)),
    ("p1", lambda o: o.Mesh.from_arrays(1, *structured(4, (0, 1, 0, 1), 0.1, 4)), on_square(0, 1)),
]


@pytest.mark.parametrize("name,make,bfn", CASES)
def test_dss_brute_force(orc, name, make, bfn):
    o = make(orc)
    groups, bnd = global_numbering(o, bfn)
    assert len(groups) == o.dof()
    rng = np.random.default_rng(0)
    a = rng.standard_normal(o.size)
    gs = a.copy()
    um = a.copy()
    mk = a.copy()
    for g, b in zip(groups, bnd):
        if len(g) > 1 and not b:
            s = 0.0
            for k in g:  # left fold from +0.0 in ascending element order
                s += a[k]
            gs[g] = s
            um[g] = a[g[0]]
        if b:
            mk[g] = 0.0
        else:
            mk[g[1:]] = 0.0
    assert np.array_equal(o.global_sum(a), gs)
    assert np.array_equal(o.unmask(a), um)
    assert np.array_equal(o.mask(a), mk)
    bvec = rng.standard_normal(o.size)
    owners = np.array([g[0] for g in groups])
    assert abs(o.dot(a, bvec) - np.dot(a[owners], bvec[owners])) < 1e-12 * np.sqrt(o.size)
    one = np.ones(o.size)
    assert o.dot(one, one) == o.dof()  # mpi_tests/dot.cpp:20-26


def test_small_mesh_topology(orc):
    o = orc.Mesh.load(3, SMALL_MESH)
    assert (o.n_nodes, o.n_elements, o.n_edges) == (58, 45, 102)
    assert o.n_edges - o.n_interior_edges == 24
    e = o.edges()
    interior = e[e[:, 0] == 0]
    assert np.all(interior[:, 1] < interior[:, 2])  # owner = first element to visit
    assert set(np.abs(interior[:, 4])) <= {1, 2, 3, 4}
    # orientation: start/delta follow the sign of the second side
    assert np.all((interior[:, 4] > 0) == (interior[:, 5] == 1))
    assert np.all((interior[:, 4] < 0) == (interior[:, 6] == -1))
    for p in (1, 2, 5):
        m = orc.Mesh.load(p, SMALL_MESH)
        assert m.dof() == 58 + 102 * (p - 1) + 45 * (p - 1) ** 2


def test_galerkin_is_assembled_operator(orc):
    # galerkin(L, u) equals mask(global_sum(element_op(unmask(u))))
    o = orc.Mesh.load(4, SMALL_MESH)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(o.size)
    ref = o.mask(o.global_sum(o.element_op(0, o.unmask(u))))
    assert np.array_equal(o.galerkin(0, u), ref)
