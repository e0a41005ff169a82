"""GPU parity of the element operators and the DSS layer against the oracle.

Bars: element operators and galerkin applies within 1e-12 relative (FP64);
global_sum / unmask / mask bit-identical; dot within 1e-13 relative.
"""
import numpy as np
import pytest

from conftest import SMALL_MESH, rel, structured

pytestmark = pytest.mark.gpu

OPS = [0, 1, 2, 3, 4, 5]  # laplace, helmholtz, mass, schro_res, schro_f, schro_l


def meshes(sem, orc):
    out = []
    for p in (3, 6):
        out.append(("small_mesh_p%d" % p, sem.load_mesh(p, SMALL_MESH), orc.Mesh.load(p, SMALL_MESH)))
    for (n, p, jit, seed) in ((8, 4, 0.1, 0), (5, 8, 0.15, 1), (3, 12, 0.1, 3), (4, 1, 0.1, 5), (3, 15, 0.1, 7)):
        xy, q = structured(n, (0, 1, 0, 1), jit, seed)
        out.append(("struct%d_p%d" % (n, p), sem.Mesh.from_arrays(p, xy, q), orc.Mesh.from_arrays(p, xy, q)))
    return out


@pytest.fixture(scope="module")
def mesh_pairs(sem, orc):
    return meshes(sem, orc)


def test_mesh_topology_matches(mesh_pairs):
    for name, m, o in mesh_pairs:
        assert (m.N, m.n_elements, m.n_nodes, m.n_edges, m.dof()) == (o.N, o.n_elements, o.n_nodes, o.n_edges,
                                                                      o.dof()), name
        X, Y = m.coords()
        Xo, Yo = o.coords()
        if m.p <= 8:  # literal GLL table: identical nodes
            assert np.array_equal(X, Xo) and np.array_equal(Y, Yo), name
        assert rel(X, Xo) < 1e-15 and rel(Y, Yo) < 1e-15, name
        for a, b in zip(m.metrics(), o.metrics()):
            assert rel(a, b) < 1e-15, name


@pytest.mark.parametrize("kind", OPS)
def test_element_op_parity(mesh_pairs, sem, kind):
    rng = np.random.default_rng(10 + kind)
    for name, m, o in mesh_pairs:
        nc = 2 if kind in (3, 4) else 1
        u = rng.standard_normal(nc * m.size)
        coef = rng.uniform(0.5, 2.0, m.size) if kind in (1, 3, 4, 5) else None
        alpha = 0.37
        y = sem.element_op(kind, u, m, coef, alpha)
        yo = o.element_op(kind, u, coef, alpha)
        assert rel(y, yo) < 1e-12, (name, kind, rel(y, yo))


def test_glaplace_parity(mesh_pairs, sem):
    rng = np.random.default_rng(1)
    for name, m, o in mesh_pairs:
        u = rng.standard_normal(m.size)
        assert rel(sem.glaplace(m, u), o.element_op(0, u)) < 1e-12, name


@pytest.mark.parametrize("kind", OPS)
def test_galerkin_parity(mesh_pairs, sem, kind):
    rng = np.random.default_rng(20 + kind)
    for name, m, o in mesh_pairs:
        nc = 2 if kind in (3, 4) else 1
        u = rng.standard_normal(nc * m.size)
        coef = rng.uniform(0.5, 2.0, m.size) if kind in (1, 3, 4, 5) else None
        y = sem.galerkin_op(kind, u, m, coef, 0.21)
        yo = o.galerkin(kind, u, coef, 0.21)
        assert rel(y, yo) < 1e-12, (name, kind, rel(y, yo))
        # masked form: non-owner copies and Dirichlet entries are exactly zero in both
        assert np.array_equal(y == 0, yo == 0), (name, kind)


def test_dss_bit_identical(mesh_pairs, sem):
    rng = np.random.default_rng(3)
    for name, m, o in mesh_pairs:
        a = rng.standard_normal(m.size)
        assert np.array_equal(sem.global_sum(a, m), o.global_sum(a)), name
        assert np.array_equal(sem.unmask(a, m), o.unmask(a)), name
        assert np.array_equal(sem.mask(a, m), o.mask(a)), name
        # signed zeros included
        z = np.where(rng.random(m.size) < 0.5, -0.0, 0.0)
        assert np.array_equal(np.signbit(sem.global_sum(z, m)), np.signbit(o.global_sum(z))), name


@pytest.mark.parametrize("kind", [0, 1, 3])
def test_galerkin_gather_equals_pair_dss(mesh_pairs, sem, kind):
    # galerkin_op assembles the element operator's output with one gather per
    # element column; the same apply assembled by the pair/corner DSS
    # (global_sum, then mask) gives the same bits, the valence-6 corner of
    # small_mesh included
    rng = np.random.default_rng(40 + kind)
    for name, m, o in mesh_pairs:
        nc = 2 if kind == 3 else 1
        u = rng.standard_normal(nc * m.size)
        coef = rng.uniform(0.5, 2.0, m.size) if kind in (1, 3) else None
        y = sem.galerkin_op(kind, u, m, coef, 0.21)
        uu = np.concatenate([sem.unmask(u[c * m.size:(c + 1) * m.size], m) for c in range(nc)])
        e = sem.element_op(kind, uu, m, coef, 0.21)
        ref = np.concatenate([sem.mask(sem.global_sum(e[c * m.size:(c + 1) * m.size], m), m) for c in range(nc)])
        assert np.array_equal(y, ref), (name, kind)


def test_dss_repeatable(mesh_pairs, sem):
    rng = np.random.default_rng(4)
    name, m, o = mesh_pairs[0]
    u = rng.standard_normal(2 * m.size)
    V = rng.uniform(-1, 1, m.size)
    y1 = sem.galerkin_op(3, u, m, V, 0.3)
    y2 = sem.galerkin_op(3, u, m, V, 0.3)
    assert np.array_equal(y1, y2)


def test_dot_parity(mesh_pairs, sem):
    rng = np.random.default_rng(5)
    for name, m, o in mesh_pairs:
        a = rng.standard_normal(m.size)
        b = rng.standard_normal(m.size)
        d, do = sem.dot(m, a, b), o.dot(a, b)
        assert abs(d - do) <= 1e-13 * max(abs(do), 1.0), name
        # known answer of mpi_tests/dot.cpp:20-26: dot(1, 1) = dof
        one = np.ones(m.size)
        assert sem.dot(m, one, one) == m.dof(), name


def test_gproj_parity(mesh_pairs, sem):
    rng = np.random.default_rng(6)
    for name, m, o in mesh_pairs:
        u = rng.standard_normal(m.size)
        assert rel(sem.gproj(u, m), o.galerkin(2, u)) < 1e-12, name
