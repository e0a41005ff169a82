"""GPU Krylov solvers and PDE drivers against the oracle and closed forms."""
import numpy as np
import pytest

from conftest import SMALL_MESH, rel, structured

pytestmark = pytest.mark.gpu


def fd_problem(n=20, k=0.0):
    # tests/pcg.cpp:8-17 / tests/minres.cpp:8-18: -u'' - k u = x^2 on [-1, 1]
    x = np.linspace(-1, 1, n)
    h = x[1] - x[0]
    return -1 / h**2, 2 / h**2 - k, -1 / h**2, x * x


@pytest.mark.parametrize("k", [0.0, 2.0])
def test_tridiag_solvers(sem, orc, k):
    lo, di, up, f = fd_problem(20, k)
    n = f.size
    A = np.diag(np.full(n, di)) + np.diag(np.full(n - 1, lo), -1) + np.diag(np.full(n - 1, up), 1)
    x = np.zeros(n)
    r = sem.minres_tridiag(lo, di, up, f, x, n, 1e-5)
    xo, ok, it, res = orc.minres_tridiag(lo, di, up, f, np.zeros(n), n, 1e-5)
    assert r.success == ok and r.n_iter == it
    assert np.linalg.norm(A @ x - f) <= 1e-5 * np.linalg.norm(f)  # tests/minres.cpp:24-28
    assert rel(x, xo) < 1e-9
    if k == 0.0:
        x = np.zeros(n)
        r = sem.pcg_tridiag(lo, di, up, f, x, n, 1e-5)
        xo, ok, it, res = orc.pcg_tridiag(lo, di, up, f, np.zeros(n), n, 1e-5)
        assert r.success == ok and r.n_iter == it
        assert np.linalg.norm(A @ x - f) <= 1e-5 * np.linalg.norm(f)  # tests/pcg.cpp:23-27
        assert rel(x, xo) < 1e-9


def test_poisson_small_mesh(sem, orc):
    # tests/poisson.cpp:11-33
    m = sem.load_mesh(6, SMALL_MESH)
    o = orc.Mesh.load(6, SMALL_MESH)
    F = m.field(lambda x, y: x * x - y * y)
    u = np.zeros(m.size)
    r = sem.poisson(u, F, m, m.dof(), 1e-10)
    uo, ok, it, res = o.poisson(np.zeros(m.size), F, m.dof(), 1e-10)
    assert r.success and ok
    assert abs(r.n_iter - it) <= 2
    assert rel(u, uo) < 1e-9


def test_poisson_config1_manufactured(sem, orc):
    # 8x8 jittered mesh on [0,1]^2, p = 4, u = sin(pi x) sin(pi y), f = 2 pi^2 u, tol 1e-12
    xy, q = structured(8, (0, 1, 0, 1), 0.1, 0)
    m = sem.Mesh.from_arrays(4, xy, q)
    o = orc.Mesh.from_arrays(4, xy, q)
    ex = m.field(lambda x, y: np.sin(np.pi * x) * np.sin(np.pi * y))
    F = 2 * np.pi**2 * ex
    u = np.zeros(m.size)
    r = sem.poisson(u, F, m, m.dof(), 1e-12)
    uo, ok, it, res = o.poisson(np.zeros(m.size), F, m.dof(), 1e-12)
    assert r.success and ok
    uu = sem.unmask(u, m)
    assert rel(uu, o.unmask(uo)) < 1e-9
    # L2 error against the exact solution (discretisation error at p = 4 on the jittered mesh)
    e = uu - ex
    err = np.sqrt(sem.dot(m, e, e) / sem.dot(m, ex, ex))
    assert err < 5e-3


def test_helmholtz_small_mesh_indefinite(sem, orc):
    # tests/helmholtz.cpp:6-31
    m = sem.load_mesh(6, SMALL_MESH)
    o = orc.Mesh.load(6, SMALL_MESH)
    F = m.field(lambda x, y: x * x - y * y)
    c = np.full(m.size, -100.0 - np.sqrt(2.0))
    u = np.zeros(m.size)
    r = sem.helmholtz(u, c, F, m, 10 * m.dof(), 1e-10)
    uo, ok, it, res = o.helmholtz(np.zeros(m.size), c, F, 10 * m.dof(), 1e-10)
    assert r.success and ok
    assert rel(u, uo) < 1e-8


def test_helmholtz_reduced_config2(sem, orc):
    # config-2 recipe at 16x16 (oracle-sized): jitter 0.15, p = 8, k^2 = 10, Gaussian forcing, tol 1e-10
    xy, q = structured(16, (0, 1, 0, 1), 0.15, 1)
    m = sem.Mesh.from_arrays(8, xy, q)
    o = orc.Mesh.from_arrays(8, xy, q)
    rng = np.random.default_rng(1)
    cen = rng.uniform(0, 1, (32, 2))
    wid = rng.uniform(0.02, 0.1, 32)
    amp = rng.standard_normal(32)
    F = m.field(lambda x, y: sum(a * np.exp(-((x - c[0])**2 + (y - c[1])**2) / (2 * s * s))
                                 for a, c, s in zip(amp, cen, wid)))
    c = np.full(m.size, 10.0)
    u = np.zeros(m.size)
    r = sem.helmholtz(u, c, F, m, 10000, 1e-10)
    uo, ok, it, res = o.helmholtz(np.zeros(m.size), c, F, 10000, 1e-10)
    assert r.success and ok
    assert abs(r.n_iter - it) <= max(3, it // 50)
    assert rel(sem.unmask(u, m), o.unmask(uo)) < 1e-9


def test_stage_solve_parity(sem, orc):
    m = sem.load_mesh(4, SMALL_MESH)
    o = orc.Mesh.load(4, SMALL_MESH)
    rng = np.random.default_rng(7)
    V = m.field(lambda x, y: 0.5 * (x * x + y * y))
    b = o.galerkin(2, rng.standard_normal(m.size))
    b = np.concatenate([b, o.galerkin(2, rng.standard_normal(m.size))])
    alpha = 0.25 * 1e-2
    z, r = sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12)
    zo, ok, it, res = o.stage_solve(V, alpha, b, None, 5000, 1e-12)
    assert r.success and ok
    zu = np.concatenate([o.unmask(zo[:m.size]), o.unmask(zo[m.size:])])
    assert rel(z, zu) < 1e-10
    # mixed precision: FP32 inner solve inside FP64 refinement, same answer
    z2, r2 = sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, sem.PREC_MIXED)
    assert r2.success
    assert rel(z2, zu) < 1e-10


def test_solver_repeatable(sem):
    m = sem.load_mesh(5, SMALL_MESH)
    F = m.field(lambda x, y: np.cos(x) * y)
    c = np.full(m.size, 3.0)
    u1 = np.zeros(m.size)
    u2 = np.zeros(m.size)
    r1 = sem.helmholtz(u1, c, F, m, 5000, 1e-12)
    r2 = sem.helmholtz(u2, c, F, m, 5000, 1e-12)
    assert r1.n_iter == r2.n_iter and r1.residual == r2.residual
    assert np.array_equal(u1, u2)


@pytest.mark.parametrize("precision", [0, 1])
def test_step_shapes_bit_identical(sem, precision):
    # every work shape of the Lanczos-step kernels (chunks per workgroup, one
    # workgroup per item or a persistent grid) gives the same bits
    xy, q = structured(20, (-2, 2, -2, 2), 0.1, 3)
    m = sem.Mesh.from_arrays(6, xy, q)
    sem.set_step_mode(m, 0)  # the two-launch step (the fused one has its own test)
    rng = np.random.default_rng(11)
    V = m.field(lambda x, y: x * x + y * y)
    b = np.concatenate([sem.gproj(rng.standard_normal(m.size), m), sem.gproj(rng.standard_normal(m.size), m)])
    ref = None
    for shape in [(1, 0, 1, 0), (2, 0, 2, 0), (1, 3, 2, 5), (2, 64, 1, 7), (1, 1, 1, 1)]:
        sem.set_step_shape(m, shape, precision)
        assert sem.get_step_shape(m, precision) == shape
        z, r = sem.stage_solve(m, V, 2e-3, b, None, 3000, 1e-11, precision)
        assert r.success
        if ref is None:
            ref = (z, r.n_iter, r.residual)
        else:
            assert np.array_equal(z, ref[0]), shape
            assert (r.n_iter, r.residual) == ref[1:], shape
    chosen, rows = sem.autotune(m, sem.OP_SCHRO_RES, precision, 12)
    assert len(rows) >= 2 and all(r[2] > 0 and r[3] > 0 for r in rows)
    assert (chosen[0], chosen[1]) in [(r[0], r[1]) for r in rows]
    z, r = sem.stage_solve(m, V, 2e-3, b, None, 3000, 1e-11, precision)
    assert np.array_equal(z, ref[0])


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_mode_bit_identical(sem, precision):
    # Lanczos-step batches replayed as captured graphs give the eager path's
    # bits, on every replay (several solves reuse one captured graph)
    xy, q = structured(16, (-2, 2, -2, 2), 0.1, 5)
    m = sem.Mesh.from_arrays(5, xy, q)
    sem.set_step_mode(m, 0)  # graphs replay the two-launch step
    rng = np.random.default_rng(13)
    V = m.field(lambda x, y: x * x + 0.5 * y * y)
    bs = [np.concatenate([sem.gproj(rng.standard_normal(m.size), m), sem.gproj(rng.standard_normal(m.size), m)])
          for _ in range(2)]
    sem.set_graph_mode(m, False)
    ref = [sem.stage_solve(m, V, 3e-3, b, None, 3000, 1e-11, precision) for b in bs]
    for steps in (3, 12):
        sem.set_graph_mode(m, True, steps)
        for rep in range(3):
            for b, (z0, r0) in zip(bs, ref):
                z, r = sem.stage_solve(m, V, 3e-3, b, None, 3000, 1e-11, precision)
                assert r.success and (r.n_iter, r.residual) == (r0.n_iter, r0.residual), (steps, rep)
                assert np.array_equal(z, z0), (steps, rep)
    with pytest.raises(sem.SemError):
        sem.set_graph_mode(m, True, 4)
    sem.set_graph_mode(m, False)
