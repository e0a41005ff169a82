"""Pins the oracle's Krylov solvers, PDE drivers and steppers: the reference's
own solver tests, direct solves, and closed-form stability functions on a
discrete eigenmode."""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import SMALL_MESH, structured
from test_oracle_dss import global_numbering, on_square


def fd(n=20, k=0.0):
    x = np.linspace(-1, 1, n)
    h = x[1] - x[0]
    A = (np.diag(np.full(n, 2.0)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)) / h**2 - k * np.eye(n)
    return -1 / h**2, 2 / h**2 - k, -1 / h**2, x * x, A


def test_minres_reference_fd(orc):
    # tests/minres.cpp:5-30 (indefinite, k = 2)
    lo, di, up, f, A = fd(20, 2.0)
    x, ok, it, res = orc.minres_tridiag(lo, di, up, f, np.zeros(20), 20, 1e-5)
    assert ok and np.linalg.norm(A @ x - f) <= 1e-5 * np.linalg.norm(f)
    # tighter tolerance: the direct solution
    x, ok, it, res = orc.minres_tridiag(lo, di, up, f, np.zeros(20), 200, 1e-14)
    assert ok and np.linalg.norm(x - np.linalg.solve(A, f)) < 1e-10 * np.linalg.norm(x)


def test_pcg_reference_fd(orc):
    # tests/pcg.cpp:5-29
    lo, di, up, f, A = fd(20)
    x, ok, it, res = orc.pcg_tridiag(lo, di, up, f, np.zeros(20), 20, 1e-5)
    assert ok and np.linalg.norm(A @ x - f) <= 1e-5 * np.linalg.norm(f)
    x, ok, it, res = orc.pcg_tridiag(lo, di, up, f, np.zeros(20), 200, 1e-13)
    assert ok and np.linalg.norm(x - np.linalg.solve(A, f)) < 1e-9 * np.linalg.norm(x)
    assert abs(res - np.sqrt(0.0 + np.dot(f - A @ x, f - A @ x))) < 1e-3 * np.linalg.norm(f)


def test_minres_one_eigenvalue_one_step(orc):
    # a multiple of the identity: MINRES converges at i = 1 with x = b / d (minres.hpp:46-99)
    b = np.linspace(1, 2, 7)
    x, ok, it, res = orc.minres_tridiag(0.0, 4.0, 0.0, b, np.zeros(7), 10, 1e-12)
    assert ok and it == 1 and np.allclose(x, b / 4, rtol=1e-15)


def test_minres_stop_test_is_r0_relative(orc):
    # minres.hpp:28,92: phibar < tol*||b - A x0|| + tol; iteration counts are 1-based,
    # and maxit + 1 on failure
    lo, di, up, f, A = fd(30, 2.0)
    x, ok, it, res = orc.minres_tridiag(lo, di, up, f, np.zeros(30), 3, 1e-14)
    assert not ok and it == 4


def test_poisson_reference_test(orc):
    # tests/poisson.cpp:11-33; the solution satisfies K u = M f on owner DOFs
    o = orc.Mesh.load(6, SMALL_MESH)
    X, Y = o.coords()
    F = X * X - Y * Y
    u, ok, it, res = o.poisson(np.zeros(o.size), F, o.dof(), 1e-10)
    assert ok
    r = o.galerkin(0, u) - o.galerkin(2, F)
    assert np.linalg.norm(r) < 1e-9 * np.linalg.norm(o.galerkin(2, F))


def test_helmholtz_reference_test(orc):
    # tests/helmholtz.cpp:6-31 (indefinite, c = -100 - sqrt(2))
    o = orc.Mesh.load(6, SMALL_MESH)
    X, Y = o.coords()
    F = X * X - Y * Y
    c = np.full(o.size, -100.0 - np.sqrt(2.0))
    u, ok, it, res = o.helmholtz(np.zeros(o.size), c, F, 10 * o.dof(), 1e-10)
    assert ok
    r = o.galerkin(1, u, c) - o.galerkin(2, F)
    assert np.linalg.norm(r) < 1e-8 * np.linalg.norm(o.galerkin(2, F))


def test_poisson_manufactured_spectral_convergence(orc):
    # u = sin(pi x) sin(pi y) on an undistorted 4 x 4 mesh of [0,1]^2: error falls spectrally with p
    errs = []
    for p in (2, 4, 6, 8):
        o = orc.Mesh.from_arrays(p, *structured(4, (0, 1, 0, 1), 0.0, 0))
        X, Y = o.coords()
        ex = np.sin(np.pi * X) * np.sin(np.pi * Y)
        u, ok, it, res = o.poisson(np.zeros(o.size), 2 * np.pi**2 * ex, o.dof(), 1e-13)
        assert ok
        e = o.unmask(u) - ex
        errs.append(np.sqrt(o.dot(e, e) / o.dot(ex, ex)))
    assert all(b < a / 10 for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 1e-7


def assembled(o, groups, bnd, kind):
    inner = [g for g, b in zip(groups, bnd) if not b]
    owners = [g[0] for g in inner]
    A = np.zeros((len(inner), len(inner)))
    for c, g in enumerate(inner):
        e = np.zeros(o.size)
        e[g] = 1.0
        A[:, c] = o.galerkin(kind, e)[owners]
    return A, inner


def tableau(scheme):
    if scheme == 1:  # Crouzeix 3-stage SDIRK4 (Hairer-Wanner II, Table 6.4 family)
        g = np.cos(np.pi / 18) / np.sqrt(3) + 0.5
        d = 1 / (6 * (2 * g - 1) ** 2)
        A = np.array([[g, 0, 0], [0.5 - g, g, 0], [2 * g, 1 - 4 * g, g]])
        b = np.array([d, 1 - 2 * d, d])
    else:  # L-stable 5-stage SDIRK4 (Hairer-Wanner II, Table 6.5)
        A = np.array([[1 / 4, 0, 0, 0, 0], [1 / 2, 1 / 4, 0, 0, 0], [17 / 50, -1 / 25, 1 / 4, 0, 0],
                      [371 / 1360, -137 / 2720, 15 / 544, 1 / 4, 0], [25 / 24, -49 / 48, 125 / 16, -85 / 12, 1 / 4]])
        b = A[-1].copy()
    return A, b


def stability(A, b, z):
    s = len(b)
    return 1 + z * b @ np.linalg.solve(np.eye(s) - z * A, np.ones(s))


@pytest.fixture(scope="module")
def eigenmode(orc):
    o = orc.Mesh.load(3, SMALL_MESH)
    groups, bnd = global_numbering(o, on_square(-1, 1))
    K, inner = assembled(o, groups, bnd, 0)
    M, _ = assembled(o, groups, bnd, 2)
    lam, phi = sla.eigh(K, M)
    k = 2
    psi = np.zeros(o.size)
    for c, g in enumerate(inner):
        psi[g] = phi[c, k]
    return o, lam[k], psi


@pytest.mark.parametrize("scheme", [1, 2])
def test_sdirk_stability_function(orc, eigenmode, scheme):
    # one step on a discrete eigenvector multiplies it by R(i lambda dt) exactly
    o, lam, psi = eigenmode
    dt = 0.8 / lam
    re, im, ok, stats, _ = o.schrodinger(scheme, psi, np.zeros(o.size), np.zeros(o.size), dt, 10000, 1e-14, dt)
    assert ok and stats[0] == 1
    R = stability(*tableau(scheme), 1j * lam * dt)
    got = (re + 1j * im)
    assert np.linalg.norm(got - R * psi) < 1e-10 * np.linalg.norm(psi)
    assert abs(R - np.exp(1j * lam * dt)) > 1e-4  # the closed form is discriminating at this dt


def test_bdf2_recurrence(orc, eigenmode):
    # implicit Euler first step, then BDF2 (schrodinger.hpp:96-102) on psi' = i lambda psi
    o, lam, psi = eigenmode
    dt = 0.5 / lam
    re, im, ok, stats, _ = o.schrodinger(0, psi, np.zeros(o.size), np.zeros(o.size), 3 * dt, 10000, 1e-14, dt)
    assert ok and stats[0] == 3
    z = 1j * lam * dt
    y0 = 1.0
    y1 = y0 / (1 - z)
    y2 = (4 / 3 * y1 - 1 / 3 * y0) / (1 - 2 / 3 * z)
    y3 = (4 / 3 * y2 - 1 / 3 * y1) / (1 - 2 / 3 * z)
    assert np.linalg.norm(re + 1j * im - y3 * psi) < 1e-10 * np.linalg.norm(psi)


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_reference_stepper_tests(orc, scheme):
    # tests/schrodinger.cpp: small_mesh, p = 3, V = 0, sin(pi x) sin(pi y), T = 0.01, dt = T/10, tol 1e-8
    o = orc.Mesh.load(3, SMALL_MESH)
    X, Y = o.coords()
    re, im, ok, stats, _ = o.schrodinger(scheme, np.sin(np.pi * X) * np.sin(np.pi * Y), np.zeros(o.size),
                                         np.zeros(o.size), 0.01, 20 * o.dof(), 1e-8, 0.001)
    assert ok and stats[0] == 10


# ----- preconditioned solvers (Precond argument of minres.hpp:19-20, pcg.hpp:59-60) -----
def test_preconditioned_identity_is_the_identity_path(orc):
    lo, di, up, f, A = fd(30, 2.0)
    x0, ok0, it0, r0 = orc.minres_tridiag(lo, di, up, f, np.zeros(30), 200, 1e-12)
    for consistent in (False, True):
        x1, ok1, it1, r1 = orc.minres_tridiag_pc(lo, di, up, f, np.zeros(30), 200, 1e-12, np.ones(30), consistent)
        assert (ok1, it1, r1) == (ok0, it0, r0) and np.array_equal(x1, x0)
    lo, di, up, f, A = fd(30)
    x0, ok0, it0, r0 = orc.pcg_tridiag(lo, di, up, f, np.zeros(30), 200, 1e-12)
    x1, ok1, it1, r1 = orc.pcg_tridiag_pc(lo, di, up, f, np.zeros(30), 200, 1e-12, np.ones(30))
    assert (ok1, it1, r1) == (ok0, it0, r0) and np.array_equal(x1, x0)


def test_diagonal_preconditioned_pcg(orc):
    # any symmetric positive definite preconditioner keeps the solution: a
    # varying positive diagonal converges to the direct solve, along other
    # iterates than the identity (so the preconditioner is applied every step)
    lo, di, up, f, A = fd(40)
    pd = 0.5 + 1.5 * np.random.default_rng(3).random(40)
    x, ok, it, res = orc.pcg_tridiag_pc(lo, di, up, f, np.zeros(40), 400, 1e-14, pd)
    assert ok and np.linalg.norm(x - np.linalg.solve(A, f)) < 1e-10 * np.linalg.norm(x)
    x4, _, _, _ = orc.pcg_tridiag_pc(lo, di, up, f, np.zeros(40), 4, 1e-14, pd)
    x4i, _, _, _ = orc.pcg_tridiag(lo, di, up, f, np.zeros(40), 4, 1e-14)
    assert np.linalg.norm(x4 - x4i) > 1e-3 * np.linalg.norm(x4i)


@pytest.mark.parametrize("c", [0.37, 5.0])
def test_minres_consistent_scalar_preconditioner(orc, c):
    # P = c I changes nothing in exact arithmetic when beta is normalised consistently
    lo, di, up, f, A = fd(30, 2.0)
    x0, ok0, it0, _ = orc.minres_tridiag(lo, di, up, f, np.zeros(30), 200, 1e-12)
    x1, ok1, it1, _ = orc.minres_tridiag_pc(lo, di, up, f, np.zeros(30), 200, 1e-12, np.full(30, c), True)
    assert ok1 and abs(it1 - it0) <= 1 and np.linalg.norm(x1 - x0) < 1e-9 * np.linalg.norm(x0)
    # the reference's normalisation (minres.hpp:63) scales beta by sqrt(c) and does not give that solution
    x2, ok2, it2, _ = orc.minres_tridiag_pc(lo, di, up, f, np.zeros(30), 30, 1e-12, np.full(30, c), False)
    assert np.linalg.norm(x2 - x0) > 1e-6 * np.linalg.norm(x0)


def test_minres_consistent_jacobi_indefinite(orc):
    lo, di, up, f, A = fd(40, 2.0)
    pd = 1.0 / np.abs(np.diag(A))
    x, ok, it, res = orc.minres_tridiag_pc(lo, di, up, f, np.zeros(40), 400, 1e-13, pd, True)
    assert ok and np.linalg.norm(x - np.linalg.solve(A, f)) < 1e-9 * np.linalg.norm(x)
