"""Pins the oracle's coefficient terms: the coefficient mass (W(u.c)W).J of
Helmholtz and of the Schrodinger L / Res / f operators
(openmp_impl/helmholtz.hpp:38-46, schrodinger.hpp:37-43, :163-207), against
exact integrals, a manufactured Helmholtz solution and the closed-form
stability function of a discrete eigenmode with a nonzero potential.

Each check is chosen so that a dropped coefficient term, a wrong sign, or a
transposed operand (c(i,j) read as c(j,i), or u and c swapped with J) fails."""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import SMALL_MESH, structured
from test_oracle_dss import global_numbering, on_square
from test_oracle_solvers import assembled, stability, tableau


def sheared_mesh(n, shear, p, orc):
    """n x n parallelogram mesh of the unit square sheared by x -> x + shear*y."""
    xy, q = structured(n, (0, 1, 0, 1), 0.0, 0)
    xy = np.asarray(xy, dtype=np.float64).reshape(-1, 2).copy()
    xy[:, 0] += shear * xy[:, 1]
    return orc.Mesh.from_arrays(p, xy.ravel(), q)


def exact_integral_sheared(f, shear, order=24):
    """Integral of f(x, y) over the sheared unit square by tensor Gauss-Legendre
    in the unsheared coordinates (x = X + shear*Y, y = Y, Jacobian 1)."""
    t, w = np.polynomial.legendre.leggauss(order)
    t = 0.5 * (t + 1.0)
    w = 0.5 * w
    Xq, Yq = np.meshgrid(t, t, indexing="ij")
    W = np.outer(w, w)
    return float(np.sum(W * f(Xq + shear * Yq, Yq)))


def coef_mass(o, c, u):
    """element-local (W(u.c)W).J alone: the Helmholtz element op minus glaplace"""
    return o.element_op(1, u, c) - o.element_op(0, u)


@pytest.mark.parametrize("p,shear", [(4, 0.0), (5, 0.35), (7, -0.6)])
def test_coefficient_mass_is_exact_quadrature(orc, p, shear):
    # On affine elements GLL quadrature of degree 2p-1 per direction is exact for
    # u*c*v of total degree <= 2p-1, so v^T M_c u summed over elements is the
    # integral of u c v over the domain. The integrand is chosen without x<->y
    # symmetry, so reading c or u transposed changes the value.
    o = sheared_mesh(3, shear, p, orc)
    X, Y = o.coords()
    fu = lambda x, y: 1.0 + x * x - 0.5 * y   # noqa: E731
    fc = lambda x, y: 2.0 + x - 3.0 * y * y    # noqa: E731
    fv = lambda x, y: 0.3 + y - x * y          # noqa: E731
    assert 2 + 2 + 2 <= 2 * p - 1
    got = float(np.dot(fv(X, Y), coef_mass(o, fc(X, Y), fu(X, Y))))
    exact = exact_integral_sheared(lambda x, y: fu(x, y) * fc(x, y) * fv(x, y), shear)
    assert abs(got - exact) < 1e-12 * max(1.0, abs(exact))
    # the plain mass op with c folded into u gives the same numbers (c is a
    # pointwise factor of u, element by element)
    assert np.allclose(coef_mass(o, fc(X, Y), fu(X, Y)), o.element_op(2, fc(X, Y) * fu(X, Y)), rtol=1e-13, atol=1e-15)
    # and a transposed coefficient (c(j,i) for c(i,j) in every element) is caught
    N = o.N
    ct = fc(X, Y).reshape(-1, N, N).transpose(0, 2, 1).ravel()
    bad = float(np.dot(fv(X, Y), coef_mass(o, ct, fu(X, Y))))
    assert abs(bad - exact) > 1e-6


def test_coefficient_mass_integrates_on_jittered_quads(orc):
    # sum_ij w_i w_j J c = integral of c for a polynomial c of total degree d with
    # d + 1 <= 2p - 1 (bilinear elements: J has degree 1 per direction); the
    # jittered interior leaves the domain [0,1]^2 unchanged
    p = 5
    o = orc.Mesh.from_arrays(p, *structured(6, (0, 1, 0, 1), 0.15, 1))
    X, Y = o.coords()
    c = 1.0 + 2.0 * X**3 - X * Y + 0.5 * Y**2
    got = float(np.sum(coef_mass(o, c, np.ones(o.size))))
    exact = 1.0 + 2.0 / 4 - 1.0 / 4 + 0.5 / 3
    assert abs(got - exact) < 1e-13


def test_schrodinger_L_is_stiffness_plus_potential_mass(orc):
    # L u = K u + M (V u) assembled (schrodinger.hpp:37-43): the potential term of L
    # is the pinned mass operator applied to V u, V non-constant
    o = orc.Mesh.from_arrays(4, *structured(5, (-1, 1, -1, 1), 0.1, 3))
    X, Y = o.coords()
    V = 3.0 + X - 2 * Y * Y
    rng = np.random.default_rng(5)
    u = o.unmask(rng.standard_normal(o.size))
    L = o.galerkin(5, u, V)
    ref = o.galerkin(0, u) + o.galerkin(2, V * u)
    assert np.max(np.abs(L - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("c0", [10.0, None])
def test_helmholtz_manufactured_spectral_convergence(orc, c0):
    # -Lap u + c u = f with u = sin(pi x) sin(pi y) on an undistorted 4x4 mesh of
    # [0,1]^2 (helmholtz.hpp:18-64 solves (K + M_c) u = M f): c = 10 (config 2's
    # k^2) and a non-constant c = 10 + 5x - 3y^2. The error falls spectrally with
    # p only if the coefficient term is right (a dropped or mis-signed c term
    # leaves an O(1) error).
    errs = []
    for p in (2, 4, 6, 8):
        o = orc.Mesh.from_arrays(p, *structured(4, (0, 1, 0, 1), 0.0, 0))
        X, Y = o.coords()
        ex = np.sin(np.pi * X) * np.sin(np.pi * Y)
        c = np.full(o.size, c0) if c0 is not None else 10.0 + 5.0 * X - 3.0 * Y * Y
        f = 2 * np.pi**2 * ex + c * ex
        u, ok, it, res = o.helmholtz(np.zeros(o.size), c, f, 10 * o.dof(), 1e-14)
        assert ok
        e = o.unmask(u) - ex
        errs.append(np.sqrt(o.dot(e, e) / o.dot(ex, ex)))
    assert all(b < a / 10 for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-7


@pytest.fixture(scope="module")
def eigenmode(orc):
    o = orc.Mesh.load(3, SMALL_MESH)
    groups, bnd = global_numbering(o, on_square(-1, 1))
    K, inner = assembled(o, groups, bnd, 0)
    M, _ = assembled(o, groups, bnd, 2)
    lam, phi = sla.eigh(K, M)
    k = 3
    psi = np.zeros(o.size)
    for c, g in enumerate(inner):
        psi[g] = phi[c, k]
    return o, lam[k], psi


@pytest.mark.parametrize("scheme", [1, 2])
@pytest.mark.parametrize("v0", [7.5, -3.25])
def test_sdirk_step_with_constant_potential(orc, eigenmode, scheme, v0):
    # With V = v0 constant, L = K + v0 M, so a discrete eigenvector of (K, M)
    # is one of L with eigenvalue lambda + v0: one step multiplies it by
    # R(i (lambda + v0) dt). Dropping the potential gives R(i lambda dt), a sign
    # error R(i (lambda - v0) dt); both differ from the closed form here.
    o, lam, psi = eigenmode
    dt = 0.6 / lam
    V = np.full(o.size, v0)
    re, im, ok, stats, _ = o.schrodinger(scheme, psi, np.zeros(o.size), V, dt, 10000, 1e-14, dt)
    assert ok and stats[0] == 1
    A, b = tableau(scheme)
    R = stability(A, b, 1j * (lam + v0) * dt)
    got = re + 1j * im
    assert np.linalg.norm(got - R * psi) < 1e-10 * np.linalg.norm(psi)
    for wrong in (lam, lam - v0):
        assert abs(stability(A, b, 1j * wrong * dt) - R) > 1e-3


def test_bdf2_with_constant_potential(orc, eigenmode):
    # implicit Euler, then BDF2 (schrodinger.hpp:91-102) on psi' = i (lambda + v0) psi
    o, lam, psi = eigenmode
    v0 = 4.0
    dt = 0.4 / lam
    V = np.full(o.size, v0)
    re, im, ok, stats, _ = o.schrodinger(0, psi, np.zeros(o.size), V, 3 * dt, 10000, 1e-14, dt)
    assert ok and stats[0] == 3
    z = 1j * (lam + v0) * dt
    y1 = 1.0 / (1 - z)
    y2 = (4 / 3 * y1 - 1 / 3) / (1 - 2 / 3 * z)
    y3 = (4 / 3 * y2 - 1 / 3 * y1) / (1 - 2 / 3 * z)
    assert np.linalg.norm(re + 1j * im - y3 * psi) < 1e-10 * np.linalg.norm(psi)


def test_sdirk_step_with_varying_potential_matches_dense(orc):
    # Non-constant V: one DIRK4S3 step equals the dense computation with the
    # assembled K, diagonal lumped M and M_V = diag(M) V on the interior nodes:
    # psi_1 = R(dt (i M^-1 (K + M_V))) psi_0, evaluated with the stage equations
    o = orc.Mesh.load(3, SMALL_MESH)
    groups, bnd = global_numbering(o, on_square(-1, 1))
    K, inner = assembled(o, groups, bnd, 0)
    M, _ = assembled(o, groups, bnd, 2)
    X, Y = o.coords()
    V = 2.0 + np.sin(1.3 * X) * Y - 0.7 * Y * Y
    own = [g[0] for g in inner]
    Vn = V[own]
    Lmat = K + np.diag(np.diag(M) * Vn)
    assert np.max(np.abs(M - np.diag(np.diag(M)))) == 0  # GLL mass is lumped
    rng = np.random.default_rng(9)
    c0 = rng.standard_normal(len(inner)) + 1j * rng.standard_normal(len(inner))
    psi0 = np.zeros(o.size, complex)
    for c, g in enumerate(inner):
        psi0[g] = c0[c]
    dt = 2e-3
    re, im, ok, stats, _ = o.schrodinger(1, psi0.real.copy(), psi0.imag.copy(), V, dt, 10000, 1e-14, dt)
    assert ok and stats[0] == 1
    A, b = tableau(1)

    def dense_step(Lm):
        Aop = 1j * np.linalg.solve(M, Lm)  # psi' = i M^-1 L psi (reference convention, no 1/2)
        s, n = len(b), len(inner)
        Z = np.zeros((s, n), complex)
        for k in range(s):
            rhs = Aop @ (c0 + sum(A[k, j] * Z[j] for j in range(k)))
            Z[k] = np.linalg.solve(np.eye(n) - A[k, k] * dt * Aop, dt * rhs)
        return c0 + sum(b[k] * Z[k] for k in range(s))

    c1 = dense_step(Lmat)
    got = (re + 1j * im)[own]
    assert np.linalg.norm(got - c1) < 1e-9 * np.linalg.norm(c1)
    # the same step with the potential dropped, or with its sign flipped, is far off
    for Lm in (K, K - np.diag(np.diag(M) * Vn)):
        assert np.linalg.norm(dense_step(Lm) - c1) > 1e-4 * np.linalg.norm(c1)
