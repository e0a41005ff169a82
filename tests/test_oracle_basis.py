"""Pins the oracle's quadrature, derivative matrix, metrics and element
operator against the reference's known-answer tests and closed forms."""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg
from scipy.special import i0


def test_gll_reference_kat(orc):
    # tests/gauss_lobatto.cpp:11-43: integral of exp(sin(pi x)) = 2 I0(1)
    exact = 2 * i0(1.0)
    errs = []
    for p in (9, 10):
        x, w = orc.gll(p)
        errs.append(abs(np.sum(np.exp(np.sin(np.pi * x)) * w) - exact))
    assert errs[0] <= 1e-4 and errs[1] <= errs[0]


@pytest.mark.parametrize("p", range(1, 16))
def test_gll_nodes_weights(orc, p):
    x, w = orc.gll(p)
    # interior nodes are the roots of P'_p; endpoints +-1
    roots = np.sort(npleg.Legendre.basis(p).deriv().roots().real) if p > 1 else np.array([])
    tol = 5e-15 if p > 8 else 1e-14  # p <= 8: the reference's 15-digit literal table
    assert x[0] == -1 and x[-1] == 1
    assert np.max(np.abs(x[1:-1] - roots), initial=0) < tol
    assert np.allclose(x, -x[::-1], atol=1e-15, rtol=0)
    # weights 2 / (p (p+1) P_p(x)^2), exact for degree <= 2p - 1
    Pp = npleg.legval(x, [0] * p + [1])
    assert np.allclose(w, 2 / (p * (p + 1) * Pp**2), rtol=1e-14, atol=0)
    assert abs(w.sum() - 2) < 1e-14
    for k in range(2 * p):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x**k) - exact) < 1e-13


def test_gll_literals_are_the_reference_table(orc):
    # gauss_lobatto.hpp:52-57, digit for digit (including the asymmetric n = 7 entry)
    assert orc.gll(3)[0][1] == -0.447213595499958
    assert orc.gll(7)[0][3] == -0.209299217902479 and orc.gll(7)[0][4] == 0.2092992179024789
    assert orc.gll(8)[0][1] == -0.899757995411460


def test_derivative_matrix_reference_kat(orc):
    # tests/derivative_matrix.cpp:14-35: D p = p' for p = -1 + 2x + 5x^2 - 3x^3 at n = 8
    x, _ = orc.gll(8)
    D = orc.dmat(x)
    p = -1 + 2 * x + 5 * x**2 - 3 * x**3
    dp = 2 + 10 * x - 9 * x**2
    assert np.max(np.abs(D @ p - dp)) <= 1e-10


@pytest.mark.parametrize("p", [2, 5, 8, 12, 15])
def test_derivative_matrix_exact_on_polynomials(orc, p):
    x, _ = orc.gll(p)
    D = orc.dmat(x)
    for k in range(p + 1):
        assert np.max(np.abs(D @ x**k - k * x ** max(k - 1, 0))) < 1e-11 * (p**2)
    assert np.max(np.abs(D.sum(axis=1))) < 1e-12 * p**2


def affine_metrics(xs, ys):
    Xxi, Yxi = (xs[1] - xs[0]) / 2, (ys[1] - ys[0]) / 2
    Xeta, Yeta = (xs[3] - xs[0]) / 2, (ys[3] - ys[0]) / 2
    J = Xxi * Yeta - Xeta * Yxi
    return J, (Yeta**2 + Xeta**2) / J, (Yeta * Yxi + Xeta * Xxi) / J, (Yxi**2 + Xxi**2) / J


def test_metrics_affine_closed_form(orc):
    xs = np.array([0.0, 2.0, 2.5, 0.5])
    ys = np.array([0.0, 0.3, 1.3, 1.0])  # parallelogram
    z, _ = orc.gll(5)
    J, A, B, C = orc.quad_metrics(xs, ys, z)
    Je, Ae, Be, Ce = affine_metrics(xs, ys)
    area = abs(0.5 * np.sum(xs * np.roll(ys, -1) - np.roll(xs, -1) * ys))
    assert np.allclose(J, Je, rtol=1e-14) and abs(Je - area / 4) < 1e-14
    assert np.allclose(A, Ae, rtol=1e-14) and np.allclose(B, Be, rtol=1e-14) and np.allclose(C, Ce, rtol=1e-14)


def test_metrics_c_index_quirk(orc):
    # Quad.hpp:53 uses Yxi[i] (not Yxi[j]); visible on a trapezoid with y1-y0 != y2-y3
    xs = np.array([0.0, 1.0, 1.0, 0.0])
    ys = np.array([0.0, 0.0, 1.4, 1.0])
    z, _ = orc.gll(4)
    J, A, B, C = orc.quad_metrics(xs, ys, z)
    Xxi = 0.25 * ((1 - z) * (xs[1] - xs[0]) + (1 + z) * (xs[2] - xs[3]))
    Yxi = 0.25 * ((1 - z) * (ys[1] - ys[0]) + (1 + z) * (ys[2] - ys[3]))
    Xeta = 0.25 * ((1 - z) * (xs[3] - xs[0]) + (1 + z) * (xs[2] - xs[1]))
    Yeta = 0.25 * ((1 - z) * (ys[3] - ys[0]) + (1 + z) * (ys[2] - ys[1]))
    i, j = np.meshgrid(np.arange(5), np.arange(5), indexing="ij")
    Jh = Xxi[j] * Yeta[i] - Xeta[i] * Yxi[j]
    C_quirk = (Yxi[i] ** 2 + Xxi[j] ** 2) / Jh
    C_kopriva = (Yxi[j] ** 2 + Xxi[j] ** 2) / Jh
    assert np.allclose(J, Jh, rtol=1e-15) and np.allclose(C, C_quirk, rtol=1e-14)
    assert np.max(np.abs(C - C_kopriva)) > 1e-3


def test_glaplace_kills_constants(orc):
    # tests/glaplace.cpp:4-24: K 1 = 0 on [-1,1]^2 at p = 5; also on a distorted quad
    for xs, ys in (([-1, 1, 1, -1], [-1, -1, 1, 1]), ([0, 1.2, 1.0, -0.1], [0, 0.1, 0.9, 1.1])):
        out = orc.glaplace_quad(np.array(xs, float), np.array(ys, float), 5, np.ones(36))
        assert np.max(np.abs(out)) < 1e-10


def stiffness_dense(orc, xs, ys, p):
    n = p + 1
    K = np.zeros((n * n, n * n))
    for k in range(n * n):
        e = np.zeros(n * n)
        e[k] = 1
        K[:, k] = orc.glaplace_quad(xs, ys, p, e)
    return K


def test_glaplace_symmetric(orc):
    xs = np.array([0.0, 1.1, 1.3, -0.2])
    ys = np.array([0.0, -0.1, 1.0, 0.8])
    K = stiffness_dense(orc, xs, ys, 4)
    assert np.max(np.abs(K - K.T)) < 1e-13 * np.max(np.abs(K))


def test_glaplace_matches_physical_stiffness_affine(orc):
    # independent assembly: K_ab = sum_q w_q |J| grad(phi_a) . grad(phi_b), grad = Jmat^{-T} grad_ref
    p = 5
    n = p + 1
    xs = np.array([0.0, 2.0, 2.5, 0.5])
    ys = np.array([0.0, 0.3, 1.3, 1.0])
    x, w = orc.gll(p)
    D = orc.dmat(x)
    Jm = np.array([[(xs[1] - xs[0]) / 2, (xs[3] - xs[0]) / 2], [(ys[1] - ys[0]) / 2, (ys[3] - ys[0]) / 2]])
    detJ = np.linalg.det(Jm)
    G = np.linalg.inv(Jm) @ np.linalg.inv(Jm).T
    K = np.zeros((n * n, n * n))
    for qa in range(n):
        for qb in range(n):
            g = np.zeros((n * n, 2))  # reference gradients of every basis function at node (qa, qb)
            for a in range(n):
                for b in range(n):
                    idx = a + b * n
                    g[idx, 0] = D[qa, a] * (b == qb)
                    g[idx, 1] = (a == qa) * D[qb, b]
            K += w[qa] * w[qb] * detJ * g @ G @ g.T
    Ko = stiffness_dense(orc, xs, ys, p)
    assert np.max(np.abs(Ko - K)) < 1e-12 * np.max(np.abs(K))


def test_mass_integrates_area(orc):
    xs = np.array([0.0, 1.1, 1.3, -0.2])
    ys = np.array([0.0, -0.1, 1.0, 0.8])
    x, w = orc.gll(6)
    J, A, B, C = orc.quad_metrics(xs, ys, x)
    area = 0.5 * np.sum(xs * np.roll(ys, -1) - np.roll(xs, -1) * ys)
    assert abs(np.sum(np.outer(w, w) * J) - area) < 1e-14
