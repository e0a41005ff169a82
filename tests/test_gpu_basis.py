"""Stored-basis FP32 MINRES (the inner solves of the mixed precision): every
Lanczos vector is kept and the solution is formed once at the end from the
recorded update coefficients, instead of the w/x recurrence of
minres.hpp:82-86 every step. The Lanczos vectors and the recurrence scalars
are the same in both forms, so a solve takes the same steps and lands on the
same iterate up to FP32 rounding; refined stage solves meet the oracle's
tolerance either way, also when the basis budget cuts every inner solve short."""
import numpy as np
import pytest

from conftest import SMALL_MESH, rel, structured

pytestmark = pytest.mark.gpu


def res_problem(sem, orc, p=4, seed=7):
    m = sem.load_mesh(p, SMALL_MESH)
    o = orc.Mesh.load(p, SMALL_MESH)
    rng = np.random.default_rng(seed)
    V = m.field(lambda x, y: 0.5 * (x * x + y * y))
    b = np.concatenate([o.galerkin(2, rng.standard_normal(m.size)), o.galerkin(2, rng.standard_normal(m.size))])
    return m, o, V, b


def both_forms(sem, m, V, alpha, b, maxit, tol):
    out = {}
    for on in (0, 1):
        sem.set_stored_basis(m, on)
        out[on] = sem.stage_solve(m, V, alpha, b, None, maxit, tol, sem.PREC_MIXED)
    sem.set_stored_basis(m, 1)
    return out


@pytest.mark.parametrize("tol", [1e-4, 1e-12])
def test_basis_equals_recurrence_small_mesh(sem, orc, tol):
    m, o, V, b = res_problem(sem, orc)
    alpha = 0.25 * 1e-2
    out = both_forms(sem, m, V, alpha, b, 5000, tol)
    (z0, r0), (z1, r1) = out[0], out[1]
    assert r0.success and r1.success
    # the same Lanczos steps (a later refinement sweep may start from a
    # residual that differs in its last FP32 bits)
    assert abs(r0.n_iter - r1.n_iter) <= 2, (r0.n_iter, r1.n_iter)
    if tol == 1e-4:
        # one inner solve: the two iterates differ by FP32 rounding only
        assert r0.n_iter == r1.n_iter
        assert rel(z1, z0) < 1e-5, rel(z1, z0)
    else:
        zo, ok, it, res = o.stage_solve(V, alpha, b, None, 5000, 1e-12)
        zu = np.concatenate([o.unmask(zo[:m.size]), o.unmask(zo[m.size:])])
        assert ok and rel(z1, zu) < 1e-10 and rel(z0, zu) < 1e-10


def test_basis_equals_recurrence_structured(sem):
    # 48 x 48 quads at degree 7 (the workload's degree), harmonic potential
    xy, q = structured(48, (-6, 6, -6, 6), 0.0, 0)
    m = sem.Mesh.from_arrays(7, xy, q)
    X, Y = m.coords()
    V = X * X + Y * Y
    rng = np.random.default_rng(3)
    b = np.concatenate([sem.galerkin_op(sem.OP_MASS, rng.standard_normal(m.size), m),
                        sem.galerkin_op(sem.OP_MASS, rng.standard_normal(m.size), m)])
    out = both_forms(sem, m, V, 0.25 * 5e-4, b, 20000, 1e-4)
    (z0, r0), (z1, r1) = out[0], out[1]
    assert r0.success and r1.success and r0.n_iter == r1.n_iter
    assert rel(z1, z0) < 1e-5, rel(z1, z0)


def test_basis_budget_cuts_inner_solves(sem, orc):
    # a budget of a few dozen vectors: every inner solve stops at the budget
    # unconverged, the FP64 refinement runs more sweeps and still meets the
    # oracle's tolerance
    m, o, V, b = res_problem(sem, orc, seed=11)
    alpha = 0.25 * 1e-2
    def sweeps(f):
        s0 = m.last_run_stats()["refinements"]  # cumulative over the mesh's solves
        out = f()
        return out, m.last_run_stats()["refinements"] - s0

    full, s_full = sweeps(lambda: sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, sem.PREC_MIXED))
    sem.set_stored_basis(m, 1, 40 * 2 * m.size * 4)
    try:
        (z, r), s_cut = sweeps(lambda: sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, sem.PREC_MIXED))
    finally:
        sem.set_stored_basis(m, 1, 8 << 30)
    zo, ok, it, res = o.stage_solve(V, alpha, b, None, 5000, 1e-12)
    zu = np.concatenate([o.unmask(zo[:m.size]), o.unmask(zo[m.size:])])
    assert r.success and ok
    assert s_cut > s_full, (s_cut, s_full)
    assert rel(z, zu) < 1e-10 and rel(full[0], zu) < 1e-10
