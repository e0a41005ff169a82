"""The fused MINRES path (one launch per Lanczos step, operator applied before
the 1/beta scaling): same recurrence as the two-launch path, so it meets the
same oracle tolerances, and its own work shapes give identical bits."""
import numpy as np
import pytest

from conftest import SMALL_MESH, rel, structured

pytestmark = pytest.mark.gpu

FP64_FUSED_OFF = ("the one-launch Lanczos step runs in FP32 only: its FP64 instantiation is switched off "
                  "after one of its work shapes faulted the GPU on a 256x256, degree-7 mesh (cause not found)")


def res_problem(sem, orc, p=4, seed=7):
    m = sem.load_mesh(p, SMALL_MESH)
    o = orc.Mesh.load(p, SMALL_MESH)
    rng = np.random.default_rng(seed)
    V = m.field(lambda x, y: 0.5 * (x * x + y * y))
    b = np.concatenate([o.galerkin(2, rng.standard_normal(m.size)), o.galerkin(2, rng.standard_normal(m.size))])
    return m, o, V, b


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_stage_solve_matches_oracle(sem, orc, precision):
    if precision == 0:
        pytest.skip(FP64_FUSED_OFF)
    m, o, V, b = res_problem(sem, orc)
    alpha = 0.25 * 1e-2
    zo, ok, it, res = o.stage_solve(V, alpha, b, None, 5000, 1e-12)
    zu = np.concatenate([o.unmask(zo[:m.size]), o.unmask(zo[m.size:])])
    sem.set_step_mode(m, 0)
    z0, r0 = sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, precision)
    sem.set_step_mode(m, 1)
    assert sem.get_step_mode(m) == 0 and sem.get_step_mode(m, sem.PREC_MIXED) == 1
    with pytest.raises(sem.SemError):
        sem.set_step_mode(m, 1, sem.PREC_FP64)
    z1, r1 = sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, precision)
    assert r1.success and r0.success and ok
    assert rel(z1, zu) < 1e-10
    if precision == 0:
        assert abs(r1.n_iter - it) <= 2 and abs(r1.n_iter - r0.n_iter) <= 2
        # phibar at the stop is a product of ~300 Givens sines; the two step
        # orders (and FMA contraction choices of the compiler) round
        # differently, which moves it by several percent on this problem. Both
        # stop below the threshold at the same step, so each lies within one
        # step's reduction of it.
        # stop test minres.hpp:92 with atol = tol: phibar < tol*||b|| + tol
        thr = 1e-12 * np.sqrt(o.dot(b[:m.size], b[:m.size]) + o.dot(b[m.size:], b[m.size:])) + 1e-12
        for r in (r0, r1):
            assert 0.5 * thr < r.residual < thr, (r.residual, thr)


@pytest.mark.skip(reason=FP64_FUSED_OFF)
def test_fused_helmholtz_and_laplace(sem, orc):
    m = sem.load_mesh(5, SMALL_MESH)
    o = orc.Mesh.load(5, SMALL_MESH)
    F = m.field(lambda x, y: np.cos(x) * y)
    c = np.full(m.size, 3.0)
    uo, ok, it, res = o.helmholtz(np.zeros(m.size), c, F, 5000, 1e-12)
    sem.set_step_mode(m, 1)
    u = np.zeros(m.size)
    r = sem.helmholtz(u, c, F, m, 5000, 1e-12)
    assert r.success and ok and abs(r.n_iter - it) <= 2
    assert rel(o.unmask(u), o.unmask(uo)) < 1e-10


@pytest.mark.skip(reason=FP64_FUSED_OFF)
def test_fused_nonconvergence_semantics(sem, orc):
    m, o, V, b = res_problem(sem, orc, seed=9)
    for mode in (0, 1):
        sem.set_step_mode(m, mode)
        z, r = sem.stage_solve(m, V, 0.5, b, None, 7, 1e-14)
        assert not r.success and r.n_iter == 8, (mode, r.n_iter)


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_shapes_bit_identical(sem, precision):
    if precision == 0:
        pytest.skip(FP64_FUSED_OFF)
    xy, q = structured(20, (-2, 2, -2, 2), 0.1, 3)
    m = sem.Mesh.from_arrays(6, xy, q)
    sem.set_step_mode(m, 1)
    rng = np.random.default_rng(11)
    V = m.field(lambda x, y: x * x + y * y)
    b = np.concatenate([sem.gproj(rng.standard_normal(m.size), m), sem.gproj(rng.standard_normal(m.size), m)])
    ref = None
    # step mode 2 (the recurrence's reduction in its own kernel after each
    # step) gives the bits of mode 1 in every shape
    for mode in (1, 2):
        sem.set_step_mode(m, mode, sem.PREC_MIXED)
        for shape in [(1, 0, 1, 0), (2, 0, 2, 0), (1, 3, 1, 0), (2, 64, 1, 0), (1, 1, 1, 0), (2, 512, 1, 0)]:
            sem.set_step_shape(m, shape, precision)
            z, r = sem.stage_solve(m, V, 2e-3, b, None, 3000, 1e-11, precision)
            assert r.success
            if ref is None:
                ref = (z, r.n_iter, r.residual)
            else:
                assert np.array_equal(z, ref[0]), (mode, shape)
                assert (r.n_iter, r.residual) == ref[1:], (mode, shape)


def test_fused_stepper_matches_two_launch(sem):
    xy, q = structured(12, (-3, 3, -3, 3), 0.0, 0)
    m = sem.Mesh.from_arrays(5, xy, q)
    X, Y = m.coords()
    V = X * X + Y * Y
    env = np.exp(-(X * X + Y * Y) / 2)
    re, im = env * np.cos(2 * X + Y), -env * np.sin(2 * X + Y)
    out = {}
    for mode in (0, 1):
        sem.set_step_mode(m, mode)
        for name, prec, guess in (("fp64", 0, 0), ("mixed_learned", 1, 2)):
            opts = sem.SolverOpts(max_iter=20000, tol=1e-12, dt=5e-3, precision=prec, guess=guess, learn_steps=4)
            st = sem.Stepper(m, 1, re, im, V, opts)
            assert st.advance(1e30, 8)
            a, b, _ = st.read()
            out[(mode, name)] = np.concatenate([a, b])
            st.close()
    for name in ("fp64", "mixed_learned"):
        assert rel(out[(1, name)], out[(0, name)]) < 1e-9, name


def test_inner_scaling_mixed_stage_solve(sem, orc):
    """Diagonal scaling of the FP32 inner solver: the same FP64 stop test and
    the same solution to the oracle's tolerance, fewer inner Lanczos steps on
    a patch with the mesh width (0.156) and degree of the 128 x 128 workload."""
    xy, q = structured(16, (-2.5, 2.5, -2.5, 2.5), 0.0, 0)
    p = 8
    m = sem.Mesh.from_arrays(p, xy, q)
    o = orc.Mesh.from_arrays(p, xy, q)
    X, Y = m.coords()
    V = X * X + Y * Y
    rng = np.random.default_rng(5)
    alpha = (np.cos(np.pi / 18.0) / np.sqrt(3.0) + 0.5) * 1e-3
    b = 1e-3 * o.galerkin(orc.SCHRO_F, rng.standard_normal(2 * m.size), V)
    zo, ok, it, _ = o.stage_solve(V, alpha, b, None, 5000, 1e-12)
    zu = np.concatenate([o.unmask(zo[:m.size]), o.unmask(zo[m.size:])])
    assert ok and sem.get_inner_scaling(m)
    z1, r1 = sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, sem.PREC_MIXED)
    sem.set_inner_scaling(m, False)
    assert not sem.get_inner_scaling(m)
    z0, r0 = sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, sem.PREC_MIXED)
    sem.set_inner_scaling(m, True)
    z2, r2 = sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, sem.PREC_MIXED)
    assert r0.success and r1.success and r2.success
    assert rel(z1, zu) < 1e-9 and rel(z0, zu) < 1e-9, (rel(z1, zu), rel(z0, zu))
    assert np.array_equal(z1, z2) and r1.n_iter == r2.n_iter  # the cached scale gives the same bits
    assert r1.n_iter < r0.n_iter, (r1.n_iter, r0.n_iter)
    # FP64 solves do not use it
    zf, rf = sem.stage_solve(m, V, alpha, b, None, 5000, 1e-12, sem.PREC_FP64)
    assert rf.success and rel(zf, zu) < 1e-9
