"""The stacked-element MFMA form of the FP32 inner solver's stage operator:
the stiffness contractions of Res (glaplace.hpp:19-25 inside
schrodinger.hpp:185-207) as v_mfma_f32_16x16x16bf16 tiles with FP32
accumulation, bf16 operands (contraction 1) or split hi + lo bf16 operands
(contraction 3). The element apply matches the FP64 oracle to the tolerance of
its operand precision; mixed-precision stage solves refined in FP64 match the
oracle to 1e-10 whatever the contraction."""
import numpy as np
import pytest

from conftest import SMALL_MESH, rel, structured

pytestmark = pytest.mark.gpu

# stated tolerances of the FP32 element apply (relative L2 against FP64)
TOL = {0: 1e-6, 1: 1e-2, 3: 2e-5}


@pytest.mark.parametrize("p", [3, 5, 7, 8, 9, 12, 15])
def test_res_apply_tolerance(sem, orc, p):
    xy, q = structured(5, (-2, 2, -2, 2), 0.1, p)
    m = sem.Mesh.from_arrays(p, xy, q)
    o = orc.Mesh.from_arrays(p, xy, q)
    X, Y = m.coords()
    V = 1.0 + X * X + 0.5 * Y
    rng = np.random.default_rng(p)
    u = rng.standard_normal(2 * m.size)
    ref = o.element_op(orc.SCHRO_RES, u, V, 0.05)
    errs = {}
    for mode in (0, 1, 3):
        sem.set_contraction(m, mode)
        assert sem.get_contraction(m) == mode
        errs[mode] = rel(sem.element_op_fp32(sem.OP_SCHRO_RES, u, m, V, 0.05), ref)
    print("p=%d: FP32 Res apply rel-L2 VALU %.2e, bf16 MFMA %.2e, split-bf16 MFMA %.2e" % (p, errs[0], errs[1], errs[3]))
    for mode, e in errs.items():
        assert e < TOL[mode], (mode, e)
    # the split form is much closer than plain bf16 (the hi*lo terms are there)
    assert errs[3] < 0.05 * errs[1]
    with pytest.raises(sem.SemError):
        sem.set_contraction(m, 2)


@pytest.mark.parametrize("mode", [1, 3])
def test_mixed_stage_solve_with_mfma(sem, orc, mode):
    m = sem.load_mesh(6, SMALL_MESH)
    o = orc.Mesh.load(6, SMALL_MESH)
    V = m.field(lambda x, y: 0.5 * (x * x + y * y) + np.sin(3 * x))
    rng = np.random.default_rng(17)
    b = np.concatenate([o.galerkin(2, rng.standard_normal(m.size)), o.galerkin(2, rng.standard_normal(m.size))])
    alpha = 0.25 * 2e-3
    zo, ok, it, _ = o.stage_solve(V, alpha, b, None, 5000, 1e-12)
    zu = np.concatenate([o.unmask(zo[:m.size]), o.unmask(zo[m.size:])])
    sem.set_contraction(m, 0)
    z0, r0 = sem.stage_solve(m, V, alpha, b, None, 20000, 1e-12, sem.PREC_MIXED)
    sem.set_contraction(m, mode)
    z1, r1 = sem.stage_solve(m, V, alpha, b, None, 20000, 1e-12, sem.PREC_MIXED)
    assert ok and r0.success and r1.success
    print("contraction %d: inner iterations %d (VALU %d), rel-L2 vs oracle %.2e" % (mode, r1.n_iter, r0.n_iter,
                                                                                   rel(z1, zu)))
    assert rel(z1, zu) < 1e-10 and rel(z0, zu) < 1e-10


def test_mfma_step_shapes_bit_identical(sem):
    # the MFMA step kernel serves every work shape (persistent requests run one
    # item per workgroup) with the same bits
    xy, q = structured(20, (-2, 2, -2, 2), 0.1, 3)
    m = sem.Mesh.from_arrays(7, xy, q)
    sem.set_contraction(m, 3)
    rng = np.random.default_rng(11)
    V = m.field(lambda x, y: x * x + y * y)
    b = np.concatenate([sem.gproj(rng.standard_normal(m.size), m), sem.gproj(rng.standard_normal(m.size), m)])
    ref = None
    for mode in (1, 2):
        sem.set_step_mode(m, mode, sem.PREC_MIXED)
        for shape in [(1, 0, 1, 0), (2, 0, 1, 0), (2, 64, 1, 0)]:
            sem.set_step_shape(m, shape, sem.PREC_MIXED)
            z, r = sem.stage_solve(m, V, 2e-3, b, None, 5000, 1e-11, sem.PREC_MIXED)
            assert r.success
            if ref is None:
                ref = (z, r.n_iter)
            else:
                assert np.array_equal(z, ref[0]) and r.n_iter == ref[1], (mode, shape)


def test_stepper_with_mfma_contractions(sem):
    xy, q = structured(12, (-3, 3, -3, 3), 0.0, 0)
    m = sem.Mesh.from_arrays(7, xy, q)
    X, Y = m.coords()
    V = X * X + Y * Y
    env = np.exp(-(X * X + Y * Y) / 2)
    re, im = env * np.cos(2 * X + Y), -env * np.sin(2 * X + Y)
    out = {}
    for mode, prec in ((None, 0), (0, 1), (1, 1), (3, 1)):
        if mode is not None:
            sem.set_contraction(m, mode)
        opts = sem.SolverOpts(max_iter=50000, tol=1e-13, dt=5e-3, precision=prec, guess=1)
        st = sem.Stepper(m, 2, re, im, V, opts)
        assert st.advance(1e30, 6)
        a, b, _ = st.read()
        out[mode] = (np.concatenate([a, b]), st.stats())
        st.close()
    for mode in (0, 1, 3):
        e = rel(out[mode][0], out[None][0])
        print("contraction %d: inner iterations %d, refinements %d, rel-L2 vs FP64 %.2e" %
              (mode, out[mode][1]["total_iters"], out[mode][1]["refinements"], e))
        assert e < 1e-9
