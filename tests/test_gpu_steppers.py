"""GPU Schrodinger steppers against the oracle (1e-9 relative L2) and the
reference's own stepper tests (tests/schrodinger.cpp)."""
import numpy as np
import pytest

from conftest import SMALL_MESH, rel

pytestmark = pytest.mark.gpu

SCHEMES = [(0, "schrodinger_bdf2"), (1, "schrodinger_dirk4s3"), (2, "schrodinger_dirk4s5")]


def packet(m):
    re = m.field(lambda x, y: np.exp(-(x * x + y * y) / 0.5) * np.cos(2 * x + y))
    im = m.field(lambda x, y: np.exp(-(x * x + y * y) / 0.5) * np.sin(2 * x + y))
    return re, im


@pytest.mark.parametrize("scheme,name", SCHEMES)
def test_reference_stepper_test(sem, scheme, name):
    # tests/schrodinger.cpp: small_mesh, p = 3, V = 0, psi0 = sin(pi x) sin(pi y), T = 0.01, dt = T/10
    m = sem.load_mesh(3, SMALL_MESH)
    V = np.zeros(m.size)
    re = m.field(lambda x, y: np.sin(np.pi * x) * np.sin(np.pi * y))
    im = np.zeros(m.size)
    opts = sem.SolverOpts(max_iter=20 * m.dof(), tol=1e-8, dt=0.001, verbose=False)
    assert getattr(sem, name)(re, im, V, m, 0.01, opts)
    assert m.last_run_stats()["steps"] == 10


@pytest.mark.parametrize("scheme,name", SCHEMES)
def test_stepper_parity(sem, orc, scheme, name):
    m = sem.load_mesh(4, SMALL_MESH)
    o = orc.Mesh.load(4, SMALL_MESH)
    V = m.field(lambda x, y: 0.5 * (x * x + y * y) + np.exp(-((x - 0.3)**2 + y * y) / 0.1))
    re, im = packet(m)
    T, dt, tol = 0.02, 0.004, 1e-13
    ro, io, ok, stats, _ = o.schrodinger(scheme, re, im, V, T, 20000, tol, dt)
    assert ok
    opts = sem.SolverOpts(max_iter=20000, tol=tol, dt=dt)
    a, b = re.copy(), im.copy()
    assert getattr(sem, name)(a, b, V, m, T, opts)
    st = m.last_run_stats()
    assert st["steps"] == int(stats[0])
    err = rel(np.concatenate([a, b]), np.concatenate([ro, io]))
    assert err < 1e-9, err
    # mixed precision (FP32 inner MINRES, FP64 refinement)
    opts.precision = sem.PREC_MIXED
    a, b = re.copy(), im.copy()
    assert getattr(sem, name)(a, b, V, m, T, opts)
    err = rel(np.concatenate([a, b]), np.concatenate([ro, io]))
    assert err < 1e-9, err


def test_callback_stops(sem):
    m = sem.load_mesh(3, SMALL_MESH)
    V = np.zeros(m.size)
    re, im = packet(m)
    seen = []

    def cb(a, b):
        seen.append(float(np.abs(a).max()))
        return len(seen) >= 3

    opts = sem.SolverOpts(max_iter=5000, tol=1e-10, dt=0.001)
    assert sem.schrodinger_dirk4s3(re, im, V, m, 1.0, opts, cb)
    assert len(seen) == 3 and m.last_run_stats()["steps"] == 3


def test_nonconvergence_is_not_an_error(sem):
    m = sem.load_mesh(3, SMALL_MESH)
    V = np.zeros(m.size)
    re, im = packet(m)
    opts = sem.SolverOpts(max_iter=2, tol=1e-14, dt=0.01)
    assert sem.schrodinger_dirk4s5(re, im, V, m, 0.05, opts) is False
