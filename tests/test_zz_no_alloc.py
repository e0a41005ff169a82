"""sem_minres_noalloc (the MINRES entry point that allocates nothing: its
workspace is a static device array) against sem_minres on device vectors:
bit-identical results on two calls in a row, the second with a smaller n, and
a clear error for sizes beyond the workspace or host vectors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fd(n, k):
    x = np.linspace(-1, 1, n)
    h = x[1] - x[0]
    A = (np.diag(np.full(n, 2 / h**2 - k)) + np.diag(np.full(n - 1, -1 / h**2), -1)
         + np.diag(np.full(n - 1, -1 / h**2), 1))
    return A, x * x + 0.3 * np.sin(5 * x)


def device_op(sem, A, n):
    def op(xp, yp, stream):
        sem.stream_synchronize(stream)
        sem.copy_to_device(yp, A @ sem.copy_to_host(xp, n))
    return op


def solve(sem, solver, A, f, maxit, tol, precond=None):
    n = f.size
    b = sem.DeviceArray.from_host(f)
    x = sem.DeviceArray.from_host(np.zeros(n))
    # a preconditioner with the consistent normalisation (the reference's
    # beta = sqrt(dot(y, y)) of minres.hpp:63 does not converge with it)
    r = sem.minres_device(n, device_op(sem, A, n), b.ptr, x.ptr, maxit, tol, precond=precond,
                          consistent=precond is not None, solver=solver)
    out = x.to_host()
    b.close()
    x.close()
    return out, r


def test_noalloc_minres_equals_minres(sem):
    for n, k, tol in ((600, 3.0, 1e-12), (150, 1.5, 1e-10)):  # the second call is smaller
        A, f = fd(n, k)
        x0, r0 = solve(sem, "minres", A, f, 5000, tol)
        x1, r1 = solve(sem, "minres_noalloc", A, f, 5000, tol)
        assert r0.success and r1.success
        assert (r1.n_iter, r1.residual) == (r0.n_iter, r0.residual)
        assert np.array_equal(x1, x0)
        assert np.linalg.norm(A @ x1 - f) < 1e-6 * np.linalg.norm(f)
    # with a preconditioner callback too
    A, f = fd(200, 2.0)
    pd = 1.0 / np.abs(np.diag(A))
    n = f.size

    def pc(rp, zp, stream):
        sem.stream_synchronize(stream)
        sem.copy_to_device(zp, pd * sem.copy_to_host(rp, n))

    x0, r0 = solve(sem, "minres", A, f, 5000, 1e-12, precond=pc)
    x1, r1 = solve(sem, "minres_noalloc", A, f, 5000, 1e-12, precond=pc)
    assert r0.success and r1.success
    assert (r1.n_iter, r1.residual) == (r0.n_iter, r0.residual) and np.array_equal(x1, x0)


def test_noalloc_limits(sem):
    n = 262144 + 1
    b = sem.DeviceArray.from_host(np.ones(n))
    x = sem.DeviceArray.from_host(np.zeros(n))
    with pytest.raises(sem.SemError, match="exceeds the workspace"):
        sem.minres_device(n, lambda xp, yp, s: None, b.ptr, x.ptr, 10, 1e-8, solver="minres_noalloc")
    b.close()
    x.close()
    with pytest.raises(sem.SemError):
        k, keep = sem._krylov(4, lambda a, c: None, None, None, True, False)
        r = sem._Results()
        sem._check(sem.lib().sem_minres_noalloc(sem.C.byref(k), sem._ptr(np.ones(4)), sem._ptr(np.zeros(4)), 10,
                                                1e-8, sem.C.byref(r)))
