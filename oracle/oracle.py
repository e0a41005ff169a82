"""ctypes wrapper of liboracle_sem.so, the FP64 CPU restatement of the
reference's OpenMP hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker. The product
(schrodingersem_amd) never imports it.

Arrays use the reference's element-local layout (e*N*N + i + j*N); two-
component vectors are [re | im].
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_sem.so")

LAPLACE, HELMHOLTZ, MASS, SCHRO_RES, SCHRO_F, SCHRO_L = range(6)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
STEP_CB = C.CFUNCTYPE(C.c_int, _dp, _dp, C.c_void_p)
_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_set_threads": (None, [C.c_int]),
            "orc_get_threads": (C.c_int, []),
            "orc_gll": (C.c_int, [C.c_int, _dp, _dp]),
            "orc_dmat": (None, [C.c_int, _dp, _dp]),
            "orc_quad_metrics": (None, [_dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp]),
            "orc_glaplace_quad": (None, [_dp, _dp, C.c_int, _dp, _dp]),
            "orc_mesh_create": (C.c_void_p, [C.c_int, _dp, C.c_int, _ip, C.c_int]),
            "orc_mesh_load": (C.c_void_p, [C.c_int, C.c_char_p]),
            "orc_mesh_free": (None, [C.c_void_p]),
            "orc_mesh_info": (None, [C.c_void_p, C.POINTER(C.c_long)]),
            "orc_mesh_gll": (None, [C.c_void_p, _dp, _dp]),
            "orc_mesh_D": (None, [C.c_void_p, _dp]),
            "orc_mesh_metrics": (None, [C.c_void_p, _dp, _dp, _dp, _dp]),
            "orc_mesh_coords": (None, [C.c_void_p, _dp, _dp]),
            "orc_mesh_edges": (None, [C.c_void_p, _ip]),
            "orc_mesh_corner_types": (None, [C.c_void_p, _ip]),
            "orc_unmask": (None, [C.c_void_p, _dp]),
            "orc_global_sum": (None, [C.c_void_p, _dp]),
            "orc_mask": (None, [C.c_void_p, _dp]),
            "orc_dot": (C.c_double, [C.c_void_p, _dp, _dp]),
            "orc_element_op": (None, [C.c_void_p, C.c_int, _dp, C.c_double, _dp, _dp]),
            "orc_galerkin": (None, [C.c_void_p, C.c_int, _dp, C.c_double, _dp, _dp]),
            "orc_stage_solve": (C.c_int, [C.c_void_p, _dp, C.c_double, _dp, _dp, C.c_int, C.c_double, _ip, _dp]),
            "orc_minres_tridiag": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                             C.c_double, _ip, _dp]),
            "orc_pcg_tridiag": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                          C.c_double, _ip, _dp]),
            "orc_minres_tridiag_pc": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                                C.c_double, _dp, C.c_int, _ip, _dp]),
            "orc_pcg_tridiag_pc": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                             C.c_double, _dp, _ip, _dp]),
            "orc_poisson": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int, C.c_double, _ip, _dp]),
            "orc_helmholtz": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_double, _ip, _dp]),
            "orc_schrodinger": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, C.c_double,
                                          C.c_double, STEP_CB, C.c_void_p, _dp, _ip, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def _f(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


def _p(a):
    return a.ctypes.data_as(_dp)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def gll(p):
    x = np.zeros(p + 1)
    w = np.zeros(p + 1)
    if lib().orc_gll(int(p), _p(x), _p(w)) != 0:
        raise ValueError(lib().orc_last_error().decode())
    return x, w


def dmat(x):
    x = _f(x)
    D = np.zeros(x.size * x.size)
    lib().orc_dmat(x.size, _p(x), _p(D))
    return D.reshape(x.size, x.size, order="F")


def quad_metrics(xs, ys, z):
    z = _f(z)
    n = z.size
    out = [np.zeros(n * n) for _ in range(4)]
    lib().orc_quad_metrics(_p(_f(xs, 4)), _p(_f(ys, 4)), n, _p(z), *[_p(a) for a in out])
    return [a.reshape(n, n, order="F") for a in out]


def glaplace_quad(xs, ys, p, u):
    n = p + 1
    u = _f(u, n * n)
    out = np.zeros(n * n)
    lib().orc_glaplace_quad(_p(_f(xs, 4)), _p(_f(ys, 4)), int(p), _p(u), _p(out))
    return out


class Mesh:
    def __init__(self, h):
        if not h:
            raise ValueError(lib().orc_last_error().decode())
        self._h = C.c_void_p(h)
        info = (C.c_long * 7)()
        lib().orc_mesh_info(self._h, info)
        (self.N, self.n_elements, self.n_nodes, self.n_edges, self._dof, self.n_interior_edges,
         self.n_boundary_nodes) = (int(v) for v in info)
        self.size = self.n_elements * self.N * self.N

    @classmethod
    def from_arrays(cls, p, xy, quads):
        xy = _f(xy)
        q = np.ascontiguousarray(quads, dtype=np.int32).ravel()
        return cls(lib().orc_mesh_create(int(p), _p(xy), xy.size // 2, q.ctypes.data_as(_ip), q.size // 4))

    @classmethod
    def load(cls, p, folder):
        return cls(lib().orc_mesh_load(int(p), folder.encode()))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _L is not None:
            _L.orc_mesh_free(h)
            self._h = None

    def dof(self):
        return self._dof

    def quadrature(self):
        x = np.zeros(self.N)
        w = np.zeros(self.N)
        lib().orc_mesh_gll(self._h, _p(x), _p(w))
        return x, w

    def D(self):
        D = np.zeros(self.N * self.N)
        lib().orc_mesh_D(self._h, _p(D))
        return D.reshape(self.N, self.N, order="F")

    def metrics(self):
        arrs = [np.zeros(self.size) for _ in range(4)]
        lib().orc_mesh_metrics(self._h, *[_p(a) for a in arrs])
        return arrs

    def coords(self):
        X = np.zeros(self.size)
        Y = np.zeros(self.size)
        lib().orc_mesh_coords(self._h, _p(X), _p(Y))
        return X, Y

    def edges(self):
        e = np.zeros(7 * self.n_edges, dtype=np.int32)
        lib().orc_mesh_edges(self._h, e.ctypes.data_as(_ip))
        return e.reshape(-1, 7)

    def corner_types(self):
        t = np.zeros(self.n_nodes, dtype=np.int32)
        lib().orc_mesh_corner_types(self._h, t.ctypes.data_as(_ip))
        return t

    def _inplace(self, fn, a):
        b = _f(a, self.size).copy()
        fn(self._h, _p(b))
        return b

    def unmask(self, a):
        return self._inplace(lib().orc_unmask, a)

    def global_sum(self, a):
        return self._inplace(lib().orc_global_sum, a)

    def mask(self, a):
        return self._inplace(lib().orc_mask, a)

    def dot(self, a, b):
        return lib().orc_dot(self._h, _p(_f(a, self.size)), _p(_f(b, self.size)))

    def element_op(self, kind, u, coef=None, alpha=0.0):
        nc = 2 if kind in (SCHRO_RES, SCHRO_F) else 1
        u = _f(u, nc * self.size)
        out = np.zeros_like(u)
        c = None if coef is None else _f(coef, self.size)
        lib().orc_element_op(self._h, kind, None if c is None else _p(c), float(alpha), _p(u), _p(out))
        return out

    def galerkin(self, kind, u, coef=None, alpha=0.0):
        nc = 2 if kind in (SCHRO_RES, SCHRO_F) else 1
        u = _f(u, nc * self.size)
        out = np.zeros_like(u)
        c = None if coef is None else _f(coef, self.size)
        lib().orc_galerkin(self._h, kind, None if c is None else _p(c), float(alpha), _p(u), _p(out))
        return out

    def stage_solve(self, V, alpha, b, z=None, max_iter=10000, tol=1e-10):
        n = 2 * self.size
        z = np.zeros(n) if z is None else _f(z, n).copy()
        it = C.c_int()
        res = C.c_double()
        ok = lib().orc_stage_solve(self._h, _p(_f(V, self.size)), float(alpha), _p(_f(b, n)), _p(z), int(max_iter),
                                   float(tol), C.byref(it), C.byref(res))
        return z, bool(ok), it.value, res.value

    def poisson(self, u, f, max_iter, tol):
        u = _f(u, self.size).copy()
        it = C.c_int()
        res = C.c_double()
        ok = lib().orc_poisson(self._h, _p(u), _p(_f(f, self.size)), int(max_iter), float(tol), C.byref(it),
                               C.byref(res))
        return u, bool(ok), it.value, res.value

    def helmholtz(self, u, c, f, max_iter, tol):
        u = _f(u, self.size).copy()
        it = C.c_int()
        res = C.c_double()
        ok = lib().orc_helmholtz(self._h, _p(u), _p(_f(c, self.size)), _p(_f(f, self.size)), int(max_iter),
                                 float(tol), C.byref(it), C.byref(res))
        return u, bool(ok), it.value, res.value

    def schrodinger(self, scheme, re, im, V, T, max_iter, tol, dt, callback=None, max_records=0):
        """scheme 0 = bdf2, 1 = dirk4s3, 2 = dirk4s5; returns (re, im, ok, stats, stage_iters)."""
        re = _f(re, self.size).copy()
        im = _f(im, self.size).copy()
        stats = np.zeros(4)
        rec = np.zeros(max(max_records, 1), dtype=np.int32)
        n = self.size
        if callback is None:
            cb = STEP_CB()
        else:
            def _cb(a, b, _u):
                return 1 if callback(np.ctypeslib.as_array(a, (n,)).copy(), np.ctypeslib.as_array(b, (n,)).copy()) else 0
            cb = STEP_CB(_cb)
        ok = lib().orc_schrodinger(self._h, int(scheme), _p(re), _p(im), _p(_f(V, n)), float(T), int(max_iter),
                                   float(tol), float(dt), cb, None, _p(stats), rec.ctypes.data_as(_ip),
                                   int(max_records))
        return re, im, bool(ok), stats, rec[:max_records]


def minres_tridiag(lo, di, up, b, x, maxit, tol):
    b = _f(b)
    x = _f(x, b.size).copy()
    it = C.c_int()
    res = C.c_double()
    ok = lib().orc_minres_tridiag(b.size, lo, di, up, _p(b), _p(x), int(maxit), float(tol), C.byref(it), C.byref(res))
    return x, bool(ok), it.value, res.value


def pcg_tridiag(lo, di, up, b, x, maxit, tol):
    b = _f(b)
    x = _f(x, b.size).copy()
    it = C.c_int()
    res = C.c_double()
    ok = lib().orc_pcg_tridiag(b.size, lo, di, up, _p(b), _p(x), int(maxit), float(tol), C.byref(it), C.byref(res))
    return x, bool(ok), it.value, res.value


def minres_tridiag_pc(lo, di, up, b, x, maxit, tol, pdiag=None, consistent=False):
    """minres.hpp with a diagonal preconditioner z = pdiag * r (None: identity)."""
    b = _f(b)
    x = _f(x, b.size).copy()
    pd = None if pdiag is None else _f(pdiag, b.size)
    it = C.c_int()
    res = C.c_double()
    ok = lib().orc_minres_tridiag_pc(b.size, lo, di, up, _p(b), _p(x), int(maxit), float(tol),
                                     None if pd is None else _p(pd), int(bool(consistent)), C.byref(it), C.byref(res))
    return x, bool(ok), it.value, res.value


def pcg_tridiag_pc(lo, di, up, b, x, maxit, tol, pdiag=None):
    """pcg.hpp with a diagonal preconditioner z = pdiag * r (None: identity)."""
    b = _f(b)
    x = _f(x, b.size).copy()
    pd = None if pdiag is None else _f(pdiag, b.size)
    it = C.c_int()
    res = C.c_double()
    ok = lib().orc_pcg_tridiag_pc(b.size, lo, di, up, _p(b), _p(x), int(maxit), float(tol),
                                  None if pd is None else _p(pd), C.byref(it), C.byref(res))
    return x, bool(ok), it.value, res.value
