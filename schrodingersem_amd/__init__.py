"""schrodingersem_amd -- Python mirror of SchrodingerSEM's hot-path API over
libsem_hip.so (HIP, gfx950).

Names and argument meaning follow the reference's namespace ``schro_omp``:
``load_mesh``, ``Mesh.dof``, ``glaplace``, ``galerkin_op``, ``gproj``,
``unmask``/``global_sum``/``mask``/``dot``, ``poisson``, ``helmholtz``,
``schrodinger_bdf2``/``schrodinger_dirk4s3``/``schrodinger_dirk4s5``,
``minres``/``pcg`` (on tridiagonal operators) and the ``SolverResults`` /
``SolverOpts`` records (solver_base.hpp:6-12, ode_base.hpp:6-13).

Element-local arrays are flat float64 arrays of length E*N*N in the
reference layout: element e's block at e*N*N, entry (i, j) at i + j*N.
Two-component operators take [re | im] (length 2*E*N*N).

There is no CPU fallback: every call runs on the GPU through the shared
library and raises if the library or the GPU is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsem_hip.so")

OP_LAPLACE, OP_HELMHOLTZ, OP_MASS, OP_SCHRO_RES, OP_SCHRO_F, OP_SCHRO_L = range(6)
PREC_FP64, PREC_MIXED = 0, 1
GUESS_REFERENCE, GUESS_EXTRAPOLATED, GUESS_LEARNED = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_fp = C.POINTER(C.c_float)


class SemError(RuntimeError):
    """A libsem_hip call returned a negative status."""


class _Results(C.Structure):
    _fields_ = [("success", C.c_int), ("n_iter", C.c_int), ("residual", C.c_double)]


class _Opts(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol", C.c_double), ("dt", C.c_double), ("verbose", C.c_int),
                ("precision", C.c_int), ("guess", C.c_int), ("learn_steps", C.c_int)]


class _Stats(C.Structure):
    _fields_ = [("steps", C.c_int), ("stage_solves", C.c_int), ("total_iters", C.c_long),
                ("refinements", C.c_long), ("last_error_estimate", C.c_double),
                ("max_true_rel_residual", C.c_double), ("mlp_train_steps", C.c_long),
                ("mlp_last_loss", C.c_double), ("learned_guess_chosen", C.c_long),
                ("learned_residual_digits", C.c_double)]


STEP_CB = C.CFUNCTYPE(C.c_int, _dp, _dp, C.c_void_p)

_lib = None


def lib():
    """Load libsem_hip.so (built in-tree by schrodingersem_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SemError("libsem_hip.so is not built; run python -m schrodingersem_amd.build")
        L = C.CDLL(LIB_PATH)
        sig = {
            "sem_last_error": (C.c_char_p, []),
            "sem_version": (C.c_int, []),
            "sem_progress_line": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_char_p,
                                            C.c_int]),
            "sem_batch_times": (C.c_int, [C.c_void_p, _dp, _ip]),
            "sem_device_info": (C.c_int, [C.c_char_p, C.c_int, _ip, _lp]),
            "sem_mesh_create": (C.c_int, [C.c_int, _dp, C.c_int, _ip, C.c_int, C.POINTER(C.c_void_p)]),
            "sem_mesh_load": (C.c_int, [C.c_int, C.c_char_p, C.POINTER(C.c_void_p)]),
            "sem_mesh_destroy": (None, [C.c_void_p]),
            "sem_mesh_info": (C.c_int, [C.c_void_p, _lp]),
            "sem_mesh_quadrature": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
            "sem_mesh_coords": (C.c_int, [C.c_void_p, _dp, _dp]),
            "sem_mesh_metrics": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
            "sem_mesh_structured": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                              C.c_ulonglong, _dp, _ip]),
            "sem_mesh_write": (C.c_int, [C.c_char_p, _dp, C.c_int, _ip, C.c_int]),
            "sem_glaplace_apply": (C.c_int, [C.c_void_p, _dp, _dp]),
            "sem_element_apply": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_double, _dp, _dp]),
            "sem_galerkin_apply": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_double, _dp, _dp]),
            "sem_gproj": (C.c_int, [C.c_void_p, _dp, _dp]),
            "sem_unmask": (C.c_int, [C.c_void_p, _dp]),
            "sem_global_sum": (C.c_int, [C.c_void_p, _dp]),
            "sem_mask": (C.c_int, [C.c_void_p, _dp]),
            "sem_dot": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
            "sem_poisson": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.POINTER(_Results)]),
            "sem_helmholtz": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_double, C.POINTER(_Results)]),
            "sem_schrodinger_bdf2": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_double, C.POINTER(_Opts), STEP_CB,
                                               C.c_void_p, _ip]),
            "sem_schrodinger_dirk4s3": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_double, C.POINTER(_Opts),
                                                  STEP_CB, C.c_void_p, _ip]),
            "sem_schrodinger_dirk4s5": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_double, C.POINTER(_Opts),
                                                  STEP_CB, C.c_void_p, _ip]),
            "sem_last_run_stats": (C.c_int, [C.c_void_p, C.POINTER(_Stats)]),
            "sem_stage_solve": (C.c_int, [C.c_void_p, _dp, C.c_double, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                          C.POINTER(_Results)]),
            "sem_minres_tridiag": (C.c_int, [C.c_long, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                             C.c_double, C.POINTER(_Results)]),
            "sem_pcg_tridiag": (C.c_int, [C.c_long, C.c_double, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                          C.c_double, C.POINTER(_Results)]),
            "sem_time_element_op": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp]),
            "sem_set_kernel_timing": (C.c_int, [C.c_void_p, C.c_int]),
            "sem_time_copy": (C.c_int, [C.c_long, C.c_int, C.c_int, _dp]),
            "sem_kernel_timing": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _dp]),
            "sem_set_step_shape": (C.c_int, [C.c_void_p, C.c_int, _ip]),
            "sem_get_step_shape": (C.c_int, [C.c_void_p, C.c_int, _ip]),
            "sem_set_graph_mode": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "sem_set_step_mode": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "sem_get_step_mode": (C.c_int, [C.c_void_p, C.c_int, _ip]),
            "sem_set_inner_scaling": (C.c_int, [C.c_void_p, C.c_int]),
            "sem_get_inner_scaling": (C.c_int, [C.c_void_p, _ip]),
            "sem_batch_set_graph_mode": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "sem_batch_set_step_mode": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "sem_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, _ip, _dp]),
            "sem_stepper_create": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, C.POINTER(_Opts),
                                             C.POINTER(C.c_void_p)]),
            "sem_stepper_advance": (C.c_int, [C.c_void_p, C.c_double, C.c_long, STEP_CB, C.c_void_p, _ip]),
            "sem_stepper_read": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
            "sem_stepper_stats": (C.c_int, [C.c_void_p, C.POINTER(_Stats)]),
            "sem_stepper_destroy": (None, [C.c_void_p]),
            "sem_stepper_set_learning": (C.c_int, [C.c_void_p, C.c_int, C.c_float]),
            "sem_stepper_set_guard": (C.c_int, [C.c_void_p, C.c_int]),
            "sem_stepper_set_training": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "sem_element_apply_fp32": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_double, _dp, _dp]),
            "sem_set_contraction": (C.c_int, [C.c_void_p, C.c_int]),
            "sem_get_contraction": (C.c_int, [C.c_void_p, _ip]),
            "sem_set_stored_basis": (C.c_int, [C.c_void_p, C.c_int, C.c_long]),
            "sem_minres": (C.c_int, [C.POINTER(_Krylov), _dp, _dp, C.c_int, C.c_double, C.POINTER(_Results)]),
            "sem_pcg": (C.c_int, [C.POINTER(_Krylov), _dp, _dp, C.c_int, C.c_double, C.POINTER(_Results)]),
            "sem_minres_noalloc": (C.c_int, [C.POINTER(_Krylov), _dp, _dp, C.c_int, C.c_double, C.POINTER(_Results)]),
            "sem_poisson_precond": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                              C.POINTER(_Results)]),
            "sem_helmholtz_precond": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                                C.POINTER(_Results)]),
            "sem_galerkin_apply_op": (C.c_int, [C.c_void_p, ELEMENT_OP_CB, C.c_void_p, _dp, _dp]),
            "sem_device_alloc": (C.c_int, [C.c_long, C.POINTER(_dp)]),
            "sem_device_free": (C.c_int, [_dp]),
            "sem_copy": (C.c_int, [_dp, _dp, C.c_long, C.c_int]),
            "sem_stream_synchronize": (C.c_int, [C.c_void_p]),
            "sem_gauss_lobatto": (C.c_int, [C.c_int, _dp, _dp]),
            "sem_derivative_matrix": (C.c_int, [C.c_int, _dp, _dp]),
            "sem_quad_metrics": (C.c_int, [_dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp]),
            "sem_stepper_checkpoint_size": (C.c_int, [C.c_void_p, _lp]),
            "sem_stepper_save": (C.c_int, [C.c_void_p, _dp, C.c_long]),
            "sem_stepper_restore": (C.c_int, [C.c_void_p, _dp, C.c_long]),
            "sem_batch_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
            "sem_batch_destroy": (None, [C.c_void_p]),
            "sem_batch_info": (C.c_int, [C.c_void_p, _ip, _ip, _ip, _lp]),
            "sem_batch_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip]),
            "sem_batch_set_step_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip]),
            "sem_batch_schrodinger_bdf2": (C.c_int, [C.c_void_p, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(_dp),
                                                     C.c_double, C.POINTER(_Opts), _ip]),
            "sem_batch_schrodinger_dirk4s3": (C.c_int, [C.c_void_p, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(_dp),
                                                     C.c_double, C.POINTER(_Opts), _ip]),
            "sem_batch_schrodinger_dirk4s5": (C.c_int, [C.c_void_p, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(_dp),
                                                     C.c_double, C.POINTER(_Opts), _ip]),
            "sem_batch_last_run_stats": (C.c_int, [C.c_void_p, C.POINTER(_Stats), _dp]),
            "sem_batch_start_bdf2": (C.c_int, [C.c_void_p, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(_dp),
                                               C.POINTER(_Opts)]),
            "sem_batch_start": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(_dp),
                                          C.POINTER(_Opts)]),
            "sem_batch_advance": (C.c_int, [C.c_void_p, C.c_double, C.c_long, _ip]),
            "sem_batch_read": (C.c_int, [C.c_void_p, C.POINTER(_dp), C.POINTER(_dp), _ip, _dp]),
            "sem_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)]),
            "sem_mlp_destroy": (None, [C.c_void_p]),
            "sem_mlp_info": (C.c_int, [C.c_void_p, _lp]),
            "sem_mlp_set_params": (C.c_int, [C.c_void_p, _fp]),
            "sem_mlp_get_params": (C.c_int, [C.c_void_p, _fp, C.c_int]),
            "sem_mlp_forward": (C.c_int, [C.c_void_p, C.c_int, _fp, _fp]),
            "sem_mlp_gradients": (C.c_int, [C.c_void_p, C.c_int, _fp, _fp, _dp]),
            "sem_mlp_adam": (C.c_int, [C.c_void_p, C.c_float]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SemError("libsem_hip error %d: %s" % (rc, lib().sem_last_error().decode(errors="replace")))


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


def _ptr(a):
    return a.ctypes.data_as(_dp)


LINOP_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, _dp, C.c_void_p)
DOT_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, _dp, _dp, C.c_void_p)
PRECOND_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, _dp, C.c_void_p)
ELEMENT_OP_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_void_p)
PRECOND_IDENTITY, PRECOND_JACOBI = 0, 1


class _Krylov(C.Structure):
    _fields_ = [("n", C.c_long), ("host", C.c_int), ("consistent", C.c_int), ("A", LINOP_CB), ("A_user", C.c_void_p),
                ("dot", DOT_CB), ("dot_user", C.c_void_p), ("precond", PRECOND_CB), ("precond_user", C.c_void_p)]


class SolverResults:
    """solver_results (solver_base.hpp:6-12)."""

    def __init__(self, r):
        self.success = bool(r.success)
        self.n_iter = int(r.n_iter)
        self.residual = float(r.residual)

    def __repr__(self):
        return "SolverResults(success=%s, n_iter=%d, residual=%.3e)" % (self.success, self.n_iter, self.residual)


class SolverOpts:
    """solver_opts (ode_base.hpp:6-13) plus the build-only precision / guess knobs."""

    def __init__(self, max_iter=1000, tol=1e-8, dt=1e-3, verbose=False, precision=PREC_FP64,
                 guess=GUESS_REFERENCE, learn_steps=0):
        self.max_iter, self.tol, self.dt, self.verbose = max_iter, tol, dt, verbose
        self.precision, self.guess, self.learn_steps = precision, guess, learn_steps

    def _c(self):
        return _Opts(int(self.max_iter), float(self.tol), float(self.dt), int(bool(self.verbose)),
                     int(self.precision), int(self.guess), int(self.learn_steps))


def structured_mesh(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=0):
    """Seeded n x n quad mesh; returns (xy[(n+1)^2, 2], quads[n^2, 4]). Host only."""
    xy = np.zeros(((n + 1) ** 2, 2))
    q = np.zeros((n * n, 4), dtype=np.int32)
    _check(lib().sem_mesh_structured(int(n), box[0], box[1], box[2], box[3], float(jitter), int(seed), _ptr(xy),
                                     q.ctypes.data_as(_ip)))
    return xy, q


def write_mesh(folder, xy, quads):
    """Write the info.txt / coordinates.txt / elements.txt files load_mesh reads."""
    xy = _f64(xy)
    q = np.ascontiguousarray(quads, dtype=np.int32).ravel()
    _check(lib().sem_mesh_write(folder.encode(), _ptr(xy), xy.size // 2, q.ctypes.data_as(_ip), q.size // 4))


class Mesh:
    """schro_omp::Mesh (openmp_impl/Mesh.hpp:16-88) resident on the GPU."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        info = (C.c_long * 6)()
        _check(lib().sem_mesh_info(self._h, info))
        self.N, self.n_elements, self.n_nodes, self.n_edges, self._dof, self.chunk = (int(v) for v in info)
        self.p = self.N - 1
        x = np.zeros(self.N)
        w = np.zeros(self.N)
        D = np.zeros(self.N * self.N)
        _check(lib().sem_mesh_quadrature(self._h, _ptr(x), _ptr(w), _ptr(D)))
        self.x, self.w = x, w
        self.D = D.reshape(self.N, self.N, order="F")
        self.size = self.n_elements * self.N * self.N

    @classmethod
    def from_arrays(cls, p, xy, quads):
        xy = _f64(xy)
        q = np.ascontiguousarray(quads, dtype=np.int32).ravel()
        h = C.c_void_p()
        _check(lib().sem_mesh_create(int(p), _ptr(xy), xy.size // 2, q.ctypes.data_as(_ip), q.size // 4,
                                     C.byref(h)))
        return cls(h.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.sem_mesh_destroy(h)
            self._h = None

    def dof(self):
        return self._dof

    def coords(self):
        """(X, Y) physical coordinates of every element-local GLL node (Quad::from_local_coo)."""
        X = np.zeros(self.size)
        Y = np.zeros(self.size)
        _check(lib().sem_mesh_coords(self._h, _ptr(X), _ptr(Y)))
        return X, Y

    def field(self, fn):
        """Sample fn(x, y) at every element-local node (the loops of tests/poisson.cpp:20-26)."""
        X, Y = self.coords()
        return np.asarray(fn(X, Y), dtype=np.float64) * np.ones_like(X)

    def metrics(self):
        arrs = [np.zeros(self.size) for _ in range(4)]
        _check(lib().sem_mesh_metrics(self._h, *[_ptr(a) for a in arrs]))
        return tuple(arrs)

    def last_run_stats(self):
        s = _Stats()
        _check(lib().sem_last_run_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}


def load_mesh(p, folder):
    """load_mesh<Mesh>(p, folder) + compute_metrics() (load_mesh.hpp:76-209)."""
    h = C.c_void_p()
    _check(lib().sem_mesh_load(int(p), folder.encode(), C.byref(h)))
    return Mesh(h.value)


def _ncomp(kind):
    return 2 if kind in (OP_SCHRO_RES, OP_SCHRO_F) else 1


def glaplace(mesh, u):
    """glaplace on every element, no assembly (glaplace.hpp:16-28)."""
    u = _f64(u, mesh.size)
    y = np.zeros_like(u)
    _check(lib().sem_glaplace_apply(mesh._h, _ptr(u), _ptr(y)))
    return y


def element_op(kind, u, mesh, coef=None, alpha=0.0):
    """The element-local operator of `kind` on every element, no assembly."""
    nc = _ncomp(kind)
    u = _f64(u, nc * mesh.size)
    y = np.zeros_like(u)
    keep = coef if coef is None else _f64(coef, mesh.size)
    cp = _ptr(keep) if keep is not None else None
    _check(lib().sem_element_apply(mesh._h, int(kind), cp, float(alpha), _ptr(u), _ptr(y)))
    return y


def galerkin_op(kind, u, mesh, coef=None, alpha=0.0):
    """galerkin_op(op, u, mesh): unmask -> element op -> global_sum -> mask (galerkin.hpp:24-35)."""
    nc = _ncomp(kind)
    u = _f64(u, nc * mesh.size)
    y = np.zeros_like(u)
    keep = coef if coef is None else _f64(coef, mesh.size)
    cp = _ptr(keep) if keep is not None else None
    _check(lib().sem_galerkin_apply(mesh._h, int(kind), cp, float(alpha), _ptr(u), _ptr(y)))
    return y


def gproj(u, mesh):
    """gproj (galerkin.hpp:42-57)."""
    u = _f64(u, mesh.size)
    y = np.zeros_like(u)
    _check(lib().sem_gproj(mesh._h, _ptr(u), _ptr(y)))
    return y


def _inplace(fn, a, mesh):
    b = _f64(a, mesh.size).copy()
    _check(fn(mesh._h, _ptr(b)))
    return b


def unmask(a, mesh):
    """unmask (openmp_impl/unmask.hpp:43-55); returns a new array."""
    return _inplace(lib().sem_unmask, a, mesh)


def global_sum(a, mesh):
    """global_sum (openmp_impl/global_sum.hpp:65-77); returns a new array."""
    return _inplace(lib().sem_global_sum, a, mesh)


def mask(a, mesh):
    """mask (openmp_impl/mask.hpp:16-28); returns a new array."""
    return _inplace(lib().sem_mask, a, mesh)


def dot(mesh, a, b):
    """dot (openmp_impl/dot.hpp:11-61)."""
    out = C.c_double()
    _check(lib().sem_dot(mesh._h, _ptr(_f64(a, mesh.size)), _ptr(_f64(b, mesh.size)), C.byref(out)))
    return out.value


def poisson(u, f, mesh, max_iter, tol):
    """poisson (openmp_impl/poisson.hpp:17-59); u (float64, length E*N*N) is updated in place."""
    if not (isinstance(u, np.ndarray) and u.dtype == np.float64 and u.flags.c_contiguous and u.size == mesh.size):
        raise ValueError("u must be a contiguous float64 array of mesh.size values")
    r = _Results()
    _check(lib().sem_poisson(mesh._h, _ptr(u), _ptr(_f64(f, mesh.size)), int(max_iter), float(tol), C.byref(r)))
    return SolverResults(r)


def helmholtz(u, c, f, mesh, max_iter, tol):
    """helmholtz (openmp_impl/helmholtz.hpp:18-64); u updated in place."""
    if not (isinstance(u, np.ndarray) and u.dtype == np.float64 and u.flags.c_contiguous and u.size == mesh.size):
        raise ValueError("u must be a contiguous float64 array of mesh.size values")
    r = _Results()
    _check(lib().sem_helmholtz(mesh._h, _ptr(u), _ptr(_f64(c, mesh.size)), _ptr(_f64(f, mesh.size)),
                               int(max_iter), float(tol), C.byref(r)))
    return SolverResults(r)


def _stepper(name, psi_real, psi_imag, potential, mesh, T, opts, callback):
    for a in (psi_real, psi_imag):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous and a.size == mesh.size):
            raise ValueError("psi arrays must be contiguous float64 arrays of mesh.size values")
    n = mesh.size
    if callback is None:
        cb = STEP_CB()
    else:
        def _cb(re, im, _user):
            return 1 if callback(np.ctypeslib.as_array(re, (n,)).copy(), np.ctypeslib.as_array(im, (n,)).copy()) else 0
        cb = STEP_CB(_cb)
    ok = C.c_int(0)
    o = opts._c()
    V = _f64(potential, n)
    _check(getattr(lib(), name)(mesh._h, _ptr(psi_real), _ptr(psi_imag), _ptr(V), float(T), C.byref(o), cb, None,
                                C.byref(ok)))
    return bool(ok.value)


def schrodinger_bdf2(psi_real, psi_imag, potential, mesh, T, opts, callback=None):
    """schrodinger_bdf2 (openmp_impl/schrodinger.hpp:22-128); psi updated in place; returns success."""
    return _stepper("sem_schrodinger_bdf2", psi_real, psi_imag, potential, mesh, T, opts, callback)


def schrodinger_dirk4s3(psi_real, psi_imag, potential, mesh, T, opts, callback=None):
    """schrodinger_dirk4s3 (openmp_impl/schrodinger.hpp:294-423)."""
    return _stepper("sem_schrodinger_dirk4s3", psi_real, psi_imag, potential, mesh, T, opts, callback)


def schrodinger_dirk4s5(psi_real, psi_imag, potential, mesh, T, opts, callback=None):
    """schrodinger_dirk4s5 (openmp_impl/schrodinger.hpp:138-285)."""
    return _stepper("sem_schrodinger_dirk4s5", psi_real, psi_imag, potential, mesh, T, opts, callback)


class Batch:
    """Independent meshes advanced together (one device mesh and one launch per
    kernel for each degree; each instance computes what it computes alone)."""

    def __init__(self, meshes):
        self.meshes = list(meshes)
        arr = (C.c_void_p * len(self.meshes))(*[m._h.value for m in self.meshes])
        h = C.c_void_p()
        _check(lib().sem_batch_create(len(self.meshes), arr, C.byref(h)))
        self._h = h

    def info(self):
        """[(degree, instances, device elements)] per degree group."""
        n = len(self.meshes)
        ng = C.c_int()
        gp = (C.c_int * n)()
        gs = (C.c_int * n)()
        ge = (C.c_long * n)()
        _check(lib().sem_batch_info(self._h, C.byref(ng), gp, gs, ge))
        return [(gp[g], gs[g], ge[g]) for g in range(ng.value)]

    def autotune(self, precision=PREC_FP64, steps=24):
        """Autotune every group's Lanczos-step shape; returns {degree: shape}."""
        groups = self.info()
        sh = (C.c_int * (4 * len(groups)))()
        _check(lib().sem_batch_autotune(self._h, int(precision), int(steps), sh))
        return {groups[g][0]: tuple(sh[4 * g:4 * g + 4]) for g in range(len(groups))}

    def set_step_shape(self, group, shape, precision=PREC_FP64):
        arr = (C.c_int * 4)(*[int(v) for v in shape])
        _check(lib().sem_batch_set_step_shape(self._h, int(group), int(precision), arr))

    def set_graph_mode(self, on, steps=12):
        _check(lib().sem_batch_set_graph_mode(self._h, int(bool(on)), int(steps)))

    def set_step_mode(self, mode, precision=PREC_FP64):
        """Lanczos step of every group: 0 two launches, 1 one fused launch, 2 fused launch plus a closing kernel."""
        _check(lib().sem_batch_set_step_mode(self._h, int(precision), int(mode)))

    def _run(self, fn, psi_real, psi_imag, potentials, T, opts):
        n = len(self.meshes)
        for k in range(n):
            for a in (psi_real[k], psi_imag[k]):
                if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                        and a.size == self.meshes[k].size):
                    raise ValueError("psi arrays must be contiguous float64 arrays of mesh.size values")
        V = [_f64(potentials[k], self.meshes[k].size) for k in range(n)]
        re = (_dp * n)(*[_ptr(a) for a in psi_real])
        im = (_dp * n)(*[_ptr(a) for a in psi_imag])
        vv = (_dp * n)(*[_ptr(a) for a in V])
        ok = (C.c_int * n)()
        o = opts._c()
        _check(getattr(lib(), fn)(self._h, re, im, vv, float(T), C.byref(o), ok))
        return [bool(v) for v in ok]

    def schrodinger_bdf2(self, psi_real, psi_imag, potentials, T, opts):
        """schrodinger_bdf2 on every instance; psi arrays updated in place; returns per-instance success."""
        return self._run("sem_batch_schrodinger_bdf2", psi_real, psi_imag, potentials, T, opts)

    def schrodinger_dirk4s3(self, psi_real, psi_imag, potentials, T, opts):
        """schrodinger_dirk4s3 on every instance; psi arrays updated in place; returns per-instance success."""
        return self._run("sem_batch_schrodinger_dirk4s3", psi_real, psi_imag, potentials, T, opts)

    def schrodinger_dirk4s5(self, psi_real, psi_imag, potentials, T, opts):
        """schrodinger_dirk4s5 on every instance; psi arrays updated in place; returns per-instance success."""
        return self._run("sem_batch_schrodinger_dirk4s5", psi_real, psi_imag, potentials, T, opts)

    def start(self, scheme, psi_real, psi_imag, potentials, opts):
        """Start a session (scheme 0 bdf2, 1 dirk4s3, 2 dirk4s5) on every instance (state stays on the GPU)."""
        n = len(self.meshes)
        arrs = [[_f64(a[k], self.meshes[k].size) for k in range(n)] for a in (psi_real, psi_imag, potentials)]
        ptrs = [(_dp * n)(*[_ptr(a) for a in group]) for group in arrs]
        o = opts._c()
        _check(lib().sem_batch_start(self._h, int(scheme), ptrs[0], ptrs[1], ptrs[2], C.byref(o)))

    def start_bdf2(self, psi_real, psi_imag, potentials, opts):
        """Start a BDF2 session on every instance (state stays on the GPU)."""
        self.start(0, psi_real, psi_imag, potentials, opts)

    def advance(self, T, max_steps=0):
        ok = C.c_int()
        _check(lib().sem_batch_advance(self._h, float(T), int(max_steps), C.byref(ok)))
        return bool(ok.value)

    def read(self):
        """(psi_real list, psi_imag list, per-instance success, t) of the session."""
        n = len(self.meshes)
        re = [np.zeros(m.size) for m in self.meshes]
        im = [np.zeros(m.size) for m in self.meshes]
        ok = (C.c_int * n)()
        t = C.c_double()
        _check(lib().sem_batch_read(self._h, (_dp * n)(*[_ptr(a) for a in re]), (_dp * n)(*[_ptr(a) for a in im]),
                                    ok, C.byref(t)))
        return re, im, [bool(v) for v in ok], t.value

    def times(self):
        """per-instance (time, completed steps) of the session"""
        n = len(self.meshes)
        t = np.zeros(n)
        st = (C.c_int * n)()
        _check(lib().sem_batch_times(self._h, _ptr(t), st))
        return t, [int(v) for v in st]

    def last_run_stats(self):
        s = _Stats()
        sec = C.c_double()
        _check(lib().sem_batch_last_run_stats(self._h, C.byref(s), C.byref(sec)))
        d = {k: getattr(s, k) for k, _ in _Stats._fields_}
        d["seconds"] = sec.value
        return d

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and _lib is not None:
            _lib.sem_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Stepper:
    """A Schrodinger stepper whose state stays on the GPU between advance() calls.

    scheme: 0 = bdf2, 1 = dirk4s3, 2 = dirk4s5 (openmp_impl/schrodinger.hpp).
    """

    def __init__(self, mesh, scheme, psi_real, psi_imag, potential, opts):
        self.mesh = mesh
        n = mesh.size
        h = C.c_void_p()
        self._opts = opts._c()
        _check(lib().sem_stepper_create(mesh._h, int(scheme), _ptr(_f64(psi_real, n)), _ptr(_f64(psi_imag, n)),
                                        _ptr(_f64(potential, n)), C.byref(self._opts), C.byref(h)))
        self._h = h

    def advance(self, T, max_steps=0, callback=None):
        n = self.mesh.size
        if callback is None:
            cb = STEP_CB()
        else:
            def _cb(re, im, _user):
                return 1 if callback(np.ctypeslib.as_array(re, (n,)).copy(),
                                     np.ctypeslib.as_array(im, (n,)).copy()) else 0
            cb = STEP_CB(_cb)
        ok = C.c_int()
        _check(lib().sem_stepper_advance(self._h, float(T), int(max_steps), cb, None, C.byref(ok)))
        return bool(ok.value)

    def read(self):
        n = self.mesh.size
        re = np.zeros(n)
        im = np.zeros(n)
        t = C.c_double()
        _check(lib().sem_stepper_read(self._h, _ptr(re), _ptr(im), C.byref(t)))
        return re, im, t.value

    def set_learning(self, epochs, lr):
        """Learned guess: Adam steps per stage solve during training, and the learning rate."""
        _check(lib().sem_stepper_set_learning(self._h, int(epochs), float(lr)))

    def stats(self):
        s = _Stats()
        _check(lib().sem_stepper_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}

    def set_training(self, loss, inputs):
        """Learned guess, before its first stage solve: loss 0 = squared error of
        the correction, 1 = assembled residual; inputs 0 = right-hand side,
        1 = the extrapolation's residual."""
        _check(lib().sem_stepper_set_training(self._h, int(loss), int(inputs)))

    def set_guard(self, on):
        """Learned guess: keep the extrapolation when its true residual is smaller (default on)."""
        _check(lib().sem_stepper_set_guard(self._h, int(bool(on))))

    def save(self):
        """Checkpoint of the session (t, step count, psi, BDF2 history, stage history, guess network)."""
        n = C.c_long()
        _check(lib().sem_stepper_checkpoint_size(self._h, C.byref(n)))
        buf = np.zeros(n.value)
        _check(lib().sem_stepper_save(self._h, _ptr(buf), n.value))
        return buf

    def restore(self, checkpoint):
        """Resume from a checkpoint of the same scheme and mesh size (see sem_stepper_restore)."""
        buf = _f64(checkpoint)
        _check(lib().sem_stepper_restore(self._h, _ptr(buf), buf.size))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and _lib is not None:
            _lib.sem_stepper_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def progress_line(t, T, dt, err=None):
    """The steppers' verbose progress line (VerboseTracker, verbose.hpp:33-44)."""
    buf = C.create_string_buffer(256)
    _check(lib().sem_progress_line(float(t), float(T), float(dt), int(err is not None),
                                   0.0 if err is None else float(err), buf, 256))
    return buf.value.decode()


def time_copy(nbytes, reps=20, variant=1):
    """Mean microseconds per launch of a device copy of nbytes (nbytes read + nbytes written)."""
    us = C.c_double()
    _check(lib().sem_time_copy(int(nbytes), int(reps), int(variant), C.byref(us)))
    return us.value


def set_kernel_timing(mesh, on):
    """0 off, 1 (or True) every fourth Lanczos step, 2 every step."""
    _check(lib().sem_set_kernel_timing(mesh._h, int(on)))


def kernel_timing(mesh, which=0):
    """(launches, total microseconds, algorithmic bytes per launch) of a Lanczos-step kernel:
    which = 0/1 operator kernel FP64/FP32, 2/3 update kernel FP64/FP32."""
    n = C.c_long()
    us = C.c_double()
    by = C.c_double()
    _check(lib().sem_kernel_timing(mesh._h, int(which), C.byref(n), C.byref(us), C.byref(by)))
    return n.value, us.value, by.value


def set_step_shape(mesh, shape, precision=PREC_FP64):
    """Work shape (op chunks/workgroup, op grid, update chunks/workgroup, update grid) of the Lanczos step."""
    arr = (C.c_int * 4)(*[int(v) for v in shape])
    _check(lib().sem_set_step_shape(mesh._h, int(precision), arr))


def set_step_mode(mesh, mode, precision=None):
    """MINRES Lanczos step: 0 two launches per step, 1 one fused launch per step,
    2 the fused step with its recurrence reduction in a separate small kernel
    (precision PREC_FP64 or PREC_MIXED; None sets PREC_MIXED, and PREC_FP64 too
    for mode 0: the fused step runs in FP32 only)."""
    if precision is None:
        precs = (PREC_FP64, PREC_MIXED) if mode == 0 else (PREC_MIXED,)
    else:
        precs = (precision,)
    for p in precs:
        _check(lib().sem_set_step_mode(mesh._h, int(p), int(mode)))


def get_step_mode(mesh, precision=PREC_FP64):
    v = C.c_int()
    _check(lib().sem_get_step_mode(mesh._h, int(precision), C.byref(v)))
    return v.value


def set_inner_scaling(mesh, on):
    """Diagonal scaling of the FP32 inner MINRES of mixed-precision Schrodinger stage solves (default on)."""
    _check(lib().sem_set_inner_scaling(mesh._h, int(bool(on))))


def get_inner_scaling(mesh):
    v = C.c_int()
    _check(lib().sem_get_inner_scaling(mesh._h, C.byref(v)))
    return bool(v.value)


def set_graph_mode(mesh, on, steps=12):
    """Replay batches of `steps` Lanczos steps (multiple of 3) as captured HIP graphs."""
    _check(lib().sem_set_graph_mode(mesh._h, int(bool(on)), int(steps)))


def get_step_shape(mesh, precision=PREC_FP64):
    arr = (C.c_int * 4)()
    _check(lib().sem_get_step_shape(mesh._h, int(precision), arr))
    return tuple(arr)


def autotune(mesh, kind=None, precision=PREC_FP64, steps=24):
    """Time the candidate Lanczos-step shapes on this mesh and keep the fastest per kernel.

    Returns (chosen shape, [(chunks per workgroup, grid, op us, update us), ...])."""
    kind = OP_SCHRO_RES if kind is None else kind
    n = C.c_int()
    cand = (C.c_int * 32)()
    us = (C.c_double * 32)()
    _check(lib().sem_autotune(mesh._h, int(kind), int(precision), int(steps), C.byref(n), cand, us))
    rows = [(cand[2 * k], cand[2 * k + 1], us[2 * k], us[2 * k + 1]) for k in range(n.value)]
    return get_step_shape(mesh, precision), rows


def stage_solve(mesh, V, alpha, b, z=None, max_iter=10000, tol=1e-10, precision=PREC_FP64):
    """One Schrodinger stage solve Res(alpha, V) z = b; returns (z, SolverResults)."""
    n = 2 * mesh.size
    b = _f64(b, n)
    z = np.zeros(n) if z is None else _f64(z, n).copy()
    r = _Results()
    _check(lib().sem_stage_solve(mesh._h, _ptr(_f64(V, mesh.size)), float(alpha), _ptr(b), _ptr(z), int(max_iter),
                                 float(tol), int(precision), C.byref(r)))
    return z, SolverResults(r)


def minres_tridiag(lo, di, up, b, x, maxit, tol):
    """minres (minres.hpp:19-107) on y_i = lo x_{i-1} + di x_i + up x_{i+1}; x updated in place."""
    b = _f64(b)
    r = _Results()
    _check(lib().sem_minres_tridiag(b.size, lo, di, up, _ptr(b), _ptr(x), int(maxit), float(tol), C.byref(r)))
    return SolverResults(r)


def pcg_tridiag(lo, di, up, b, x, maxit, tol):
    """pcg (pcg.hpp:59-97) on a tridiagonal operator; x updated in place."""
    b = _f64(b)
    r = _Results()
    _check(lib().sem_pcg_tridiag(b.size, lo, di, up, _ptr(b), _ptr(x), int(maxit), float(tol), C.byref(r)))
    return SolverResults(r)


def time_element_op(mesh, kind, precision=PREC_FP64, reps=50):
    """Mean microseconds per launch of the element-operator kernel and its algorithmic bytes."""
    us = C.c_double()
    by = C.c_double()
    _check(lib().sem_time_element_op(mesh._h, int(kind), int(precision), int(reps), C.byref(us), C.byref(by)))
    return us.value, by.value


class MlpNet:
    """The guess network's kernels (forward, MSE backward, Adam) on host arrays."""

    def __init__(self, n_in, h1, h2, n_out, seed=1):
        h = C.c_void_p()
        _check(lib().sem_mlp_create(int(n_in), int(h1), int(h2), int(n_out), int(seed), C.byref(h)))
        self._h = h
        info = (C.c_long * 9)()
        _check(lib().sem_mlp_info(h, info))
        self.in_pad, self.out_pad, self.nparams = int(info[0]), int(info[1]), int(info[2])
        self.offsets = [int(v) for v in info[3:9]]
        self.dims = (n_in, h1, h2, n_out)

    def params(self, gradients=False):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().sem_mlp_get_params(self._h, p.ctypes.data_as(_fp), int(gradients)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().sem_mlp_set_params(self._h, p.ctypes.data_as(_fp)))

    @staticmethod
    def _padded(A, width, pad, what):
        """rows x width or rows x pad float32 array as a contiguous rows x pad array (zero padding)."""
        A = np.asarray(A, dtype=np.float32)
        if A.ndim != 2 or A.shape[1] not in (width, pad):
            raise ValueError("%s must have %d (or padded %d) columns, got shape %s" % (what, width, pad, A.shape))
        if A.shape[1] == pad:
            return np.ascontiguousarray(A)
        out = np.zeros((A.shape[0], pad), dtype=np.float32)
        out[:, :width] = A
        return out

    def forward(self, X):
        X = self._padded(X, self.dims[0], self.in_pad, "X")
        Y = np.zeros((X.shape[0], self.out_pad), dtype=np.float32)
        _check(lib().sem_mlp_forward(self._h, X.shape[0], X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp)))
        return Y

    def gradients(self, X, T):
        X = self._padded(X, self.dims[0], self.in_pad, "X")
        T = self._padded(T, self.dims[3], self.out_pad, "T")
        if T.shape[0] != X.shape[0]:
            raise ValueError("X and T must have the same number of rows")
        loss = C.c_double()
        _check(lib().sem_mlp_gradients(self._h, X.shape[0], X.ctypes.data_as(_fp), T.ctypes.data_as(_fp),
                                       C.byref(loss)))
        return loss.value, self.params(gradients=True)

    def adam(self, lr):
        _check(lib().sem_mlp_adam(self._h, float(lr)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and _lib is not None:
            _lib.sem_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def device_info():
    name = C.create_string_buffer(256)
    ncu = C.c_int()
    hbm = C.c_long()
    _check(lib().sem_device_info(name, 256, C.byref(ncu), C.byref(hbm)))
    return name.value.decode(), ncu.value, hbm.value


# ------------------------------------------------------------ generic solvers --
def _krylov(n, A, dot, precond, host, consistent):
    """sem_krylov with Python callables; host=True: they take and fill numpy arrays,
    host=False: they receive DevicePtr views (device pointers, library stream)."""
    keep = []

    def wrap(fn, kind):
        if fn is None:
            return kind()
        if kind is LINOP_CB or kind is PRECOND_CB:
            def cb(_u, x, y, stream):
                try:
                    if host:
                        fn(np.ctypeslib.as_array(x, (n,)), np.ctypeslib.as_array(y, (n,)))
                    else:
                        fn(x, y, stream)
                    return 0
                except Exception:  # noqa: BLE001 - reported to the library as a failed callback
                    return 1
        else:
            def cb(_u, a, b, out, stream):
                try:
                    out[0] = fn(np.ctypeslib.as_array(a, (n,)), np.ctypeslib.as_array(b, (n,))) if host \
                        else fn(a, b, stream)
                    return 0
                except Exception:  # noqa: BLE001
                    return 1
        f = kind(cb)
        keep.append(f)
        return f

    k = _Krylov(n, int(bool(host)), int(bool(consistent)), wrap(A, LINOP_CB), None, wrap(dot, DOT_CB), None,
                wrap(precond, PRECOND_CB), None)
    return k, keep


def minres(x, A, b, maxit, tol, dot=None, precond=None, consistent=False):
    """minres.hpp:19-107 on the GPU over host arrays: A(x, out) fills out = A x,
    dot(a, b) returns a float, precond(r, out) fills out = P r (None: Euclidean
    dot, IdentityPreconditioner). x is updated in place."""
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
        raise ValueError("x must be a contiguous float64 array")
    k, keep = _krylov(x.size, A, dot, precond, True, consistent)
    r = _Results()
    _check(lib().sem_minres(C.byref(k), _ptr(_f64(b, x.size)), _ptr(x), int(maxit), float(tol), C.byref(r)))
    del keep
    return SolverResults(r)


def pcg(x, A, b, maxit, tol, dot=None, precond=None):
    """pcg.hpp:59-97 on the GPU over host arrays (see minres)."""
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
        raise ValueError("x must be a contiguous float64 array")
    k, keep = _krylov(x.size, A, dot, precond, True, False)
    r = _Results()
    _check(lib().sem_pcg(C.byref(k), _ptr(_f64(b, x.size)), _ptr(x), int(maxit), float(tol), C.byref(r)))
    del keep
    return SolverResults(r)


def minres_device(n, A, b_dev, x_dev, maxit, tol, dot=None, precond=None, consistent=False, solver="minres"):
    """The same on device vectors (pointers from DeviceArray): A(x_ptr, y_ptr, stream);
    solver "minres", "pcg" or "minres_noalloc" (the entry point that allocates nothing)."""
    k, keep = _krylov(n, A, dot, precond, False, consistent)
    r = _Results()
    fn = {"minres": lib().sem_minres, "pcg": lib().sem_pcg, "minres_noalloc": lib().sem_minres_noalloc}[solver]
    _check(fn(C.byref(k), b_dev, x_dev, int(maxit), float(tol), C.byref(r)))
    del keep
    return SolverResults(r)


class DeviceArray:
    """n doubles of device memory (sem_device_alloc), with host copies in and out."""

    def __init__(self, n):
        self.n = int(n)
        p = _dp()
        _check(lib().sem_device_alloc(self.n, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = _f64(a)
        d = cls(a.size)
        _check(lib().sem_copy(d.ptr, _ptr(a), a.size, 0))
        return d

    def to_host(self):
        a = np.zeros(self.n)
        _check(lib().sem_copy(_ptr(a), self.ptr, self.n, 1))
        return a

    def close(self):
        if getattr(self, "ptr", None) is not None and _lib is not None:
            _lib.sem_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


def copy_to_host(ptr, n):
    a = np.zeros(n)
    _check(lib().sem_copy(_ptr(a), ptr, n, 1))
    return a


def copy_to_device(ptr, a):
    a = _f64(a)
    _check(lib().sem_copy(ptr, _ptr(a), a.size, 0))


def stream_synchronize(stream):
    _check(lib().sem_stream_synchronize(stream))


def poisson_precond(u, f, mesh, max_iter, tol, precond=PRECOND_JACOBI):
    """poisson (openmp_impl/poisson.hpp:17-59) with a preconditioner; u in/out."""
    r = _Results()
    _check(lib().sem_poisson_precond(mesh._h, _ptr(u), _ptr(_f64(f, mesh.size)), int(max_iter), float(tol),
                                     int(precond), C.byref(r)))
    return SolverResults(r)


def helmholtz_precond(u, c, f, mesh, max_iter, tol, precond=PRECOND_JACOBI):
    """helmholtz (openmp_impl/helmholtz.hpp:18-64) with a preconditioner; u in/out."""
    r = _Results()
    _check(lib().sem_helmholtz_precond(mesh._h, _ptr(u), _ptr(_f64(c, mesh.size)), _ptr(_f64(f, mesh.size)),
                                       int(max_iter), float(tol), int(precond), C.byref(r)))
    return SolverResults(r)


def galerkin_apply_op(mesh, op, u):
    """galerkin_op (galerkin.hpp:24-35) with a host element operator op(u, y)
    on (n_elements, N*N) arrays in the element-local layout; the library runs
    unmask, global_sum and mask on the GPU and calls op between them."""
    n = mesh.size

    def cb(_u, ne, N, ud, yd, stream):
        try:
            stream_synchronize(stream)
            uu = copy_to_host(ud, ne * N * N).reshape(ne, N * N)
            yy = np.zeros_like(uu)
            op(uu, yy)
            copy_to_device(yd, yy.ravel())
            return 0
        except Exception:  # noqa: BLE001
            return 1

    f = ELEMENT_OP_CB(cb)
    din = DeviceArray.from_host(_f64(u, n))
    dout = DeviceArray(n)
    _check(lib().sem_galerkin_apply_op(mesh._h, f, None, din.ptr, dout.ptr))
    out = dout.to_host()
    din.close()
    dout.close()
    return out


def gauss_lobatto(n):
    """gauss_lobatto<double>(n): n + 1 GLL nodes and weights (host only)."""
    x = np.zeros(n + 1)
    w = np.zeros(n + 1)
    _check(lib().sem_gauss_lobatto(int(n), _ptr(x), _ptr(w)))
    return x, w


def derivative_matrix(x):
    x = _f64(x)
    D = np.zeros(x.size * x.size)
    _check(lib().sem_derivative_matrix(x.size, _ptr(x), _ptr(D)))
    return D.reshape(x.size, x.size, order="F")


def quad_metrics(xs, ys, z):
    z = _f64(z)
    n = z.size
    out = [np.zeros(n * n) for _ in range(4)]
    _check(lib().sem_quad_metrics(_ptr(_f64(xs, 4)), _ptr(_f64(ys, 4)), n, _ptr(z), *[_ptr(a) for a in out]))
    return [a.reshape(n, n, order="F") for a in out]


CONTRACT_VALU, CONTRACT_BF16, CONTRACT_BF16X3 = 0, 1, 3


def element_op_fp32(kind, u, mesh, coef=None, alpha=0.0):
    """The element operator in FP32 (no assembly); Res follows set_contraction."""
    nc = _ncomp(kind)
    u = _f64(u, nc * mesh.size)
    out = np.zeros_like(u)
    c = None if coef is None else _f64(coef, mesh.size)
    _check(lib().sem_element_apply_fp32(mesh._h, int(kind), None if c is None else _ptr(c), float(alpha), _ptr(u),
                                        _ptr(out)))
    return out


def set_contraction(mesh, mode):
    """Stiffness contractions of the FP32 inner solver: 0 VALU, 1 bf16 MFMA, 3 split-bf16 MFMA."""
    _check(lib().sem_set_contraction(mesh._h, int(mode)))


def set_stored_basis(mesh, on, budget_bytes=0):
    """FP32 one-launch MINRES: keep every Lanczos vector (on, default) or update w/x each step."""
    _check(lib().sem_set_stored_basis(mesh._h, int(bool(on)), int(budget_bytes)))


def get_contraction(mesh):
    v = C.c_int()
    _check(lib().sem_get_contraction(mesh._h, C.byref(v)))
    return v.value
