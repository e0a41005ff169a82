"""Geometry of the FP32 operator applies: the factors gA, gB, gC, gJ are formed
in the kernels from each element's bilinear map (the metric functions of
Quad.hpp:39-54 are linear in the GLL node) instead of loaded. Every element
operator in FP32 matches the FP64 oracle (stored FP64 factors) to the FP32
tolerance, on jittered meshes, on the irregular small_mesh and on strongly
skewed quads, where a wrong term or index of the metric would show at once."""
import numpy as np
import pytest

from conftest import SMALL_MESH, rel, structured

pytestmark = pytest.mark.gpu

TOL = 1e-6  # relative L2 of an FP32 apply against FP64
KINDS = ["LAPLACE", "HELMHOLTZ", "MASS", "SCHRO_RES", "SCHRO_F", "SCHRO_L"]


def _check_all_kinds(sem, orc, m, o, seed):
    X, Y = m.coords()
    V = 1.0 + X * X + 0.5 * Y
    rng = np.random.default_rng(seed)
    sem.set_contraction(m, 0)
    errs = {}
    for name in KINDS:
        ks, ko = getattr(sem, "OP_" + name), getattr(orc, name)
        nc = 2 if name in ("SCHRO_RES", "SCHRO_F") else 1
        u = rng.standard_normal(nc * m.size)
        coef = None if name in ("LAPLACE", "MASS") else V
        alpha = 0.05 if name == "SCHRO_RES" else 0.0
        ref = o.element_op(ko, u, coef, alpha)
        errs[name] = rel(sem.element_op_fp32(ks, u, m, coef, alpha), ref)
    print({k: "%.1e" % v for k, v in errs.items()})
    for name, e in errs.items():
        assert e < TOL, (name, e)


@pytest.mark.parametrize("p", [2, 4, 7, 8, 11])
def test_fp32_kinds_jittered(sem, orc, p):
    xy, q = structured(6, (-3, 1, -2, 2), 0.15, p)
    _check_all_kinds(sem, orc, sem.Mesh.from_arrays(p, xy, q), orc.Mesh.from_arrays(p, xy, q), p)


def test_fp32_kinds_small_mesh(sem, orc):
    _check_all_kinds(sem, orc, sem.load_mesh(5, SMALL_MESH), orc.Mesh.load(5, SMALL_MESH), 3)


def test_fp32_kinds_skewed_quads(sem, orc):
    # a 2 x 2 patch of parallelograms and trapezoids: X_eta, Y_xi and the
    # z-dependent parts of the map are all nonzero
    xy = np.array([[0.0, 0.0], [1.0, 0.2], [2.3, 0.1],
                   [0.4, 1.0], [1.6, 1.3], [2.5, 1.1],
                   [0.3, 2.2], [1.2, 2.6], [2.9, 2.4]])
    q = np.array([[0, 1, 4, 3], [1, 2, 5, 4], [3, 4, 7, 6], [4, 5, 8, 7]], dtype=np.int32)
    for p in (3, 7):
        _check_all_kinds(sem, orc, sem.Mesh.from_arrays(p, xy, q), orc.Mesh.from_arrays(p, xy, q), 10 + p)
