"""CPU checks of the C-ABI library: it loads, exports every function that
include/sem.h declares, the host-only entry points work without a GPU, and
GPU entry points fail loudly (no CPU fallback) when no GPU is present."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, SMALL_MESH

HEADER = os.path.join(ROOT, "include", "sem.h")


def declared_functions():
    with open(HEADER) as fh:
        txt = fh.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(sem_[a-z0-9_]+)\s*\(", txt))
    return sorted(n for n in names if not n.endswith("_cb"))


def test_library_loads_and_exports_header(sem):
    lib = ctypes.CDLL(sem.LIB_PATH, mode=os.RTLD_NOW)
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_structured_mesh_generator(sem):
    xy, q = sem.structured_mesh(8, (0, 1, 0, 1), 0.1, 0)
    assert xy.shape == (81, 2) and q.shape == (64, 4)
    h = 1 / 8
    g = np.stack(np.meshgrid(np.arange(9) * h, np.arange(9) * h), -1).reshape(-1, 2)
    d = np.abs(xy - g)
    assert d.max() <= 0.1 * h + 1e-15
    border = (g[:, 0] == 0) | (g[:, 0] == 1) | (g[:, 1] == 0) | (g[:, 1] == 1)
    assert np.all(d[border] == 0)
    assert np.all(d[~border].max(axis=0) > 0.05 * h)
    # counter-clockwise quads: positive shoelace area, total area 1
    P = xy[q]
    area = 0.5 * np.sum(P[:, :, 0] * np.roll(P[:, :, 1], -1, 1) - np.roll(P[:, :, 0], -1, 1) * P[:, :, 1], axis=1)
    assert np.all(area > 0) and abs(area.sum() - 1) < 1e-14
    # seeded and reproducible; golden fixture of the config-1 mesh
    xy2, _ = sem.structured_mesh(8, (0, 1, 0, 1), 0.1, 0)
    assert np.array_equal(xy, xy2)
    gold = np.loadtxt(os.path.join(ROOT, "tests", "golden", "config1_mesh", "coordinates.txt"))
    assert np.array_equal(gold, xy)


def test_mesh_write_roundtrip(sem, orc, tmp_path):
    xy, q = sem.structured_mesh(5, (-1, 1, -1, 1), 0.1, 3)
    sem.write_mesh(str(tmp_path), xy, q)
    m = orc.Mesh.load(3, str(tmp_path))
    assert m.n_elements == 25 and m.n_nodes == 36 and m.dof() == 36 + 60 * 2 + 25 * 4


def test_gpu_entry_points_fail_loudly_without_gpu(sem):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(sem.SemError, match="HIP|device"):
        sem.load_mesh(3, SMALL_MESH)


def test_mlp_host_arrays_are_padded_or_rejected(sem):
    """The network binding takes rows x n_in (zero-padded here) or rows x in_pad
    arrays and rejects any other width before the library reads them."""
    P = sem.MlpNet._padded(np.ones((3, 5), np.float64), 5, 8, "X")
    assert P.shape == (3, 8) and P.dtype == np.float32 and P.flags.c_contiguous
    assert np.array_equal(P[:, :5], np.ones((3, 5))) and not P[:, 5:].any()
    Q = sem.MlpNet._padded(np.arange(16, dtype=np.float32).reshape(2, 8), 5, 8, "X")
    assert np.array_equal(Q, np.arange(16, dtype=np.float32).reshape(2, 8))
    for bad in (np.ones((3, 6)), np.ones(5)):
        with pytest.raises(ValueError):
            sem.MlpNet._padded(bad, 5, 8, "X")


@pytest.mark.parametrize("t,T,dt,err", [(0.0123, 0.1, 1e-3, None), (0.05, 0.1, 5e-4, 3.25e-7), (0.1, 0.1, 1e-2, 1.0),
                                        (0.11, 0.1, 1e-3, None)])
def test_progress_line_is_verbose_tracker(sem, t, T, dt, err):
    # VerboseTracker (verbose.hpp:9-44): std::fixed t and std::scientific error,
    # both at setprecision(int(2 - log10 dt)) (schrodinger.hpp:28), then '\r'
    prec = int(2 - np.log10(dt))
    q = min(int(30 * (t / T)), 30)
    # out[p+1] = '>' and then out[n+1] = ']': at the end of the run the bracket wins
    bar = "[" + "-" * q + ">" + " " * (30 - q - 1) + "]" if q < 30 else "[" + "-" * 30 + "]"
    want = "%.*f :: %s" % (prec, t, bar)
    if err is not None:
        want += " :: error ~ %.*e" % (prec, err)
    assert sem.progress_line(t, T, dt, err) == want + "\r"


@pytest.mark.parametrize("p", [1, 4, 7, 8, 10, 13])
def test_host_basis_routines_match_oracle(sem, orc, p):
    # gauss_lobatto / derivative_matrix / Quad::compute_metrics are host-only entry points
    x, w = sem.gauss_lobatto(p)
    xo, wo = orc.gll(p)
    assert np.array_equal(x, xo) if p <= 8 else np.max(np.abs(x - xo)) < 5e-15
    assert np.max(np.abs(w - wo)) < 1e-14
    D = sem.derivative_matrix(x)
    assert np.max(np.abs(D - orc.dmat(x))) < 1e-12 * (p + 1) ** 2
    xs = np.array([0.0, 1.2, 1.0, -0.1])
    ys = np.array([0.0, 0.1, 0.9, 1.1])
    got = sem.quad_metrics(xs, ys, x)
    want = orc.quad_metrics(xs, ys, x)
    for a, b in zip(got, want):
        assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))
