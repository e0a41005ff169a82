"""Batched uneven instances: every instance of a batch computes bit for bit
what it computes alone, and matches the oracle's schrodinger_bdf2."""
import numpy as np
import pytest

from conftest import rel, structured

pytestmark = pytest.mark.gpu


def instance(sem, n, p, seed):
    rng = np.random.default_rng(100 + seed)
    xy, q = structured(n, (-3, 3, -3, 3), 0.1, seed)
    m = sem.Mesh.from_arrays(p, xy, q)
    cx, cy = rng.uniform(-1, 1, 2)
    kx, ky = rng.uniform(-2, 2, 2)
    V = m.field(lambda x, y: 0.5 * (x * x + y * y) + 2.0 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / 0.5))
    env = m.field(lambda x, y: np.exp(-((x + cx) ** 2 + (y + cy) ** 2) / 0.8))
    X, Y = m.coords()
    re = env * np.cos(kx * X + ky * Y)
    im = env * np.sin(kx * X + ky * Y)
    return m, xy, q, V, re, im


SPECS = [(5, 4, 1), (8, 4, 2), (3, 6, 3), (6, 4, 4), (4, 6, 5), (2, 8, 6)]


def test_batch_matches_single_runs_bitwise(sem):
    dt, steps = 2e-3, 4
    T = (steps - 0.5) * dt
    inst = [instance(sem, n, p, s) for n, p, s in SPECS]
    opts = sem.SolverOpts(max_iter=5000, tol=1e-12, dt=dt)
    single = []
    for m, _, _, V, re, im in inst:
        a, b = re.copy(), im.copy()
        assert sem.schrodinger_bdf2(a, b, V, m, T, opts)
        assert m.last_run_stats()["steps"] == steps
        single.append((a, b))
    batch = sem.Batch([x[0] for x in inst])
    groups = batch.info()
    assert sorted(g[0] for g in groups) == [4, 6, 8]
    assert sum(g[1] for g in groups) == len(SPECS)
    for shapes in (None, "tuned"):
        if shapes == "tuned":
            tuned = batch.autotune(sem.PREC_FP64, 6)
            assert set(tuned) == {4, 6, 8}
        re = [x[4].copy() for x in inst]
        im = [x[5].copy() for x in inst]
        ok = batch.schrodinger_bdf2(re, im, [x[3] for x in inst], T, opts)
        assert all(ok)
        st = batch.last_run_stats()
        assert st["steps"] == steps and st["seconds"] > 0
        for k in range(len(inst)):
            assert np.array_equal(re[k], single[k][0]), k
            assert np.array_equal(im[k], single[k][1]), k
    batch.close()


def test_batch_shapes_and_graphs_bitwise(sem):
    dt, steps = 2e-3, 3
    T = (steps - 0.5) * dt
    inst = [instance(sem, n, 4, s) for n, s in ((7, 11), (4, 12), (9, 13))]
    opts = sem.SolverOpts(max_iter=5000, tol=1e-12, dt=dt)
    batch = sem.Batch([x[0] for x in inst])
    out = []
    for shape, graphs in (((1, 0, 1, 0), False), ((1, 0, 1, 0), True), ((2, 5, 1, 3), True), ((1, 2, 2, 0), False),
                          ((1, 2, 2, 0), True)):
        batch.set_step_shape(0, shape)
        batch.set_graph_mode(graphs, 6)
        re = [x[4].copy() for x in inst]
        im = [x[5].copy() for x in inst]
        assert all(batch.schrodinger_bdf2(re, im, [x[3] for x in inst], T, opts))
        out.append((re, im))
    for re, im in out[1:]:
        for k in range(len(inst)):
            assert np.array_equal(re[k], out[0][0][k]) and np.array_equal(im[k], out[0][1][k])


def test_batch_oracle_parity(sem, orc):
    dt, steps, tol = 2e-3, 3, 1e-13
    T = (steps - 0.5) * dt
    inst = [instance(sem, n, p, s) for n, p, s in ((6, 4, 21), (4, 6, 22))]
    opts = sem.SolverOpts(max_iter=5000, tol=tol, dt=dt)
    batch = sem.Batch([x[0] for x in inst])
    re = [x[4].copy() for x in inst]
    im = [x[5].copy() for x in inst]
    assert all(batch.schrodinger_bdf2(re, im, [x[3] for x in inst], T, opts))
    for k, (m, xy, q, V, r0, i0) in enumerate(inst):
        o = orc.Mesh.from_arrays(m.p, xy, q)
        ro, io, ok, stats, _ = o.schrodinger(0, r0, i0, V, T, 5000, tol, dt)
        assert ok and int(stats[0]) == steps
        err = rel(np.concatenate([re[k], im[k]]), np.concatenate([ro, io]))
        assert err < 1e-9, (k, err)


def test_batch_session_matches_one_shot(sem):
    dt, steps = 2e-3, 4
    T = (steps - 0.5) * dt
    inst = [instance(sem, n, p, s) for n, p, s in ((5, 4, 31), (3, 6, 32), (6, 4, 33))]
    opts = sem.SolverOpts(max_iter=5000, tol=1e-12, dt=dt)
    batch = sem.Batch([x[0] for x in inst])
    re = [x[4].copy() for x in inst]
    im = [x[5].copy() for x in inst]
    assert all(batch.schrodinger_bdf2(re, im, [x[3] for x in inst], T, opts))
    batch.start_bdf2([x[4] for x in inst], [x[5] for x in inst], [x[3] for x in inst], opts)
    assert batch.advance(T, 2)
    _, _, _, t = batch.read()
    assert abs(t - 2 * dt) < 1e-15
    assert batch.advance(T, 0)
    r2, i2, ok, t = batch.read()
    assert all(ok) and batch.last_run_stats()["steps"] == steps
    for k in range(len(inst)):
        assert np.array_equal(r2[k], re[k]) and np.array_equal(i2[k], im[k])


@pytest.mark.skip(reason="the one-launch Lanczos step runs in FP32 only (its FP64 instantiation is switched off "
                         "after a work shape faulted the GPU on a 256x256, degree-7 mesh; cause not found); "
                         "batched BDF2 solves in FP64")
def test_batch_fused_step_bitwise(sem):
    """The one-launch Lanczos step over a batch: per-instance recurrences and
    stop tests (instances stop at different steps; the kernel skips the items
    of stopped ones), each instance equal bit for bit to its single run with
    the same step, in one-item and persistent work shapes."""
    dt, steps = 2e-3, 3
    T = (steps - 0.5) * dt
    inst = [instance(sem, n, p, s) for n, p, s in ((7, 4, 41), (3, 4, 42), (9, 4, 43), (5, 6, 44))]
    opts = sem.SolverOpts(max_iter=5000, tol=1e-12, dt=dt)
    single = []
    for m, _, _, V, re, im in inst:
        sem.set_step_mode(m, 1, sem.PREC_FP64)
        sem.set_graph_mode(m, False)
        a, b = re.copy(), im.copy()
        assert sem.schrodinger_bdf2(a, b, V, m, T, opts)
        single.append((a, b))
    batch = sem.Batch([x[0] for x in inst])
    batch.set_graph_mode(False)
    batch.set_step_mode(1)
    for shape in ((2, 0, 1, 0), (1, 3, 1, 0), (2, 64, 1, 0)):
        for g in range(len(batch.info())):
            batch.set_step_shape(g, shape)
        re = [x[4].copy() for x in inst]
        im = [x[5].copy() for x in inst]
        assert all(batch.schrodinger_bdf2(re, im, [x[3] for x in inst], T, opts))
        for k in range(len(inst)):
            assert np.array_equal(re[k], single[k][0]) and np.array_equal(im[k], single[k][1]), (shape, k)
    batch.close()
