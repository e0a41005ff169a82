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


SDIRK_SPECS = [(5, 4, 51), (3, 6, 52), (7, 4, 53), (4, 6, 54), (2, 8, 55)]


@pytest.mark.parametrize("scheme", [1, 2])
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_sdirk_matches_single_runs_bitwise(sem, scheme, precision):
    """Batched DIRK4S3 / DIRK4S5 (schrodinger.hpp:294-423, :138-285) with the
    extrapolated guess, FP64 or mixed precision (per-instance refinement, the
    one-launch FP32 step replayed as graphs in the batch, direct launches in
    the single runs): every instance equal bit for bit to its single run."""
    dt, steps = 2e-3, 4
    T = (steps - 0.5) * dt
    inst = [instance(sem, n, p, s) for n, p, s in SDIRK_SPECS]
    opts = sem.SolverOpts(max_iter=20000, tol=1e-12, dt=dt, precision=precision, guess=1)
    fn = (None, "schrodinger_dirk4s3", "schrodinger_dirk4s5")[scheme]
    single, iters = [], 0
    for m, _, _, V, re, im in inst:
        # batches run the FP32 inner MINRES with the w/x recurrence; the single
        # runs use it too (a stored-basis solve rounds its iterate differently)
        sem.set_stored_basis(m, 0)
        a, b = re.copy(), im.copy()
        assert getattr(sem, fn)(a, b, V, m, T, opts)
        st = m.last_run_stats()
        assert st["steps"] == steps
        iters += st["total_iters"]
        single.append((a, b))
    batch = sem.Batch([x[0] for x in inst])
    re = [x[4].copy() for x in inst]
    im = [x[5].copy() for x in inst]
    ok = getattr(batch, fn)(re, im, [x[3] for x in inst], T, opts)
    assert all(ok)
    st = batch.last_run_stats()
    assert st["steps"] == steps and st["stage_solves"] == steps * (3 if scheme == 1 else 5) * 3
    if precision == 1:
        assert st["refinements"] > 0
    print("scheme %d precision %d: batch of %d, %d Lanczos steps in the batch, %d in the single runs"
          % (scheme, precision, len(inst), st["total_iters"], iters))
    for k in range(len(inst)):
        assert np.array_equal(re[k], single[k][0]) and np.array_equal(im[k], single[k][1]), k
    batch.close()


def test_batch_learned_guess(sem, orc):
    """Batched DIRK4S5 with the learned guess in mixed precision: one network
    per degree group, trained on the first steps' stages, then frozen; every
    instance meets its own stop test, so it agrees with its single FP64 run to
    the tolerance, and the oracle to 1e-9."""
    dt, steps, tol = 2e-3, 8, 1e-13
    T = (steps - 0.5) * dt
    inst = [instance(sem, n, 4, s) for n, s in ((6, 61), (4, 62), (5, 63))]
    batch = sem.Batch([x[0] for x in inst])
    re = [x[4].copy() for x in inst]
    im = [x[5].copy() for x in inst]
    opts = sem.SolverOpts(max_iter=20000, tol=tol, dt=dt, precision=1, guess=2, learn_steps=4)
    assert all(batch.schrodinger_dirk4s5(re, im, [x[3] for x in inst], T, opts))
    st = batch.last_run_stats()
    assert st["steps"] == steps and st["mlp_train_steps"] > 0 and st["learned_guess_chosen"] > 0
    ref = sem.SolverOpts(max_iter=20000, tol=tol, dt=dt)
    for k, (m, xy, q, V, r0, i0) in enumerate(inst):
        a, b = r0.copy(), i0.copy()
        assert sem.schrodinger_dirk4s5(a, b, V, m, T, ref)
        e = rel(np.concatenate([re[k], im[k]]), np.concatenate([a, b]))
        if k == 0:
            o = orc.Mesh.from_arrays(m.p, xy, q)
            ro, io, oko, _, _ = o.schrodinger(2, r0, i0, V, T, 20000, tol, dt)
            assert oko
            eo = rel(np.concatenate([re[k], im[k]]), np.concatenate([ro, io]))
            print("learned batch instance 0: rel-L2 vs oracle %.2e" % eo)
            assert eo < 1e-9
        print("learned batch instance %d: rel-L2 vs single FP64 run %.2e" % (k, e))
        assert e < 1e-9
    print("learned batch: %d Adam steps, %d of %d stage solves from the network's guess"
          % (st["mlp_train_steps"], st["learned_guess_chosen"], st["stage_solves"]))
    batch.close()


@pytest.mark.parametrize("scheme,precision", [(0, 0), (1, 0), (2, 1)])
def test_batch_instance_failure(sem, scheme, precision):
    """One instance's solve fails (a shared max_iter that its larger mesh
    cannot meet): it stops there as its single run does (BDF2 keeps the
    unconverged iterate and counts the step, schrodinger.hpp:96-114; the SDIRK
    schemes keep psi of the step's start, :222-226), with its own time and
    step count, while the other instance of its degree group runs to T, equal
    bit for bit to its single run."""
    dt, steps = 2e-3, 3
    T = (steps - 0.5) * dt
    small = instance(sem, 2, 4, 71)
    large = instance(sem, 14, 4, 72)
    fns = ("schrodinger_bdf2", "schrodinger_dirk4s3", "schrodinger_dirk4s5")
    fn = fns[scheme]

    def single(x, maxit):
        m, _, _, V, re, im = x
        sem.set_stored_basis(m, 0)  # the batch's FP32 inner MINRES form (w/x recurrence)
        a, b = re.copy(), im.copy()
        ok = getattr(sem, fn)(a, b, V, m, T, sem.SolverOpts(max_iter=maxit, tol=1e-12, dt=dt, precision=precision))
        return ok, a, b, m.last_run_stats()["steps"]

    maxit = None
    for cand in (10, 20, 40, 80, 160, 320, 640):
        if single(small, cand)[0] and not single(large, cand)[0]:
            maxit = cand
            break
    assert maxit is not None, "no shared iteration limit separates the two instances"
    ok_s, rs, is_, ns = single(small, maxit)
    ok_l, rl, il, nl = single(large, maxit)
    assert ok_s and not ok_l and ns == steps
    batch = sem.Batch([small[0], large[0]])
    assert len(batch.info()) == 1
    re = [small[4].copy(), large[4].copy()]
    im = [small[5].copy(), large[5].copy()]
    opts = sem.SolverOpts(max_iter=maxit, tol=1e-12, dt=dt, precision=precision)
    batch.start(scheme, re, im, [small[3], large[3]], opts)
    assert not batch.advance(T)
    r2, i2, ok, t = batch.read()
    assert ok == [True, False]
    assert np.array_equal(r2[0], rs) and np.array_equal(i2[0], is_)
    assert np.array_equal(r2[1], rl) and np.array_equal(i2[1], il)
    tt, nn = batch.times()
    print("scheme %d: shared max_iter %d; the large instance stopped after %d steps (single run %d)"
          % (scheme, maxit, nn[1], nl))
    assert nn == [steps, nl] and abs(tt[0] - steps * dt) < 1e-15 and abs(tt[1] - nl * dt) < 1e-15
    batch.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_fused_step_graphs_bitwise(sem, mode):
    """The one-launch FP32 Lanczos step replayed as captured graphs gives the
    direct launches' bits and iteration counts (mixed-precision stage solves
    on a plain mesh, graph batches of 6 and 12 steps)."""
    xy, q = structured(9, (-2, 2, -2, 2), 0.1, 5)
    m = sem.Mesh.from_arrays(6, xy, q)
    V = m.field(lambda x, y: x * x + y * y)
    rng = np.random.default_rng(21)
    b = np.concatenate([sem.gproj(rng.standard_normal(m.size), m), sem.gproj(rng.standard_normal(m.size), m)])
    sem.set_step_mode(m, mode, sem.PREC_MIXED)
    sem.set_stored_basis(m, 0)  # graph replay runs the w/x recurrence; so do the direct launches here
    out = []
    for graphs, g in ((False, 12), (True, 12), (True, 6), (True, 9)):
        sem.set_graph_mode(m, graphs, g)
        z, r = sem.stage_solve(m, V, 2e-3, b, None, 5000, 1e-12, sem.PREC_MIXED)
        assert r.success
        out.append((z, r.n_iter))
    for z, n in out[1:]:
        assert np.array_equal(z, out[0][0]) and n == out[0][1]
    sem.set_graph_mode(m, False)
