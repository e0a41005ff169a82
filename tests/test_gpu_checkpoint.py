"""Checkpoint and resume of stepper sessions: a run split by save/restore
through a host array equals the one-shot run bit for bit, for BDF2 (whose
history, schrodinger.hpp:85-102, the checkpoint carries), for extrapolated and
learned stage guesses, and in mixed precision."""
import numpy as np
import pytest

from conftest import SMALL_MESH, structured

pytestmark = pytest.mark.gpu


def problem(sem, p=4):
    m = sem.load_mesh(p, SMALL_MESH)
    X, Y = m.coords()
    V = 1.0 + 0.5 * (X * X + Y * Y) + np.sin(2 * X) * Y
    env = np.exp(-((X - 0.2) ** 2 + Y ** 2) * 4)
    return m, V, env * np.cos(3 * X), env * np.sin(3 * X)


def run(sem, m, scheme, V, re, im, opts, splits, epochs=None):
    """advance through the step counts in `splits`, saving after each part and
    resuming in a fresh session created from zeros"""
    st = sem.Stepper(m, scheme, re, im, V, opts)
    if epochs:
        st.set_learning(epochs, 3e-3)
    for k, n in enumerate(splits):
        assert st.advance(1e30, n)
        if k + 1 < len(splits):
            ck = st.save()
            st.close()
            st = sem.Stepper(m, scheme, np.zeros(m.size), np.zeros(m.size), V, opts)
            if epochs:
                st.set_learning(epochs, 3e-3)
            st.restore(ck)
    a, b, t = st.read()
    s = st.stats()
    st.close()
    return a, b, t, s


@pytest.mark.parametrize("scheme,guess,precision", [(0, 0, 0), (0, 1, 1), (1, 1, 0), (2, 1, 1), (1, 2, 1), (2, 2, 0)])
def test_split_run_equals_one_shot(sem, scheme, guess, precision):
    m, V, re, im = problem(sem)
    opts = sem.SolverOpts(max_iter=20000, tol=1e-12, dt=2e-3, precision=precision, guess=guess, learn_steps=3)
    a1, b1, t1, _ = run(sem, m, scheme, V, re, im, opts, [7])
    a2, b2, t2, s2 = run(sem, m, scheme, V, re, im, opts, [2, 3, 2])
    assert t1 == t2
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    if guess == 2:
        assert s2["learned_guess_chosen"] > 0 or s2["mlp_train_steps"] == 0


def test_restore_into_other_options(sem):
    # a learned mixed-precision session's checkpoint resumes an FP64 session with
    # reference guesses: same psi and time as a session created from the state
    m, V, re, im = problem(sem)
    lo = sem.SolverOpts(max_iter=20000, tol=1e-12, dt=2e-3, precision=1, guess=2, learn_steps=2)
    st = sem.Stepper(m, 1, re, im, V, lo)
    assert st.advance(1e30, 4)
    ck = st.save()
    a0, b0, t0 = st.read()
    st.close()
    fo = sem.SolverOpts(max_iter=20000, tol=1e-12, dt=2e-3)
    s1 = sem.Stepper(m, 1, np.zeros(m.size), np.zeros(m.size), V, fo)
    s1.restore(ck)
    a, b, t = s1.read()
    assert t == t0 and np.array_equal(a, a0) and np.array_equal(b, b0)
    assert s1.advance(1e30, 2)
    a1, b1, _ = s1.read()
    s2 = sem.Stepper(m, 1, a0, b0, V, fo)
    assert s2.advance(1e30, 2)
    a2, b2, _ = s2.read()
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    # mismatched scheme or a truncated array is an error, not a silent resume
    s3 = sem.Stepper(m, 0, re, im, V, fo)
    with pytest.raises(sem.SemError):
        s3.restore(ck)
    with pytest.raises(sem.SemError):
        s1.restore(ck[:100])
    for s in (s1, s2, s3):
        s.close()


def test_bdf2_resume_continues_the_history(sem):
    # resuming BDF2 from psi alone restarts with implicit Euler (the reference's
    # only option); resuming from the checkpoint continues BDF2
    xy, q = structured(6, (-1, 1, -1, 1), 0.1, 8)
    m = sem.Mesh.from_arrays(5, xy, q)
    X, Y = m.coords()
    V = X * X + Y * Y
    re, im = np.exp(-4 * (X * X + Y * Y)), np.zeros(m.size)
    opts = sem.SolverOpts(max_iter=20000, tol=1e-12, dt=5e-3)
    a1, b1, _, _ = run(sem, m, 0, V, re, im, opts, [6])
    a2, b2, _, _ = run(sem, m, 0, V, re, im, opts, [3, 3])
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    st = sem.Stepper(m, 0, re, im, V, opts)
    assert st.advance(1e30, 3)
    ar, br, _ = st.read()
    st.close()
    a3, b3, _, _ = run(sem, m, 0, V, ar, br, opts, [3])  # history lost: implicit Euler restart
    assert not np.array_equal(a1, a3)
