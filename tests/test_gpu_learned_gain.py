"""What the trained guess network gains over the method without it.

The learned-guess DIRK4S5 workload (256 x 256 quads of [-12, 12]^2, degree 7,
harmonic well plus 64 bumps, dt = 1e-3 mapped to the reference's convention)
in mixed precision at MINRES tolerance 1e-12 (the reference's stop test
||r|| < tol ||b|| + tol, on the true FP64 residual): a learned session trains
its network on the stage solutions of the first 50 steps and is checkpointed;
from that one checkpoint the extrapolated-guess fallback and the frozen
network each run the same 10 steps. The network's run must take fewer inner
MINRES iterations and less wall time (network inference and the residual
guard included), and both runs must agree to 1e-9 relative L2.
"""
import time

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _run(sem, m, V, ck, guess, steps):
    from schrodingersem_amd import configs
    opts = sem.SolverOpts(max_iter=100000, tol=TOL, dt=configs.DT_REF, precision=sem.PREC_MIXED, guess=guess,
                          learn_steps=50)
    st = sem.Stepper(m, 2, np.zeros(m.size), np.zeros(m.size), V, opts)
    st.restore(ck)
    assert st.advance(1e30, 1)  # first step after the restore: buffers and graphs set up
    s0 = st.stats()
    t0 = time.perf_counter()
    assert st.advance(1e30, steps)
    sec = time.perf_counter() - t0
    s1 = st.stats()
    a, b, _ = st.read()
    st.close()
    d = {k: s1[k] - s0[k] for k in ("stage_solves", "total_iters", "learned_guess_chosen",
                                    "learned_residual_digits")}
    d["seconds"] = sec
    return d, np.concatenate([a, b])


def test_learned_guess_gain_dirk4s5_256x256_p7(sem):
    from schrodingersem_amd import configs
    m, xy, q, V, re, im = configs.config5(sem, 256, 7)
    opts = sem.SolverOpts(max_iter=100000, tol=TOL, dt=configs.DT_REF, precision=sem.PREC_MIXED,
                          guess=sem.GUESS_LEARNED, learn_steps=50)
    tr = sem.Stepper(m, 2, re, im, V, opts)
    t0 = time.perf_counter()
    assert tr.advance(1e30, 50)
    train_s = time.perf_counter() - t0
    ts = tr.stats()
    assert ts["mlp_train_steps"] > 0 and ts["learned_guess_chosen"] == 0
    ck = tr.save()
    tr.close()
    steps = 10
    ext, psi_e = _run(sem, m, V, ck, sem.GUESS_EXTRAPOLATED, steps)
    lrn, psi_l = _run(sem, m, V, ck, sem.GUESS_LEARNED, steps)
    print("training window %.2f s; %d frozen steps: extrapolated %d iterations in %.3f s, learned %d iterations "
          "in %.3f s (network used %d of %d stage solves, %.2f digits of initial residual saved per solve)"
          % (train_s, steps, ext["total_iters"], ext["seconds"], lrn["total_iters"], lrn["seconds"],
             lrn["learned_guess_chosen"], lrn["stage_solves"],
             lrn["learned_residual_digits"] / max(1, lrn["learned_guess_chosen"])))
    assert ext["stage_solves"] == lrn["stage_solves"] == 5 * steps
    assert lrn["learned_guess_chosen"] > 0
    # the start is the minimal-residual point between the two guesses: never worse
    assert lrn["learned_residual_digits"] > 0
    assert lrn["total_iters"] < 0.9 * ext["total_iters"]
    assert lrn["seconds"] < ext["seconds"]
    assert rel(psi_l, psi_e) < 1e-9
