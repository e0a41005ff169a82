"""Writes the golden fixtures of the full-size Schrodinger workloads from the
oracle (the CPU restatement of the reference's FP64 OpenMP path):

  config3_dirk4s3_step.npz  one DIRK4S3 step, 128 x 128 quads of [-10,10]^2, p = 8
  config5_dirk4s5_step.npz  one DIRK4S5 step, 256 x 256 quads of [-12,12]^2, p = 7

Inputs are those of schrodingersem_amd.configs (V_ref, conjugated packet,
dt_ref = 5e-4), MINRES tolerance 1e-13, zero stage guesses (the reference's
schrodinger_dirk4s3 / _dirk4s5, openmp_impl/schrodinger.hpp:138-285, 294-423).
The full wavefunction (2 x 2.7 M or 2 x 8.4 M doubles) is too large to commit,
so each fixture holds a size-independent summary of psi = [re | im] after the
step: 32 Gaussian projections (configs.sketch), every 4099th entry, the
norm, and the oracle's iteration counts.

Usage: python tests/golden/make_config_fixtures.py [config3] [config5]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import schrodingersem_amd as sem  # noqa: E402  (host-only mesh generator)
from schrodingersem_amd import configs  # noqa: E402
from oracle import oracle  # noqa: E402

STRIDE = 4099
TOL = 1e-13


def summarize(psi):
    return dict(sketch=configs.sketch(psi), samples=psi[::STRIDE].copy(), norm=np.linalg.norm(psi))


def one_step(name, gen, scheme, n, p):
    oracle.set_threads(os.cpu_count() or 1)
    t0 = time.time()
    o, xy, q, V, re, im = gen(sem, n, p, make=oracle.Mesh.from_arrays)
    dt = configs.DT_REF
    rec = 8
    a, b, ok, stats, iters = o.schrodinger(scheme, re, im, V, dt, 100000, TOL, dt, max_records=rec)
    if not ok:
        raise RuntimeError("%s: oracle step did not converge" % name)
    psi = np.concatenate([a, b])
    s = summarize(psi)
    out = os.path.join(HERE, name + ".npz")
    np.savez(out, n=n, p=p, scheme=scheme, dt=dt, tol=TOL, stride=STRIDE, stage_iters=iters[:int(stats[1])],
             psi0_norm=np.linalg.norm(np.concatenate([re, im])), **s)
    print("%s: %d stage solves, iterations %s, %.1f s -> %s" % (name, stats[1], iters[:int(stats[1])],
                                                               time.time() - t0, out), flush=True)


def main(argv):
    which = argv or ["config3", "config5"]
    if "config3" in which:
        one_step("config3_dirk4s3_step", configs.config3, 1, 128, 8)
    if "config5" in which:
        one_step("config5_dirk4s5_step", configs.config5, 2, 256, 7)


if __name__ == "__main__":
    main(sys.argv[1:])
