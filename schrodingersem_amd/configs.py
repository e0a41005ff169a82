"""Synthetic inputs of the five benchmark workloads, shared by bench.py and the
tests. Every generator is seeded; the mesh comes from the library's own
SplitMix64 structured-mesh generator, the fields from numpy's PCG64.

The reference integrates psi_t = +i(-Lap + V) psi (openmp_impl/schrodinger.hpp:
175-176, 199-200). The Schrodinger workloads are stated for
i psi_t = (-Lap/2 + V) psi, which maps onto the reference's equation exactly
with V_ref = 2V, dt_ref = dt/2 and psi conjugated on input and output; the
generators below return V_ref and the conjugated psi, and DT_REF.
"""
import numpy as np

DT = 1e-3          # physical time step of the Schrodinger workloads
DT_REF = DT / 2.0  # the same step in the reference's convention


def _bumps(X, Y, rng, count, box, heights=(-5.0, 5.0), widths=(0.3, 1.5)):
    h = rng.uniform(heights[0], heights[1], count)
    w = rng.uniform(widths[0], widths[1], count)
    c = rng.uniform(-box, box, (count, 2))
    V = np.zeros_like(X)
    for k in range(count):
        V = V + h[k] * np.exp(-((X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2) / (2 * w[k] ** 2))
    return V


def _packet(X, Y):
    """Gaussian wavepacket of width 1 at the origin with momentum (2, 1), conjugated."""
    env = np.exp(-(X * X + Y * Y) / 2.0)
    return env * np.cos(2 * X + Y), -env * np.sin(2 * X + Y)


def _mesh(sem, p, xy, q, make):
    return make(p, xy, q) if make is not None else sem.Mesh.from_arrays(p, xy, q)


def config1(sem, n=8, p=4, make=None):
    """Poisson: 8 x 8 quads of [0,1]^2 jittered by U(-0.1h, 0.1h) (seed 0), degree 4,
    u = sin(pi x) sin(pi y), f = 2 pi^2 u."""
    xy, q = sem.structured_mesh(n, (0.0, 1.0, 0.0, 1.0), 0.1, 0)
    m = _mesh(sem, p, xy, q, make)
    X, Y = m.coords()
    ex = np.sin(np.pi * X) * np.sin(np.pi * Y)
    return m, xy, q, ex, 2 * np.pi ** 2 * ex


def gaussian_forcing(X, Y, seed, count=32):
    """sum of `count` Gaussians, centres U([0,1]^2), widths U(0.02, 0.1) (the
    standard deviation of exp(-r^2 / (2 s^2))), amplitudes N(0, 1)."""
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0, 1, (count, 2))
    wid = rng.uniform(0.02, 0.1, count)
    amp = rng.standard_normal(count)
    F = np.zeros_like(X)
    for a, c, s in zip(amp, cen, wid):
        F = F + a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * s * s))
    return F


def config2(sem, n=64, p=8, make=None):
    """Helmholtz -Lap u + k^2 u = f, k^2 = 10: n x n quads of [0,1]^2 jittered by
    U(-0.15h, 0.15h) (seed 1), degree 8, f = 32 random Gaussians (seed 1)."""
    xy, q = sem.structured_mesh(n, (0.0, 1.0, 0.0, 1.0), 0.15, 1)
    m = _mesh(sem, p, xy, q, make)
    X, Y = m.coords()
    return m, xy, q, np.full(m.size, 10.0), gaussian_forcing(X, Y, 1)


def config3(sem, n=128, p=8, make=None):
    """DIRK4S3 Schrodinger: n x n quads of [-10,10]^2, degree 8, V = 0.5|x|^2 plus
    16 Gaussian bumps (seed 2). Returns (mesh, xy, quads, V_ref, re, im)."""
    xy, q = sem.structured_mesh(n, (-10.0, 10.0, -10.0, 10.0), 0.0, 0)
    m = _mesh(sem, p, xy, q, make)
    X, Y = m.coords()
    V = 0.5 * (X * X + Y * Y) + _bumps(X, Y, np.random.default_rng(2), 16, 10.0)
    re, im = _packet(X, Y)
    return m, xy, q, 2.0 * V, re, im


def config4(sem, n_inst=64, box=8.0, make=None, only=None):
    """64 independent BDF2 instances, seeds 3+i: n x n quads of [-box, box]^2 with n
    uniform in [16, 96], degree uniform in {4, 6, 8, 10, 12}, jitter 0.1h,
    V = 0.5|x|^2 plus 1-8 Gaussian bumps, a Gaussian wavepacket with random
    centre, width and momentum. Returns a list of dicts (index, mesh, xy, quads, V, re, im, n, p);
    `only` (a set of indices) builds just those instances, with the same inputs."""
    inst = []
    for i in range(n_inst):
        if only is not None and i not in only:
            continue
        seed = 3 + i
        rng = np.random.default_rng(seed)
        n = int(rng.integers(16, 97))
        p = int(rng.choice([4, 6, 8, 10, 12]))
        xy, q = sem.structured_mesh(n, (-box, box, -box, box), 0.1, seed)
        m = _mesh(sem, p, xy, q, make)
        X, Y = m.coords()
        V = 0.5 * (X * X + Y * Y)
        for _ in range(int(rng.integers(1, 9))):
            c = rng.uniform(-box / 2, box / 2, 2)
            h, w = rng.uniform(-5, 5), rng.uniform(0.3, 1.5)
            V = V + h * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * w * w))
        c = rng.uniform(-2, 2, 2)
        k = rng.uniform(-2, 2, 2)
        w = rng.uniform(0.5, 1.5)
        env = np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * w * w))
        inst.append(dict(index=i, mesh=m, xy=xy, quads=q, V=V, re=env * np.cos(k[0] * X + k[1] * Y),
                         im=env * np.sin(k[0] * X + k[1] * Y), n=n, p=p))
    return inst


def config4_params(n_inst=64):
    """(n, p) of every config-4 instance without building meshes."""
    out = []
    for i in range(n_inst):
        rng = np.random.default_rng(3 + i)
        n = int(rng.integers(16, 97))
        out.append((n, int(rng.choice([4, 6, 8, 10, 12]))))
    return out


def config5(sem, n=256, p=7, make=None):
    """DIRK4S5 with the learned guess: n x n quads of [-12,12]^2, degree 7,
    V = 0.5|x|^2 plus 64 Gaussian bumps (seed 4). Returns (mesh, xy, quads, V_ref, re, im)."""
    xy, q = sem.structured_mesh(n, (-12.0, 12.0, -12.0, 12.0), 0.0, 0)
    m = _mesh(sem, p, xy, q, make)
    X, Y = m.coords()
    V = 0.5 * (X * X + Y * Y) + _bumps(X, Y, np.random.default_rng(4), 64, 12.0)
    re, im = _packet(X, Y)
    return m, xy, q, 2.0 * V, re, im


def sketch(v, k=32, seed=20261204):
    """k Gaussian projections of v (a size-independent summary of a long
    vector): ||S e||^2 / k estimates ||e||^2 for e = v - w with the same S."""
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    rng = np.random.default_rng(seed)
    out = np.empty(k)
    for r in range(k):
        out[r] = rng.standard_normal(v.size) @ v
    return out
