"""The stage-guess network's HIP kernels (MFMA bf16 GEMMs, GELU, MSE backward,
Adam) against the numpy restatement in oracle/mlp_oracle.py, and the learned
guess inside the SDIRK stepper against the extrapolated guess."""
import numpy as np
import pytest

from conftest import rel, structured

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(sem):
    return sem.MlpNet(7 * 25, 256, 256, 2 * 25, seed=3)  # the p = 4 network shape


def data(net, rows, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((rows, net.in_pad), np.float32)
    X[:, :net.dims[0]] = rng.standard_normal((rows, net.dims[0]))
    T = np.zeros((rows, net.out_pad), np.float32)
    T[:, :net.dims[3]] = rng.standard_normal((rows, net.dims[3]))
    return X, T


def randomize(net, seed):
    from oracle import mlp_oracle as mo
    rng = np.random.default_rng(seed)
    p = net.params()
    p = (p + 0.05 * rng.standard_normal(p.size)).astype(np.float32)
    # keep the padding rows/columns of the weights at zero
    W1, b1, W2, b2, W3, b3 = mo.unpack(p, net.dims, net.in_pad, net.out_pad, net.offsets)
    W1[net.dims[0]:, :] = 0
    W3[:, net.dims[3]:] = 0
    b3[net.dims[3]:] = 0
    net.set_params(p)
    return p


def test_forward_matches_numpy(net):
    from oracle import mlp_oracle as mo
    p = randomize(net, 1)
    X, _ = data(net, 1000, 2)
    Y = net.forward(X)
    Yo, _ = mo.forward(p, X, net.dims, net.in_pad, net.out_pad, net.offsets)
    assert rel(Y[:, :net.dims[3]], Yo[:, :net.dims[3]]) < 1e-3


def test_gradients_match_numpy(net):
    from oracle import mlp_oracle as mo
    p = randomize(net, 3)
    X, T = data(net, 4500, 4)  # several column-sum row blocks and four split-K slices (ragged last one)
    loss, g = net.gradients(X, T)
    lo, go = mo.gradients(p, X, T, net.dims, net.in_pad, net.out_pad, net.offsets)
    assert abs(loss - lo) < 1e-4 * lo
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    bounds = list(net.offsets) + [net.nparams]
    for k in range(6):
        a, b = g[bounds[k]:bounds[k + 1]], go[bounds[k]:bounds[k + 1]]
        assert rel(a, b) < 3e-3, (names[k], rel(a, b))
    # bit-reproducible
    loss2, g2 = net.gradients(X, T)
    assert loss2 == loss and np.array_equal(g, g2)


def test_adam_matches_numpy(net):
    from oracle import mlp_oracle as mo
    p = randomize(net, 5)
    X, T = data(net, 300, 6)
    m = np.zeros(net.nparams, np.float32)
    v = np.zeros(net.nparams, np.float32)
    q = p.copy()
    for t in range(1, 4):
        _, g = net.gradients(X, T)
        net.adam(1e-3)
        q, m, v = mo.adam(q, g, m, v, t, 1e-3)
        assert rel(net.params(), q) < 1e-6
        net.set_params(q)


def test_training_reduces_loss(sem):
    net = sem.MlpNet(40, 64, 64, 8, seed=7)
    rng = np.random.default_rng(8)
    X = np.zeros((4096, net.in_pad), np.float32)
    X[:, :40] = rng.standard_normal((4096, 40))
    M = rng.standard_normal((40, 8)) / np.sqrt(40)
    T = np.zeros((4096, net.out_pad), np.float32)
    T[:, :8] = np.tanh(X[:, :40] @ M)
    l0, _ = net.gradients(X, T)
    for _ in range(300):
        net.gradients(X, T)
        net.adam(3e-3)
    l1, _ = net.gradients(X, T)
    assert l1 < 0.05 * l0, (l0, l1)


def test_learned_guess_in_stepper(sem):
    # 12 x 12 mesh of [-3,3]^2 at p = 4, DIRK4S3: train on 8 steps, then compare
    # the frozen network with the extrapolated guess over the next 8 steps
    xy, q = structured(12, (-3, 3, -3, 3), 0.0, 0)
    m = sem.Mesh.from_arrays(4, xy, q)
    X, Y = m.coords()
    V = 2 * (0.5 * (X * X + Y * Y))
    env = np.exp(-(X * X + Y * Y) / 2)
    re, im = env * np.cos(2 * X + Y), -env * np.sin(2 * X + Y)
    res = {}
    for name, guess in (("ext", sem.GUESS_EXTRAPOLATED), ("learned", sem.GUESS_LEARNED)):
        opts = sem.SolverOpts(max_iter=20000, tol=1e-12, dt=5e-3, guess=guess, learn_steps=8)
        st = sem.Stepper(m, 1, re, im, V, opts)
        assert st.advance(1e30, 8)
        s0 = st.stats()
        assert st.advance(1e30, 8)
        s1 = st.stats()
        a, b, _ = st.read()
        res[name] = (s1["total_iters"] - s0["total_iters"], np.concatenate([a, b]), s1)
    opts = sem.SolverOpts(max_iter=20000, tol=1e-14, dt=5e-3)
    st = sem.Stepper(m, 1, re, im, V, opts)
    assert st.advance(1e30, 16)
    a, b, _ = st.read()
    truth = np.concatenate([a, b])
    for name in res:
        assert rel(res[name][1], truth) < 1e-9, name
    assert res["learned"][2]["mlp_train_steps"] > 0
    # every frozen stage solve starts from the minimal-residual point between
    # the extrapolation and the network's guess (never a larger residual than
    # the extrapolation's), and the run takes no more iterations than the
    # fallback; the gain at the workload's size is in test_gpu_learned_gain.py
    assert res["learned"][2]["learned_guess_chosen"] > 0
    assert res["learned"][2]["learned_residual_digits"] > 0
    assert res["ext"][2]["learned_guess_chosen"] == 0
    print("iterations over the 8 frozen steps: extrapolated %d, learned %d" % (res["ext"][0], res["learned"][0]))
    assert res["learned"][0] <= res["ext"][0]
