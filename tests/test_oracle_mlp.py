"""Pins the numpy restatement of the guess network (oracle/mlp_oracle.py)
against PyTorch's autograd and Adam on the CPU. The restatement rounds GEMM
operands to bfloat16 as the GPU does, so it is compared with an unrounded
float64 autograd at a tolerance (a few percent) that a wrong transpose, a
missing GELU derivative or a wrong loss scale would exceed by far."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import mlp_oracle as mo


def setup(rows=48, dims=(20, 32, 24, 10), in_pad=32, out_pad=16, seed=0):
    rng = np.random.default_rng(seed)
    n_in, h1, h2, n_out = dims
    sizes = [in_pad * h1, h1, h1 * h2, h2, h2 * out_pad, out_pad]
    off = [0]
    for s in sizes:
        off.append(off[-1] + s)
    p = (0.3 * rng.standard_normal(off[-1])).astype(np.float32)
    W1 = p[off[0]:off[1]].reshape(in_pad, h1)
    W1[n_in:] = 0.0  # padding rows carry nothing
    X = np.zeros((rows, in_pad), np.float32)
    X[:, :n_in] = rng.standard_normal((rows, n_in))
    T = np.zeros((rows, out_pad), np.float32)
    T[:, :n_out] = rng.standard_normal((rows, n_out))
    return p, X, T, dims, in_pad, out_pad, off[:6]


def torch_loss(p, X, T, dims, in_pad, out_pad, off):
    n_in, h1, h2, n_out = dims
    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    W1 = pt[off[0]:off[1]].reshape(in_pad, h1)
    b1 = pt[off[1]:off[2]]
    W2 = pt[off[2]:off[3]].reshape(h1, h2)
    b2 = pt[off[3]:off[4]]
    W3 = pt[off[4]:off[5]].reshape(h2, out_pad)
    b3 = pt[off[5]:off[5] + out_pad]
    x = torch.tensor(X.astype(np.float64))
    a1 = torch.nn.functional.gelu(x @ W1 + b1)
    a2 = torch.nn.functional.gelu(a1 @ W2 + b2)
    y = a2 @ W3 + b3
    t = torch.tensor(T.astype(np.float64))
    loss = ((y[:, :n_out] - t[:, :n_out]) ** 2).mean()
    loss.backward()
    return float(loss.detach()), pt.grad.detach().numpy(), y.detach().numpy()


def test_forward_and_gradients_match_autograd():
    p, X, T, dims, in_pad, out_pad, off = setup()
    Y, _ = mo.forward(p, X, dims, in_pad, out_pad, off)
    loss, g = mo.gradients(p, X, T, dims, in_pad, out_pad, off)
    tl, tg, ty = torch_loss(p, X, T, dims, in_pad, out_pad, off)
    n_out = dims[3]
    assert np.linalg.norm(Y[:, :n_out] - ty[:, :n_out]) < 0.02 * np.linalg.norm(ty[:, :n_out])
    assert abs(loss - tl) < 0.02 * tl
    ends = list(off) + [off[5] + out_pad]
    for k in range(6):
        a, b = g[ends[k]:ends[k + 1]], tg[ends[k]:ends[k + 1]]
        assert np.linalg.norm(a - b) < 0.05 * np.linalg.norm(b) + 1e-12, k


def test_gradient_checks_catch_a_planted_mistake():
    # a wrong loss scale (mean over rows only) must fail the same tolerance
    p, X, T, dims, in_pad, out_pad, off = setup(seed=3)
    _, g = mo.gradients(p, X, T, dims, in_pad, out_pad, off)
    _, tg, _ = torch_loss(p, X, T, dims, in_pad, out_pad, off)
    wrong = g * dims[3]
    assert np.linalg.norm(wrong - tg) > 0.5 * np.linalg.norm(tg)


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    p = rng.standard_normal(50).astype(np.float32)
    m = np.zeros(50, np.float32)
    v = np.zeros(50, np.float32)
    pt = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 6):
        g = rng.standard_normal(50).astype(np.float32)
        p, m, v = mo.adam(p, g, m, v, t, 3e-3)
        pt.grad = torch.tensor(g)
        opt.step()
    assert np.allclose(p, pt.detach().numpy(), rtol=1e-5, atol=1e-6)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7 + 2**-8, -3.140625, 65504.0], np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie goes to even (mantissa ...0)
    assert r[2] == 1.0 + 2**-6              # tie goes to even (mantissa ...10)
    assert r[3] == -3.140625
    assert r[4] == 65536.0
