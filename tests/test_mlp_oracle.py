"""Pins of the learned scheduler's CPU oracle (oracle/mlp_ref.py).

The network has no KPP counterpart; its oracle is pinned by closed forms,
an independent BF16 rounding (torch's), and central finite differences of
its own float64 loss, chosen so that a transposed weight, a dropped bias
term or a wrong GELU derivative fails.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_ref as R


def test_bf16_rounding_matches_torch_and_ties_to_even():
    assert R.bf16(1.0) == 1.0
    assert R.bf16(1.0 + 2.0 ** -8) == 1.0                      # tie -> even (down)
    assert R.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6      # tie -> even (up)
    assert R.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    x = np.random.default_rng(0).standard_normal(100000) * 10.0 ** np.random.default_rng(1).uniform(-8, 8, 100000)
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).double().numpy()
    assert np.array_equal(R.bf16(x), ref)


def test_gelu_closed_forms():
    assert R.gelu(0.0) == 0.0
    assert abs(R.gelu(10.0) - 10.0) < 1e-12
    assert abs(R.gelu(-10.0)) < 1e-12
    assert R.gelu_grad(0.0) == 0.5
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (R.gelu(z + h) - R.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - R.gelu_grad(z))) < 1e-8
    # the tanh form is within 1e-3 of the exact x*Phi(x)
    from scipy.special import erf
    exact = 0.5 * z * (1 + erf(z / np.sqrt(2)))
    assert np.max(np.abs(exact - R.gelu(z))) < 1e-3


def test_layout_counts():
    w_off, b_off, total = R.layout(96)
    assert total == 128 * 96 + 128 + 2 * (128 * 128 + 128) + 16 * 128 + 16
    assert w_off[0] == 0 and b_off[0] == 128 * 96 and w_off[1] == b_off[0] + 128
    layers = [(np.arange(128 * 96.0).reshape(128, 96), np.arange(128.0)),
              (np.ones((128, 128)), np.zeros(128)), (np.ones((128, 128)), np.zeros(128)),
              (np.full((16, 128), 2.0), np.full(16, 3.0))]
    back = R.unpack(R.pack(layers, 96), 96)
    for (W, b), (W2, b2) in zip(layers, back):
        assert np.array_equal(W, W2) and np.array_equal(b, b2)


def _random_net(in_pad, nin, seed):
    rng = np.random.default_rng(seed)
    layers = []
    for no, ki, used in [(128, in_pad, nin), (128, 128, 128), (128, 128, 128), (16, 128, 128)]:
        W = np.zeros((no, ki))
        W[:, :used] = rng.uniform(-1, 1, (no, used)) * np.sqrt(3.0 / used)
        layers.append((W, rng.uniform(-0.2, 0.2, no)))
    return R.pack(layers, in_pad)


def test_loss_closed_form():
    O = np.zeros((3, 16))
    O[:, 0] = [1.0, 2.0, 3.0]
    y = np.array([1.0, 0.0, 1.0])
    L = R.loss(O, y, np.array([0, 3, 7]), np.array([1.0, 1.0, 0.0]))
    assert abs(L - (0 + 4 + 0 + 3 * np.log(8))) < 1e-12


def test_gradient_matches_finite_differences():
    in_pad, nin = 32, 11
    flat = _random_net(in_pad, nin, 3)
    rng = np.random.default_rng(4)
    n = 12
    X = np.zeros((n, in_pad))
    X[:, :nin] = rng.standard_normal((n, nin))
    y = rng.standard_normal(n)
    bk = rng.integers(0, 8, n)
    wr = (rng.uniform(size=n) > 0.2).astype(float)
    L0, g = R.grads(flat, X, y, bk, wr, in_pad, emulate=False)

    def mean_loss(f):
        O, _, _ = R.forward(f, X, in_pad, emulate=False)
        return R.loss(O, y, bk, wr) / n

    assert abs(L0 / n - mean_loss(flat)) < 1e-12
    w_off, b_off, total = R.layout(in_pad)
    idx = list(rng.choice(np.arange(w_off[0], w_off[0] + 128 * in_pad).reshape(128, in_pad)[:, :nin].ravel(), 20))
    for l in range(1, 4):
        idx += list(rng.integers(w_off[l], b_off[l], 15))
    for l in range(4):
        idx += list(rng.integers(b_off[l], b_off[l] + (128 if l < 3 else 9), 5))
    h = 1e-6
    for i in idx:
        e = np.zeros(total)
        e[i] = h
        fd = (mean_loss(flat + e) - mean_loss(flat - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padded inputs and outputs get no gradient
    Wg = R.unpack(g, in_pad)
    assert np.all(Wg[0][0][:, nin:] == 0)
    assert np.all(Wg[3][0][9:] == 0) and np.all(Wg[3][1][9:] == 0)


def test_emulated_gradient_close_to_exact():
    """BF16 rounding of stored values changes gradients only at the 1e-2 level."""
    in_pad, nin = 96, 81
    flat = _random_net(in_pad, nin, 5)
    rng = np.random.default_rng(6)
    n = 256
    X = np.zeros((n, in_pad))
    X[:, :nin] = R.bf16(rng.standard_normal((n, nin)))
    y = rng.standard_normal(n)
    bk = rng.integers(0, 8, n)
    wr = np.ones(n)
    _, g0 = R.grads(flat, X, y, bk, wr, in_pad, emulate=False)
    _, g1 = R.grads(flat, X, y, bk, wr, in_pad, emulate=True)
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)


def test_adam_first_step_is_signed_lr():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    w1, m1, m2 = R.adam(w, g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(w1, w - 0.01 * np.sign(g), rtol=0, atol=1e-7)


def test_bucket_centres():
    from kpp_amd import learned
    rng = np.random.default_rng(4)
    nstp = np.concatenate([rng.integers(5, 20, 5000), rng.integers(30, 90, 3000)])
    edges = np.quantile(nstp, np.linspace(0, 1, 9)[1:-1]) + 0.5
    c = learned.bucket_centres(nstp, edges)
    b = np.searchsorted(edges, nstp, side="right")
    assert c.dtype == np.float32 and len(c) == 8
    assert np.all(np.diff(c) > 0)
    for k in range(8):
        if (b == k).any():
            assert abs(c[k] - nstp[b == k].mean()) < 1e-3 or k > 0 and c[k] > c[k - 1]
    # an empty bucket sits between its edges
    c2 = learned.bucket_centres(np.array([1, 2, 3, 50, 60]), np.array([1.5, 2.5, 3.5, 10.5, 20.5, 30.5, 55.5]))
    assert np.all(np.diff(c2) > 0) and 10.5 <= c2[4] <= 20.5
