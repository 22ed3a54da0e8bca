"""Pins of the learned-prior oracle (CPU): known-answer SplitMix64 outputs,
bf16 rounding against torch's conversion, gradients against central finite
differences of the float64 network, Adam against its closed-form first
step and torch.optim.Adam, features against hand-computed values."""
import math

import numpy as np
import pytest

from oracle import prior as P


def test_splitmix64_known_answers():
    # SplitMix64 seeded with 0 (Steele, Lea & Flood 2014 reference stream)
    s = 0
    out = []
    for _ in range(3):
        s, z = P.splitmix64(s)
        out.append(z)
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_bf16_round_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 0.0,
                     3.0e-39, 65504.0], np.float32)
    r = P.bf16(vals)
    assert r[0] == one
    assert r[1] == one                     # tie -> even (1.0)
    assert r[2] == one + 2 * ulp           # tie -> even mantissa
    assert r[3] == one + ulp               # above the tie
    assert r[4] == np.float32(-2.5)


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-3, 5000),
                        rng.normal(0, 1e4, 5000)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(P.bf16(x), ref)


def _rand_net(seed):
    rng = np.random.default_rng(seed)
    p = P.init_params(seed).astype(np.float64)
    p[P.OFF_B1:P.OFF_W2] = rng.normal(0, 0.1, P.H)
    p[P.OFF_B2:P.OFF_W3] = rng.normal(0, 0.1, P.H)
    p[P.OFF_B3] = 0.3
    return p, rng


@pytest.mark.parametrize("pos_weight", [1.0, 4.0])
def test_gradient_matches_finite_differences(pos_weight):
    p, rng = _rand_net(11)
    E = 48
    f = rng.normal(0, 1, (E, P.NF))
    y = (rng.random(E) < 0.3).astype(np.float64)
    g, loss = P.backward(p, f, y, pos_weight, bf16_round=False)
    assert math.isclose(loss, P.loss_of(p, f, y, pos_weight), rel_tol=1e-12)
    h = 1e-6
    fd = np.empty(P.NPARAM)
    for i in range(P.NPARAM):
        pp = p.copy()
        pp[i] += h
        lp = P.loss_of(pp, f, y, pos_weight)
        pp[i] -= 2 * h
        lm = P.loss_of(pp, f, y, pos_weight)
        fd[i] = (lp - lm) / (2 * h)
    assert np.abs(fd - g).max() < 1e-8 * max(1.0, np.abs(g).max() * 100)
    # every block of the gradient carries signal (a dropped term shows up)
    for lo, hi in ((P.OFF_W1, P.OFF_B1), (P.OFF_B1, P.OFF_W2), (P.OFF_W2, P.OFF_B2),
                   (P.OFF_B2, P.OFF_W3), (P.OFF_W3, P.OFF_B3), (P.OFF_B3, P.NPARAM)):
        assert np.abs(g[lo:hi]).max() > 1e-6


def test_loss_closed_form_at_zero_logits():
    # all weights zero, b3 = 0: z = 0, loss = mean_e w_e log 2, dL/db3 = mean w_e (1/2 - y)
    p = np.zeros(P.NPARAM)
    y = np.array([1, 0, 0, 1, 0], np.float64)
    f = np.ones((5, P.NF))
    g, loss = P.backward(p, f, y, 3.0, bf16_round=False)
    w = np.where(y > 0.5, 3.0, 1.0)
    assert math.isclose(loss, float(np.mean(w) * math.log(2.0)), rel_tol=1e-12)
    assert math.isclose(g[P.OFF_B3], float(np.mean(w * (0.5 - y))), rel_tol=1e-12)
    assert np.all(g[:P.OFF_B3] == 0)


def test_bf16_forward_rounds_inputs_and_weights():
    # a network that is the identity on feature 0 through single ReLU paths:
    # the bf16 path must equal the float64 path on the bf16-rounded feature
    p = np.zeros(P.NPARAM, np.float32)
    p[P.OFF_W1 + 0 * P.NF + 0] = 1.0           # h1[0] = relu(f0)
    p[P.OFF_W2 + 0 * P.H + 0] = 1.0            # h2[0] = relu(h1[0])
    p[P.OFF_W3 + 0] = 1.0                      # z = h2[0]
    f = np.zeros((3, P.NF), np.float32)
    f[:, 0] = [1 + 2 ** -8 + 2 ** -12, 0.3, -2.0]
    z, _ = P.forward(p, f, bf16_round=True)
    np.testing.assert_array_equal(z, np.maximum(P.bf16(f[:, 0]), 0))


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(5)
    p0 = rng.normal(0, 1, 200).astype(np.float32)
    g = rng.normal(0, 1e-2, 200).astype(np.float32)
    p1, m, v = P.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, 1e-3)
    expect = p0.astype(np.float64) - 1e-3 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p1, expect, rtol=0, atol=2e-7)


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    p0 = rng.normal(0, 1, 300).astype(np.float32)
    t = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([t], lr=2e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for step in range(1, 6):
        g = rng.normal(0, 0.05, 300).astype(np.float32)
        t.grad = torch.from_numpy(g.copy())
        opt.step()
        p, m, v = P.adam(p, m, v, g, step, 2e-3)
    np.testing.assert_allclose(p, t.detach().numpy(), rtol=0, atol=3e-6)


def test_init_params_glorot():
    a = P.init_params(7)
    b = P.init_params(7)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, P.init_params(8))
    W1, b1, W2, b2, w3, b3 = P.unpack(a)
    for W, fi, fo in ((W1, 12, 64), (W2, 64, 64), (w3, 64, 1)):
        bound = math.sqrt(6.0 / (fi + fo))
        assert np.abs(W).max() <= bound
        assert np.abs(W).max() > 0.9 * bound
        # uniform(-b, b): mean has standard deviation b / sqrt(3 * count)
        assert abs(float(W.mean())) < 4 * bound / math.sqrt(3 * W.size)
    assert not b1.any() and not b2.any() and b3 == 0


def test_features_hand_computed():
    # depot (0,0); nodes at (3,4), (6,8), (-4,0), (0,-2); Euclidean float distances
    x = np.array([0, 3, 6, -4, 0], np.float32)
    y = np.array([0, 4, 8, 0, -2], np.float32)
    n = 5
    d = np.sqrt((x[:, None] - x[None]) ** 2 + (y[:, None] - y[None]) ** 2).astype(np.float32)
    dem = np.array([0, 10, 20, 30, 40], np.int32)
    k = 2
    # candidate lists: nearest non-self nodes (excluding the depot as the reference does)
    nn = np.array([[4, 3], [2, 4], [1, 4], [4, 1], [3, 1]], np.int32)
    f = P.features(x, y, dem, 100, d, nn, k)
    assert f.shape == (n * k, 12)
    kd = np.array([d[0, 3], d[1, 4], d[2, 4], d[3, 1], d[4, 1]], np.float32)
    mkd = np.float32(kd.astype(np.float64).sum() / n)
    R = np.float32(10.0)
    e = 1 * k + 0          # edge (1, 2), rank 0
    np.testing.assert_allclose(f[e, 0], 5.0 / mkd, rtol=1e-7)
    assert f[e, 1] == 0.0 and f[e + 1, 1] == 0.5
    np.testing.assert_allclose(f[e, 2:4], [0.1, 0.2], rtol=1e-7)
    np.testing.assert_allclose(f[e, 4:6], [math.atan2(4, 3) / math.pi] * 2, rtol=1e-6)
    np.testing.assert_allclose(f[e, 6:8], [0.5, 1.0], rtol=1e-7)
    np.testing.assert_allclose(f[e, 8:10], [kd[1] / mkd, kd[2] / mkd], rtol=1e-6)
    np.testing.assert_allclose(f[e, 10], (5.0 + 10.0 - 5.0) / 10.0, rtol=1e-7)
    assert f[e, 11] == 1.0                 # 1 is in nn[2]
    e3 = 3 * k + 1         # edge (3, 1): 3 is not in nn[1] = {2, 4}
    assert f[e3, 11] == 0.0
    np.testing.assert_allclose(f[3 * k, 4], 1.0, rtol=1e-7)     # angle of (-4,0) is pi
    np.testing.assert_allclose(f[4 * k, 4], -0.5, rtol=1e-7)    # angle of (0,-2) is -pi/2


def test_edge_labels():
    nn = np.array([[1, 2], [2, 3], [1, 3], [2, 1]])
    tour = [0, 1, 2, 0, 3, 0]
    lab = P.edge_labels(nn, 2, tour).reshape(4, 2)
    np.testing.assert_array_equal(lab, [[1, 1], [1, 0], [1, 0], [0, 0]])
