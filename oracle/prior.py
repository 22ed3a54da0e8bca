"""CPU oracle of the learned edge prior (test infrastructure only).

The reference has no learned component; this module restates the build's
own definition so the HIP kernels can be checked against it:

* ``features``: 12 features per candidate edge (i, nn[i][r]), r < k.
* ``forward`` / ``backward``: the 12 -> 64 -> 64 -> 1 ReLU MLP with the
  kernels' BF16 roundings (weights, input features, the first hidden layer
  and the back-propagated pre-activations rounded to bfloat16 with
  round-to-nearest-even; products accumulated exactly and rounded once to
  fp32 here, in fp32 MFMA accumulators on the device).
* ``adam``: Adam (Kingma & Ba, 2015) in fp32.
* ``init_params``: Glorot-uniform weights from a SplitMix64 stream.

Pins: the gradient formulas are checked against central finite
differences of the unrounded float64 network; rounding can be switched off
(``bf16=False``) so the same code is the float64 network; Adam's first
step has the closed form w - lr * g / (|g| + eps); features have closed
forms on hand-placed points.  Only tests, the smoke check and the bench's
CPU leg import this module.
"""
from __future__ import annotations

import numpy as np

NF = 12          # features per edge
NFP = 16         # padded to one MFMA K tile
H = 64           # hidden width
# parameter vector layout (fp32): W1[H][NF], b1[H], W2[H][H], b2[H], w3[H], b3
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * NF
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1

_M64 = (1 << 64) - 1


def splitmix64(state: int):
    """One SplitMix64 step (Steele, Lea & Flood, OOPSLA 2014): returns
    (new_state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform W1, W2, w3 (bound sqrt(6/(fan_in+fan_out))), zero
    biases.  Draw t of the stream (t = 0, 1, ... over W1, W2, w3 in layout
    order) is u = (out >> 40) * 2^-24 in [0,1), weight = (2u - 1) * bound,
    all in fp32."""
    p = np.zeros(NPARAM, np.float32)
    s = seed & _M64
    for off, cnt, fi, fo in ((OFF_W1, H * NF, NF, H), (OFF_W2, H * H, H, H), (OFF_W3, H, H, 1)):
        bound = np.float32(np.sqrt(np.float32(6.0) / np.float32(fi + fo), dtype=np.float32))
        for t in range(cnt):
            s, z = splitmix64(s)
            u = np.float32(z >> 40) * np.float32(2.0 ** -24)
            p[off + t] = (np.float32(2.0) * u - np.float32(1.0)) * bound
    return p


def unpack(p: np.ndarray):
    """Views of the parameter vector (keeps its dtype: float64 for the
    finite-difference network)."""
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(H, NF), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def bf16(x) -> np.ndarray:
    """Round fp32 to bfloat16, nearest-even, returned as fp32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


# --------------------------------------------------------------- features --

def node_stats(x, y, dist, nn, k):
    """Per node: radius and polar angle / pi around the depot (node 0), the
    distance to its k-th candidate; per instance: the largest radius R and
    the mean k-th-candidate distance."""
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    dx = x - x[0]
    dy = y - y[0]
    rad = np.sqrt(dx * dx + dy * dy).astype(np.float32)
    ang = (np.arctan2(dy, dx).astype(np.float32) / np.float32(np.pi)).astype(np.float32)
    n = len(x)
    kd = dist[np.arange(n), nn[:, k - 1]].astype(np.float32)
    R = np.float32(rad.max()) if rad.max() > 0 else np.float32(1.0)
    mkd = np.float32(np.float64(kd).sum() / n)
    if mkd <= 0:
        mkd = np.float32(1.0)
    return rad, ang, kd, R, mkd


def features(x, y, demand, capacity, dist, nn, k) -> np.ndarray:
    """[n*k, 12] float32 features of the edges (i, nn[i][r]), i-major.

    0 d(i,j)/mkd, 1 r/k, 2 dem_i/Q, 3 dem_j/Q, 4 angle_i, 5 angle_j,
    6 rad_i/R, 7 rad_j/R, 8 kd_i/mkd, 9 kd_j/mkd,
    10 savings (d(0,i) + d(0,j) - d(i,j))/R, 11 mutual (i in nn[j][:k])."""
    dist = np.asarray(dist, np.float32)
    nn = np.asarray(nn, np.int64)[:, :k]
    n = dist.shape[0]
    rad, ang, kd, R, mkd = node_stats(x, y, dist, nn, k)
    dem = np.asarray(demand, np.float32)
    Q = np.float32(capacity)
    i = np.repeat(np.arange(n), k)
    r = np.tile(np.arange(k), n)
    j = nn[i, r]
    dij = dist[i, j]
    f = np.empty((n * k, NF), np.float32)
    f[:, 0] = dij / mkd
    f[:, 1] = r.astype(np.float32) / np.float32(k)
    f[:, 2] = dem[i] / Q
    f[:, 3] = dem[j] / Q
    f[:, 4] = ang[i]
    f[:, 5] = ang[j]
    f[:, 6] = rad[i] / R
    f[:, 7] = rad[j] / R
    f[:, 8] = kd[i] / mkd
    f[:, 9] = kd[j] / mkd
    f[:, 10] = ((dist[0, i] + dist[0, j]) - dij) / R
    mutual = (nn[j] == i[:, None]).any(axis=1)
    f[:, 11] = mutual.astype(np.float32)
    return f


def edge_labels(nn, k, tour) -> np.ndarray:
    """1 where the candidate edge (i, nn[i][r]) joins consecutive nodes of
    the tour (either direction), else 0; [n*k] float32."""
    nn = np.asarray(nn, np.int64)[:, :k]
    n = nn.shape[0]
    adj = np.zeros((n, n), bool)
    t = np.asarray(tour, np.int64)
    adj[t[:-1], t[1:]] = True
    adj[t[1:], t[:-1]] = True
    i = np.repeat(np.arange(n), k)
    return adj[i, nn.ravel()].astype(np.float32)


# -------------------------------------------------------------------- MLP --

def forward(p, feat, bf16_round=True):
    """Returns (logits [E], cache).  feat [E, 12].  With bf16_round the
    kernels' roundings apply; without, it is the float64 network."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if bf16_round:
        q, dt = bf16, np.float32
    else:
        q, dt = (lambda v: np.asarray(v, np.float64)), np.float64
    X = q(np.asarray(feat, np.float32).T if bf16_round else np.asarray(feat, np.float64).T)
    W1q, W2q, w3q = q(W1), q(W2), q(w3)
    f64 = lambda v: np.asarray(v, np.float64)
    z1 = (f64(W1q) @ f64(X) + f64(b1)[:, None]).astype(dt)
    h1 = q(np.maximum(z1, 0))
    z2 = (f64(W2q) @ f64(h1) + f64(b2)[:, None]).astype(dt)
    h2 = np.maximum(z2, 0)
    # layer 3 on the vector units: fp32 products summed over neurons
    z3 = (f64(w3q) @ f64(h2) + np.float64(b3)).astype(dt)
    cache = dict(X=X, W1q=W1q, W2q=W2q, w3q=w3q, z1=z1, h1=h1, z2=z2, h2=h2)
    return z3, cache


def bce_grad(z, label, pos_weight):
    """Per-edge dL/dz of L = mean_e w_e (softplus(z_e) - y_e z_e) with
    w_e = pos_weight for positives, 1 otherwise; and L itself."""
    z = np.asarray(z, np.float64)
    y = np.asarray(label, np.float64)
    w = np.where(y > 0.5, pos_weight, 1.0)
    E = len(z)
    sig = 1.0 / (1.0 + np.exp(-z))
    loss = np.sum(w * (np.logaddexp(0.0, z) - y * z)) / E
    return w * (sig - y) / E, loss


def backward(p, feat, label, pos_weight, bf16_round=True):
    """Gradient of the weighted BCE loss w.r.t. the parameter vector, with
    the kernels' roundings; returns (grad [NPARAM] float32/64, loss)."""
    z3, c = forward(p, feat, bf16_round)
    g, loss = bce_grad(z3, label, pos_weight)
    q = bf16 if bf16_round else (lambda v: np.asarray(v, np.float64))
    dt = np.float32 if bf16_round else np.float64
    g = g.astype(dt)
    h2, h1, z2, z1, X = c["h2"], c["h1"], c["z2"], c["z1"], c["X"]
    grad = np.zeros(NPARAM, np.float64)
    grad[OFF_W3:OFF_B3] = np.asarray(h2, np.float64) @ np.asarray(g, np.float64)
    grad[OFF_B3] = np.sum(np.asarray(g, np.float64))
    dz2 = (np.asarray(c["w3q"], dt)[:, None] * g[None, :]) * (z2 > 0)
    dz2 = dz2.astype(dt)
    grad[OFF_B2:OFF_W3] = np.sum(np.asarray(dz2, np.float64), axis=1)
    dz2q = q(dz2)
    grad[OFF_W2:OFF_B2] = (np.asarray(dz2q, np.float64) @ np.asarray(h1, np.float64).T).ravel()
    dh1 = (np.asarray(c["W2q"], np.float64).T @ np.asarray(dz2q, np.float64)).astype(dt)
    dz1 = (dh1 * (z1 > 0)).astype(dt)
    grad[OFF_B1:OFF_W2] = np.sum(np.asarray(dz1, np.float64), axis=1)
    dz1q = q(dz1)
    grad[OFF_W1:OFF_B1] = (np.asarray(dz1q, np.float64) @ np.asarray(X, np.float64).T).ravel()
    return grad.astype(dt), loss


def loss_of(p, feat, label, pos_weight):
    """Float64 loss of the unrounded network (finite-difference target)."""
    z3, _ = forward(np.asarray(p, np.float64), feat, bf16_round=False)
    return bce_grad(z3, label, pos_weight)[1]


def adam(p, m, v, grad, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One fp32 Adam step (t counts from 1); returns new (p, m, v).

    m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
    p -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
    b1, b2 are taken as fp32; 1 - b and 1 - b^t are formed in float64 from
    the fp32 value and rounded once to fp32."""
    f = np.float32
    b1, b2 = f(beta1), f(beta2)
    om1, om2 = f(1.0 - np.float64(b1)), f(1.0 - np.float64(b2))
    c1, c2 = f(1.0 - np.float64(b1) ** t), f(1.0 - np.float64(b2) ** t)
    p = np.asarray(p, f).copy()
    g = np.asarray(grad, f)
    m = (b1 * np.asarray(m, f) + om1 * g).astype(f)
    v = (b2 * np.asarray(v, f) + om2 * (g * g)).astype(f)
    mh = (m / c1).astype(f)
    vh = (v / c2).astype(f)
    p = (p - f(lr) * (mh / (np.sqrt(vh) + f(eps)))).astype(f)
    return p, m, v
