"""GPU parity of the learned prior (features, MFMA forward / backward, Adam,
training, dense prior) against the CPU oracle in oracle/prior.py.

Tolerances: features are fp32 formulas of the same inputs (atan2 may differ
by an ulp, so 1e-6); logits and gradients differ from the oracle only in
fp32 accumulation order (MFMA vs float64 sums rounded once), which can flip
a bf16 rounding of a hidden activation now and then -- hence an absolute
logit bound of 2e-3 with a median below 1e-5, and a gradient bound of 2e-3
of the largest component.  Adam is elementwise fp32 with the same
operation order: bit-exact."""
import numpy as np
import pytest

from helpers import log
from oracle import prior as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca():
    import cvrp_aco_ocl_amd as ca
    from cvrp_aco_ocl_amd import prior
    ca.prior = prior
    return ca


def _net_params(seed):
    rng = np.random.default_rng(seed)
    p = P.init_params(seed)
    p[P.OFF_B1:P.OFF_W2] = rng.normal(0, 0.1, P.H).astype(np.float32)
    p[P.OFF_B2:P.OFF_W3] = rng.normal(0, 0.1, P.H).astype(np.float32)
    p[P.OFF_B3] = np.float32(-0.7)
    return p


def test_init_params_equal_oracle(ca):
    for seed in (1, 12345):
        with ca.prior.PriorNet(seed) as net:
            np.testing.assert_array_equal(net.params(), P.init_params(seed))


@pytest.mark.parametrize("which", ["uniform", "clustered"])
def test_features_match_oracle(ca, which):
    from cvrp_aco_ocl_amd import instances
    inst = instances.uniform_100(3) if which == "uniform" else instances.clustered(4, customers=300)
    k = 16
    nn = inst.nn_lists(20)
    log(f"features {which} n={inst.n} k={k}")
    got = ca.prior.features(inst, nn, k)
    ref = P.features(inst.x, inst.y, inst.demand, inst.capacity, inst.dist, nn, k)
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(got[:, [1, 2, 3, 11]], ref[:, [1, 2, 3, 11]])


def _features(seed, E):
    """Oracle features of the first E candidate edges of clustered instances
    (as many instances as needed)."""
    from cvrp_aco_ocl_amd import instances
    out, have, s = [], 0, seed
    while have < E:
        inst = instances.clustered(s, customers=400)
        nn = inst.nn_lists(16)
        out.append(P.features(inst.x, inst.y, inst.demand, inst.capacity, inst.dist, nn, 16))
        have += len(out[-1])
        s += 1000
    return np.concatenate(out)[:E]


@pytest.mark.parametrize("E", [1, 15, 16, 17, 1000, 6400])
def test_forward_matches_oracle(ca, E):
    p = _net_params(5)
    f = _features(1, E)
    log(f"forward E={E}")
    with ca.prior.PriorNet(params=p) as net:
        z = net.forward(f)
    zr, _ = P.forward(p, f)
    err = np.abs(z - zr)
    assert err.max() < 2e-3, err.max()
    assert np.median(err) < 1e-5
    assert np.abs(zr).max() > 0.1  # a live network


@pytest.mark.parametrize("E,pw", [(16, 1.0), (1000, 4.0), (4099, 4.0), (16384, 2.0)])
def test_gradient_matches_oracle(ca, E, pw):
    p = _net_params(9)
    f = _features(2, E)
    rng = np.random.default_rng(E)
    y = (rng.random(E) < 0.15).astype(np.float32)
    log(f"gradient E={E} pos_weight={pw}")
    with ca.prior.PriorNet(params=p) as net:
        g, loss = net.gradient(f, y, pw)
    gr, lr = P.backward(p, f, y, pw)
    assert abs(loss - lr) <= 1e-5 * abs(lr)
    scale = np.abs(gr).max()
    assert np.abs(g - gr).max() <= 2e-3 * scale, (np.abs(g - gr).argmax(), np.abs(g - gr).max(), scale)
    # blocks individually (relative to their own scale)
    for lo, hi in ((P.OFF_W1, P.OFF_B1), (P.OFF_B1, P.OFF_W2), (P.OFF_W2, P.OFF_B2),
                   (P.OFF_B2, P.OFF_W3), (P.OFF_W3, P.OFF_B3)):
        s = np.abs(gr[lo:hi]).max()
        assert np.abs(g[lo:hi] - gr[lo:hi]).max() <= 1e-2 * s + 1e-9, (lo, hi)


def test_adam_bit_exact(ca):
    p = _net_params(3)
    rng = np.random.default_rng(1)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    ref = p.copy()
    with ca.prior.PriorNet(params=p) as net:
        for t in range(1, 5):
            g = rng.normal(0, 1e-2, P.NPARAM).astype(np.float32)
            net.adam_step(g, 1e-3)
            ref, m, v = P.adam(ref, m, v, g, t, 1e-3)
            np.testing.assert_array_equal(net.params(), ref)


def test_train_reduces_loss_and_is_deterministic(ca):
    # labels from a fixed rule on the features: short, mutual, low-rank edges
    f = _features(3, 6 * 2048)
    y = ((f[:, 0] < 0.6) & (f[:, 11] > 0.5) & (f[:, 1] < 0.3)).astype(np.float32)
    log(f"train E={len(f)} positives={int(y.sum())}")
    runs = []
    for _ in range(2):
        with ca.prior.PriorNet(seed=4) as net:
            losses = net.train(f, y, epochs=40, batch=2048, lr=5e-3, lr_final=5e-4, pos_weight=2.0,
                               shuffle_seed=9)
            runs.append((losses, net.params(), net.forward(f)))
    losses, params, z = runs[0]
    np.testing.assert_array_equal(params, runs[1][1])
    assert losses[-1] < 0.4 * losses[0], losses
    # ranking quality: AUC of the learned logits
    pos, neg = z[y > 0.5], z[y < 0.5]
    auc = (np.searchsorted(np.sort(neg), pos).sum()) / (len(pos) * len(neg))
    assert auc > 0.97, auc


def test_build_prior(ca):
    from cvrp_aco_ocl_amd import instances
    inst = instances.uniform_100(5)
    k = 16
    nn = inst.nn_lists(20)
    p = _net_params(6)
    with ca.prior.PriorNet(params=p) as net:
        dense = net.build(inst, nn, k, exponent=2.0, floor=0.1)
        z = net.forward(ca.prior.features(inst, nn, k))
    sig = 1.0 / (1.0 + np.exp(-z.astype(np.float64)))
    expect = np.full((inst.n, inst.n), 0.1 * 0.1, np.float64)
    i = np.repeat(np.arange(inst.n), k)
    expect[i, nn[:, :k].ravel()] = sig ** 2
    np.testing.assert_allclose(dense, expect, rtol=2e-6, atol=0)


def test_apply_equals_dense_prior(ca):
    """aco_prior_apply on a context equals creating it with the dense prior
    of aco_prior_build: same total_info after init, same colony afterwards."""
    from cvrp_aco_ocl_amd import instances
    insts = [instances.uniform_100(7), instances.clustered(8, customers=250)]
    p = _net_params(10)
    k = 16
    with ca.prior.PriorNet(params=p) as net:
        dense = [net.build(it, it.nn_lists(k), k, exponent=2.0, floor=0.05) for it in insts]
        runs = []
        for mode in ("dense", "apply"):
            kw = dict(n_ants=16, k=20, seed=5, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=7,
                      tau0_rule=ca.TAU0_NN_TOUR)
            with ca.GpuACO(insts, prior=dense if mode == "dense" else None, **kw) as g:
                if mode == "apply":
                    net.apply(g, k=k, exponent=2.0, floor=0.05)
                g.init_aco()
                ti0 = [g.pheromone(b)[1] for b in range(len(insts))]
                g.run(3)
                runs.append((ti0, [g.ants(b) for b in range(len(insts))]))
    for b in range(len(insts)):
        np.testing.assert_array_equal(runs[0][0][b], runs[1][0][b])
        for x, y in zip(runs[0][1][b], runs[1][1][b]):
            np.testing.assert_array_equal(x, y)
    # the prior changed the heuristic
    with ca.GpuACO(insts, n_ants=16, k=20, seed=5, tau0_rule=ca.TAU0_NN_TOUR) as g:
        g.init_aco()
        assert not np.array_equal(g.pheromone(0)[1], runs[0][0][0])


def test_labelled_edges_of_training_instances(ca):
    """The training set builder: features of every candidate edge (lists from
    the batched GPU builder) equal the per-instance ones, labels mark the
    plain colony's best tours."""
    import argparse
    from cvrp_aco_ocl_amd import instances as I, prior as PR, train_prior as TP
    a = argparse.Namespace(batch=3, iterations=5, ants=16, k=16)
    f, y, costs, _ = TP.labelled_edges(3, 7, a, "test")
    insts = [I.training(i, seed=7) for i in range(3)]
    assert f.shape == (sum(it.n for it in insts) * 16, 12) and y.shape == (f.shape[0],)
    assert set(np.unique(y)) <= {0.0, 1.0} and y.sum() > 0
    off = 0
    for it in insts:
        ref = PR.features(it, it.nn_lists(16), 16)
        np.testing.assert_array_equal(f[off: off + ref.shape[0]], ref)
        off += ref.shape[0]
    assert len(costs) == 3 and all(c > 0 for c in costs)
