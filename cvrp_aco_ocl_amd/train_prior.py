"""Train the learned edge prior and compare it with the plain colony.

    python -m cvrp_aco_ocl_amd.train_prior train --instances 4000
    python -m cvrp_aco_ocl_amd.train_prior eval --instances 64

``train``: generated instances (instances.training, n-1 in [100,500]) are
solved by the plain colony in batches on the GPU; the edges of each best
tour label that instance's k=16 candidate edges; the network is trained on
the GPU by the library's own kernels (aco_prior_train) and its weights are
written to data/prior_weights.f32 with a summary in data/prior_train.json.

``eval``: the evaluation instances (instances.evaluation) are solved
batched by the plain colony for ``--iterations`` (its final best is the
reference cost C*) and by the colony with the prior (the prior's build time
counts towards its clock); it reports the wall-clock time to reach
1.02 C* and the gap to C* at equal wall-clock budgets.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import instances as I
from . import prior as PR
from .solver import Instance

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
# exponent of the prior probability in eta (eta * p^gamma), chosen on held-out
# validation instances (0.25 .. 2 tried; 0.5 was the only one ahead at 1 s)
PRIOR_EXPONENT = 0.5


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def auc(score, label):
    """P(score of a positive > score of a negative), ties counted half."""
    neg = np.sort(score[label < 0.5])
    pos = score[label > 0.5]
    lo = np.searchsorted(neg, pos, side="left")
    hi = np.searchsorted(neg, pos, side="right")
    return float((lo.sum() + 0.5 * (hi - lo).sum()) / (len(pos) * len(neg)))


def labelled_edges(count, seed, a, tag):
    """Features and labels of the candidate edges of training instances
    0..count-1 of ``seed`` (labels: the plain colony's best tour)."""
    t0 = time.perf_counter()
    feats, labels, costs = [], [], []
    for lo in range(0, count, a.batch):
        chunk = [I.training(i, seed=seed) for i in range(lo, min(count, lo + a.batch))]
        best = PR.solve_plain(chunk, a.iterations, n_ants=a.ants, k=a.k, seed=seed + 1 + lo)
        lists = Instance.nn_lists_batch(chunk, a.k)   # the chunk's candidate lists in one GPU pass
        for inst, nn, (tour, length) in zip(chunk, lists, best):
            feats.append(PR.features(inst, nn, a.k))
            labels.append(PR.edge_labels(nn, a.k, tour))
            costs.append(length)
        log(f"{tag}: labelled {lo + len(chunk)}/{count} instances ({time.perf_counter() - t0:.1f}s)")
    return np.concatenate(feats), np.concatenate(labels), costs, time.perf_counter() - t0


def cmd_train(a):
    f, y, costs, t_label = labelled_edges(a.instances, a.seed, a, "train")
    log(f"training set: {len(f)} edges, {int(y.sum())} positive ({y.mean():.4f})")
    t1 = time.perf_counter()
    net = PR.PriorNet(seed=a.seed + 7)
    losses = net.train(f, y, epochs=a.epochs, batch=a.train_batch, lr=a.lr, lr_final=a.lr_final,
                       pos_weight=a.pos_weight, shuffle_seed=a.seed + 11)
    t_train = time.perf_counter() - t1
    z = net.forward(f)
    train_auc = auc(z, y)
    # baseline: the candidate rank alone (lower rank scores higher)
    rank_auc = auc(-f[:, 1], y)
    log(f"trained {a.epochs} epochs in {t_train:.2f}s: loss {losses[0]:.4f} -> {losses[-1]:.4f}, "
        f"AUC {train_auc:.4f} (candidate rank alone {rank_auc:.4f})")
    n_edges, n_pos = int(len(f)), int(y.sum())
    del f, y
    # held-out instances (another generator seed)
    fv, yv, _, _ = labelled_edges(a.validation, a.seed + 1, a, "validation")
    val_auc, val_rank_auc = auc(net.forward(fv), yv), auc(-fv[:, 1], yv)
    log(f"validation: AUC {val_auc:.4f} (candidate rank alone {val_rank_auc:.4f})")
    os.makedirs(DATA, exist_ok=True)
    net.save(a.out)
    summary = {
        "instances": a.instances, "colony_iterations": a.iterations, "ants": a.ants, "k": a.k,
        "edges": n_edges, "positives": n_pos, "epochs": a.epochs,
        "batch": a.train_batch, "lr": a.lr, "lr_final": a.lr_final, "pos_weight": a.pos_weight,
        "epoch_loss": [round(float(v), 6) for v in losses], "train_auc": round(train_auc, 5),
        "rank_only_auc": round(rank_auc, 5), "validation_instances": a.validation,
        "validation_auc": round(val_auc, 5), "validation_rank_only_auc": round(val_rank_auc, 5),
        "label_seconds": round(t_label, 2),
        "train_seconds": round(t_train, 3), "seed": a.seed,
        "mean_best_cost": float(np.mean(costs)),
    }
    with open(a.summary, "w") as fh:
        json.dump(summary, fh, indent=1)
    print(json.dumps(summary))


def compare(insts, iterations, net, ants=64, k=16, seed=1, floor=0.05,
            budgets=(0.25, 0.5, 1.0, 2.0), exponent=None, rerank_pool=0):
    """Plain colony for ``iterations`` (C* = its final best), then the prior
    colony until every instance is within 2% of C* or ``iterations``
    (``rerank_pool``: the prior also chooses the candidates, run_colony)."""
    plain = PR.run_colony(insts, iterations, None, n_ants=ants, k=k, seed=seed)
    cstar = [r[-1][1] for r in plain["records"]]
    targets = [1.02 * c for c in cstar]
    pri = PR.run_colony(insts, iterations, net, n_ants=ants, k=k, seed=seed, floor=floor,
                        targets=targets, prior_exponent=exponent, rerank_pool=rerank_pool)
    t_plain = [PR.time_to_target(r, t) for r, t in zip(plain["records"], targets)]
    t_prior = [PR.time_to_target(r, t, pri["build_s"]) for r, t in zip(pri["records"], targets)]
    gaps = {}
    for bud in budgets:
        gp = [PR.best_at(r, bud) / c - 1 for r, c in zip(plain["records"], cstar)]
        gq = [PR.best_at(r, bud, pri["build_s"]) / c - 1 for r, c in zip(pri["records"], cstar)]
        gaps[str(bud)] = {"plain": float(np.mean(gp)), "prior": float(np.mean(gq))}
    fin = lambda v: [x for x in v if np.isfinite(x)]
    reached = fin(t_prior)
    out = {
        "instances": len(insts), "iterations": iterations, "ants": ants, "k": k,
        "prior_build_s": round(pri["build_s"], 4), "prior_iterations": pri["iterations"],
        "plain_time_to_2pct_s": {"mean": float(np.mean(t_plain)), "median": float(np.median(t_plain))},
        "plain_all_reached_s": float(np.max(t_plain)),
        "prior_time_to_2pct_s": {"mean": float(np.mean(t_prior)) if len(reached) == len(t_prior) else None,
                                 "median": float(np.median(t_prior)),
                                 "reached": len(reached)},
        "prior_all_reached_s": float(np.max(t_prior)) if len(reached) == len(t_prior) else None,
        "plain_wall_s": round(plain["wall_s"], 3),
        "plain_ant_solutions_per_s": round(plain["ant_solutions"] / plain["wall_s"], 1),
        "prior_exponent": exponent,
        "rerank_pool": rerank_pool,
        "gap_at_budget": gaps,
        "mean_cstar": float(np.mean(cstar)),
    }
    out["ratio_mean"] = (out["prior_time_to_2pct_s"]["mean"] / out["plain_time_to_2pct_s"]["mean"]
                         if out["prior_time_to_2pct_s"]["mean"] is not None else None)
    out["ratio_median"] = out["prior_time_to_2pct_s"]["median"] / out["plain_time_to_2pct_s"]["median"]
    return out


def cmd_eval(a):
    net = PR.PriorNet.load(a.weights)
    insts = I.evaluation(a.instances)
    res = compare(insts, a.iterations, net, ants=a.ants, k=a.k, floor=a.floor, exponent=a.exponent,
                  rerank_pool=a.rerank_pool)
    print(json.dumps(res))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--instances", type=int, default=4000)
    t.add_argument("--validation", type=int, default=128)
    t.add_argument("--iterations", type=int, default=200)
    t.add_argument("--ants", type=int, default=64)
    t.add_argument("--k", type=int, default=16)
    t.add_argument("--batch", type=int, default=128, help="instances per colony batch")
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--train-batch", type=int, default=16384)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--lr-final", type=float, default=2e-4)
    t.add_argument("--pos-weight", type=float, default=3.0)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", default=PR.WEIGHTS_PATH)
    t.add_argument("--summary", default=os.path.join(DATA, "prior_train.json"))
    e = sub.add_parser("eval")
    e.add_argument("--instances", type=int, default=64)
    e.add_argument("--iterations", type=int, default=2000)
    e.add_argument("--ants", type=int, default=64)
    e.add_argument("--k", type=int, default=16)
    e.add_argument("--floor", type=float, default=0.05)
    e.add_argument("--exponent", type=float, default=PRIOR_EXPONENT)
    e.add_argument("--weights", default=PR.WEIGHTS_PATH)
    e.add_argument("--rerank-pool", type=int, default=0,
                   help="the prior also chooses the k candidates among this many nearest (0: off)")
    a = ap.parse_args(argv)
    {"train": cmd_train, "eval": cmd_eval}[a.cmd](a)


if __name__ == "__main__":
    main()
