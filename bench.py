"""Benchmark: ACO iterations per second on one MI355X, as ant-solutions/s.

Workload (the reference's metric configuration that fits one GPU): the
481-node clustered instance -- 480 customers in 8 Gaussian clusters
(sigma 6) in [0,1000]^2, random depot, demand U{1..100}, capacity for 30
routes, nint int32 distances -- with 480 ants, alpha 1, beta 2, rho 0.1,
6 ranks, roulette over the k=30 nearest candidates (Philox draws), and
2-opt + swap + relocate local search.  One step = one whole iteration
(construct -> local search -> statistics -> fused pheromone pass), replayed
as one hipGraph, inputs resident in HBM.

Prints ONE JSON line.  Extra legs: per-kernel device time from HIP event
pairs on the library's stream (roofline of the dominant kernel), the
best-so-far trajectory of a full run (time to a 2% gap), and the CPU
oracle timed on this host (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ant-solutions/s + time-to-2%-gap, learned-prior vs plain ACO, 1 MI355X"
STAGES = {0: "construct", 1: "local_search", 2: "iteration"}
HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(customers=480, clusters=8, sigma=6.0, routes=30, m=480, k=30, ls=7,
                iterations_full=1000)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--full-iterations", type=int, default=WORKLOAD["iterations_full"],
                    help="length of the best-so-far run for time-to-gap (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline sample budget (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--prior-instances", type=int, default=64,
                    help="learned-prior vs plain leg: evaluation instances (0 = skip)")
    ap.add_argument("--prior-iterations", type=int, default=2000,
                    help="plain iterations that fix the reference cost C*")
    ap.add_argument("--mlp-probe", type=int, default=0, metavar="INSTANCES",
                    help="only run the learned prior's MLP on that many evaluation instances")
    ap.add_argument("--batch-instances", type=int, default=256,
                    help="instances of the batched leg (0 = skip)")
    ap.add_argument("--n1000-seconds", type=float, default=10.0,
                    help="GPU vs CPU oracle on the 1001-node instance (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 10 device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes) and exit")
    return ap.parse_args()


def copy_probe(mib: int) -> None:
    import torch
    nbytes = mib << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes,
                      "launches": 10}))


def mlp_probe(instances_count: int) -> None:
    """The learned prior's dense layers on the evaluation set's candidate
    edges (instances of 500 customers, k = 16): 3 forward passes over every
    edge, then one training epoch (batches of 16384), for a kernel trace and
    MFMA counters.  Labels are the candidate rank < 2 (the timing does not
    depend on them)."""
    import cvrp_aco_ocl_amd as ca
    from cvrp_aco_ocl_amd import instances, prior as PR
    insts = instances.evaluation(instances_count)
    lists = ca.Instance.nn_lists_batch(insts, 16)
    feat = np.concatenate([PR.features(it, nn, 16) for it, nn in zip(insts, lists)])
    label = (feat[:, 1] < 2.0 / 16.0).astype(np.float32)
    net = PR.PriorNet.load()
    t0 = time.perf_counter()
    for _ in range(3):
        net.forward(feat)
    t_fwd = (time.perf_counter() - t0) / 3
    t0 = time.perf_counter()
    net.train(feat, label, epochs=1, batch=16384)
    t_train = time.perf_counter() - t0
    net.close()
    print(json.dumps({"mlp_probe_edges": int(len(feat)), "forward_wall_s": round(t_fwd, 4),
                      "train_epoch_wall_s": round(t_train, 4),
                      "forward_flop_per_edge": 2 * (12 * 64 + 64 * 64 + 64)}))


def make_solver(ca, inst, seed):
    return ca.GpuACO(inst, n_ants=WORKLOAD["m"], k=WORKLOAD["k"], seed=seed + 1,
                     selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, ls_flags=WORKLOAD["ls"],
                     use_graph=True)


def measured_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` on this workload from the committed
    FETCH_SIZE / WRITE_SIZE profile (calibrated against a plain copy), or
    None when there is no measurement for it."""
    path = os.path.join(ROOT, "profiles", "traffic_clustered481.md")
    key = f"bench: {kernel}_traffic_bytes="
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key):
                    return int(line[len(key):].strip())
    except OSError:
        pass
    return None


def construct_bytes(inst, sizes, m, k, fallbacks):
    """Algorithmic bytes of one construction launch: per step k candidate
    ids + k total_info + the chosen distance + the tour write (8k+8 B), per
    ant the demand and depot-distance rows (8n B), per full-row fallback
    the total_info row (4n B)."""
    steps = int(np.sum(sizes[:m] - 1))
    n = inst.n
    return steps * (8 * k + 8) + m * 8 * n + fallbacks * 4 * n


def pheromone_bytes(inst):
    """Fused pass: read tau and heur, write tau and total_info (16 n^2 B)."""
    return 16 * inst.n * inst.n


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): this build measures one "
              "GPU's share of the work; sharding across GPUs is a design, not a measured path",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.mlp_probe:
        mlp_probe(args.mlp_probe)
        return
    import cvrp_aco_ocl_amd as ca
    from cvrp_aco_ocl_amd import instances

    inst = instances.clustered(args.seed, customers=WORKLOAD["customers"],
                               nclust=WORKLOAD["clusters"], sigma=WORKLOAD["sigma"],
                               routes=WORKLOAD["routes"])
    m, k = WORKLOAD["m"], WORKLOAD["k"]
    print(f"# workload {inst.name}: n={inst.n} Q={inst.capacity} m={m} k={k}", file=sys.stderr,
          flush=True)

    # ---- timed region: K iterations, graph replay ----
    g = make_solver(ca, inst, args.seed)
    g.init_aco()
    tune = g.autotune(reps=3)   # local-search launch shape; results are the same for every shape
    shape = g.launch_shape()
    print(f"# autotune: {tune} -> {shape}", file=sys.stderr, flush=True)
    g.run(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(args.steps)
    g.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = m * args.steps / elapsed
    cnt0 = g.counters()
    print(f"# timed {args.steps} iterations: {ms_per_step:.3f} ms/iter, {value:.0f} ant-sol/s",
          file=sys.stderr, flush=True)

    # ---- per-kernel device time (HIP events on the library stream) ----
    # (per-phase launches, so each kernel is timed on its own)
    if shape[5]:
        g.set_launch_shape(*shape[:5], 0)
    g.profile_enable(True)
    fb_before = g.counters()["fallbacks"]
    ev_before = g.counters()["ls_evals"]
    size_sum = 0
    cons_bytes = 0
    for _ in range(args.profile_steps):
        fb0 = g.counters()["fallbacks"]
        g.run_aco_iteration()
        _, sizes, _ = g.ants()
        # construct bytes need this iteration's tours before local search moved
        # nodes around: sizes do not change under local search
        cons_bytes += construct_bytes(inst, sizes, m, k, g.counters()["fallbacks"] - fb0)
        size_sum += int(np.sum(sizes[:m]))
    kern = {}
    for name in ("construct", "local_search", "statistics", "pheromone", "gather"):
        ms, launches = g.profile(name)
        if launches:
            kern[name] = {"avg_ms": ms / launches, "launches": launches}
    cnt1 = g.counters()
    ls_evals = cnt1["ls_evals"] - ev_before
    g.profile_enable(False)
    g.close()

    avg = {n: v["avg_ms"] for n, v in kern.items()}
    dominant = max(avg, key=avg.get)
    cons_bpl = cons_bytes / args.profile_steps
    ph_bpl = pheromone_bytes(inst)
    per_kernel = {}
    for name, ms in avg.items():
        entry = {"avg_ms": round(ms, 4)}
        if name == "construct":
            entry["algorithmic_bytes"] = int(cons_bpl)
            entry["GBps"] = round(cons_bpl / (ms * 1e-3) / 1e9, 2)
        if name == "pheromone":
            entry["algorithmic_bytes"] = ph_bpl
            entry["GBps"] = round(ph_bpl / (ms * 1e-3) / 1e9, 2)
        if name == "local_search":
            entry["move_evals_per_s"] = round(ls_evals / args.profile_steps / (ms * 1e-3), 1)
        per_kernel[name] = entry
    if dominant == "construct":
        achieved = cons_bpl / (avg["construct"] * 1e-3) / 1e9
        traffic_alg = int(cons_bpl)
    elif dominant == "pheromone":
        achieved = ph_bpl / (avg["pheromone"] * 1e-3) / 1e9
        traffic_alg = ph_bpl
    else:
        # local search: 16 B per evaluated move (the four distance / edge
        # words a 2-opt, swap or relocate delta reads)
        traffic_alg = int(16 * ls_evals / max(args.profile_steps, 1))
        achieved = traffic_alg / (avg[dominant] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(dominant) if args.seed == 0 else None,
                "algorithmic_bytes_per_launch": traffic_alg}

    # ---- best-so-far trajectory: time to 2% of the run's final best ----
    quality = None
    if args.full_iterations > 0:
        g = make_solver(ca, inst, args.seed)
        g.set_launch_shape(*shape)
        t0 = time.perf_counter()
        g.init_aco()
        g.run(args.full_iterations)
        g.synchronize()
        wall = time.perf_counter() - t0
        rec = g.bsf_records()
        g.close()
        best = rec[-1][1]
        target = 1.02 * best
        hit = next(r for r in rec if r[1] <= target)
        quality = {"iterations": args.full_iterations, "best": best, "wall_s": round(wall, 3),
                   "time_to_2pct_gap_s": round(hit[2], 4), "iter_to_2pct_gap": hit[0],
                   "records": len(rec)}
        print(f"# full run {args.full_iterations} it: best {best} in {wall:.2f}s, 2% gap at "
              f"it {hit[0]} ({hit[2]:.3f}s)", file=sys.stderr, flush=True)

    # ---- CPU baseline: the oracle on this host's cores ----
    cpu = None
    if args.cpu_seconds > 0:
        import oracle
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        P = oracle.Problem(inst.dist, inst.demand, inst.capacity, inst.max_dist,
                           inst.service_time, k=k, tie=0)
        col = oracle.Colony(P, m, selection=oracle.SEL_KNN, rng=oracle.RNG_PHILOX,
                            ls_flags=WORKLOAD["ls"], seed=args.seed + 1, threads=threads)
        col.init()
        its = 0
        t0 = time.perf_counter()
        while True:
            col.iteration()
            its += 1
            el = time.perf_counter() - t0
            if el >= args.cpu_seconds or its >= 200:
                break
        cpu = {"value": round(m * its / el, 2), "unit": "ant-solutions/s", "cores": threads,
               "kind": "port",
               "sample": f"{its} whole iterations (480 ants each, same workload, OpenMP over ants) "
                         f"in {el:.2f}s"}
        print(f"# cpu baseline: {cpu['value']} ant-sol/s on {threads} threads", file=sys.stderr,
              flush=True)

    # ---- learned prior vs plain: 64 clustered n=501 instances, batched ----
    prior_leg = None
    if args.prior_instances > 0:
        from cvrp_aco_ocl_amd import prior as PR, train_prior as TP
        net = PR.PriorNet.load()
        ev = instances.evaluation(args.prior_instances)
        print(f"# prior leg: {len(ev)} instances, plain {args.prior_iterations} iterations",
              file=sys.stderr, flush=True)
        prior_leg = TP.compare(ev, args.prior_iterations, net, ants=64, k=16,
                               exponent=TP.PRIOR_EXPONENT, budgets=(0.5, 1.0, 2.0, 5.0, 10.0))
        prior_leg["workload"] = ("64 clustered instances (500 customers, 8 clusters, sigma 6, "
                                 "30 routes; seeds 100000+), 64 ants each, k=16 Philox roulette, "
                                 "2-opt+swap+relocate, one batched context; C* = plain best after "
                                 f"{args.prior_iterations} iterations; prior time includes "
                                 "aco_prior_apply")
        net.close()
        print(f"# prior leg: plain median {prior_leg['plain_time_to_2pct_s']['median']:.3f}s, "
              f"prior median {prior_leg['prior_time_to_2pct_s']['median']:.3f}s",
              file=sys.stderr, flush=True)

    # ---- 1001-node instance: GPU vs the CPU oracle on all host cores ----
    n1000 = None
    if args.n1000_seconds > 0:
        big = instances.mixed_1000(0)
        gb = ca.GpuACO(big, n_ants=256, k=32, seed=1, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                       ls_flags=7)
        gb.init_aco()
        tuned = gb.autotune(reps=2)
        gb.run(3)
        gb.synchronize()
        t0 = time.perf_counter()
        gb.run(20)
        gb.synchronize()
        g_rate = 256 * 20 / (time.perf_counter() - t0)
        gb_shape = gb.launch_shape()
        gb.close()
        import oracle
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        Pb = oracle.Problem(big.dist, big.demand, big.capacity, big.max_dist, big.service_time,
                            k=32, tie=0)
        colb = oracle.Colony(Pb, 256, selection=oracle.SEL_KNN, rng=oracle.RNG_PHILOX, ls_flags=7,
                             seed=1, threads=threads)
        colb.init()
        its, t0 = 0, time.perf_counter()
        while True:
            colb.iteration()
            its += 1
            el = time.perf_counter() - t0
            if el >= args.n1000_seconds or its >= 50:
                break
        c_rate = 256 * its / el
        n1000 = {"workload": "1001 nodes (500 uniform + 500 clustered, 5% heavy demands, ~12 "
                             "customers/route), 256 ants, k=32, full local search",
                 "gpu_ant_solutions_per_s": round(g_rate, 1),
                 "cpu_ant_solutions_per_s": round(c_rate, 2), "cpu_threads": threads,
                 "cpu_sample": f"{its} iterations in {el:.2f}s",
                 "ratio": round(g_rate / c_rate, 2), "launch_shape": gb_shape,
                 "autotune_ms": {f"{STAGES[st]}{sh}": round(ms, 3) for sh, st, ms in tuned}}
        print(f"# n=1001: GPU {g_rate:.0f} vs CPU {c_rate:.1f} ant-sol/s ({g_rate / c_rate:.1f}x)",
              file=sys.stderr, flush=True)

    # ---- 256 uneven instances batched in one context (parity-test config,
    # reported for its throughput; not the metric's workload) ----
    batch = None
    if args.batch_instances > 0:
        binst = instances.batch(0, count=args.batch_instances)
        ns = [it.n for it in binst]
        t0 = time.perf_counter()
        gq = ca.GpuACO(binst, n_ants=64, k=16, seed=1, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                       ls_flags=7)
        gq.init_aco()
        gq.synchronize()
        create_s = time.perf_counter() - t0
        tuned = gq.autotune(reps=2)
        gq.run(2)
        gq.synchronize()
        K = 10
        t0 = time.perf_counter()
        gq.run(K)
        gq.synchronize()
        el = time.perf_counter() - t0
        bshape = gq.launch_shape()
        gq.close()
        batch = {"workload": f"{len(binst)} instances, n {min(ns)}..{max(ns)} (uniform or 4-12 "
                             "clusters, demands 1..50, 5-40 routes), 64 ants each, k=16 Philox "
                             "roulette, 2-opt+swap+relocate, one batched context",
                 "ant_solutions_per_s": round(len(binst) * 64 * K / el, 1),
                 "ms_per_iteration": round(1e3 * el / K, 3), "create_and_init_s": round(create_s, 3),
                 "launch_shape": bshape,
                 "iteration_ms": {f"{'fused' if sh[5] else 'per-phase'}{sh}": round(ms, 3)
                                  for sh, st, ms in tuned if st == 2 and ms > 0}}
        print(f"# batch: {batch['ant_solutions_per_s']:.0f} ant-sol/s, {batch['ms_per_iteration']:.2f} "
              f"ms/iter, shape {bshape}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "ant-solutions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"clustered-481: {WORKLOAD['customers']} customers, 8 Gaussian clusters "
                        f"sigma 6, Q for 30 routes, nint int32 distances; m=480, k=30 roulette_knn "
                        f"(Philox), 2-opt+swap+relocate, alpha 1 beta 2 rho 0.1, 6 ranks",
            "n": inst.n, "ants": m, "k": k, "capacity": inst.capacity, "seed": args.seed,
            "step": "one ACO iteration (hipGraph replay)",
            "launch_shape": {"ls_waves": shape[0], "ls_ring": shape[1], "ls_prefetch": shape[2],
                             "construct_ants_per_block": shape[3], "ls_row16": shape[4],
                             "fused_iteration": shape[5]},
        },
        "roofline": roofline,
        "kernels": per_kernel,
        "ls_move_evals_per_iter": int(ls_evals / max(args.profile_steps, 1)),
        "quality": quality,
        "learned_prior": prior_leg,
        "n1001": n1000,
        "batch256": batch,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
