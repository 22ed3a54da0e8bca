"""Benchmark: ACO iterations on one MI355X, as ant-solutions/s.

Workload (the largest single-GPU configuration of the metric): the
1001-node mixed instance -- 500 uniform and 500 clustered customers (10
Gaussian clusters, sigma 25) in [0,1000]^2, depot at the centre, demand
U{1..100} with 5% heavy outliers in [80,100], capacity for ~12 customers per
route, nint int32 distances -- with 256 ants, alpha 1, beta 2, rho 0.1, 6
ranks, roulette over the k=32 nearest candidates (Philox draws) and full
local search (2-opt + swap + relocate).  One step = one whole iteration
(construct -> local search -> statistics -> fused pheromone pass), replayed
as one hipGraph, inputs resident in HBM.

Prints ONE JSON line.  Besides the timed steps: per-kernel device time from
HIP event pairs on the library's stream (roofline of the dominant kernel
against the LDS read bandwidth measured here by libcvrpprobe.so), the
best-so-far trajectory of a 2000-iteration run (time to a 2% gap), the same
workload under the reference's rules (I-Roulette over the full row), the
CPU oracle on this host (reference rules on the CPU share, on one thread,
and the candidate-roulette comparator), and the other configurations as
extra keys: the 481-node clustered instance, 256 batched instances, the
learned prior against plain eta on 64 batched instances, and a 3001-node
instance (past the local search's LDS limit) improved by sub-problem
decomposition rounds.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ant-solutions/s + time-to-2%-gap, learned-prior vs plain ACO, 1 MI355X"
STAGES = {0: "construct", 1: "local_search", 2: "iteration"}
HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(m=256, k=32, ls=7, iterations_full=2000)
CLUSTERED = dict(customers=480, clusters=8, sigma=6.0, routes=30, m=480, k=30, ls=7)
KERNELS = ("construct", "local_search", "statistics", "pheromone")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--full-iterations", type=int, default=WORKLOAD["iterations_full"],
                    help="length of the best-so-far run for time-to-gap (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline sample budget per leg (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = this process's CPU share (OMP_NUM_THREADS if set, else the "
                         "CPUs it may run on)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--shape", default="",
                    help="launch shape waves,ring,prefetch,apb,row16,fused instead of autotuning "
                         "(profiling runs, so the trace holds only the measured shape)")
    ap.add_argument("--reference-rules-steps", type=int, default=5,
                    help="iterations of the workload under I-Roulette on the GPU (0 = skip)")
    ap.add_argument("--clustered-steps", type=int, default=50,
                    help="481-node clustered leg: timed iterations (0 = skip)")
    ap.add_argument("--prior-instances", type=int, default=64,
                    help="learned-prior vs plain leg: evaluation instances (0 = skip)")
    ap.add_argument("--prior-iterations", type=int, default=2000,
                    help="plain iterations that fix the reference cost C*")
    ap.add_argument("--batch-instances", type=int, default=256,
                    help="instances of the batched leg (0 = skip)")
    ap.add_argument("--decompose-iterations", type=int, default=10,
                    help="3001-node decomposition leg: master iterations, a round after each "
                         "(0 = skip)")
    ap.add_argument("--mlp-probe", type=int, default=0, metavar="INSTANCES",
                    help="only run the learned prior's MLP on that many evaluation instances")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 10 device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes) and exit")
    ap.add_argument("--lds-probe", action="store_true",
                    help="only run the LDS bandwidth microbenchmark and exit")
    ap.add_argument("--roof-probe", action="store_true",
                    help="only run the roof microbenchmarks (LDS read, HBM copy, BF16 MFMA) and exit")
    ap.add_argument("--ls-phases", action="store_true",
                    help="only split the workload's local search into its phases (2-opt, swap, "
                         "relocate, all three) on the same constructed tours and exit")
    return ap.parse_args()


# ---------------------------------------------------------------- probes ----

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


def lds_peak():
    """Aggregate LDS read bandwidth of this device (GB/s), measured by
    libcvrpprobe.so: 2 blocks of 256 threads per CU, 2048 trips of 16
    sixteen-byte reads per thread (about 0.6 ms per launch), best of 5."""
    path = os.path.join(ROOT, "cvrp_aco_ocl_amd", "libcvrpprobe.so")
    lib = ctypes.CDLL(path)
    lib.probe_lds_bandwidth.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
    gbps, ms = ctypes.c_double(), ctypes.c_double()
    rc = lib.probe_lds_bandwidth(5, 2048, 2, ctypes.byref(gbps), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"probe_lds_bandwidth failed ({rc})")
    return gbps.value, ms.value


def _probe_lib():
    lib = ctypes.CDLL(os.path.join(ROOT, "cvrp_aco_ocl_amd", "libcvrpprobe.so"))
    D = ctypes.POINTER(ctypes.c_double)
    lib.probe_hbm_copy.argtypes = [ctypes.c_int, ctypes.c_int, D, D]
    lib.probe_mfma_bf16.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D]
    return lib


def measured_peaks() -> dict:
    """The device's roofs measured by libcvrpprobe.so in this run, beside the
    published MI355X figures: LDS read bandwidth (the local search's roof),
    HBM copy bandwidth (2 x 1 GiB buffers, read + write bytes; the pheromone
    pass's roof) and the dense BF16 matrix-core rate of the two gfx950 BF16
    MFMA shapes (4 waves per SIMD, 4 independent chains each; the prior MLP's
    roof).  Best of 5 launches each, about a second in all."""
    lib = _probe_lib()
    lds, _ = lds_peak()
    v, ms = ctypes.c_double(), ctypes.c_double()
    if lib.probe_hbm_copy(5, 1024, ctypes.byref(v), ctypes.byref(ms)) != 0:
        raise RuntimeError("probe_hbm_copy failed")
    hbm, hbm_ms = v.value, ms.value
    mfma = {}
    for shape, name in ((0, "16x16x32"), (1, "32x32x16")):
        if lib.probe_mfma_bf16(5, 2048, shape, 4, ctypes.byref(v), ctypes.byref(ms)) != 0:
            raise RuntimeError("probe_mfma_bf16 failed")
        mfma[name] = {"TFLOPs": round(v.value, 1), "best_launch_ms": round(ms.value, 4)}
    return {"lds_read_GBps": round(lds, 1), "hbm_copy_GBps": round(hbm, 1),
            "hbm_copy_best_launch_ms": round(hbm_ms, 4), "mfma_bf16": mfma,
            "published": {"hbm_GBps": HBM_PEAK_GBS, "mfma_bf16_dense_TFLOPs": 2500.0}}


def ls_phases(shape) -> None:
    """Local-search time of the workload per phase: the tours of iteration 6
    searched with 2-opt alone, swap alone (after the 2-opt), relocate alone
    (after both), and all three, each from the same start; HIP-event time of
    the one launch, evaluated moves and applied moves per ant."""
    import cvrp_aco_ocl_amd as ca
    from cvrp_aco_ocl_amd import instances
    inst = instances.mixed_1000(0)
    m, k = WORKLOAD["m"], WORKLOAD["k"]
    mk = lambda flags: ca.GpuACO(inst, n_ants=m, k=k, seed=1, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX,
                                 ls_flags=flags, use_graph=False)
    g = mk(7)
    g.init_aco()
    g.run(5)
    g.construct_solutions()
    start = tuple(a.copy() for a in g.ants())
    g.close()
    state, out = start, {}
    for flags, name in ((1, "2opt"), (2, "swap"), (4, "relocate"), (7, "all")):
        h = mk(flags)
        h.init_aco()
        if shape:
            h.set_launch_shape(*shape[:5])
        h.set_ants(*(start if flags == 7 else state))
        h.profile_enable(True)
        h.local_search()
        h.update_statistics()
        ms, _ = h.profile("local_search")
        c = h.counters()
        after = tuple(a.copy() for a in h.ants())
        h.close()
        out[name] = {"ms": round(ms, 4), "evals_per_ant": round(c["ls_evals"] / m, 1),
                     "moves_per_ant": round(c["ls_moves"] / m, 1)}
        if flags in (1, 2):
            state = after
    print(json.dumps({"ls_phases": out, "shape": list(shape) if shape else None}))


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


# --------------------------------------------------------------- helpers ----

def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of ``kernel`` from the committed FETCH_SIZE /
    WRITE_SIZE profile of ``workload`` (calibrated against a plain copy), or
    None when there is no measurement for it."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
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
    """Algorithmic bytes of one candidate-roulette construction launch: per
    step k candidate ids + k total_info + the chosen distance + the tour
    write (8k+8 B), per ant the demand and depot-distance rows (8n B), per
    full-row fallback the total_info row (4n B)."""
    steps = int(np.sum(sizes[:m] - 1))
    n = inst.n
    return steps * (8 * k + 8) + m * 8 * n + fallbacks * 4 * n


def pheromone_bytes(n, k):
    """Fused pass: read tau and heur, write tau and total_info (16 n^2 B),
    and the candidate rows' total_info: read the ids, write the values
    (8 n k B)."""
    return 16 * n * n + 8 * n * k


def cpu_share():
    """Threads this process may use: OMP_NUM_THREADS when the host sets a
    share, else the CPUs in its affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return (int(env) if env.isdigit() and int(env) > 0 else aff), aff


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return None


def profile_iterations(g, inst, m, k, steps):
    """Per-kernel device time (HIP event pairs on the library's stream) over
    `steps` per-phase iterations, with the algorithmic bytes of each."""
    g.profile_enable(True)
    ev_before = g.counters()["ls_evals"]
    cons_bytes = 0
    for _ in range(steps):
        fb0 = g.counters()["fallbacks"]
        g.run_aco_iteration()
        _, sizes, _ = g.ants()
        # local search does not change tour sizes, so the construction's
        # bytes follow from the sizes after the iteration
        cons_bytes += construct_bytes(inst, sizes, m, k, g.counters()["fallbacks"] - fb0)
    kern = {}
    for name in KERNELS:
        ms, launches = g.profile(name)
        if launches:
            kern[name] = ms / launches
    ls_evals = (g.counters()["ls_evals"] - ev_before) / steps
    g.profile_enable(False)
    per = {}
    for name, ms in kern.items():
        e = {"avg_ms": round(ms, 4)}
        if name == "construct":
            e["algorithmic_bytes"] = int(cons_bytes / steps)
            e["GBps"] = round(cons_bytes / steps / (ms * 1e-3) / 1e9, 2)
        elif name == "pheromone":
            e["algorithmic_bytes"] = pheromone_bytes(inst.n, k)
            e["GBps"] = round(pheromone_bytes(inst.n, k) / (ms * 1e-3) / 1e9, 2)
        elif name == "local_search":
            e["move_evals_per_launch"] = int(ls_evals)
            e["move_evals_per_s"] = round(ls_evals / (ms * 1e-3), 1)
            e["algorithmic_lds_bytes"] = int(16 * ls_evals)
        per[name] = e
    return kern, per, ls_evals


def time_to_gap(records, cstar, frac=0.02):
    for it, length, t in records:
        if length <= (1.0 + frac) * cstar:
            return it, t
    return None, None


# ------------------------------------------------------------------- legs ----

def leg_quality(ca, inst, shape, iterations, seed):
    """2000-iteration plain run: C* = its final best; time and iteration at
    which the best-so-far first came within 2% of C*."""
    g = ca.GpuACO(inst, n_ants=WORKLOAD["m"], k=WORKLOAD["k"], seed=seed + 1, selection=ca.SEL_KNN,
                  rng=ca.RNG_PHILOX, ls_flags=WORKLOAD["ls"], use_graph=True)
    g.set_launch_shape(*shape)
    t0 = time.perf_counter()
    g.init_aco()
    g.run(iterations)
    g.synchronize()
    wall = time.perf_counter() - t0
    rec = g.bsf_records()
    g.close()
    cstar = rec[-1][1]
    it, t = time_to_gap(rec, cstar)
    log(f"{iterations} iterations: C* {cstar} in {wall:.2f}s, 2% gap at iteration {it} ({t:.3f}s)")
    return {"iterations": iterations, "cstar": cstar, "wall_s": round(wall, 3),
            "time_to_2pct_gap_s": round(t, 4), "iter_to_2pct_gap": it, "records": len(rec)}


def leg_reference_rules_gpu(ca, inst, steps, seed):
    """The workload under the reference's rules on the GPU: I-Roulette over
    the full row with Park-Miller lane streams (kernel.cl:325-373), same
    local search."""
    g = ca.GpuACO(inst, n_ants=WORKLOAD["m"], seed=seed + 1, selection=ca.SEL_IROULETTE,
                  rng=ca.RNG_PARK_MILLER, ls_flags=WORKLOAD["ls"], use_graph=True)
    g.init_aco()
    g.autotune(reps=2)
    g.run(2)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(steps)
    g.synchronize()
    el = time.perf_counter() - t0
    g.profile_enable(True)
    g.run_aco_iteration()
    cons_ms = g.profile("construct")[0]
    ls_ms = g.profile("local_search")[0]
    g.close()
    rate = WORKLOAD["m"] * steps / el
    log(f"reference rules on the GPU: {rate:.0f} ant-sol/s ({1e3 * el / steps:.2f} ms/iteration)")
    return {"ant_solutions_per_s": round(rate, 1), "ms_per_iteration": round(1e3 * el / steps, 3),
            "construct_ms": round(cons_ms, 3), "local_search_ms": round(ls_ms, 3),
            "rules": "I-Roulette over the full row, Park-Miller lane streams, 2-opt+swap+relocate"}


def _cpu_rate(oracle, inst, m, selection, rng, k, threads, seconds, max_iters):
    P = oracle.Problem(inst.dist, inst.demand, inst.capacity, inst.max_dist, inst.service_time,
                       k=k, tie=0)
    col = oracle.Colony(P, m, selection=selection, rng=rng, ls_flags=WORKLOAD["ls"], seed=1,
                        tau0_rule=1, threads=threads)
    col.init()
    its, t0 = 0, time.perf_counter()
    while True:
        col.iteration()
        its += 1
        el = time.perf_counter() - t0
        if el >= seconds or its >= max_iters:
            break
    return m * its / el, its, el


def leg_cpu(inst, args):
    """The CPU oracle on this host: the reference's rules (I-Roulette over
    the full row, Park-Miller, the same local search) with OpenMP over ants
    on this process's CPU share and on one thread, and the candidate
    roulette (the GPU path's rule) as an algorithm-matched comparator."""
    import oracle
    share, affinity = cpu_share()
    threads = args.cpu_threads or share
    nproc = os.cpu_count()
    ref_rate, ref_its, ref_el = _cpu_rate(oracle, inst, WORKLOAD["m"], oracle.SEL_IROULETTE,
                                          oracle.RNG_PM, 0, threads, args.cpu_seconds, 100)
    log(f"cpu reference rules, {threads} threads: {ref_rate:.1f} ant-sol/s")
    one_m = 16
    one_rate, one_its, one_el = _cpu_rate(oracle, inst, one_m, oracle.SEL_IROULETTE, oracle.RNG_PM,
                                          0, 1, max(2.0, args.cpu_seconds / 2), 100)
    log(f"cpu reference rules, 1 thread: {one_rate:.2f} ant-sol/s")
    knn_rate, knn_its, knn_el = _cpu_rate(oracle, inst, WORKLOAD["m"], oracle.SEL_KNN,
                                          oracle.RNG_PHILOX, WORKLOAD["k"], threads,
                                          max(2.0, args.cpu_seconds / 2), 100)
    log(f"cpu candidate roulette, {threads} threads: {knn_rate:.1f} ant-sol/s")
    return {
        "value": round(ref_rate, 2), "unit": "ant-solutions/s", "cores": threads, "kind": "port",
        "sample": f"{ref_its} whole iterations of the same workload under the reference's rules "
                  f"(I-Roulette over the full row, Park-Miller, 2-opt+swap+relocate; 256 ants, "
                  f"OpenMP over ants) in {ref_el:.2f}s",
        "nproc": nproc, "affinity_cpus": affinity,
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": cpu_model(),
        "one_thread": {"value": round(one_rate, 3),
                       "sample": f"{one_its} iterations of {one_m} ants in {one_el:.2f}s"},
        "per_core_times_nproc_estimate": round(one_rate * (nproc or 1), 1),
        "knn_comparator": {"value": round(knn_rate, 2), "cores": threads,
                           "sample": f"{knn_its} iterations (k=32 candidate roulette, Philox) in "
                                     f"{knn_el:.2f}s"},
    }


def leg_clustered(ca, instances, args):
    inst = instances.clustered(args.seed, customers=CLUSTERED["customers"],
                               nclust=CLUSTERED["clusters"], sigma=CLUSTERED["sigma"],
                               routes=CLUSTERED["routes"])
    m, k = CLUSTERED["m"], CLUSTERED["k"]
    mk = lambda: ca.GpuACO(inst, n_ants=m, k=k, seed=args.seed + 1, selection=ca.SEL_KNN,
                           rng=ca.RNG_PHILOX, ls_flags=CLUSTERED["ls"], use_graph=True)
    g = mk()
    g.init_aco()
    tune = g.autotune(reps=3)
    shape = g.launch_shape()
    g.run(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(args.clustered_steps)
    g.synchronize()
    el = time.perf_counter() - t0
    if shape[5]:
        g.set_launch_shape(*shape[:5], 0)
    kern, per, _ = profile_iterations(g, inst, m, k, 10)
    g.close()
    out = {"workload": "481 nodes: 480 customers in 8 Gaussian clusters (sigma 6), random depot, "
                       "Q for 30 routes; 480 ants, k=30 Philox roulette, 2-opt+swap+relocate",
           "ant_solutions_per_s": round(m * args.clustered_steps / el, 1),
           "ms_per_iteration": round(1e3 * el / args.clustered_steps, 4),
           "launch_shape": shape, "kernels": per,
           "autotune_ms": {f"{STAGES[st]}{sh}": round(ms, 3) for sh, st, ms in tune}}
    if args.full_iterations > 0:
        g = mk()
        g.set_launch_shape(*shape)
        t0 = time.perf_counter()
        g.init_aco()
        g.run(2000)
        g.synchronize()
        wall = time.perf_counter() - t0
        rec = g.bsf_records()
        g.close()
        it, t = time_to_gap(rec, rec[-1][1])
        out["quality"] = {"iterations": 2000, "cstar": rec[-1][1], "wall_s": round(wall, 3),
                          "time_to_2pct_gap_s": round(t, 4), "iter_to_2pct_gap": it}
    log(f"clustered 481: {out['ant_solutions_per_s']:.0f} ant-sol/s")
    return out


def leg_decompose(ca, instances, args):
    """3001 nodes, past the local-search LDS limit: master colony without
    local search (256 ants, k=16) against the same master with a
    decomposition round after every iteration (n/50 sub-problems capped by
    the route count, 75 iterations of 64 ants with 2-opt+swap+relocate per
    round, all sub-colonies of a round in one batched context)."""
    from cvrp_aco_ocl_amd import decompose as dc
    inst = instances.clustered(args.seed + 11, customers=3000, nclust=12, sigma=40.0, routes=40)
    it = args.decompose_iterations
    master = dict(n_ants=256, k=16, ls_flags=0, seed=args.seed + 4)
    t0 = time.perf_counter()
    with ca.GpuACO(inst, selection=ca.SEL_KNN, rng=ca.RNG_PHILOX, **master) as g:
        g.init_aco()
        g.run(it)
        g.synchronize()
        plain_len = g.best()[1]
    plain_wall = time.perf_counter() - t0
    res = dc.solve_decomposed(inst, it, sub_iterations=75, master=master, sub=dict(n_ants=64, k=16,
                                                                                   seed=args.seed + 6))
    rounds = res["rounds"]
    out = {"workload": "3001 nodes: 3000 customers in 12 Gaussian clusters (sigma 40), Q for 40 "
                       "routes; master 256 ants k=16 without local search (the local search's LDS "
                       "rows do not fit), a decomposition round after each iteration",
           "iterations": it, "plain_length": round(float(plain_len), 1), "plain_wall_s": round(plain_wall, 3),
           "decomposed_length": round(res["length"], 1), "decomposed_wall_s": round(res["wall_s"], 3),
           "sub_problems_per_round": rounds[0][2] if rounds else 0,
           "replaced_per_round": [r[3] for r in rounds],
           "wall_s_per_round": round((res["wall_s"] - plain_wall) / max(1, len(rounds)), 4),
           "valid": bool(inst.check_solution(res["tour"]))}
    log(f"decomposition 3001: plain {plain_len:.0f} -> decomposed {res['length']:.0f} "
        f"in {res['wall_s']:.2f}s ({len(rounds)} rounds)")
    return out


def leg_batch(ca, instances, args):
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
    out = {"workload": f"{len(binst)} instances, n {min(ns)}..{max(ns)} (uniform or 4-12 "
                       "clusters, demands 1..50, 5-40 routes), 64 ants each, k=16 Philox "
                       "roulette, 2-opt+swap+relocate, one batched context",
           "ant_solutions_per_s": round(len(binst) * 64 * K / el, 1),
           "ms_per_iteration": round(1e3 * el / K, 3), "create_and_init_s": round(create_s, 3),
           "launch_shape": bshape,
           "iteration_ms": {f"{'fused' if sh[5] else 'per-phase'}{sh}": round(ms, 3)
                            for sh, st, ms in tuned if st == 2 and ms > 0}}
    log(f"batch: {out['ant_solutions_per_s']:.0f} ant-sol/s, {out['ms_per_iteration']:.2f} ms/iter")
    return out


def leg_prior(instances, args):
    from cvrp_aco_ocl_amd import prior as PR, train_prior as TP
    net = PR.PriorNet.load()
    ev = instances.evaluation(args.prior_instances)
    log(f"prior leg: {len(ev)} instances, plain {args.prior_iterations} iterations")
    out = TP.compare(ev, args.prior_iterations, net, ants=64, k=16, exponent=TP.PRIOR_EXPONENT,
                     budgets=(0.5, 1.0, 2.0, 5.0, 10.0))
    out["workload"] = ("64 clustered instances (500 customers, 8 clusters, sigma 6, 30 routes; "
                       "seeds 100000+), 64 ants each, k=16 Philox roulette, 2-opt+swap+relocate, "
                       f"one batched context; C* = plain best after {args.prior_iterations} "
                       "iterations; prior time includes aco_prior_apply")
    net.close()
    log(f"prior leg: plain median {out['plain_time_to_2pct_s']['median']:.3f}s, prior median "
        f"{out['prior_time_to_2pct_s']['median']:.3f}s")
    return out


# ------------------------------------------------------------------- main ----

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
    if args.lds_probe:
        gbps, ms = lds_peak()
        print(json.dumps({"lds_read_GBps": round(gbps, 1), "best_launch_ms": round(ms, 4)}))
        return
    if args.roof_probe:
        print(json.dumps({"peaks_measured": measured_peaks()}))
        return
    if args.mlp_probe:
        mlp_probe(args.mlp_probe)
        return
    if args.ls_phases:
        ls_phases([int(v) for v in args.shape.split(",")] if args.shape else None)
        return
    import cvrp_aco_ocl_amd as ca
    from cvrp_aco_ocl_amd import instances

    inst = instances.mixed_1000(args.seed)
    m, k = WORKLOAD["m"], WORKLOAD["k"]
    log(f"workload {inst.name}: n={inst.n} Q={inst.capacity} m={m} k={k}")

    # ---- timed region: K iterations, graph replay ----
    g = ca.GpuACO(inst, n_ants=m, k=k, seed=args.seed + 1, selection=ca.SEL_KNN,
                  rng=ca.RNG_PHILOX, ls_flags=WORKLOAD["ls"], use_graph=True)
    g.init_aco()
    if args.shape:
        g.set_launch_shape(*[int(v) for v in args.shape.split(",")])
        tune = []
    else:
        tune = g.autotune(reps=3)   # launch shapes; results are the same for every shape
    shape = g.launch_shape()
    log(f"autotune -> {shape}")
    g.run(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(args.steps)
    g.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = m * args.steps / elapsed
    log(f"timed {args.steps} iterations: {ms_per_step:.3f} ms/iter, {value:.0f} ant-sol/s")

    # ---- per-kernel device time (per-phase launches, HIP events) ----
    if shape[5]:
        g.set_launch_shape(*shape[:5], 0)
    kern, per_kernel, ls_evals = profile_iterations(g, inst, m, k, args.profile_steps)
    g.close()
    dominant = max(kern, key=kern.get)
    peaks = measured_peaks()
    peak_lds = peaks["lds_read_GBps"]
    log(f"LDS read peak {peak_lds:.0f} GB/s; dominant kernel {dominant} {kern[dominant]:.3f} ms")
    if dominant == "local_search":
        # 16 B per evaluated move (the four distance / edge words of a 2-opt,
        # swap or relocate delta), all read from LDS
        alg = int(16 * ls_evals)
        achieved = alg / (kern[dominant] * 1e-3) / 1e9
        roofline = {"bound": "lds", "kernel": dominant, "achieved": round(achieved, 2),
                    "peak": round(peak_lds, 1), "unit": "GB/s", "frac": round(achieved / peak_lds, 5),
                    "traffic": measured_traffic(dominant, "mixed1001") if args.seed == 0 else None,
                    "algorithmic_bytes_per_launch": alg,
                    "peak_source": "LDS read bandwidth measured in this run (libcvrpprobe.so)"}
    else:
        alg = per_kernel[dominant].get("algorithmic_bytes", 0)
        achieved = alg / (kern[dominant] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                    "traffic": measured_traffic(dominant, "mixed1001") if args.seed == 0 else None,
                    "algorithmic_bytes_per_launch": alg}
    ph = per_kernel.get("pheromone")
    hbm_pass = None
    if ph:
        hbm_pass = {"bound": "hbm", "kernel": "pheromone", "achieved": ph["GBps"],
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ph["GBps"] / HBM_PEAK_GBS, 5),
                    "frac_of_measured_copy": round(ph["GBps"] / peaks["hbm_copy_GBps"], 5),
                    "note": "4 MB per matrix at this size: resident in the last-level cache"}

    quality = leg_quality(ca, inst, shape, args.full_iterations, args.seed) \
        if args.full_iterations > 0 else None
    ref_gpu = leg_reference_rules_gpu(ca, inst, args.reference_rules_steps, args.seed) \
        if args.reference_rules_steps > 0 else None
    cpu = leg_cpu(inst, args) if args.cpu_seconds > 0 else None
    if cpu is not None:
        cpu["gpu_over_cpu"] = {"candidate_roulette_gpu_vs_reference_rules_cpu": round(value / cpu["value"], 2),
                               "candidate_roulette_gpu_vs_candidate_roulette_cpu":
                                   round(value / cpu["knn_comparator"]["value"], 2)}
        if ref_gpu is not None:
            cpu["gpu_over_cpu"]["reference_rules_gpu_vs_reference_rules_cpu"] = round(
                ref_gpu["ant_solutions_per_s"] / cpu["value"], 2)
    clustered = leg_clustered(ca, instances, args) if args.clustered_steps > 0 else None
    batch = leg_batch(ca, instances, args) if args.batch_instances > 0 else None
    decomposed = leg_decompose(ca, instances, args) if args.decompose_iterations > 0 else None
    prior_leg = leg_prior(instances, args) if args.prior_instances > 0 else None

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
            "workload": "mixed-1001: 1000 customers (500 uniform, 500 in 10 Gaussian clusters "
                        "sigma 25) in [0,1000]^2, central depot, demand U{1..100} with 5% in "
                        "[80,100], Q for ~12 customers per route, nint int32 distances; m=256, "
                        "k=32 roulette_knn (Philox), 2-opt+swap+relocate, alpha 1 beta 2 rho 0.1, "
                        "6 ranks",
            "n": inst.n, "ants": m, "k": k, "capacity": inst.capacity, "seed": args.seed,
            "step": "one ACO iteration (hipGraph replay)",
            "launch_shape": {"ls_waves": shape[0], "ls_ring": shape[1], "ls_prefetch": shape[2],
                             "construct_ants_per_block": shape[3], "ls_row16": shape[4],
                             "fused_iteration": shape[5]},
        },
        "roofline": roofline,
        "roofline_pheromone_pass": hbm_pass,
        "peaks_measured": peaks,
        "kernels": per_kernel,
        "ls_move_evals_per_iter": int(ls_evals),
        "autotune_ms": {f"{STAGES[st]}{sh}": round(ms, 3) for sh, st, ms in tune},
        "quality": quality,
        "reference_rules_gpu": ref_gpu,
        "cpu_baseline": cpu,
        "clustered481": clustered,
        "decomposed3001": decomposed,
        "batch256": batch,
        "learned_prior": prior_leg,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
