#!/usr/bin/env python3
"""Benchmark: SEM time stepping on one MI355X.

Workload (default): configuration 2 -- 2-D elastic (Elastic2D) on 500x500
order-4 quads of 10 m, 8,008,002 DOFs, vertical Ricker source, 100
receivers, dt = 2e-4.  A "step" is one explicit Newmark time step of the
whole wavefield (stiffness + source + mass inverse + update + receivers).

value      DOF-updates/s of the fused path (batched colour kernels in hipGraph
           step blocks), inputs resident in HBM, K timed steps
plain      the same K steps through the plain one-shot-per-launch path
accuracy   receiver relative L2 of the FP32 HIP path against the FP64 oracle
           on a bounded sample of the same workload (same h, material law,
           source and receiver line on a smaller square)
roofline   the fused kernel's algorithmic HBM bytes per launch over its
           average duration (HIP events on the batch stream)
cpu_baseline  the FP64 oracle (a port of the reference's CPU path) on the
           bounded sample, 1 thread, on this host

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "DOF-updates/s, batched vs plain FP32 @≤1e-5 rel-L2; p50/p99 shot latency"
HBM_PEAK_GBS = 8000.0


def build_problem(sem, W, path):
    M = sem.Mesh(W["dim"], W["order"], W["physics"], W["cells"], W["coords"], W["par"])
    B = sem.Batch()
    k = B.add_shot(M, W["nsteps"], W["dt"])
    for s in W["sources"]:
        B.add_source(k, s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
    for r in W["receivers"]:
        B.add_receiver(k, r)
    B.prepare(path)
    return M, B, k


def fused_bytes_per_step(M, dim, physics, order):
    """Algorithmic HBM bytes of one fused step (all colour launches).

    per element : nd (4 B gather-map word + 4c B gathered u) + per-element
                  Jacobian factors + vertex material (affine meshes)
    per element-node beyond a node's first/last colour: 8c B accumulator
                  read-modify-write; first colour writes 4c, last reads 4c
    per node    : 4 B inverse mass + c (16 B read a_, v, kv, ku + 20 B write
                  u, v, a_, ku, kv); u comes from the gather
    """
    nd, c, E, G = M.nd, M.c, M.E, M.G
    nv = 2 ** dim
    kpar = 1 if physics == 0 else (4 if dim == 2 else 6)
    ng = dim * (dim + 1) // 2 if physics == 0 else dim * dim + 1
    per_elem = nd * (4 + 4 * c) + 4 * (ng + nv * kpar)
    acc = 8 * c * (E * nd - G)
    per_node = 4 + c * (16 + 20)
    return E * per_elem + acc + G * per_node


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--plain-steps", type=int, default=0, help="0 = same as --steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=500, help="elements per side")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs one process on one MI355X; "
              "--gpus %d is not supported (use --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from salvus_amd import sem, workloads

    total = args.warmup + args.steps
    W = workloads.cfg2(n=args.n, nsteps=total)
    G = (args.n * 4 + 1) ** 2
    dofs = 2 * G

    # ---------------- fused (batched) path
    M, B, k = build_problem(sem, W, sem.PATH_FUSED)
    B.run(args.warmup)
    t0 = time.perf_counter()
    B.run(args.steps)
    fused_s = time.perf_counter() - t0
    value = dofs * args.steps / fused_s
    cfg = B.config()
    seis_fused = B.seismograms(k)

    # kernel timing (eager launches bracketed by HIP events, same stream)
    W2 = workloads.cfg2(n=args.n, nsteps=60)
    M2, B2, _ = build_problem(sem, W2, sem.PATH_FUSED)
    B2.run(10)
    os.environ["SEM_TIME_KERNELS"] = "1"
    B2.run(40)
    os.environ["SEM_TIME_KERNELS"] = "0"
    kms, nlaunch = B2.kernel_time()
    step_kernel_s = kms / 1e3 / 40
    bytes_step = fused_bytes_per_step(M2, 2, 1, 4)
    per_launch_bytes = bytes_step / M2.ncolours
    avg_launch_s = kms / 1e3 / max(nlaunch, 1)
    achieved = per_launch_bytes / avg_launch_s / 1e9
    del B2, M2

    # ---------------- plain path
    psteps = args.plain_steps or args.steps
    Wp = workloads.cfg2(n=args.n, nsteps=args.warmup + psteps)
    Mp, Bp, kp = build_problem(sem, Wp, sem.PATH_PLAIN)
    Bp.run(min(args.warmup, 20))
    t0 = time.perf_counter()
    Bp.run(psteps)
    plain_s = time.perf_counter() - t0
    plain_value = dofs * psteps / plain_s
    del Bp, Mp

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fused_s * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Gaussian-correlated +-5% Elastic2D model)",
        "config": {"workload": W["name"] + ", vertical Ricker 30 Hz, 100 receivers, dt 2e-4",
                   "elements": len(W["cells"]), "order": 4, "dofs": dofs,
                   "path": "fused", "elems_per_block": cfg["elems_per_block"],
                   "graph_block_steps": cfg["block_steps"]},
        "plain": {"value": plain_value, "steps": psteps, "ms_per_step": plain_s * 1e3 / psteps},
        "batched_vs_plain": value / plain_value,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_fused<2,4,elastic,affine>",
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "avg_launch_us": avg_launch_s * 1e6,
                     "kernel_ms_per_step": step_kernel_s * 1e3},
    }

    # ---------------- accuracy sample + CPU baseline (FP64 oracle, 1 thread)
    if not args.no_cpu:
        from oracle import pyoracle as po
        from tests.problems import rel_l2
        Ws = workloads.cfg2(n=40, nsteps=500)
        S = po.Sim(2, 4, 1, Ws["cells"], Ws["coords"], Ws["par"])
        s = Ws["sources"][0]
        S.add_source(s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
        for r in Ws["receivers"]:
            S.add_receiver(r)
        t0 = time.perf_counter()
        ref = S.run(Ws["dt"], Ws["nsteps"])
        cpu_s = time.perf_counter() - t0
        Gs = (40 * 4 + 1) ** 2
        Ms, Bs, ks = build_problem(sem, Ws, sem.PATH_FUSED)
        Bs.run(Ws["nsteps"])
        err = rel_l2(Bs.seismograms(ks), ref)
        Mp2, Bp2, kp2 = build_problem(sem, Ws, sem.PATH_PLAIN)
        Bp2.run(Ws["nsteps"])
        err_p = rel_l2(Bp2.seismograms(kp2), ref)
        out["accuracy"] = {"receiver_rel_l2_fused": err, "receiver_rel_l2_plain": err_p,
                           "bar": 1e-5,
                           "sample": "cfg2 law on 40x40 quads (10 m), 500 steps, 100 receivers"}
        out["cpu_baseline"] = {"value": 2 * Gs * Ws["nsteps"] / cpu_s, "unit": "DOF-updates/s",
                               "cores": 1, "kind": "port",
                               "sample": "FP64 oracle, cfg2 law on 40x40 quads, 500 steps (%.1f s)"
                               % cpu_s}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
