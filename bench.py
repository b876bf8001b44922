#!/usr/bin/env python3
"""Benchmark: batched SEM time stepping on one MI355X.

Workloads (--workload):
  cfg4 (default)  3-D elastic (Elastic3D), order-4 hexes, 32 shots on four
                  meshes of 32^3, 48^3, 64^3 and 96^3 elements (8 shots each,
                  seeded moment-tensor Ricker sources at 15 Hz, 3 surface
                  receivers per shot), seeded smooth vp/vs/rho, vp/vs = 1.73:
                  2.0 G DOF-updates per step
  cfg3            3-D acoustic, 128^3 order-4 hexes (135 M nodes), layered vp,
                  Ricker 20 Hz, 1024 surface receivers, one shot
  cfg2            2-D elastic, 500^2 order-4 quads (8 M DOFs), vertical force,
                  100 receivers, one shot
A "step" advances every shot by one explicit Newmark time step (stiffness +
sources + mass inverse + update + receivers).

value        DOF-updates/s of the batched path: every shot in one set of tile
             kernel launches per colour, hipGraph-captured step blocks, fields
             resident in HBM, K timed steps
plain        the same shots through the plain one-shot-per-launch path
             (gather / stiffness / scatter / update kernels per shot)
accuracy     receiver and final-wavefield relative L2 of both FP32 HIP paths
             against FP64 oracle runs of the workload's full length
             (committed fixtures: the 32^3 meshes of cfg4 / cfg3, cfg2 at
             full size)
roofline     the tile kernel's algorithmic HBM bytes per launch over its
             average duration (HIP events on the batch stream); traffic from
             the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the
             same workload when present
cpu_baseline the FP64 oracle (a port of the reference's CPU path) on the same
             mesh for a bounded number of steps, 1 thread and every host
             core, on this host
serving      (cfg4) a configuration-5 request stream through the
             continuous-batching server and the plain FIFO server

Prints one JSON line.
"""
import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "DOF-updates/s, batched vs plain FP32 @≤1e-5 rel-L2; p50/p99 shot latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def elem_floats(dim, physics):
    ng = dim * (dim + 1) // 2 if physics == 0 else dim * dim + 1
    kp = 1 if physics == 0 else (4 if dim == 2 else 6)
    return ng + (2 ** dim) * kp


def tile_bytes_per_step(M, S, dim, physics):
    """Algorithmic HBM bytes of one step of one group on the tile kernel.

    per tile      : 4 B per lattice slot word + the elements' affine factors
                    and vertex material
    per shot      : u gather, 4c B per mesh slot (tile halos included);
                    accumulator 4c B per slot read (not its first colour) and
                    per slot written (not its last colour); per node 4 B
                    inverse mass + c (8 B read w, kc + 12 B write u, w, kc),
                    w = v + dt/2 a_ being the Newmark state the update reads
                    and kc the packed pair of bf16 Kahan compensations of u
                    and w (a_, read out only, is stored once per graph block)
    """
    c = M.c
    per_tile = 4 * M.tile_slots + 4 * M.ntiles * M.tile_elems * elem_floats(dim, physics)
    per_shot = (4 * c * M.tile_valid_slots + 4 * c * (M.tile_acc_reads + M.tile_acc_writes)
                + M.G * (4 + 20 * c))
    return per_tile + S * per_shot


def element_bytes_per_step(M, S, dim, physics):
    """Algorithmic HBM bytes of one step of one group in the element-by-element
    formulation (the operator's own data flow, independent of tiling):

    per element : n^d (4 B gather index + 12c S B: u gather, a scatter
                  read-modify-write) + the element's factors, here 4 B x
                  (affine factors + 2^d vertex material) per element instead of
                  per GLL point
    per node    : S (4 B inverse mass + c (4 B a + 16 B read u, v, a_, kc
                  + 16 B write)), kc the packed Kahan compensations
    """
    c, nd = M.c, M.nd
    return M.E * (nd * (4 + 12 * c * S) + 4 * elem_floats(dim, physics)) + M.G * S * (4 + 36 * c)


def setup(sem, workloads, name, nsteps, sizes, shots_per_mesh):
    """(dim, physics, meshes, shot specs, description) of a workload; a shot
    spec is (mesh index, dt, sources, receivers)."""
    if name == "cfg4":
        W = workloads.cfg4(sizes=sizes, shots_per_mesh=shots_per_mesh, nsteps=nsteps)
        meshes = [sem.Mesh(3, 4, 1, mw["cells"], mw["coords"], mw["par"]) for mw in W["meshes"]]
        shots = [(i, mw["dt"], [s], mw["receivers"])
                 for i, mw in enumerate(W["meshes"]) for s in mw["shots"]]
        return 3, 1, meshes, shots, W["name"] + ", order 4, moment-tensor Ricker 15 Hz, " \
                                                "3 receivers per shot"
    W = workloads.cfg3(nsteps=nsteps) if name == "cfg3" else workloads.cfg2(nsteps=nsteps)
    meshes = [sem.Mesh(W["dim"], W["order"], W["physics"], W["cells"], W["coords"], W["par"])]
    return W["dim"], W["physics"], meshes, [(0, W["dt"], W["sources"], W["receivers"])], \
        W["name"] + ", %d receivers" % len(W["receivers"])


def make_batch(sem, meshes, shots, path, nsteps, spg=0):
    B = sem.Batch()
    ids = []
    for mi, dt, sources, receivers in shots:
        k = B.add_shot(meshes[mi], nsteps, dt)
        for s in sources:
            if s["kind"] == "moment":
                B.add_moment_source(k, s["loc"], s["M"], s["amp"], s["f0"], s["t0"])
            else:
                B.add_source(k, s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
        for r in receivers:
            B.add_receiver(k, r)
        ids.append(k)
    B.prepare(path, 0, spg)
    return B, ids


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def profiled_traffic(name):
    """HBM bytes per tile-kernel launch from the committed rocprofv3 summary
    of this workload (FETCH_SIZE x fetch factor + WRITE_SIZE x write factor),
    or None."""
    p = ROOT / "profiles" / ("tile_kernel_%s_pmc.txt" % name)
    if not p.exists():
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="cfg4", choices=("cfg4", "cfg3", "cfg2"))
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--sizes", default="32,48,64,96", help="cfg4 mesh sizes")
    ap.add_argument("--shots", type=int, default=8, help="cfg4 shots per mesh")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="pick each mesh's tile shape with the autotuner (default: the "
                         "library's default shape, so that every tile launch runs one kernel)")
    ap.add_argument("--spg", type=int, default=0, help="shots per group (0: autotuned)")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=400.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs one process on one MI355X; "
              "--gpus %d is not supported (use --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    # graph blocks of the timed steps (at most 50): the block graph is
    # captured when the batch is prepared and the warm-up runs its own tail
    # graph, so the timed region is whole-block replays
    os.environ["SEM_GRAPH_STEPS"] = str(max(1, min(50, args.steps)))
    from salvus_amd import sem, workloads

    timing_steps = 10
    total = args.warmup + args.steps + timing_steps
    sizes = tuple(int(x) for x in args.sizes.split(","))
    t0 = time.perf_counter()
    dim, physics, meshes, shots, desc = setup(sem, workloads, args.workload, total, sizes,
                                              args.shots)
    per_mesh = [sum(1 for s in shots if s[0] == i) for i in range(len(meshes))]
    if args.tune:  # tile shape (elements per workgroup) per mesh, timed on
        for M, k in zip(meshes, per_mesh):  # as many shots as a group will hold
            M.autotune(reps=5, shots=max(1, min(k, args.spg or 4)))
    log("meshes ready (%.1f s): %s" % (time.perf_counter() - t0,
                                        ", ".join("E=%d tiles %s" % (M.E, M.tile_shape)
                                                  for M in meshes)))
    dofs = sum(M.c * M.G * k for M, k in zip(meshes, per_mesh))

    # ---------------- batched path
    B, ids = make_batch(sem, meshes, shots, sem.PATH_FUSED, total, args.spg)
    cfg = B.config()
    lanes = B.lanes()
    log("batched path prepared: %s, %d lane(s)" % (cfg, lanes))
    B.run(args.warmup)
    t0 = time.perf_counter()
    B.run(args.steps)
    fused_s = time.perf_counter() - t0
    value = dofs * args.steps / fused_s
    log("batched: %.2f ms/step, %.3g DOF-updates/s" % (fused_s * 1e3 / args.steps, value))

    # kernel timing: eager launches bracketed by HIP events on the batch stream
    os.environ["SEM_TIME_KERNELS"] = "1"
    B.run(timing_steps)
    os.environ["SEM_TIME_KERNELS"] = "0"
    kms, nlaunch = B.kernel_time()
    launches_per_step = nlaunch / timing_steps
    avg_launch_s = kms / 1e3 / max(nlaunch, 1)
    S = max(per_mesh)
    groups = [(M, k) for M, k in zip(meshes, per_mesh)]
    step_bytes = sum(element_bytes_per_step(M, k, dim, physics) for M, k in groups)
    per_launch_bytes = step_bytes / launches_per_step
    achieved = per_launch_bytes / avg_launch_s / 1e9
    tile_launch_bytes = sum(tile_bytes_per_step(M, k, dim, physics) for M, k in groups) \
        / launches_per_step
    tile_achieved = tile_launch_bytes / avg_launch_s / 1e9
    B.close()
    del B
    traffic = profiled_traffic(args.workload)
    shape = ",".join(map(str, meshes[-1].tile_shape or ()))
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
        "dtype": "f32",
        "data": "synthetic (seeded models and sources)",
        "config": {"workload": desc, "meshes": [int(M.E) for M in meshes], "shots": len(ids),
                   "dof_updates_per_step": dofs,
                   "path": "batched (%s)" % ("tile kernel" if cfg["tiled_groups"] == cfg["groups"]
                                             else "element kernel" if cfg["tiled_groups"] == 0
                                             else "tile and element kernels"),
                   "tile_shapes": [list(M.tile_shape) if M.tile_shape else None for M in meshes],
                   "shots_per_group": cfg["shots_per_group"], "groups": cfg["groups"],
                   "graph_block_steps": cfg["block_steps"], "lanes": lanes},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ("k_tile<%d,4,%s,%s>" % (dim, ("acoustic", "elastic")[physics], shape)
                                if cfg["tiled_groups"] else
                                "k_fused<%d,4,%s,%s>" % (dim, ("acoustic", "elastic")[physics],
                                                         "affine" if meshes[-1].affine else "pointwise")),
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "bytes_model": "element formulation: per element n^d (4 + 12 c S) + 4 x "
                                    "(affine factors + vertex material); per node S (4 + 36 c)",
                     "tile_bytes_per_launch": tile_launch_bytes,
                     "tile_achieved_GBs": tile_achieved,
                     "tile_frac": tile_achieved / HBM_PEAK_GBS,
                     "note": "achieved and frac price the element formulation's bytes (an "
                             "equivalent bandwidth, which can exceed the peak); the tile kernel "
                             "moves tile_bytes_per_launch, and HBM carries traffic per launch",
                     "avg_launch_us": avg_launch_s * 1e6,
                     "launches_per_step": launches_per_step,
                     "launch_note": "a launch is one colour of one kernel shape: with several "
                                    "lanes, its per-lane dispatches run side by side and count once",
                     "kernel_ms_per_step": kms / timing_steps},
    }
    if traffic is not None:
        out["roofline"]["traffic_source"] = "profiles/tile_kernel_%s_pmc.txt" % args.workload

    # ---------------- plain path
    if not args.no_plain:
        Bp, _ = make_batch(sem, meshes, shots, sem.PATH_PLAIN, 2 + args.plain_steps)
        Bp.run(2)
        t0 = time.perf_counter()
        Bp.run(args.plain_steps)
        plain_s = time.perf_counter() - t0
        plain_value = dofs * args.plain_steps / plain_s
        Bp.close()
        del Bp
        out["plain"] = {"value": plain_value, "steps": args.plain_steps,
                        "ms_per_step": plain_s * 1e3 / args.plain_steps}
        out["batched_vs_plain"] = value / plain_value
        log("plain: %.2f ms/step, %.3g DOF-updates/s" % (plain_s * 1e3 / args.plain_steps,
                                                         plain_value))
    for M in meshes:
        M.close()

    # ---------------- accuracy at full length + CPU baseline (FP64 oracle)
    if not args.no_cpu:
        from oracle import pyoracle as po
        from tests.golden.make_fullrun_fixtures import input_digest, problem
        from tests.problems import rel_l2
        fx_name = {"cfg4": "cfg4-32", "cfg3": "cfg3-32", "cfg2": "cfg2"}[args.workload]
        F = np.load(ROOT / "tests" / "golden" / ("fullrun_%s.npz" % fx_name))
        fdim, fphys, fcells, fcoords, fpar, fdt, fsteps, fshots, frecs, _ = problem(fx_name)
        assert input_digest(fcells, fcoords, fpar) == str(F["digest"]), "fixture inputs differ"
        Mf = sem.Mesh(fdim, 4, fphys, fcells, fcoords, fpar)
        errs, field_errs = {}, {}
        for name, path in (("batched", sem.PATH_FUSED), ("plain", sem.PATH_PLAIN)):
            Bs, ks = make_batch(sem, [Mf], [(0, fdt, srcs, frecs) for srcs in fshots], path, fsteps)
            Bs.run(fsteps)
            errs[name] = max(rel_l2(Bs.seismograms(k), F["seis%d" % i].astype(np.float64))
                             for i, k in enumerate(ks))
            field_errs[name] = max(rel_l2(Bs.get_field(k, "u").ravel()[::int(F["stride"])],
                                          F["u_sample%d" % i]) for i, k in enumerate(ks))
            Bs.close()
        Mf.close()
        out["accuracy"] = {"receiver_rel_l2_batched": errs["batched"],
                           "receiver_rel_l2_plain": errs["plain"],
                           "final_u_rel_l2_batched": field_errs["batched"],
                           "final_u_rel_l2_plain": field_errs["plain"], "bar": 1e-5,
                           "sample": "against FP64 oracle runs of the full length (fixture "
                                     "tests/golden/fullrun_%s.npz): %s, %d steps, %d shot(s), "
                                     "worst shot; final u on every %d-th DOF"
                                     % (fx_name, fx_name, fsteps, len(fshots), int(F["stride"]))}
        log("accuracy vs FP64 over %d steps: receivers %.2e / %.2e (batched / plain)"
            % (fsteps, errs["batched"], errs["plain"]))
        # CPU baseline: the oracle on this workload's mesh (the 32^3 sub-box
        # for the 3-D workloads, the full 500^2 mesh for cfg2), a bounded
        # number of steps, 1 thread and every host core
        ksteps = {"cfg4": 15, "cfg3": 40, "cfg2": 50}[args.workload]
        src0 = fshots[0]
        times = {}
        nthr = min(os.cpu_count() or 1, 16)
        for nt in (1, nthr):
            O = po.Sim(fdim, 4, fphys, fcells, fcoords, fpar)
            for sp in src0:
                if sp["kind"] == "moment":
                    O.add_moment_source(sp["loc"], sp["M"], sp["amp"], sp["f0"], sp["t0"])
                else:
                    O.add_source(sp["loc"], sp["amp"], sp["f0"], sp["t0"], sp["dir"])
            for r in frecs:
                O.add_receiver(r)
            t0 = time.perf_counter()
            O.run(fdt, ksteps, nthreads=nt)
            times[nt] = time.perf_counter() - t0
            dof_s = O.c * O.G * ksteps
            del O
        desc = "FP64 oracle, %s mesh (%d elements, %d DOFs), %d steps of shot 0" % (
            fx_name, len(fcells), dof_s // ksteps, ksteps)
        out["cpu_baseline"] = {"value": dof_s / times[1], "unit": "DOF-updates/s", "cores": 1,
                               "kind": "port", "sample": "%s (%.1f s)" % (desc, times[1])}
        out["cpu_baseline_all_cores"] = {"value": dof_s / times[nthr], "unit": "DOF-updates/s",
                                         "cores": nthr, "kind": "port",
                                         "sample": "the same, colour-parallel (%.1f s)" % times[nthr]}
        out["cpu_host"] = {"nproc": os.cpu_count(), "model": cpu_model()}
        log("CPU oracle: %.3g DOF-updates/s on 1 core, %.3g on %d" %
            (dof_s / times[1], dof_s / times[nthr], nthr))

    # ---------------- serving sample: continuous batching vs plain FIFO
    if args.workload == "cfg4" and not args.no_serve:
        from salvus_amd import serve
        specs, reqs = serve.draw_stream(args.serve_requests, args.serve_rate, 5,
                                        models_per_category=2, size_scale=0.5, steps=(50, 300))
        models = [serve.build_model(s["dim"], s["physics"], s["n"], s["seed"]) for s in specs]
        log("serving models: %s" % [(s["dim"], s["physics"], s["n"]) for s in specs])
        sb, _ = serve.run(models, reqs)
        sp, _ = serve.run(models, reqs, path=sem.PATH_PLAIN, max_active=1)
        out["serving"] = {
            "batched": sb, "plain": sp,
            "sample": "%d Poisson requests at %g/s, configuration-5 mix with sizes scaled 0.5 "
                      "and 50-300 steps, 2 models per category" % (len(reqs), args.serve_rate)}
        for mm in models:
            mm["mesh"].close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
