"""Serving workload (configuration 5) and its driver.

A Poisson stream of shot requests at a fixed arrival rate over a mix of
models: 40 % 2-D acoustic quads (100^2-800^2 elements), 30 % 2-D elastic
quads (200^2-1000^2) and 30 % 3-D acoustic/elastic hexes (32^3-96^3), order 4,
500-3000 steps each, sizes drawn log-uniformly from a seed.  Models are
device-resident meshes (a model pool per category, each size drawn
log-uniformly; requests pick a model of their category); a request is one
shot (Ricker force or moment-tensor source at a seeded position, 3 receivers).

run() serves the stream through sem.Server and reports per-request latency
(completion - arrival: p50, p99) and aggregate DOF-updates/s (all requests'
node-component updates over the makespan), for the batched server
(continuous batching, all active requests in the same launches) or the plain
one-shot-per-launch server (one request at a time, FIFO).

The plain server on a whole configuration-5 stream runs for tens of
minutes; --plain-model gives its latencies from a FIFO schedule of measured
per-model service times (set-up plus steps, each model timed alone on the
plain server), and --plain-prefix N checks that model against the live plain
server on the first N requests.

python -m salvus_amd.serve [--requests N] [--rate R] [--scale S] ...
"""
import argparse
import json
import time

import numpy as np

from . import meshgen, sem, workloads

CATEGORIES = (  # share, dim, physics (None: acoustic or elastic), size range
    (0.4, 2, 0, (100, 800)),
    (0.3, 2, 1, (200, 1000)),
    (0.3, 3, None, (32, 96)),
)
H2, H3 = 10.0, 15.0


def _loguniform(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def build_model(dim, physics, n, seed):
    """Mesh, dt and receivers of one model (seeded smooth material)."""
    rng = np.random.default_rng(seed)
    if dim == 2:
        L = n * H2
        cells, coords = meshgen.box_quad(n, n, L, L)
        d = 0.05 * workloads.gaussian_field((n + 1, n + 1), 50.0 / H2, rng).ravel()
        vp = 3000.0 * (1 + d)
        if physics == 0:
            par = vp[cells][..., None]
        else:
            vs, rho = vp / 1.732, 2500.0 * (1 + d)
            c11 = rho * vp ** 2
            par = np.stack([rho, c11, rho * (vp ** 2 - 2 * vs ** 2), c11, rho * vs ** 2],
                           axis=1)[cells]
        dt = 0.4 * H2 * workloads.GLL4_MIN_SPACING / vp.max()
        recs = [(0.25 * L, 0.75 * L), (0.5 * L, 0.75 * L), (0.75 * L, 0.75 * L)]
    else:
        if physics == 0:
            L = n * H3
            cells, coords = meshgen.box_hex(n, n, n, L, L, L)
            d = 0.03 * workloads.gaussian_field((n + 1,) * 3, 100.0 / H3, rng).ravel()
            vp = (2000.0 + 3000.0 * coords[:, 2] / L) * (1 + d)
            par = vp[cells][..., None]
            dt = 0.4 * H3 * workloads.GLL4_MIN_SPACING / vp.max()
        else:
            cells, coords, par, dt, L = workloads.elastic3d_mesh(n, H3, seed)
        recs = [(0.5 * L, 0.5 * L, L), (0.25 * L, 0.75 * L, L), (0.75 * L, 0.25 * L, L)]
    M = sem.Mesh(dim, 4, physics, cells, coords, par)
    return dict(mesh=M, dim=dim, physics=physics, n=n, L=L, dt=dt, recs=recs)


def draw_stream(nreq, rate, seed, models_per_category=4, size_scale=1.0, steps=(500, 3000)):
    """Model specs and the request stream (arrival times in seconds)."""
    rng = np.random.default_rng(seed)
    specs = []
    for ci, (share, dim, phys, (lo, hi)) in enumerate(CATEGORIES):
        for k in range(models_per_category):
            p = phys if phys is not None else int(k % 2)
            n = max(2, int(round(_loguniform(rng, lo, hi) * size_scale)))
            specs.append(dict(category=ci, dim=dim, physics=p, n=n, seed=seed * 100 + len(specs)))
    shares = np.array([c[0] for c in CATEGORIES])
    t = 0.0
    reqs = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate)
        ci = int(rng.choice(len(CATEGORIES), p=shares))
        models = [j for j, s in enumerate(specs) if s["category"] == ci]
        reqs.append(dict(model=int(rng.choice(models)), arrival=t,
                         nsteps=int(rng.integers(steps[0], steps[1] + 1)),
                         pos=rng.uniform(0.3, 0.7, 3), mt=rng.normal(0.0, 1.0, (3, 3)),
                         seed=int(rng.integers(1 << 30))))
    return specs, reqs


def _sources(model, req):
    L, dim = model["L"], model["dim"]
    loc = tuple(req["pos"][:dim] * L)
    f0 = 3000.0 / (8 * (H2 if dim == 2 else H3)) / 2
    if model["physics"] == 1 and dim == 3:
        M = 0.5 * (req["mt"] + req["mt"].T) * 1e15
        return [dict(kind="moment", loc=loc, M=M, amp=1.0, f0=f0, t0=1.2 / f0)]
    c = 1 if model["physics"] == 0 else dim
    direction = (0.0, 1.0)[:c] if c == 2 else (1.0,)
    return [dict(kind="force", loc=loc, amp=1e6, f0=f0, t0=1.2 / f0, dir=direction)]


def run(models, reqs, path=sem.PATH_FUSED, max_active=0, block_steps=32, hbm_budget=0,
        progress=None):
    """Serves the stream in real time; returns per-request latencies and the
    aggregate throughput.  progress(status) is called every ~10 s."""
    S = sem.Server(path=path, hbm_budget=hbm_budget, max_active=max_active,
                   block_steps=block_steps)
    t_start = S.clock()
    ids = []
    for r in reqs:
        m = models[r["model"]]
        ids.append(S.submit(m["mesh"], r["nsteps"], m["dt"], _sources(m, r), m["recs"],
                            arrival=t_start + r["arrival"]))
    S.drain(progress=progress)
    lat, dofs, t_last, grouped = [], 0, t_start, []
    for r, i in zip(reqs, ids):
        info = S.request(i)
        if info["state"] != "done":
            raise sem.SemError("request %d ended in state %s" % (i, info["state"]))
        lat.append(info["done"] - info["arrival"])
        grouped.append(info["group_shots"])
        m = models[r["model"]]["mesh"]
        dofs += m.c * m.G * r["nsteps"]
        t_last = max(t_last, info["done"])
    results = [S.result(i) for i in ids]
    S.close()
    lat = np.array(lat)
    span = t_last - (t_start + reqs[0]["arrival"])
    grouped = np.array(grouped)
    return dict(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), dof_updates_per_s=dofs / span, makespan_s=span,
                requests=len(reqs), dof_updates=dofs,
                in_multi_shot_groups=int((grouped > 1).sum()),
                mean_group_shots=float(grouped.mean())), results


def plain_service_times(models, reqs, short=(100, 400)):
    """Measured service time of every model on the plain one-shot-per-launch
    server: one request alone at two step counts (admission, set-up, steps,
    retirement), fitted as a + b * nsteps per model."""
    fit = {}
    for mi in sorted({r["model"] for r in reqs}):
        m = models[mi]
        r0 = next(r for r in reqs if r["model"] == mi)
        t = []
        for n in short:
            S = sem.Server(path=sem.PATH_PLAIN, max_active=1)
            t0 = S.clock()
            i = S.submit(m["mesh"], n, m["dt"], _sources(m, r0), m["recs"], arrival=t0)
            S.drain()
            t.append(S.request(i)["done"] - t0)
            S.close()
        b = (t[1] - t[0]) / (short[1] - short[0])
        fit[mi] = (t[0] - b * short[0], b)
    return fit


def fifo_schedule(models, reqs, fit):
    """Latencies of a one-at-a-time FIFO server with the fitted service
    times: start = max(arrival, previous finish), finish = start + a + b n."""
    lat, dofs, t_free = [], 0, 0.0
    for r in reqs:
        a, b = fit[r["model"]]
        start = max(r["arrival"], t_free)
        t_free = start + a + b * r["nsteps"]
        lat.append(t_free - r["arrival"])
        m = models[r["model"]]["mesh"]
        dofs += m.c * m.G * r["nsteps"]
    lat = np.array(lat)
    span = t_free - reqs[0]["arrival"]
    return dict(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), dof_updates_per_s=dofs / span, makespan_s=span,
                requests=len(reqs), dof_updates=dofs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=20.0, help="arrivals per second")
    ap.add_argument("--scale", type=float, default=1.0, help="model size scale")
    ap.add_argument("--steps", default="500,3000")
    ap.add_argument("--models", type=int, default=4, help="models per category")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--plain", action="store_true", help="also serve through the plain path")
    ap.add_argument("--plain-model", action="store_true",
                    help="plain FIFO server on the whole stream from measured per-model "
                         "service times")
    ap.add_argument("--plain-prefix", type=int, default=0,
                    help="serve the first N requests live through the plain server and "
                         "compare with the service-time model on the same N")
    ap.add_argument("--no-batched", action="store_true")
    args = ap.parse_args()
    lo, hi = (int(x) for x in args.steps.split(","))
    specs, reqs = draw_stream(args.requests, args.rate, args.seed, args.models, args.scale,
                              (lo, hi))
    models = [build_model(s["dim"], s["physics"], s["n"], s["seed"]) for s in specs]

    def progress(tag):
        return lambda st: print("[serve %s] done %d, active %d, pending %d" %
                                (tag, st["done"], st["active"], st["pending"]), flush=True)
    out = {}
    if not args.no_batched:
        out["batched"] = run(models, reqs, progress=progress("batched"))[0]
        print("[serve] batched " + json.dumps(out["batched"]), flush=True)
    if args.plain:
        out["plain"] = run(models, reqs, path=sem.PATH_PLAIN, max_active=1,
                           progress=progress("plain"))[0]
    if args.plain_model or args.plain_prefix:
        fit = plain_service_times(models, reqs)
        out["plain_service_fit"] = {str(k): dict(setup_s=v[0], step_s=v[1]) for k, v in fit.items()}
        if args.plain_model:
            out["plain_model"] = fifo_schedule(models, reqs, fit)
            print("[serve] plain model " + json.dumps(out["plain_model"]), flush=True)
        if args.plain_prefix:
            pre = reqs[:args.plain_prefix]
            out["plain_prefix"] = dict(
                measured=run(models, pre, path=sem.PATH_PLAIN, max_active=1,
                             progress=progress("plain prefix"))[0],
                model=fifo_schedule(models, pre, fit))
            print("[serve] plain prefix " + json.dumps(out["plain_prefix"]), flush=True)
    out["models"] = [dict(dim=s["dim"], physics=s["physics"], n=s["n"]) for s in specs]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
