"""Serving loop on the GPU: continuous batching of shot requests.

Bars: every request's seismograms are the same (bit for bit) alone and in a
mix of concurrently active requests on different meshes and physics;
requests arriving while others run are admitted between blocks and overlap
them on the device; admission respects the HBM budget; a served shot agrees
with the FP64 oracle within the run tolerance; the plain one-shot-per-launch
server agrees with the batched one within the same tolerance."""
import time

import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import sem
from tests import problems as pb

pytestmark = pytest.mark.gpu

TOL_RUN = 1e-5

# dim, physics, elements per side, steps
CASES = [(2, 0, 12, 90), (2, 1, 10, 70), (3, 0, 4, 60), (3, 1, 4, 50)]


def _problems():
    return [pb.make(d, ph, nel, order=4, seed=30 + i, nsteps=ns)
            for i, (d, ph, nel, ns) in enumerate(CASES)]


def _mesh(P):
    return sem.Mesh(P["dim"], P["order"], P["physics"], P["cells"], P["coords"], P["par"])


def _sources(P, scale=1.0):
    s = P["src"]
    return [dict(kind="force", loc=np.asarray(s["loc"]) * scale, amp=s["amp"], f0=s["f0"],
                 t0=s["t0"], dir=s["dir"])]


def _alone(M, P, scale=1.0, **kw):
    S = sem.Server(block_steps=16, **kw)
    r = S.submit(M, P["nsteps"], P["dt"], _sources(P, scale), P["recs"])
    S.drain()
    assert S.request(r)["state"] == "done"
    out = S.result(r)
    S.close()
    return out


def test_results_same_alone_and_in_mix():
    Ps = _problems()
    Ms = [_mesh(P) for P in Ps]
    shots = [(M, P, 1.0) for M, P in zip(Ms, Ps)] + [(Ms[0], Ps[0], 0.9), (Ms[3], Ps[3], 1.1)]
    alone = [_alone(M, P, f) for M, P, f in shots]
    S = sem.Server(block_steps=16)
    ids = [S.submit(M, P["nsteps"], P["dt"], _sources(P, f), P["recs"]) for M, P, f in shots]
    st = S.pump()
    assert st["active"] == len(shots)  # all admitted in the first block
    S.drain()
    for rid, a in zip(ids, alone):
        info = S.request(rid)
        assert info["state"] == "done" and info["steps"] == S.requests[rid][1]
        np.testing.assert_array_equal(S.result(rid), a)
    assert np.abs(alone[0]).max() > 0


def test_late_arrivals_join_running_requests():
    """A request arriving while another runs is admitted at the next block
    boundary and finishes while the first one is still active or after it,
    with the same result as alone."""
    Ps = _problems()
    M0, M2 = _mesh(Ps[0]), _mesh(Ps[2])
    ref = _alone(M2, Ps[2])
    S = sem.Server(block_steps=8)
    a = S.submit(M0, 400, Ps[0]["dt"], _sources(Ps[0]), Ps[0]["recs"])
    S.pump()
    b = S.submit(M2, Ps[2]["nsteps"], Ps[2]["dt"], _sources(Ps[2]), Ps[2]["recs"],
                 arrival=S.clock())
    S.drain()
    ia, ib = S.request(a), S.request(b)
    assert ib["admit"] > ia["admit"] and ib["admit"] < ia["done"]
    np.testing.assert_array_equal(S.result(b), ref)


def test_admission_by_hbm_budget():
    """With a budget that holds one request, requests run one after another;
    results are unchanged."""
    Ps = _problems()
    M = _mesh(Ps[2])
    probe = sem.Server(block_steps=16)
    r = probe.submit(M, Ps[2]["nsteps"], Ps[2]["dt"], _sources(Ps[2]), Ps[2]["recs"])
    need = probe.request(r)["bytes"]
    probe.close()
    ref = _alone(M, Ps[2])
    S = sem.Server(block_steps=16, hbm_budget=int(need * 1.5))
    ids = [S.submit(M, Ps[2]["nsteps"], Ps[2]["dt"], _sources(Ps[2]), Ps[2]["recs"])
           for _ in range(3)]
    peak = 0
    while True:
        st = S.pump()
        peak = max(peak, st["active"])
        assert st["bytes"] <= need * 1.5
        if st["active"] == 0 and st["pending"] == 0:
            break
    assert peak == 1
    infos = [S.request(i) for i in ids]
    assert infos[0]["done"] <= infos[1]["admit"] and infos[1]["done"] <= infos[2]["admit"]
    for i in ids:
        np.testing.assert_array_equal(S.result(i), ref)
    with pytest.raises(sem.SemError):
        S.submit(M, 10 ** 9, Ps[2]["dt"], _sources(Ps[2]), Ps[2]["recs"])  # over the budget


@pytest.mark.parametrize("case", [1, 3])
def test_served_shot_matches_oracle(case):
    P = _problems()[case]
    M = _mesh(P)
    O = po.Sim(P["dim"], P["order"], P["physics"], P["cells"], P["coords"], P["par"])
    s = P["src"]
    O.add_source(s["loc"], s["amp"], s["f0"], s["t0"], s["dir"])
    for r in P["recs"]:
        O.add_receiver(r)
    ref = O.run(P["dt"], P["nsteps"])
    got = _alone(M, P)
    assert pb.rel_l2(got, ref) < TOL_RUN
    plain = _alone(M, P, path=sem.PATH_PLAIN, max_active=1)
    assert pb.rel_l2(plain, ref) < TOL_RUN


def test_cancel_pending_and_active():
    """A cancelled active request leaves at once (its memory released), a
    cancelled pending one never starts; the others finish with their alone
    results."""
    Ps = _problems()
    Ms = [_mesh(P) for P in Ps]
    alone = [_alone(M, P) for M, P in zip(Ms, Ps)]
    S = sem.Server(block_steps=8)
    ids = [S.submit(M, P["nsteps"], P["dt"], _sources(P), P["recs"]) for M, P in zip(Ms, Ps)]
    late = S.submit(Ms[0], Ps[0]["nsteps"], Ps[0]["dt"], _sources(Ps[0]), Ps[0]["recs"],
                    arrival=S.clock() + 3600.0)
    st = S.pump()
    assert st["active"] == 4 and st["pending"] == 1
    held = st["bytes"]
    S.cancel(ids[1])
    S.cancel(late)
    st = S.pump()
    assert st["active"] == 3 and st["pending"] == 0
    assert st["bytes"] == held - S.request(ids[1])["bytes"]
    with pytest.raises(sem.SemError):
        S.cancel(ids[1])
    S.drain()
    assert S.request(ids[1])["state"] == "cancelled" and S.request(late)["state"] == "cancelled"
    assert S.request(late)["admit"] < 0
    for k in (0, 2, 3):
        assert S.request(ids[k])["state"] == "done"
        np.testing.assert_array_equal(S.result(ids[k]), alone[k])


def test_multi_shot_groups():
    """Arrived requests on one mesh with step counts within the spread are
    admitted as one group and share its launches; a shorter one is done at
    its own count; every result equals the request served alone, and equals
    a server that keeps every request alone."""
    Ps = _problems()
    P, Q = Ps[1], Ps[3]
    M, MQ = _mesh(P), _mesh(Q)
    # (mesh, problem, steps, source scale): the first three group (70, 66, 72:
    # within 10 % of 72), 40 steps is outside the spread, the hex is on its own mesh
    shots = [(M, P, 70, 1.0), (M, P, 66, 0.9), (MQ, Q, 50, 1.0), (M, P, 72, 1.1), (M, P, 40, 1.0)]
    alone = [_alone(m, dict(p, nsteps=n), f) for m, p, n, f in shots]
    out = {}
    for cap in (4, 1):
        S = sem.Server(block_steps=8, shots_per_group=cap)
        ids = [S.submit(m, n, p["dt"], _sources(p, f), p["recs"]) for m, p, n, f in shots]
        st = S.pump()
        assert st["active"] == len(shots)
        groups = [S.request(i)["group_shots"] for i in ids]
        assert groups == ([3, 3, 1, 3, 1] if cap == 4 else [1] * 5)
        S.drain()
        infos = [S.request(i) for i in ids]
        for i, inf in zip(ids, infos):
            assert inf["state"] == "done" and inf["steps"] == S.requests[i][1]
        if cap == 4:
            # the 66-step request is copied out before its 72-step group ends
            assert infos[1]["done"] <= infos[3]["done"]
        out[cap] = [S.result(i) for i in ids]
        S.close()
    for k in range(len(shots)):
        np.testing.assert_array_equal(out[4][k], alone[k])
        np.testing.assert_array_equal(out[1][k], alone[k])
    assert np.abs(alone[1]).max() > 0


def test_cancel_inside_a_group():
    """Cancelling one request of a multi-shot group leaves the others'
    results unchanged; the group's memory goes when its last request ends."""
    P = _problems()[1]
    M = _mesh(P)
    alone = _alone(M, dict(P, nsteps=64))
    S = sem.Server(block_steps=8)
    a = S.submit(M, 64, P["dt"], _sources(P), P["recs"])
    b = S.submit(M, 64, P["dt"], _sources(P, 0.9), P["recs"])
    st = S.pump()
    assert S.request(a)["group_shots"] == 2 and st["active"] == 2
    held = st["bytes"]
    S.cancel(b)
    st = S.pump()
    assert st["active"] == 1 and st["bytes"] == held  # shared with a: kept until the group ends
    S.drain()
    assert S.request(b)["state"] == "cancelled"
    np.testing.assert_array_equal(S.result(a), alone)
    assert S.pump()["bytes"] == 0


def _full_size_models():
    """One model per configuration-5 category at the top of its size range:
    800^2 acoustic quads, 1000^2 elastic quads, 96^3 elastic hexes."""
    from salvus_amd import serve
    return [serve.build_model(2, 0, 800, 501), serve.build_model(2, 1, 1000, 502),
            serve.build_model(3, 1, 96, 503)]


def test_full_size_requests_same_alone_and_in_mix():
    """Configuration 5 at its largest sizes: the three requests, arriving one
    after another while the earlier ones run, share the device; each result
    equals the same request served alone, bit for bit."""
    from salvus_amd import serve
    models = _full_size_models()
    rng = np.random.default_rng(5)
    # sources next to each model's first receiver, so that a few dozen steps
    # already record the wave
    reqs = []
    for k, n in enumerate((96, 80, 64)):
        m = models[k]
        pos = np.zeros(3)
        pos[:m["dim"]] = np.asarray(m["recs"][0]) / m["L"] - 0.4 / m["n"]
        reqs.append(dict(model=k, arrival=0.0, nsteps=n, pos=pos, mt=rng.normal(0.0, 1.0, (3, 3)),
                         seed=k))
    alone = []
    for r in reqs:
        m = models[r["model"]]
        S = sem.Server(block_steps=16)
        i = S.submit(m["mesh"], r["nsteps"], m["dt"], serve._sources(m, r), m["recs"])
        S.drain()
        alone.append(S.result(i))
        S.close()
    S = sem.Server(block_steps=16)
    ids = []
    for k, r in enumerate(reqs):
        m = models[r["model"]]
        ids.append(S.submit(m["mesh"], r["nsteps"], m["dt"], serve._sources(m, r), m["recs"],
                            arrival=S.clock()))
        st = S.pump()  # the next request arrives while this one runs
        assert st["active"] == k + 1
    S.drain()
    admits = [S.request(i)["admit"] for i in ids]
    assert admits[1] < S.request(ids[0])["done"] and admits[2] < S.request(ids[0])["done"]
    for i, a in zip(ids, alone):
        assert S.request(i)["state"] == "done"
        assert np.abs(a).max() > 0
        np.testing.assert_array_equal(S.result(i), a)
    for m in models:
        m["mesh"].close()
