"""Host-side logic of the serving driver and the benchmark's byte models
(no GPU): the request stream is reproducible and respects the configuration-5
ranges; source records have the C ABI layout; the roofline byte formulas
count what they state on a hand-checked case."""
import types

import numpy as np
import pytest

import bench
from salvus_amd import sem, serve


def test_stream_reproducible_and_in_range():
    specs, reqs = serve.draw_stream(400, 20.0, seed=5, models_per_category=4)
    specs2, reqs2 = serve.draw_stream(400, 20.0, seed=5, models_per_category=4)
    assert specs == specs2
    assert [r["arrival"] for r in reqs] == [r["arrival"] for r in reqs2]
    arr = np.array([r["arrival"] for r in reqs])
    assert np.all(np.diff(arr) > 0)
    assert abs(len(reqs) / arr[-1] - 20.0) < 4.0  # Poisson rate
    assert all(500 <= r["nsteps"] <= 3000 for r in reqs)
    ranges = {0: (100, 800), 1: (200, 1000), 2: (32, 96)}
    for s in specs:
        lo, hi = ranges[s["category"]]
        assert lo <= s["n"] <= hi
        assert s["dim"] == (3 if s["category"] == 2 else 2)
    cats = np.bincount([specs[r["model"]]["category"] for r in reqs], minlength=3) / len(reqs)
    assert np.allclose(cats, [0.4, 0.3, 0.3], atol=0.08)
    assert {s["physics"] for s in specs if s["category"] == 2} == {0, 1}


def test_size_scale():
    specs, _ = serve.draw_stream(10, 5.0, seed=1, models_per_category=3, size_scale=0.5)
    for s in specs:
        lo, hi = {0: (100, 800), 1: (200, 1000), 2: (32, 96)}[s["category"]]
        assert 0.5 * lo - 1 <= s["n"] <= 0.5 * hi + 1


def test_source_record_layout():
    r = sem.source_record(dict(kind="force", loc=(1.0, 2.0), amp=3.0, f0=4.0, t0=5.0,
                               dir=(0.0, 1.0)))
    assert r.shape == (sem.SOURCE_DOUBLES,)
    assert list(r[:10]) == [0, 1, 2, 0, 3, 4, 5, 0, 1, 0]
    M = np.arange(9.0).reshape(3, 3)
    r = sem.source_record(dict(kind="moment", loc=(1.0, 2.0, 3.0), M=M, amp=1.0, f0=2.0,
                               t0=0.5))
    assert r[0] == 1 and list(r[1:4]) == [1, 2, 3]
    assert list(r[10:19]) == list(M.ravel())


def test_byte_models_on_a_hand_case():
    # one 2-D acoustic element of order 1 (4 nodes), 1 tile of 1x1 elements
    M = types.SimpleNamespace(c=1, nd=4, E=1, G=4, ntiles=1, tile_elems=1, tile_slots=4,
                              tile_valid_slots=4, tile_acc_reads=0, tile_acc_writes=0)
    # element formulation: 4 nodes x (4 + 12) + 4 x (3 factors + 4 vertex vp) + 4 x (4 + 36)
    assert bench.element_bytes_per_step(M, 1, 2, 0) == 4 * 16 + 4 * 7 + 4 * 40
    # tile algorithm: 4 words + 4 x 7 element floats + (4 u + 4 x (4 mi + 8 read w, kc
    # + 12 write u, w, kc))
    assert bench.tile_bytes_per_step(M, 1, 2, 0) == 16 + 28 + 16 + 4 * 24
    # S shots scale the per-shot parts only
    assert bench.tile_bytes_per_step(M, 3, 2, 0) == 16 + 28 + 3 * (16 + 4 * 24)


def test_fifo_schedule_of_service_times():
    """The plain server's FIFO schedule: a request starts at its arrival or
    when the one before it finishes, whichever is later, and takes the
    model's set-up plus steps times the per-step time."""
    M = types.SimpleNamespace(c=1, G=100)
    models = [dict(mesh=M)]
    fit = {0: (1.0, 0.01)}  # 1 s set-up, 10 ms per step
    reqs = [dict(model=0, arrival=0.0, nsteps=100),   # 0 -> 2
            dict(model=0, arrival=0.5, nsteps=200),   # 2 -> 5
            dict(model=0, arrival=10.0, nsteps=100)]  # 10 -> 12 (idle before)
    out = serve.fifo_schedule(models, reqs, fit)
    lat = np.array([2.0, 4.5, 2.0])
    assert out["makespan_s"] == 12.0
    assert out["dof_updates"] == 100 * 400
    assert abs(out["mean_s"] - lat.mean()) < 1e-12
    assert abs(out["p50_s"] - np.percentile(lat, 50)) < 1e-12
    assert abs(out["dof_updates_per_s"] - 100 * 400 / 12.0) < 1e-9


def test_split_coupled_parts():
    """The fluid and solid parts of a mixed quad mesh keep their cells'
    vertex coordinates and order, share exactly the interface vertices, and
    match the coupled oracle's own split."""
    from salvus_amd import meshgen
    cells, coords = meshgen.box_quad(6, 4, 60.0, 40.0)
    fluid = coords[cells].mean(axis=1)[:, 0] < 30.0
    (cf, xf, _), (cs, xs, _) = sem.split_coupled(cells, coords, fluid, None)
    np.testing.assert_array_equal(xf[cf], coords[cells[fluid]])
    np.testing.assert_array_equal(xs[cs], coords[cells[~fluid]])
    assert cf.max() + 1 == len(xf) and cs.max() + 1 == len(xs)
    shared = {tuple(p) for p in xf} & {tuple(p) for p in xs}
    assert len(shared) == 5 and all(p[0] == 30.0 for p in shared)  # the line x = 30
    with pytest.raises(sem.SemError):
        sem.split_coupled(cells, coords, fluid[:-1], None)
