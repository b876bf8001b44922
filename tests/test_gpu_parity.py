"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Bars: gather maps and colourings bit-exact; FP32 stiffness and mass within
FP32 rounding of FP64 (relative 2e-6); receiver seismograms and final
wavefields within relative L2 1e-5 of the FP64 oracle; batched results
identical (bitwise) to the same shot run alone; graph replay identical to
eager launches."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import meshgen, sem
from tests import problems as pb
from tests import refmeshes as rm

pytestmark = pytest.mark.gpu

TOL_OP = 2e-6    # one FP32 operator application vs FP64
TOL_RUN = 1e-5   # receiver / wavefield relative L2 over a run (the metric's accuracy bar)


def oracle_for(P, quirks=po.DEFAULT_QUIRKS):
    S = po.Sim(P["dim"], P["order"], P["physics"], P["cells"], P["coords"], P["par"], quirks)
    return S


def mesh_for(P, quirks=sem.DEFAULT_QUIRKS):
    return sem.Mesh(P["dim"], P["order"], P["physics"], P["cells"], P["coords"], P["par"], quirks)


CASES = [(2, 0), (2, 1), (3, 0), (3, 1)]


@pytest.mark.parametrize("dim,physics", CASES)
def test_mesh_maps_bit_exact(dim, physics):
    P = pb.make(dim, physics, 5 if dim == 3 else 9, order=4)
    M = mesh_for(P)
    S = oracle_for(P)
    np.testing.assert_array_equal(M.conn(), S.conn())
    ocol, _ = po.color_elements(dim, P["cells"], len(P["coords"]))
    np.testing.assert_array_equal(M.colours(), ocol)
    assert M.check_colouring() == 0
    assert M.affine
    mi = M.inverse_mass().astype(np.float64)
    ref = S.inverse_mass()
    assert np.max(np.abs(mi - ref) / np.abs(ref)) < 1e-6


# Non-parallelogram quads run the reference's quad operator (QuadP1.h:20-27
# Jacobian, J^-1 at the output node, TensorQuad.cpp:383-390) by default and
# the consistent one with the switch off; the oracle restates both.
QUAD_MODES = [sem.DEFAULT_QUIRKS, sem.DEFAULT_QUIRKS & ~sem.QUIRK_QUAD_JACOBIAN]


@pytest.mark.parametrize("dim,physics", CASES)
@pytest.mark.parametrize("order", [3, 4])
@pytest.mark.parametrize("distorted", [False, True])
def test_stiffness_operator(dim, physics, order, distorted):
    P = pb.make(dim, physics, 4 if dim == 3 else 7, order=order, seed=order, distorted=distorted)
    for quirks in (QUAD_MODES if dim == 2 else QUAD_MODES[:1]):
        M = mesh_for(P, quirks)
        S = oracle_for(P, quirks)
        assert M.affine == (not distorted)
        rng = np.random.default_rng(11)
        xyz = S.node_coords()
        u = np.stack([np.sin(0.05 * xyz @ rng.normal(size=dim)) + 0.1 * rng.normal(size=S.G)
                      for _ in range(S.c)])
        ref = S.apply_stiffness(u)
        mi = M.inverse_mass().astype(np.float64)
        assert np.max(np.abs(mi - S.inverse_mass()) / S.inverse_mass()) < 1e-6
        for path in (sem.PATH_FUSED, sem.PATH_PLAIN):
            ku = M.apply_stiffness(u, path).astype(np.float64)
            assert pb.rel_l2(ku, ref) < TOL_OP, (quirks, path, pb.rel_l2(ku, ref))


@pytest.mark.parametrize("physics", [0, 1])
@pytest.mark.parametrize("quirks", QUAD_MODES)
@pytest.mark.parametrize("path", [sem.PATH_FUSED, sem.PATH_PLAIN])
def test_distorted_quad_run(physics, quirks, path):
    """A run on non-parallelogram quads in both quad-operator modes."""
    P = pb.make(2, physics, 10, order=4, seed=8, distorted=True, nsteps=250)
    seis, seis_ref, u, u_ref, _, _ = _run_both(P, path, quirks=quirks)
    e_s, e_u = pb.rel_l2(seis, seis_ref), pb.rel_l2(u, u_ref)
    print("distorted quads physics %d quirks %d path %d: receivers %.3g, u %.3g"
          % (physics, quirks, path, e_s, e_u))
    assert e_s < TOL_RUN and e_u < TOL_RUN


def _run_both(P, path, quirks=po.DEFAULT_QUIRKS, init=None):
    S = oracle_for(P, quirks)
    M = mesh_for(P, quirks)
    src = P["src"]
    S.add_source(src["loc"], src["amp"], src["f0"], src["t0"], src["dir"])
    for r in P["recs"]:
        S.add_receiver(r)
    prob = sem.Order2Newmark(M, P["dt"], P["nsteps"], path=path)
    prob.add_source(src["loc"], src["amp"], src["f0"], src["t0"], src["dir"])
    for r in P["recs"]:
        prob.add_receiver(r)
    if init is not None:
        S.set_field("u", init)
        prob.set_field("u", init)
    seis_ref = S.run(P["dt"], P["nsteps"])
    prob.takeTimeStep(P["nsteps"])
    prob.mesh = M
    return (prob.seismograms(), seis_ref, prob.get_field("u"), S.get_field("u"), prob, S)


@pytest.mark.parametrize("dim,physics", CASES)
@pytest.mark.parametrize("path", [sem.PATH_FUSED, sem.PATH_PLAIN])
def test_run_receivers_and_wavefield(dim, physics, path):
    P = pb.make(dim, physics, 5 if dim == 3 else 10, order=4, seed=3, nsteps=300)
    seis, seis_ref, u, u_ref, prob, S = _run_both(P, path)
    assert np.abs(seis_ref).max() > 0
    e_seis, e_u = pb.rel_l2(seis, seis_ref), pb.rel_l2(u, u_ref)
    print("dim %d physics %d path %d: receivers rel-L2 %.3e, wavefield %.3e" % (dim, physics, path, e_seis, e_u))
    assert e_seis < TOL_RUN and e_u < TOL_RUN
    steps, t = prob.batch.progress(prob.shot)
    assert steps == P["nsteps"] and t == pytest.approx(S.time(), rel=1e-15)


def test_distorted_hex_run():
    P = pb.make(3, 1, 4, order=4, seed=5, distorted=True, nsteps=200)
    seis, seis_ref, u, u_ref, _, _ = _run_both(P, sem.PATH_FUSED)
    assert pb.rel_l2(seis, seis_ref) < TOL_RUN and pb.rel_l2(u, u_ref) < TOL_RUN


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("path", [sem.PATH_FUSED, sem.PATH_PLAIN])
def test_newmark_quirk_matches_oracle(dim, path):
    """Order2Newmark.cpp:93-96's list {ax, ay, ax, a} on Elastic2D and
    Elastic3D: x advanced twice per step, y once, z never."""
    P = pb.make(dim, 1, 4 if dim == 3 else 8, order=3, seed=6, nsteps=120)
    q = po.QUIRK_NEWMARK3D | po.QUIRK_E3D_SOURCE
    seis, seis_ref, u, u_ref, _, _ = _run_both(P, path, quirks=q)
    assert pb.rel_l2(u, u_ref) < TOL_RUN and pb.rel_l2(seis, seis_ref) < TOL_RUN
    if dim == 3:
        assert np.abs(u[2]).max() == 0.0  # z never advances under the reference's list


def test_quad_eigenfunction_on_gpu():
    """test_Scalar_2D.cpp:238-308 through the HIP path (FP32)."""
    cells, coords = rm.quad_4x4()
    L, vp = 1e5, 5800.0
    M = sem.Mesh(2, 3, sem.ACOUSTIC, cells, coords, np.full(len(cells) * 4, vp))
    xyz = M.node_coords()
    un = np.prod(np.sin(np.pi / L * (xyz - L)), axis=1)
    M.set_dirichlet(np.any(np.isclose(xyz, 0, atol=1e-6 * L) | np.isclose(xyz, L, atol=1e-6 * L), axis=1))
    nsteps = 2439
    prob = sem.Order2Newmark(M, 1e-2, nsteps)
    prob.set_field("u", un)
    worst, t = 0.0, 0.0
    for k in range(nsteps):
        prob.takeTimeStep(1)
        t += 1e-2
        exact = np.cos(np.pi / L * np.sqrt(2) * t * vp) * un
        worst = max(worst, np.abs(exact - prob.get_field("u")[0]).max())
    assert worst <= 0.001288 * 1.01


def test_batched_identical_to_alone():
    """Shots sharing a mesh (interleaved group) and a shot on another mesh,
    advanced in one batch, equal the same shots run alone, bit for bit."""
    P1 = pb.make(3, 0, 4, order=4, seed=7, nsteps=80)
    P2 = pb.make(3, 0, 5, order=4, seed=8, nsteps=60)
    M1, M2 = mesh_for(P1), mesh_for(P2)
    locs = [P1["src"]["loc"] * f for f in (1.0, 0.8, 1.1)]

    def build(shots):
        B = sem.Batch()
        ids = []
        for (M, P, loc) in shots:
            k = B.add_shot(M, P["nsteps"], P["dt"])
            B.add_source(k, loc, P["src"]["amp"], P["src"]["f0"], P["src"]["t0"])
            for r in P["recs"]:
                B.add_receiver(k, r)
            ids.append(k)
        return B, ids

    mix = [(M1, P1, l) for l in locs] + [(M2, P2, P2["src"]["loc"])]
    B, ids = build(mix)
    B.prepare(sem.PATH_FUSED, 0, 4)
    assert B.config()["groups"] == 2
    B.run(80)
    for (M, P, loc), k in zip(mix, ids):
        A, (a,) = build([(M, P, loc)])
        A.prepare(sem.PATH_FUSED, B.config()["elems_per_block"], 1)
        A.run(80)
        np.testing.assert_array_equal(A.seismograms(a), B.seismograms(k))
        np.testing.assert_array_equal(A.get_field(a, "u"), B.get_field(k, "u"))


@pytest.mark.parametrize("no_graph", ["0", "1"])
def test_lanes_identical(monkeypatch, no_graph):
    """Groups on two meshes split over two or four streams (SEM_LANES: each
    part's colour launches on its own stream, joined once per step) equal
    the one-stream schedule bit for bit, in graph replays and eagerly."""
    P1 = pb.make(3, 1, 4, order=4, seed=17, nsteps=60)
    P2 = pb.make(3, 1, 5, order=4, seed=18, nsteps=60)
    M1, M2 = mesh_for(P1), mesh_for(P2)
    monkeypatch.setenv("SEM_NO_GRAPH", no_graph)
    monkeypatch.setenv("SEM_GRAPH_STEPS", "7")
    out = []
    for lanes in ("1", "2", "4"):
        monkeypatch.setenv("SEM_LANES", lanes)
        B = sem.Batch()
        ids = []
        for M, P in ((M1, P1), (M1, P1), (M2, P2), (M2, P2), (M2, P2)):
            k = B.add_shot(M, P["nsteps"], P["dt"])
            B.add_source(k, P["src"]["loc"] * (1.0 + 0.05 * len(ids)), P["src"]["amp"], P["src"]["f0"],
                         P["src"]["t0"], P["src"]["dir"])
            for r in P["recs"]:
                B.add_receiver(k, r)
            ids.append(k)
        B.prepare(sem.PATH_FUSED, 0, 1)
        assert B.config()["groups"] == 5
        B.run(25)
        B.run(35)
        out.append([(B.seismograms(k), B.get_field(k, "u")) for k in ids])
        B.close()
    for other in out[1:]:
        for (s1, u1), (s2, u2) in zip(out[0], other):
            np.testing.assert_array_equal(s1, s2)
            np.testing.assert_array_equal(u1, u2)


@pytest.mark.parametrize("block", ["", "7"])
def test_graph_replay_identical_to_eager(monkeypatch, block):
    """Runs of 40 steps: with the default block (50) each is one captured
    40-step tail graph; with blocks of 7, five full-block replays and a
    5-step tail."""
    P = pb.make(2, 1, 8, order=4, seed=9, nsteps=160)
    out = []
    monkeypatch.setenv("SEM_GRAPH_STEPS", block)
    for no_graph in ("1", "0", "0"):
        monkeypatch.setenv("SEM_NO_GRAPH", no_graph)
        M = mesh_for(P)
        prob = sem.Order2Newmark(M, P["dt"], P["nsteps"])
        prob.add_source(P["src"]["loc"], P["src"]["amp"], P["src"]["f0"], P["src"]["t0"], P["src"]["dir"])
        for r in P["recs"]:
            prob.add_receiver(r)
        prob.initializeGlobalDofs()
        for _ in range(4):  # several replays with host work between them
            prob.takeTimeStep(40)
            _ = prob.get_field("v")
        out.append((prob.seismograms(), prob.get_field("u")))
    for s, u in out[1:]:
        np.testing.assert_array_equal(s, out[0][0])
        np.testing.assert_array_equal(u, out[0][1])


def test_moment_tensor_source_matches_oracle():
    P = pb.make(3, 1, 4, order=4, seed=10, nsteps=150)
    M = mesh_for(P)
    S = oracle_for(P)
    Mt = np.array([[1.0, 0.2, -0.4], [0.2, -0.6, 0.3], [-0.4, 0.3, 0.5]]) * 1e6
    src = P["src"]
    S.add_moment_source(src["loc"], Mt, 1.0, src["f0"], src["t0"])
    for r in P["recs"]:
        S.add_receiver(r)
    prob = sem.Order2Newmark(M, P["dt"], P["nsteps"])
    prob.add_moment_source(src["loc"], Mt, 1.0, src["f0"], src["t0"])
    for r in P["recs"]:
        prob.add_receiver(r)
    ref = S.run(P["dt"], P["nsteps"])
    prob.takeTimeStep(P["nsteps"])
    assert pb.rel_l2(prob.seismograms(), ref) < TOL_RUN


@pytest.mark.parametrize("dim,physics,nel", [(2, 0, 13), (2, 1, 9), (3, 0, 5), (3, 1, 6)])
def test_tile_kernel_matches_element_kernel_and_oracle(dim, physics, nel, monkeypatch):
    """Structured meshes take the tile kernel (partial tiles included); the
    element-colour kernel (SEM_NO_TILES=1) and the oracle agree with it."""
    P = pb.make(dim, physics, nel, order=4, seed=12, nsteps=150)
    monkeypatch.setenv("SEM_NO_TILES", "0")
    seis_t, ref, u_t, u_ref, prob, _ = _run_both(P, sem.PATH_FUSED)
    assert prob.mesh.ntiles > 0
    monkeypatch.setenv("SEM_NO_TILES", "1")
    seis_e, _, u_e, _, prob_e, _ = _run_both(P, sem.PATH_FUSED)
    assert prob_e.mesh.ntiles == 0
    assert pb.rel_l2(seis_t, ref) < TOL_RUN and pb.rel_l2(u_t, u_ref) < TOL_RUN
    assert pb.rel_l2(seis_t, seis_e) < TOL_RUN


def test_distorted_mesh_uses_element_kernel():
    P = pb.make(3, 0, 3, order=4, seed=13, distorted=True)
    assert mesh_for(P).ntiles == 0


@pytest.mark.parametrize("tile", ["2,2,2", "4,2,2", "4,4,2", "3,1,2"])
def test_tile_shapes(tile, monkeypatch):
    """Every compiled tile shape (partial edge tiles on a 5^3 mesh) matches
    the oracle; a shape with no compiled kernel falls back to the element
    kernel."""
    monkeypatch.setenv("SEM_TILE", tile)
    P = pb.make(3, 1, 5, order=3, seed=14, nsteps=100)
    seis, ref, u, u_ref, prob, _ = _run_both(P, sem.PATH_FUSED)
    compiled = tile != "3,1,2"
    assert (prob.mesh.ntiles > 0) == compiled
    if compiled:
        assert prob.mesh.tile_shape == tuple(int(x) for x in tile.split(","))
    assert pb.rel_l2(seis, ref) < TOL_RUN and pb.rel_l2(u, u_ref) < TOL_RUN


@pytest.mark.parametrize("dim,physics,order,tile", [
    (2, 0, 3, "8,8"), (2, 1, 5, "16,8"), (2, 0, 6, "16,8"), (3, 0, 2, "4,2,2"),
    (3, 1, 5, "2,2,2"), (3, 0, 5, "4,2,2"), (3, 0, 4, "4,4,2"),
    (3, 1, 4, "4,4,2"), (3, 1, 4, "4,2,2"), (3, 1, 4, "2,2,2")])
def test_tile_orders(dim, physics, order, tile, monkeypatch):
    """Tile kernels of other orders and shapes agree with the oracle."""
    monkeypatch.setenv("SEM_TILE", tile)
    P = pb.make(dim, physics, 5 if dim == 3 else 11, order=order, seed=15 + order, nsteps=80,
                cfl=0.12)
    seis, ref, u, u_ref, prob, _ = _run_both(P, sem.PATH_FUSED)
    assert prob.mesh.ntiles > 0
    assert pb.rel_l2(seis, ref) < TOL_RUN and pb.rel_l2(u, u_ref) < TOL_RUN


@pytest.mark.parametrize("dim,tile", [(2, "16,8"), (3, "4,4,2"), (3, "4,2,2")])
def test_acoustic_occupancy_variants_identical(dim, tile, monkeypatch):
    """The normal and the high-occupancy acoustic tile kernels (chosen per
    launch) give bit-identical results, and both match the oracle."""
    monkeypatch.setenv("SEM_TILE", tile)
    P = pb.make(dim, 0, 5 if dim == 3 else 18, order=4, seed=31 + dim, nsteps=100)
    out = {}
    for mode in ("normal", "high"):
        monkeypatch.setenv("SEM_TILE_OCCUPANCY", mode)
        seis, ref, u, u_ref, prob, _ = _run_both(P, sem.PATH_FUSED)
        assert prob.mesh.ntiles > 0
        assert pb.rel_l2(seis, ref) < TOL_RUN and pb.rel_l2(u, u_ref) < TOL_RUN
        out[mode] = (seis, u)
    np.testing.assert_array_equal(out["high"][0], out["normal"][0])
    np.testing.assert_array_equal(out["high"][1], out["normal"][1])


@pytest.mark.parametrize("dim,physics", CASES)
def test_diagonal_geometry_variant_identical(dim, physics, monkeypatch):
    """Axis-aligned lattices take the tile kernel that keeps only the
    diagonal of J^-1 (or of G); the terms it drops are exact zeros, so its
    results are those of the general tile kernel, and both match the
    oracle."""
    P = pb.make(dim, physics, 5 if dim == 3 else 10, order=4, seed=41 + dim, nsteps=150)
    out = {}
    for general in ("0", "1"):
        monkeypatch.setenv("SEM_TILE_GENERAL", general)
        seis, ref, u, u_ref, prob, _ = _run_both(P, sem.PATH_FUSED)
        assert prob.mesh.ntiles > 0 and prob.mesh.tile_diagonal
        assert pb.rel_l2(seis, ref) < TOL_RUN and pb.rel_l2(u, u_ref) < TOL_RUN
        out[general] = (seis, u)
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    np.testing.assert_array_equal(out["0"][1], out["1"][1])


@pytest.mark.parametrize("dim,order", [(2, o) for o in range(1, 11)] + [(3, o) for o in range(1, 8)])
@pytest.mark.parametrize("physics", [0, 1])
def test_every_compiled_order(dim, order, physics):
    """Every order the element kernel and the plain path are compiled for
    (2-D 1-10, 3-D 1-7): the stiffness operator on a distorted mesh (element
    kernel, pointwise geometry) and on an affine one (tile kernel where one
    is compiled, else the element kernel), both paths, and the inverse mass,
    against the oracle (measured up to 3e-7 relative L2 at every order)."""
    tol = TOL_OP
    nel = 3 if dim == 3 else 4
    for distorted in (True, False):
        P = pb.make(dim, physics, nel, order=order, seed=order + 20, distorted=distorted)
        M = mesh_for(P)
        S = oracle_for(P)
        mi = M.inverse_mass().astype(np.float64)
        assert np.max(np.abs(mi - S.inverse_mass()) / S.inverse_mass()) < 1e-6
        rng = np.random.default_rng(order)
        xyz = S.node_coords()
        u = np.stack([np.sin(0.05 * xyz @ rng.normal(size=dim)) + 0.1 * rng.normal(size=S.G)
                      for _ in range(S.c)])
        ref = S.apply_stiffness(u)
        for path in (sem.PATH_FUSED, sem.PATH_PLAIN):
            err = pb.rel_l2(M.apply_stiffness(u, path).astype(np.float64), ref)
            print("dim %d order %d physics %d distorted %d path %d tiles %s: %.2e"
                  % (dim, order, physics, distorted, path, M.tile_shape, err), flush=True)
            assert err < tol, (distorted, path, err)
