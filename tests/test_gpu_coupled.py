"""GPU parity of the 2-D fluid-solid coupled run against the FP64 coupled
oracle (oracle/coupled.py), through the C ABI.

The fluid and the solid elements of one quad mesh become two meshes of one
batch, joined by a coupling: the combined inverse mass at the shared nodes
(Order2Newmark.cpp:32-66) and the edge loads of ElasticToAcoustic.cpp:17-41 /
AcousticToElastic2D.cpp:17-60 from the velocities after the previous step.
Bar: each medium's final displacement within relative 1e-5 (max norm) of
FP64 on both HIP paths, for a straight and a staircase interface, plain and
distorted cells, a source on either side; an uncoupled shot on the fluid
mesh in the same batch stays the single-physics result."""
import numpy as np
import pytest

from oracle import coupled as cp
from oracle import pyoracle as po
from salvus_amd import meshgen, sem

pytestmark = pytest.mark.gpu

H = 10.0
TOL = 1e-5


def _materials(cells, fluid):
    nf, ns = int(fluid.sum()), int((~fluid).sum())
    par_f = np.full((nf, 4, 1), 1500.0)
    # fluid RHO varies per vertex: rho_0 is the neighbour cell's mean
    rho_f = 1000.0 + 37.0 * (np.arange(4 * nf).reshape(nf, 4) % 7)
    rho, vp, vs = 2500.0, 3000.0, 1732.0
    c11, c55, c13 = rho * vp ** 2, rho * vs ** 2, rho * (vp ** 2 - 2 * vs ** 2)
    par_s = np.tile(np.array([rho, c11, c13, c11, c55]), (ns, 4, 1))
    return par_f, rho_f, par_s


def _problem(interface, distorted):
    nx, ny = 8, 6
    cells, coords = meshgen.box_quad(nx, ny, nx * H, ny * H)
    if distorted:
        coords = coords + 0.7 * np.sin(coords[:, ::-1] / 7.0)
    centres = coords[cells].mean(axis=1)
    if interface == "straight":
        fluid = centres[:, 0] < 4 * H
    else:  # a staircase: the fluid/solid boundary has corners and T-junctions
        fluid = centres[:, 0] < (3 + (centres[:, 1] > 3 * H)) * H
    return cells, coords, fluid


def _part(cells, coords, idx):
    fluid = np.zeros(len(cells), dtype=bool)
    fluid[idx] = True
    return sem.split_coupled(cells, coords, fluid, None)[0][:2]


@pytest.mark.parametrize("path", [sem.PATH_FUSED, sem.PATH_PLAIN])
@pytest.mark.parametrize("interface,distorted,side,quirk", [
    ("straight", False, "f", False), ("straight", False, "s", False),
    ("staircase", False, "f", False), ("staircase", True, "s", False),
    ("straight", False, "s", True)])
def test_coupled_run(path, interface, distorted, side, quirk):
    """quirk: the reference's Newmark field list {ax, ay, ax, a} on both
    media (the solid's x advanced twice per step, Order2Newmark.cpp:93-96).
    That update is unstable (the oracle grows about 1.3x per step), so its
    case runs 40 steps from a source next to the interface."""
    cells, coords, fluid = _problem(interface, distorted)
    par_f, rho_f, par_s = _materials(cells, fluid)
    order, nsteps = 4, (40 if quirk else 1500)
    dt = 0.3 * H * 0.17 / 3000.0
    quirks = sem.DEFAULT_QUIRKS | (sem.QUIRK_NEWMARK3D if quirk else 0)
    C = cp.CoupledSim(order, cells, coords, fluid, par_f, rho_f, par_s, newmark_quirk=quirk)
    if side == "f":
        loc, f0 = np.array([12.0, 21.0]), 1500.0 / (8 * H)
        src = dict(loc=loc, amp=1.0, f0=f0, t0=1.2 / f0, direction=None)
    elif not quirk:
        loc, f0 = np.array([63.0, 27.0]), 3000.0 / (8 * H)
        src = dict(loc=loc, amp=1e9, f0=f0, t0=1.2 / f0, direction=[0.6, 1.0])
    else:
        loc, f0 = np.array([44.0, 27.0]), 3000.0 / (8 * H)
        src = dict(loc=loc, amp=1e9, f0=f0, t0=0.3 / f0, direction=[0.6, 1.0])
    C.add_source(side, src["loc"], src["amp"], src["f0"], src["t0"], src["direction"])
    ref = C.run(dt, nsteps)
    assert np.abs(ref["f"]).max() > 0 and np.abs(ref["s"]).max() > 0  # both media moved

    cf, xf = _part(cells, coords, np.flatnonzero(fluid))
    cs, xs = _part(cells, coords, np.flatnonzero(~fluid))
    mf = sem.Mesh(2, order, sem.ACOUSTIC, cf, xf, par_f, quirks)
    ms = sem.Mesh(2, order, sem.ELASTIC, cs, xs, par_s, quirks)
    cpl = sem.Coupling(mf, ms, rho_f)
    assert cpl.shared_nodes == len(C.f2s)
    B = sem.Batch()
    kf = B.add_shot(mf, nsteps, dt)
    ks = B.add_shot(ms, nsteps, dt)
    kfree = B.add_shot(mf, nsteps, dt)  # same mesh, not coupled
    B.add_source(kf if side == "f" else ks, src["loc"], src["amp"], src["f0"], src["t0"],
                 src["direction"])
    B.add_source(kfree, [12.0, 21.0], 1.0, 1500.0 / (8 * H), 1.2 * 8 * H / 1500.0)
    B.couple(cpl, kf, ks)
    B.prepare(path)
    B.run(nsteps)
    for key, k in (("f", kf), ("s", ks)):
        got = B.get_field(k, "u").reshape(ref[key].shape)
        err = np.abs(got - ref[key]).max() / np.abs(ref[key]).max()
        print("%s path %d %s: max rel err %.2e" % (interface, path, key, err), flush=True)
        assert err < TOL, (key, err)
    # the uncoupled shot on the fluid mesh: the single-physics oracle
    S = po.Sim(2, order, po.ACOUSTIC, cf, xf, par_f,
               po.DEFAULT_QUIRKS | (po.QUIRK_NEWMARK3D if quirk else 0))
    S.add_source(np.array([12.0, 21.0]), 1.0, 1500.0 / (8 * H), 1.2 * 8 * H / 1500.0)
    S.run(dt, nsteps)
    ref1 = S.get_field("u")
    got1 = B.get_field(kfree, "u").reshape(ref1.shape)
    assert np.abs(got1 - ref1).max() / np.abs(ref1).max() < TOL
    B.close()
    cpl.close()


def test_coupling_errors():
    cells, coords, fluid = _problem("straight", False)
    par_f, rho_f, par_s = _materials(cells, fluid)
    cf, xf = _part(cells, coords, np.flatnonzero(fluid))
    cs, xs = _part(cells, coords, np.flatnonzero(~fluid))
    mf = sem.Mesh(2, 4, sem.ACOUSTIC, cf, xf, par_f)
    ms = sem.Mesh(2, 4, sem.ELASTIC, cs, xs, par_s)
    with pytest.raises(sem.SemError):
        sem.Coupling(ms, mf, rho_f)  # physics swapped
    cpl = sem.Coupling(mf, ms, rho_f)
    B = sem.Batch()
    kf = B.add_shot(mf, 10, 1e-4)
    ks = B.add_shot(ms, 12, 1e-4)
    with pytest.raises(sem.SemError):
        B.couple(cpl, kf, ks)  # step counts differ
    ks2 = B.add_shot(ms, 10, 1e-4)
    with pytest.raises(sem.SemError):
        B.couple(cpl, ks2, kf)  # shots on the wrong meshes
    B.couple(cpl, kf, ks2)
    with pytest.raises(sem.SemError):
        B.couple(cpl, kf, ks2)  # already coupled
    B.close()
    cpl.close()


def test_mixed_exodus_model_coupled_run(tmp_path):
    """One quad Exodus file whose elemental "fluid" flag marks a staircase
    fluid region (ExodusModel.cpp:473-518), read through exodus.read_mesh /
    read_model, flagged and served by model.coupled_parameters, cut by
    split_coupled and coupled: within the run tolerance of the FP64 coupled
    oracle given the same arrays, on the tile path."""
    from salvus_amd import exodus, model
    from tests.test_exodus import write_exodus
    cells, coords, fluid = _problem("staircase", False)
    par_f, rho_f, par_s = _materials(cells, fluid)
    E = len(cells)
    ev = {"fluid": fluid.astype(float)}
    fi, si = np.flatnonzero(fluid), np.flatnonzero(~fluid)
    for v in range(4):
        vp = np.zeros(E)
        vp[fi] = par_f[:, v, 0] + 10.0 * (fi % 5)  # VP varies per cell
        ev["VP_%d" % v] = vp
        for q, name in enumerate(("RHO", "C11", "C13", "C33", "C55")):
            a = np.zeros(E)
            a[si] = par_s[:, v, q]
            if name == "RHO":
                a[fi] = rho_f[:, v]
            ev["%s_%d" % (name, v)] = a
    path = tmp_path / "mixed.e"
    write_exodus(path, coords, [(cells + 1, "QUAD4")], elem_vars=ev)
    d, rc, rxyz, _ = exodus.read_mesh(path)
    _, centres, variables = exodus.read_model(path)
    flag, pf, ps, rho = model.coupled_parameters(rc, rxyz, centres, variables)
    np.testing.assert_array_equal(flag, fluid)
    np.testing.assert_array_equal(rho, rho_f)
    (cf, xf, pf), (cs, xs, ps) = sem.split_coupled(rc, rxyz, flag, (pf, ps))
    mf = sem.Mesh(2, 4, sem.ACOUSTIC, cf, xf, pf)
    ms = sem.Mesh(2, 4, sem.ELASTIC, cs, xs, ps)
    cpl = sem.Coupling(mf, ms, rho)
    nsteps, dt = 400, 0.3 * H * 0.17 / 3000.0
    f0 = 1500.0 / (8 * H)
    C = cp.CoupledSim(4, rc, rxyz, flag, pf, rho, ps)
    C.add_source("f", np.array([12.0, 21.0]), 1.0, f0, 1.2 / f0, None)
    ref = C.run(dt, nsteps)
    assert np.abs(ref["s"]).max() > 0  # the wave crossed into the solid
    B = sem.Batch()
    kf = B.add_shot(mf, nsteps, dt)
    ks = B.add_shot(ms, nsteps, dt)
    B.add_source(kf, [12.0, 21.0], 1.0, f0, 1.2 / f0)
    B.couple(cpl, kf, ks)
    B.prepare(sem.PATH_FUSED)
    B.run(nsteps)
    for key, k in (("f", kf), ("s", ks)):
        got = B.get_field(k, "u").reshape(ref[key].shape)
        err = np.abs(got - ref[key]).max() / np.abs(ref[key]).max()
        print("exodus mixed %s: max rel err %.2e" % (key, err), flush=True)
        assert err < TOL, (key, err)
    # the same problem in one call from the file: bit for bit
    mf2, ms2, cpl2 = sem.coupled_from_exodus(path)
    assert (mf2.E, mf2.G, ms2.E, ms2.G) == (mf.E, mf.G, ms.E, ms.G)
    B2 = sem.Batch()
    kf2 = B2.add_shot(mf2, nsteps, dt)
    ks2 = B2.add_shot(ms2, nsteps, dt)
    B2.add_source(kf2, [12.0, 21.0], 1.0, f0, 1.2 / f0)
    B2.couple(cpl2, kf2, ks2)
    B2.prepare(sem.PATH_FUSED)
    B2.run(nsteps)
    for k, k2 in ((kf, kf2), (ks, ks2)):
        np.testing.assert_array_equal(B2.get_field(k2, "u"), B.get_field(k, "u"))
    B2.close()
    cpl2.close()
    B.close()
    cpl.close()
