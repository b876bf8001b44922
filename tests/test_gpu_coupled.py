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
