"""GPU parity on meshes beyond the structured boxes of the benchmark
workloads, through the C ABI against the FP64 oracle:

* the hex eigenfunction of test_Scalar_3D.cpp:152-235 (Dirichlet walls on
  every face, orders 3/4/5) held to the reference's own error bounds on the
  tile kernel, the element kernel and the plain path;
* the 7-hex cell list of test_TensorHex.cpp:424-494 (mixed face and edge
  orientations, Mesh.cpp:90-278) and box meshes whose cells are listed in
  rotated vertex orders (each element's reference axes turned by a proper
  rotation): gather map bit-exact, stiffness operator and a short run;
* GLL node coordinates (buildNodalPoints, QuadP1.cpp:176-216,
  HexP1.cpp:312-354) against the oracle's.
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import meshgen, sem
from tests import problems as pb
from tests import refmeshes as rm

pytestmark = pytest.mark.gpu

TOL_OP = 2e-6
TOL_RUN = 1e-5
L, VP = 1e5, 5800.0  # test_Model.cpp:35


def eigen_error(order, path, monkeypatch, no_tiles=False):
    """Worst |u - exact| over one second of the 3-D standing wave (dt 1e-2,
    Dirichlet on all faces; test_Scalar_3D.cpp:104-145) on a 3^3 box of the
    reference test's extent (hex_eigenfunction.e itself is not available)."""
    monkeypatch.setenv("SEM_NO_TILES", "1" if no_tiles else "0")
    cells, coords = meshgen.box_hex(3, 3, 3, L, L, L)
    M = sem.Mesh(3, order, sem.ACOUSTIC, cells, coords, np.full(len(cells) * 8, VP))
    xyz = M.node_coords()
    un = np.prod(np.sin(np.pi / L * (xyz - (5e4 + L / 2))), axis=1)  # test_Scalar_3D.cpp:31-37
    M.set_dirichlet(np.any(np.isclose(xyz, 0, atol=1e-6 * L) | np.isclose(xyz, L, atol=1e-6 * L),
                           axis=1))
    dt = 1e-2
    prob = sem.Order2Newmark(M, dt, 120, path=path)
    prob.set_field("u", un)
    worst, t = 0.0, 0.0
    omega = np.pi / L * np.sqrt(3) * VP
    while True:  # step, compare, stop once t > 1 s (test_Scalar_3D.cpp:104-145)
        prob.takeTimeStep(1)
        t += dt
        worst = max(worst, np.abs(np.cos(omega * t) * un - prob.get_field("u")[0]).max())
        if t > 1.0:
            break
    return worst, M


@pytest.mark.parametrize("order,bound", [(3, 0.00048205), (4, 0.000489815), (5, 0.000486752)])
@pytest.mark.parametrize("kernel", ["tile", "element", "plain"])
def test_hex_eigenfunction_reference_bound(order, bound, kernel, monkeypatch):
    path = sem.PATH_PLAIN if kernel == "plain" else sem.PATH_FUSED
    err, M = eigen_error(order, path, monkeypatch, no_tiles=kernel == "element")
    if kernel == "tile":
        assert M.tile_shape is not None  # Dirichlet rows through the tile lattice words
    print("order %d %s: worst error %.6g (bound %.6g)" % (order, kernel, err, bound * 1.01))
    assert err <= bound * 1.01  # test_Scalar_3D.cpp:213-232


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", ["tile", "element", "plain"])
def test_elastic_standing_p_wave(dim, kernel, monkeypatch):
    """u_x = sin(pi x/L) cos(omega t), Dirichlet on the x faces: exact for the
    elastic operators when the coupling stiffness C13 (2-D) / c12 and c13
    (3-D) vanishes, so the run stays within the start-up amplitude omega dt/2
    of the closed form and keeps the other components at zero; and it equals
    the FP64 oracle's run within the run tolerance.  On 4^3 hexes the tile
    path takes the order-4 4x2x2 elastic tile, whose Dirichlet flags are read
    through the padded LDS lattice."""
    monkeypatch.setenv("SEM_NO_TILES", "1" if kernel == "element" else "0")
    if dim == 3:
        monkeypatch.setenv("SEM_TILE", "4,2,2")
    Lb, order, n_el = 1000.0, 4, 4
    if dim == 2:
        rho, c11 = 2000.0, 2000.0 * 3000.0 ** 2
        cells, coords = meshgen.box_quad(n_el, n_el, Lb, Lb)
        par = np.tile([rho, c11, 0.0, 1.1 * c11, 0.3 * c11], 4 * len(cells))
    else:
        rho, vph = 2000.0, 3000.0
        c11 = rho * vph ** 2
        # c12 = c11 - 2 c66 = 0 (VSH^2 = VPH^2/2); c13 = ETA (c11 - 2 c44) = 0 (ETA = 0)
        cells, coords = meshgen.box_hex(n_el, n_el, n_el, Lb, Lb, Lb)
        par = np.tile([rho, 3300.0, vph, 1400.0, vph / np.sqrt(2), 0.0], 8 * len(cells))
    M = sem.Mesh(dim, order, sem.ELASTIC, cells, coords, par)
    if kernel == "tile":
        assert M.ntiles > 0 and (dim == 2 or M.tile_shape == (4, 2, 2))
    xyz = M.node_coords()
    ux = np.sin(np.pi * xyz[:, 0] / Lb)
    u0 = np.zeros((dim, M.G))
    u0[0] = ux
    wall = np.isclose(xyz[:, 0], 0, atol=1e-9 * Lb) | np.isclose(xyz[:, 0], Lb, atol=1e-9 * Lb)
    M.set_dirichlet(wall)
    omega, dt, nsteps = np.pi / Lb * np.sqrt(c11 / rho), 2e-4, 600
    prob = sem.Order2Newmark(M, dt, nsteps, path=sem.PATH_PLAIN if kernel == "plain" else sem.PATH_FUSED)
    prob.set_field("u", u0)
    S = po.Sim(dim, order, po.ELASTIC, cells, coords, par)
    S.set_field("u", u0)
    S.set_dirichlet(wall)
    worst = 0.0
    for _ in range(nsteps // 50):
        prob.takeTimeStep(50)
        S.run(dt, 50)
        u, ref = prob.get_field("u"), S.get_field("u")
        worst = max(worst, np.abs(u[0] - np.cos(omega * prob.time()) * ux).max())
        assert np.abs(u[1:]).max() < 1e-6
        assert np.abs(u - ref).max() <= TOL_RUN * np.abs(ref).max()
    print("dim %d %s: worst error %.3g (bound %.3g)" % (dim, kernel, worst, omega * dt / 2 * 1.05))
    assert worst < omega * dt / 2 * 1.05


# ---------------------------------------------------------------- rotated cells
QUAD_SIGNS = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]])            # QuadP1.h:98-108
HEX_SIGNS = np.array([[-1, -1, -1], [-1, 1, -1], [1, 1, -1], [1, -1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])  # HexP1.cpp:123-133


def proper_rotations(dim):
    out = []
    for perm in itertools.permutations(range(dim)):
        for signs in itertools.product((-1, 1), repeat=dim):
            R = np.zeros((dim, dim), dtype=int)
            for a in range(dim):
                R[a, perm[a]] = signs[a]
            if round(np.linalg.det(R)) == 1:
                out.append(R)
    return out


def rotate_cells(cells, dim, rng):
    """Lists every cell's vertices in the order of a randomly turned
    reference frame (a proper rotation: det J stays positive).  Returns the
    new cells and, per cell, the permutation new position -> old position."""
    S = QUAD_SIGNS if dim == 2 else HEX_SIGNS
    key = {tuple(s): i for i, s in enumerate(S)}
    rots = proper_rotations(dim)
    perms = np.array([[key[tuple(R @ s)] for s in S] for R in rots])
    pick = perms[rng.integers(0, len(rots), size=len(cells))]
    return np.take_along_axis(cells, pick, axis=1), pick


def run_pair(dim, physics, order, cells, coords, par, nsteps, path, src_loc, rec_locs, dt):
    S = po.Sim(dim, order, physics, cells, coords, par)
    M = sem.Mesh(dim, order, physics, cells, coords, par)
    c = 1 if physics == sem.ACOUSTIC else dim
    direction = np.array([0.3, 1.0, -0.5])[:c]
    f0 = 3000.0 / 80.0
    S.add_source(src_loc, 1e3, f0, 1.2 / f0, direction)
    for r in rec_locs:
        S.add_receiver(r)
    ref = S.run(dt, nsteps)
    prob = sem.Order2Newmark(M, dt, nsteps, path=path)
    prob.add_source(src_loc, 1e3, f0, 1.2 / f0, direction)
    for r in rec_locs:
        prob.add_receiver(r)
    prob.takeTimeStep(nsteps)
    return M, S, prob.seismograms(), ref, prob.get_field("u"), S.get_field("u")


def check_mesh(dim, physics, order, cells, coords, par, src_loc, rec_locs, dt, nsteps=120):
    S = po.Sim(dim, order, physics, cells, coords, par)
    M = sem.Mesh(dim, order, physics, cells, coords, par)
    np.testing.assert_array_equal(M.conn(), S.conn())  # Mesh.cpp:361-452 closure numbering
    assert M.check_colouring() == 0
    rng = np.random.default_rng(3)
    xyz = S.node_coords()
    u = np.stack([np.sin(xyz @ rng.normal(size=dim) / np.ptp(xyz)) + 0.1 * rng.normal(size=S.G)
                  for _ in range(S.c)])
    ref = S.apply_stiffness(u)
    for path in (sem.PATH_FUSED, sem.PATH_PLAIN):
        ku = M.apply_stiffness(u, path).astype(np.float64)
        assert pb.rel_l2(ku, ref) < TOL_OP, (path, pb.rel_l2(ku, ref))
    for path in (sem.PATH_FUSED, sem.PATH_PLAIN):
        _, _, seis, seis_ref, uu, uu_ref = run_pair(dim, physics, order, cells, coords, par, nsteps,
                                                    path, src_loc, rec_locs, dt)
        assert np.abs(seis_ref).max() > 0
        e_s, e_u = pb.rel_l2(seis, seis_ref), pb.rel_l2(uu, uu_ref)
        print("dim %d physics %d path %d: receivers %.3g, u %.3g" % (dim, physics, path, e_s, e_u))
        assert e_s < TOL_RUN and e_u < TOL_RUN
    return M


@pytest.mark.parametrize("physics", [0, 1])
def test_seven_hex_mesh(physics):
    """test_TensorHex.cpp:424-494's cell list (order 3, as there): faces and
    edges shared in mixed orientations."""
    cells, coords = rm.seven_hex()
    coords = coords * 100.0
    rng = np.random.default_rng(4)
    par = pb.vertex_material(3, physics, cells, coords, rng)
    M = check_mesh(3, physics, 3, cells, coords, par, src_loc=(37.0, 58.0, 41.0),
                   rec_locs=[(150.0, 50.0, 50.0), (50.0, -50.0, 50.0), (-70.0, 30.0, 80.0)],
                   dt=4e-4)
    assert M.tile_shape is None


@pytest.mark.parametrize("dim,physics", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_rotated_cells(dim, physics):
    h, nel = 10.0, (8 if dim == 2 else 4)
    Lb = nel * h
    rng = np.random.default_rng(10 + dim + physics)
    if dim == 2:
        cells, coords = meshgen.box_quad(nel, nel, Lb, Lb)
    else:
        cells, coords = meshgen.box_hex(nel, nel, nel, Lb, Lb, Lb)
    rot, pick = rotate_cells(np.asarray(cells), dim, rng)
    assert not np.array_equal(rot, cells)
    par = pb.vertex_material(dim, physics, rot, coords, rng)
    src = np.array([0.53, 0.47, 0.51])[:dim] * Lb
    recs = [np.array([0.2 + 0.2 * k, 0.8, 0.35])[:dim] * Lb for k in range(4)]
    dt = 0.3 * h * (1 - 0.6546536707079772) / 2 / (3000.0 * 1.05)
    check_mesh(dim, physics, 4, rot, coords, par, src, recs, dt)


@pytest.mark.parametrize("case", ["quad", "quad-distorted", "hex", "hex-distorted", "seven-hex"])
def test_node_coords_match_oracle(case):
    """GLL node coordinates per DOF (buildNodalPoints).  The oracle evaluates
    the reference's 15-digit GLL literals, the library its own FP64 GLL
    nodes; the two agree to the literals' rounding (1e-15 of the extent)."""
    if case == "seven-hex":
        cells, coords = rm.seven_hex()
        dim, order = 3, 3
    else:
        dim = 2 if case.startswith("quad") else 3
        P = pb.make(dim, 0, 6 if dim == 2 else 3, order=4, seed=12, distorted="distorted" in case)
        cells, coords, order = P["cells"], P["coords"], 4
    par = np.full(len(cells) * 2 ** dim, 3000.0)
    M = sem.Mesh(dim, order, 0, cells, coords, par)
    S = po.Sim(dim, order, 0, cells, coords, par)
    got, ref = M.node_coords(), S.node_coords()
    assert got.shape == ref.shape
    scale = np.ptp(np.asarray(coords))
    assert np.abs(got - ref).max() <= 1e-14 * scale
