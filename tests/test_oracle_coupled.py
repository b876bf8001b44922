"""The coupled 2-D fluid-solid oracle (oracle/coupled.py, host only).

Without an interface it must be the single-physics oracle; its edge
quadrature and normals are checked against geometry; with an interface the
coupling must move the other medium only once the wave arrives, and the
solid's coupling load is the same in x and y (the reference adds it to both
rows without a normal, AcousticToElastic2D.cpp:54-55)."""
import numpy as np

from oracle import coupled as cp
from oracle import pyoracle as po
from salvus_amd import meshgen

H = 10.0


def _materials(cells, fluid):
    nf, ns = int(fluid.sum()), int((~fluid).sum())
    par_f = np.full((nf, 4, 1), 1500.0)
    rho_f = np.full((nf, 4), 1000.0)
    rho, vp, vs = 2500.0, 3000.0, 1732.0
    c11, c55, c13 = rho * vp ** 2, rho * vs ** 2, rho * (vp ** 2 - 2 * vs ** 2)
    par_s = np.tile(np.array([rho, c11, c13, c11, c55]), (ns, 4, 1))
    return par_f, rho_f, par_s


def _box(nx, ny):
    return meshgen.box_quad(nx, ny, nx * H, ny * H)


def test_all_fluid_is_the_scalar_oracle():
    cells, coords = _box(6, 6)
    fluid = np.ones(len(cells), dtype=bool)
    par_f, rho_f, par_s = _materials(cells, fluid)
    C = cp.CoupledSim(4, cells, coords, fluid, par_f, rho_f, par_s)
    assert C.edges == {"f": [], "s": []}
    S = po.Sim(2, 4, po.ACOUSTIC, cells, coords, par_f)
    loc, f0 = np.array([31.0, 27.0]), 1500.0 / (8 * H)
    C.add_source("f", loc, 1.0, f0, 1.2 / f0)
    S.add_source(loc, 1.0, f0, 1.2 / f0)
    dt = 0.3 * H * 0.17 / 1500.0
    u = C.run(dt, 150)["f"]
    S.run(dt, 150)
    ref = S.get_field("u")
    assert np.abs(ref).max() > 0
    np.testing.assert_allclose(u, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_all_solid_is_the_elastic_oracle():
    cells, coords = _box(5, 5)
    fluid = np.zeros(len(cells), dtype=bool)
    par_f, rho_f, par_s = _materials(cells, fluid)
    C = cp.CoupledSim(4, cells, coords, fluid, par_f, rho_f, par_s)
    S = po.Sim(2, 4, po.ELASTIC, cells, coords, par_s)
    loc, f0 = np.array([22.0, 26.0]), 3000.0 / (8 * H)
    C.add_source("s", loc, 1e3, f0, 1.2 / f0, [0.3, 1.0])
    S.add_source(loc, 1e3, f0, 1.2 / f0, [0.3, 1.0])
    dt = 0.3 * H * 0.17 / 3000.0
    u = C.run(dt, 120)["s"]
    S.run(dt, 120)
    ref = S.get_field("u")
    assert np.abs(ref).max() > 0
    np.testing.assert_allclose(u, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_edge_quadrature_and_normals():
    cells, coords = _box(4, 3)
    coords = coords + 0.7 * np.sin(coords[:, ::-1] / 7.0)  # a gently distorted mesh
    fluid = np.arange(len(cells)) % 4 < 2
    par_f, rho_f, par_s = _materials(cells, fluid)
    C = cp.CoupledSim(4, cells, coords, fluid, par_f, rho_f, par_s)
    w = po.gll(4)[1]
    for e in range(len(cells)):
        centre = coords[cells[e]].mean(axis=0)
        for i in range(4):
            half, nrm = C.edge_geometry(e, i)
            a, b = coords[cells[e][i]], coords[cells[e][(i + 1) % 4]]
            assert abs(np.linalg.norm(nrm) - 1.0) < 1e-14
            assert abs(np.dot(nrm, b - a)) < 1e-12           # normal to the edge
            assert np.dot(nrm, (a + b) / 2 - centre) > 0      # pointing out of the element
            assert abs(half * w.sum() - np.linalg.norm(b - a)) < 1e-12  # integral of 1 = length
    # the edge nodes run from vertex i to vertex i+1 (reference vertex order)
    n = 5
    corner = {0: 0, 1: n - 1, 2: n * n - 1, 3: n * (n - 1)}
    for i in range(4):
        nodes = cp._edge_nodes(i, n)
        assert nodes[0] == corner[i] and nodes[-1] == corner[(i + 1) % 4]


def test_interface_coupling():
    cells, coords = _box(8, 4)
    centres = coords[cells].mean(axis=1)
    fluid = centres[:, 0] < 4 * H  # left half fluid, right half solid
    par_f, rho_f, par_s = _materials(cells, fluid)
    C = cp.CoupledSim(4, cells, coords, fluid, par_f, rho_f, par_s)
    assert len(C.f2s) == 4 * 4 + 1  # the interface line x = 40 of order-4 nodes
    assert C.edges["s"] and all(rho0 == 1000.0 for _, _, rho0 in C.edges["s"])
    # shared nodes: one inverse mass for both fields
    F, S = C.part["f"], C.part["s"]
    for gf, gs in C.f2s.items():
        assert F["mi"][gf] == S["mi"][gs]
    f0 = 1500.0 / (8 * H)
    C.add_source("f", np.array([12.0, 21.0]), 1.0, f0, 1.2 / f0)
    dt = 0.3 * H * 0.17 / 3000.0
    # the discrete support grows by one element per step: after two steps
    # the source (in the second element column) has not reached x = 40
    C.run(dt, 2)
    assert np.abs(F["u"]).max() > 0 and np.abs(S["u"]).max() == 0.0
    C.run(dt, 600)
    assert np.abs(S["u"]).max() > 1e-15  # the fluid moved the solid (potential vs displacement units)
    s = C.surface()["s"]
    assert np.abs(s).max() > 0
    np.testing.assert_array_equal(s[0], s[1])  # same load in x and y
