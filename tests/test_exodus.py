"""Exodus II files (host only): meshes read as Mesh::read takes them
(DMPlexCreateExodusFromFile: all blocks, reference vertex order) and models
as ExodusModel::read holds them (first block, elemental variables of block id
1 at time step 1).  The files are written with scipy's netCDF writer in the
Exodus II layout, independently of the reader under test; the hexahedron
vertex permutation is pinned to the reference's vertex table
(test_HexP1.cpp:21-48)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from salvus_amd import exodus, meshgen, model, sem

HEX_REF = np.array([[-1, -1, -1], [-1, 1, -1], [1, 1, -1], [1, -1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
# Exodus HEX8 canonical node order: bottom face counter-clockwise seen from +z
EXO_HEX8 = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


def _chars(strings, width=33):
    a = np.zeros((len(strings), width), dtype="S1")
    for i, s in enumerate(strings):
        for j, ch in enumerate(s.encode()):
            a[i, j] = bytes([ch])
    return a


def write_exodus(path, xyz, blocks, elem_vars=None, side_sets=(), large=True, block_ids=None):
    """blocks: list of (1-based connectivity, elem_type); elem_vars: {name:
    values over block 1}; side_sets: [(name, 1-based elements, sides)]."""
    f = netcdf_file(str(path), "w", version=2)
    n, dim = xyz.shape
    f.createDimension("time_step", None)  # scipy: the record dimension comes first
    f.createDimension("len_name", 33)
    f.createDimension("num_dim", dim)
    f.createDimension("num_nodes", n)
    f.createDimension("num_elem", sum(len(c) for c, _ in blocks))
    f.createDimension("num_el_blk", len(blocks))
    t = f.createVariable("time_whole", "d", ("time_step",))
    t[0] = 0.0
    if large:
        for k, a in enumerate("xyz"[:dim]):
            f.createVariable("coord" + a, "d", ("num_nodes",))[:] = xyz[:, k]
    else:
        f.createVariable("coord", "d", ("num_dim", "num_nodes"))[:] = xyz.T
    ids = block_ids or list(range(1, len(blocks) + 1))
    prop = f.createVariable("eb_prop1", "i", ("num_el_blk",))
    prop[:] = ids
    prop.name = b"ID"
    for b, (conn, etype) in enumerate(blocks, start=1):
        f.createDimension("num_el_in_blk%d" % b, len(conn))
        f.createDimension("num_nod_per_el%d" % b, conn.shape[1])
        v = f.createVariable("connect%d" % b, "i", ("num_el_in_blk%d" % b, "num_nod_per_el%d" % b))
        v[:] = conn
        v.elem_type = etype.encode()
    if elem_vars:
        names = list(elem_vars)
        f.createDimension("num_elem_var", len(names))
        f.createVariable("name_elem_var", "c", ("num_elem_var", "len_name"))[:] = _chars(names)
        for i, name in enumerate(names, start=1):
            v = f.createVariable("vals_elem_var%deb1" % i, "d", ("time_step", "num_el_in_blk1"))
            v[0] = elem_vars[name]
    if side_sets:
        f.createDimension("num_side_sets", len(side_sets))
        f.createVariable("ss_names", "c", ("num_side_sets", "len_name"))[:] = _chars([s[0] for s in side_sets])
        for s, (_, el, sd) in enumerate(side_sets, start=1):
            f.createDimension("num_side_ss%d" % s, len(el))
            f.createVariable("elem_ss%d" % s, "i", ("num_side_ss%d" % s,))[:] = el
            f.createVariable("side_ss%d" % s, "i", ("num_side_ss%d" % s,))[:] = sd
    f.close()


def test_hex8_order_becomes_the_reference_vertex_order(tmp_path):
    path = tmp_path / "cube.e"
    write_exodus(path, EXO_HEX8, [(np.arange(1, 9)[None, :], "HEX8")])
    dim, cells, xyz, ss = exodus.read_mesh(path)
    assert dim == 3 and cells.shape == (1, 8) and ss == []
    np.testing.assert_array_equal(xyz[cells[0]], HEX_REF)


def test_quad_mesh_blocks_side_sets_and_model(tmp_path):
    cells, coords = meshgen.box_quad(3, 2, 30.0, 20.0)
    E = len(cells)
    path = tmp_path / "quads.e"
    rng = np.random.default_rng(2)
    vp = {"VP_%d" % v: rng.uniform(1000, 2000, E // 2) for v in range(4)}
    write_exodus(path, coords, [(cells[: E // 2] + 1, "QUAD4"), (cells[E // 2:] + 1, "QUAD")],
                 elem_vars=vp, side_sets=[("absorbing_boundary", [1, 4], [4, 1])], large=False)
    dim, got_cells, xyz, ss = exodus.read_mesh(path)
    assert dim == 2
    np.testing.assert_array_equal(got_cells, cells)  # blocks in order, QUAD4 order kept
    np.testing.assert_array_equal(xyz, coords)
    assert len(ss) == 1 and ss[0][0] == "absorbing_boundary"
    np.testing.assert_array_equal(ss[0][1], [0, 3])
    np.testing.assert_array_equal(ss[0][2], [4, 1])
    # the model: first block only, as ExodusModel holds it
    mdim, centres, variables = exodus.read_model(path)
    assert mdim == 2 and set(variables) == set(vp)
    np.testing.assert_allclose(centres, model.element_centres(cells[: E // 2], coords))
    for k in vp:
        np.testing.assert_array_equal(variables[k], vp[k])
    # served to the first block's own elements, the variables come back
    par = model.vertex_parameters(2, 0, cells[: E // 2], coords, centres, variables)
    np.testing.assert_array_equal(par[:, :, 0], np.stack([vp["VP_%d" % v] for v in range(4)], axis=1))


def test_hex_mesh_feeds_the_topology(tmp_path):
    cells, coords = meshgen.box_hex(2, 2, 1, 20.0, 20.0, 10.0)
    exo = cells[:, [0, 3, 2, 1, 4, 5, 6, 7]] + 1  # the reference's order back to Exodus HEX8
    path = tmp_path / "hexes.e"
    names = ["RHO", "VPV", "VPH", "VSV", "VSH", "ETA"]
    ev = {"%s_%d" % (n, v): np.full(len(cells), 1.0 + q + 0.1 * v) for q, n in enumerate(names)
          for v in range(8)}
    write_exodus(path, coords, [(exo, "HEX8")], elem_vars=ev)
    dim, got, xyz, _ = exodus.read_mesh(path)
    np.testing.assert_array_equal(got, cells)
    conn, colour, info = sem.topology(3, 4, got, len(xyz))
    assert conn.shape == (len(cells), 125)
    _, centres, variables = exodus.read_model(path)
    par = model.vertex_parameters(3, 1, got, xyz, centres, variables)
    assert par.shape == (len(cells), 8, 6) and par[0, 7, 5] == pytest.approx(6.7)


def test_errors(tmp_path):
    h5 = tmp_path / "nc4.e"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(exodus.ExodusError, match="netCDF-4"):
        exodus.read_mesh(h5)
    junk = tmp_path / "junk.e"
    junk.write_bytes(b"not a mesh")
    with pytest.raises(exodus.ExodusError, match="not a classic"):
        exodus.read_model(junk)
    tri = tmp_path / "tri.e"
    write_exodus(tri, np.array([[0.0, 0], [1, 0], [0, 1]]), [(np.array([[1, 2, 3]]), "TRI3")])
    with pytest.raises(exodus.ExodusError, match="QUAD4"):
        exodus.read_mesh(tri)
    other = tmp_path / "ids.e"
    cells, coords = meshgen.box_quad(1, 1)
    write_exodus(other, coords, [(cells + 1, "QUAD4")], elem_vars={"VP_0": np.ones(1)}, block_ids=[7])
    with pytest.raises(exodus.ExodusError, match="id 1"):
        exodus.read_model(other)
    plain = tmp_path / "novars.e"
    write_exodus(plain, coords, [(cells + 1, "QUAD4")])
    assert exodus.read_model(plain)[2] == {}


def _mixed_model_file(path, cells, coords, fluid):
    """A quad model file with a per-element fluid flag, fluid VP and RHO and
    solid RHO, C11, C13, C33, C55 (zeros where a family does not apply)."""
    E = len(cells)
    idx = np.arange(E, dtype=float)
    ev = {"fluid": fluid.astype(float)}
    for v in range(4):
        ev["VP_%d" % v] = np.where(fluid, 1500.0 + idx + 0.25 * v, 0.0)
        ev["RHO_%d" % v] = np.where(fluid, 1000.0 + 3 * idx + v, 2500.0 + v)
        for k, name in enumerate(("C11", "C13", "C33", "C55")):
            ev["%s_%d" % (name, v)] = np.where(fluid, 0.0, 1e10 * (k + 1) + 1e6 * idx + v)
    write_exodus(path, coords, [(cells + 1, "QUAD4")], elem_vars=ev)
    return ev


def test_mixed_model_flags_and_coupled_parameters(tmp_path):
    """ExodusModel::getElementType (ExodusModel.cpp:473-518): the "fluid"
    variable of the nearest model element, 1 fluid, 0 elastic, other values
    and a missing flag errors; the coupled inputs (fluid VP, solid RHO and
    C-tensor, fluid vertex RHO for rho_0) come back per split part and feed
    split_coupled."""
    cells, coords = meshgen.box_quad(6, 4, 60.0, 40.0)
    cent = model.element_centres(cells, coords)
    fluid = (cent[:, 0] < 30.0) | (cent[:, 1] > 30.0)  # an L-shaped fluid
    path = tmp_path / "mixed.e"
    ev = _mixed_model_file(path, cells, coords, fluid)
    dim, centres, variables = exodus.read_model(path)
    got, par_f, par_s, rho = model.coupled_parameters(cells, coords, centres, variables)
    np.testing.assert_array_equal(got, fluid)
    for v in range(4):
        np.testing.assert_array_equal(par_f[:, v, 0], ev["VP_%d" % v][fluid])
        np.testing.assert_array_equal(rho[:, v], ev["RHO_%d" % v][fluid])
        for q, name in enumerate(("RHO", "C11", "C13", "C33", "C55")):
            np.testing.assert_array_equal(par_s[:, v, q], ev["%s_%d" % (name, v)][~fluid])
    (cf, xf, pf), (cs, xs, ps) = sem.split_coupled(cells, coords, got, (par_f, par_s))
    np.testing.assert_array_equal(xf[cf], coords[cells[fluid]])
    np.testing.assert_array_equal(xs[cs], coords[cells[~fluid]])
    assert pf is par_f and ps is par_s
    # a flag read as 1.7 truncates to fluid; 2 is no physics; no flag at all
    v2 = dict(variables, fluid=np.where(fluid, 1.7, 0.2))
    np.testing.assert_array_equal(model.element_fluid(2, cells, coords, centres, v2), fluid)
    with pytest.raises(ValueError, match="unknown physics"):
        model.element_fluid(2, cells, coords, centres, dict(variables, fluid=np.full(len(cells), 2.0)))
    with pytest.raises(KeyError, match="Fluid flag"):
        model.element_fluid(2, cells, coords, centres, {k: a for k, a in variables.items() if k != "fluid"})
    # no RHO for rho_0: the reference's missing-parameter error
    no_rho = {k: a for k, a in variables.items() if not k.startswith("RHO_")}
    with pytest.raises(KeyError, match="RHO"):
        model.coupled_parameters(cells, coords, centres, no_rho)
    # hexahedra: all fluid or all elastic, never both
    hc, hx = meshgen.box_hex(2, 1, 1, 20.0, 10.0, 10.0)
    hcent = model.element_centres(hc, hx)
    assert model.element_fluid(3, hc, hx, hcent, {"fluid": np.ones(2)}).all()
    with pytest.raises(ValueError, match="hexahedral"):
        model.element_fluid(3, hc, hx, hcent, {"fluid": np.array([1.0, 0.0])})


def test_coupled_from_exodus_refusals(tmp_path):
    """sem.coupled_from_exodus refuses, before any device work, a hexahedral
    file (the coupling is 2-D) and a model whose flag marks every element
    alike."""
    hc, hx = meshgen.box_hex(2, 1, 1, 20.0, 10.0, 10.0)
    hexes = tmp_path / "hex.e"
    write_exodus(hexes, hx, [(hc[:, [0, 3, 2, 1, 4, 5, 6, 7]] + 1, "HEX8")], elem_vars={"fluid": np.ones(2)})
    with pytest.raises(sem.SemError, match="2-D"):
        sem.coupled_from_exodus(hexes)
    cells, coords = meshgen.box_quad(3, 2, 30.0, 20.0)
    allfluid = tmp_path / "fluid.e"
    _mixed_model_file(allfluid, cells, coords, np.ones(len(cells), dtype=bool))
    with pytest.raises(sem.SemError, match="every element fluid"):
        sem.coupled_from_exodus(allfluid)
