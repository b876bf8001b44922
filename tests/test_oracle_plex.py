"""Pins plex-lite (gather map, scatter order, global numbering): DMPlex point numbering against the
reference's test_Mesh.cpp facts, the closure gather map against the 7-hex
known-answer test of test_TensorHex.cpp:400-603, and the invariants that
hold for any correct gather map."""
import numpy as np
import pytest

from oracle import pyoracle as po
from salvus_amd import meshgen
from tests import refmeshes as rm


def test_quad_mesh_numbering():
    cells, coords = rm.quad_4x4()
    P = po.Plex(2, 4, cells, 25)
    assert (P.E, P.V, P.NEd) == (16, 25, 40)
    assert P.cell_cones()[0].tolist() == [41, 42, 43, 44]  # test_Mesh.cpp:84-85
    cones = P.cell_cones()
    assert sorted(np.nonzero((cones == 42).any(1))[0].tolist()) == [0, 1]  # :90
    assert sorted(np.nonzero((cones == 43).any(1))[0].tolist()) == [0, 4]  # :91
    assert (cones == 41).any(1).sum() == 1 and (cones == 44).any(1).sum() == 1  # :88-89
    assert rm.boundary_point_sets(P, coords, 2) == rm.QUAD_4x4_BOUNDARY


def test_hex_mesh_numbering():
    cells, coords = rm.hex_2x2x2()
    P = po.Plex(3, 4, cells, 27)
    assert (P.E, P.V, P.NF) == (8, 27, 36)
    assert P.cell_cones()[0].tolist() == [35, 36, 37, 38, 39, 40]  # test_Mesh.cpp:192-193
    cones = P.cell_cones()
    for face, other in ((36, 1), (38, 4), (39, 2)):  # :199-201
        assert sorted(np.nonzero((cones == face).any(1))[0].tolist()) == [0, other]
    for face in (35, 37, 40):  # :196-198
        assert (cones == face).any(1).sum() == 1
    assert rm.boundary_point_sets(P, coords, 3) == rm.HEX_2x2x2_BOUNDARY


def test_section_offsets_follow_point_order():
    cells, coords = rm.hex_2x2x2()
    N = 4
    P = po.Plex(3, N, cells, 27)
    off = P.offsets()
    sizes = ([(N - 1) ** 3] * P.E + [1] * P.V + [(N - 1) ** 2] * P.NF + [N - 1] * P.NEd)
    np.testing.assert_array_equal(off, np.concatenate([[0], np.cumsum(sizes)[:-1]]))
    assert P.nglobal == sum(sizes) == (2 * N + 1) ** 3


def test_seven_hex_closure_kat():
    """test_TensorHex.cpp:400-603: ones written on a face of each neighbour
    light exactly the matching face of the central element (order 3)."""
    cells, coords = rm.seven_hex()
    order = 3
    P = po.Plex(3, order, cells, 32)
    conn = P.conn()
    insert_element = [1, 2, 3, 4, 5, 6]
    insert_faces = [1, 0, 3, 2, 5, 4]
    lit_face = {1: 0, 2: 1, 3: 2, 4: 3, 5: 4, 6: 5}  # bot, top, fnt, bck, rgt, lft
    for el, fa in zip(insert_element, insert_faces):
        g = np.zeros(P.nglobal)
        g[conn[el][rm.hex_face_nodes(order, fa)]] = 1.0
        expect = np.zeros((order + 1) ** 3)
        expect[rm.hex_face_nodes(order, lit_face[el])] = 1.0
        np.testing.assert_array_equal(g[conn[0]], expect)


def _coordinate_consistency(dim, order, cells, coords):
    nv = 4 if dim == 2 else 8
    S = po.Sim(dim, order, po.ACOUSTIC, cells, coords, np.ones(len(cells) * nv))
    xyz, conn = S.node_coords(), S.conn()
    scale = np.abs(coords).max() + 1
    for e in range(len(cells)):
        pts = po.nodal_points(order, coords[cells[e]])
        np.testing.assert_allclose(xyz[conn[e]], pts, atol=1e-12 * scale)
    assert np.unique(conn).size == S.G


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_gather_map_coordinate_consistency(order):
    _coordinate_consistency(3, order, *rm.seven_hex())
    _coordinate_consistency(3, order, *meshgen.box_hex(3, 2, 4, 3.0, 2.0, 4.0))
    _coordinate_consistency(2, order, *meshgen.box_quad(5, 3, 5.0, 3.0))
    _coordinate_consistency(2, order, *rm.quad_4x4())
    _coordinate_consistency(3, order, *rm.hex_2x2x2())


def test_unique_node_count_structured():
    for (nx, ny, nz, N) in ((3, 4, 5, 4), (2, 2, 2, 3)):
        P = po.Plex(3, N, meshgen.box_hex(nx, ny, nz)[0], (nx + 1) * (ny + 1) * (nz + 1))
        assert P.nglobal == (nx * N + 1) * (ny * N + 1) * (nz * N + 1)
    P = po.Plex(2, 4, meshgen.box_quad(50, 50)[0], 51 * 51)
    assert P.nglobal == 40401  # configuration 1: 50x50 quads, order 4


@pytest.mark.parametrize("dim", [2, 3])
def test_colouring_is_race_free_and_parity_on_boxes(dim):
    if dim == 2:
        cells, coords = meshgen.box_quad(7, 5)
        shape = (5, 7)
    else:
        cells, coords = meshgen.box_hex(5, 4, 3)
        shape = (3, 4, 5)
    col, n = po.color_elements(dim, cells, len(coords))
    assert n == 2 ** dim
    P = po.Plex(dim, 4, cells, len(coords))
    conn = P.conn()
    for c in range(n):
        nodes = conn[col == c].ravel()
        assert np.unique(nodes).size == nodes.size  # no node shared within a colour
    # greedy in x-fastest order gives the 2^d parity colouring on a box
    idx = np.indices(shape).reshape(dim, -1)[::-1]  # (i, j[, k])
    parity = sum((idx[a] % 2) << a for a in range(dim))
    np.testing.assert_array_equal(col, parity)
