"""Small meshes whose DMPlex numbering the reference's own tests pin.

quad_4x4 and hex_2x2x2 are the reference's test meshes quad_eigenfunction.e
and small_hex_mesh_to_test_sources.e (external files, not present), rebuilt
from the facts the tests state about them: element 0's vertex coordinates,
cone and neighbours, and the boundary point sets (test_Mesh.cpp:80-136,
:162-270).  seven_hex is the inline cell list of test_TensorHex.cpp:424-494.
"""
import numpy as np


def quad_4x4():
    # 16 cells x-fastest; 25 vertices numbered y-fastest (vertex 5*ix+iy);
    # element 0 = (0,0),(25000,0),(25000,25000),(0,25000).
    v = lambda ix, iy: 5 * ix + iy  # noqa: E731
    cells = [[v(cx, cy), v(cx + 1, cy), v(cx + 1, cy + 1), v(cx, cy + 1)]
             for cy in range(4) for cx in range(4)]
    coords = np.zeros((25, 2))
    for ix in range(5):
        for iy in range(5):
            coords[v(ix, iy)] = (25000.0 * ix, 25000.0 * iy)
    return np.array(cells, dtype=np.int32), coords


QUAD_4x4_BOUNDARY = {  # test_Mesh.cpp:120-129 (side set -> points)
    0: [16, 17, 18, 19, 20, 44, 56, 65, 74],
    1: [36, 37, 38, 39, 40, 52, 61, 70, 79],
    2: [16, 21, 26, 31, 36, 41, 45, 48, 51],
    3: [20, 25, 30, 35, 40, 73, 76, 78, 80],
}


def hex_2x2x2():
    # 8 cells x,y,z; vertex 9*ix+3*iy+iz; element 0 vertices as test_Mesh.cpp:180-189.
    v = lambda ix, iy, iz: 9 * ix + 3 * iy + iz  # noqa: E731
    offs = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)]
    cells = [[v(cx + o[0], cy + o[1], cz + o[2]) for o in offs]
             for cz in range(2) for cy in range(2) for cx in range(2)]
    coords = np.zeros((27, 3))
    for ix in range(3):
        for iy in range(3):
            for iz in range(3):
                coords[v(ix, iy, iz)] = (5e4 * ix, 5e4 * iy, 5e4 * iz)
    return np.array(cells, dtype=np.int32), coords


HEX_2x2x2_BOUNDARY = {  # test_Mesh.cpp:226-268; sides x0,x1,y0,y1,z0,z1
    0: [8, 9, 10, 11, 12, 13, 14, 15, 16, 35, 46, 55, 64, 71, 72, 73, 74, 91, 92, 93, 104, 105,
        106, 117, 118],
    1: [26, 27, 28, 29, 30, 31, 32, 33, 34, 41, 51, 60, 68, 83, 84, 85, 86, 99, 100, 101, 112,
        113, 114, 122, 123],
    2: [8, 9, 10, 17, 18, 19, 26, 27, 28, 40, 45, 59, 63, 71, 78, 80, 81, 86, 88, 89, 104, 109,
        110, 114, 115],
    3: [14, 15, 16, 23, 24, 25, 32, 33, 34, 50, 54, 67, 70, 92, 95, 97, 98, 100, 102, 103, 118,
        119, 121, 122, 124],
    4: [8, 11, 14, 17, 20, 23, 26, 29, 32, 37, 42, 48, 52, 74, 75, 79, 80, 83, 87, 88, 93, 94,
        97, 99, 102],
    5: [10, 13, 16, 19, 22, 25, 28, 31, 34, 57, 61, 66, 69, 105, 108, 110, 111, 113, 115, 116,
        117, 120, 121, 123, 124],
}


def seven_hex():
    cells = np.array([0, 1, 2, 3, 4, 5, 6, 7,
                      12, 13, 14, 15, 0, 3, 2, 1,
                      4, 7, 6, 5, 8, 9, 10, 11,
                      16, 0, 3, 19, 17, 18, 5, 4,
                      1, 20, 23, 2, 7, 6, 22, 21,
                      3, 2, 25, 24, 5, 27, 26, 6,
                      28, 29, 1, 0, 31, 4, 7, 30], dtype=np.int32).reshape(7, 8)
    coords = np.array([0, 0, 0, 0, 1, 0, 1, 1, 0, 1, 0, 0, 0, 0, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1,
                       0, 0, 2, 1, 0, 2, 1, 1, 2, 0, 1, 2, 0, 0, -1, 0, 1, -1, 1, 1, -1, 1, 0, -1,
                       0, -1, 0, 0, -1, 1, 1, -1, 1, 1, -1, 0, 0, 2, 0, 0, 2, 1, 1, 2, 1, 1, 2, 0,
                       2, 0, 0, 2, 1, 0, 2, 1, 1, 2, 0, 1, -1, 0, 0, -1, 1, 0, -1, 1, 1, -1, 0, 1],
                      dtype=np.float64).reshape(32, 3)
    return cells, coords


def hex_face_nodes(order, face):
    """Element-local nodes of hex face `face` (Hexahedra.cpp:269-317:
    0 bottom t=-1, 1 top, 2 front s=-1, 3 back, 4 right r=+1, 5 left)."""
    n, N = order + 1, order
    out = []
    for p in range(n ** 3):
        i, j, k = p % n, (p // n) % n, p // (n * n)
        hit = {0: k == 0, 1: k == N, 2: j == 0, 3: j == N, 4: i == N, 5: i == 0}[face]
        if hit:
            out.append(p)
    return out


def boundary_point_sets(plex, coords, dim):
    """Points (vertices, edges, faces) lying in each axis-aligned boundary
    plane x0,x1,y0,y1(,z0,z1) of a box mesh."""
    E, V = plex.E, plex.V
    ec = plex.edge_cones()
    fc = plex.face_cones() if dim == 3 else np.zeros((0, 4), dtype=np.int32)
    lo, hi = coords.min(axis=0), coords.max(axis=0)
    sets = {}
    for a in range(dim):
        for side, val in enumerate((lo[a], hi[a])):
            on = coords[:, a] == val
            s = set((E + np.nonzero(on)[0]).tolist())
            for k, f in enumerate(fc):
                if on[f - E].all():
                    s.add(E + V + k)
            for k, (p, q) in enumerate(ec):
                if on[p - E] and on[q - E]:
                    s.add(E + V + plex.NF + k)
            sets[2 * a + side] = sorted(s)
    return sets
