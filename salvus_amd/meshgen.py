"""Structured quad/hex box meshes as cell lists in the reference's vertex
order (QuadP1.h:98-108: v0(-1,-1) v1(+1,-1) v2(+1,+1) v3(-1,+1); HexP1.cpp:
123-133: v0(-1,-1,-1) v1(-1,+1,-1) v2(+1,+1,-1) v3(+1,-1,-1) v4(-1,-1,+1)
v5(+1,-1,+1) v6(+1,+1,+1) v7(-1,+1,+1)), i.e. the input that
Mesh::read(dim, numCells, numVerts, numVertsPerElem, cells, coords)
(Mesh.cpp:31-51) takes.  Cells are numbered x-fastest, vertices x-fastest.
"""
import numpy as np


def box_quad(nx, ny, lx=1.0, ly=1.0, x0=0.0, y0=0.0):
    """nx*ny quads on [x0,x0+lx]x[y0,y0+ly]; returns (cells[E,4] int32,
    coords[V,2] float64)."""
    xs = x0 + lx * np.arange(nx + 1, dtype=np.float64) / nx
    ys = y0 + ly * np.arange(ny + 1, dtype=np.float64) / ny
    X, Y = np.meshgrid(xs, ys, indexing="xy")  # [ny+1, nx+1]
    coords = np.stack([X.ravel(), Y.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i, j = i.ravel(), j.ravel()
    vid = lambda a, b: a + (nx + 1) * b  # noqa: E731
    cells = np.stack([vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)], axis=1)
    return cells.astype(np.int32), coords


def box_hex(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, x0=0.0, y0=0.0, z0=0.0):
    """nx*ny*nz hexes; returns (cells[E,8] int32, coords[V,3] float64)."""
    xs = x0 + lx * np.arange(nx + 1, dtype=np.float64) / nx
    ys = y0 + ly * np.arange(ny + 1, dtype=np.float64) / ny
    zs = z0 + lz * np.arange(nz + 1, dtype=np.float64) / nz
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()

    def vid(a, b, c):
        return a + (nx + 1) * (b + (ny + 1) * c)

    cells = np.stack([vid(i, j, k), vid(i, j + 1, k), vid(i + 1, j + 1, k), vid(i + 1, j, k),
                      vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1),
                      vid(i, j + 1, k + 1)], axis=1)
    return cells.astype(np.int32), coords
