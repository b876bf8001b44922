"""Exodus II mesh and model files, read the way the reference reads them.

Exodus II files are netCDF files; the classic and 64-bit-offset netCDF
formats are read with scipy.io.netcdf_file (netCDF-4 / HDF5-based Exodus
files are refused with a clear error).  Names follow the Exodus II layout:
num_dim, num_nodes, num_elem, coordx/coordy/coordz (or coord[num_dim]),
eb_prop1 (element block ids), connect<b> (1-based node ids of block b,
attribute elem_type), name_elem_var, vals_elem_var<i>eb<b>[time_step],
ss_names, elem_ss<s>, side_ss<s>.

* read_model follows ExodusModel::read (ExodusModel.cpp:207-260, :190-205,
  :414-420, :350-378, :262-300): connectivity of the first element block,
  the elemental variables of the block with id 1 at time step 1, and model
  element centres as vertex means of that connectivity -- what
  model.vertex_parameters takes.
* read_mesh follows Mesh::read (Mesh.cpp:53-85, DMPlexCreateExodusFromFile):
  every element block in order, 0-based vertices, in the reference's vertex
  order.  Quadrilaterals keep the Exodus QUAD4 order (counter-clockwise, as
  QuadP1.h:98-108).  Exodus HEX8 lists the bottom face counter-clockwise
  seen from +z while the reference's hexahedra list it clockwise
  (HexP1.cpp:123-133), so PETSc's Exodus reader inverts hexahedra: nodes 2
  and 4 are exchanged.  Side sets are returned as (name, elements, sides),
  0-based elements and Exodus side numbers.
"""
import numpy as np
from scipy.io import netcdf_file


class ExodusError(ValueError):
    pass


def _open(path):
    with open(path, "rb") as f:
        magic = f.read(4)
    if magic[:3] == b"\x89HD":
        raise ExodusError("%s is a netCDF-4 (HDF5-based) Exodus file; only the classic and 64-bit "
                          "offset netCDF formats are read" % path)
    if magic[:3] != b"CDF" or magic[3] not in (1, 2):
        raise ExodusError("%s is not a classic or 64-bit offset netCDF file" % path)
    return netcdf_file(str(path), "r", mmap=False)


def _strings(var):
    """Rows of a netCDF char array as str (NUL padding stripped)."""
    a = np.asarray(var[:])
    if a.ndim == 1:
        a = a[None, :]
    return [b"".join(bytes(c) for c in row).split(b"\x00")[0].decode().strip() for row in a]


def _dim(nc, name, default=0):
    d = nc.dimensions.get(name)
    return default if d is None else int(d)


def _coords(nc):
    dim = _dim(nc, "num_dim")
    if dim not in (2, 3):
        raise ExodusError("num_dim must be 2 or 3, got %d" % dim)
    v = nc.variables
    if "coordx" in v:
        cols = [np.asarray(v["coord" + a][:], dtype=np.float64) for a in "xyz"[:dim]]
    elif "coord" in v:
        c = np.asarray(v["coord"][:], dtype=np.float64)
        cols = [c[k] for k in range(dim)]
    else:
        raise ExodusError("no nodal coordinates (coordx/coordy[/coordz] or coord)")
    return dim, np.stack(cols, axis=1)


def _block_ids(nc):
    nblk = _dim(nc, "num_el_blk")
    if nblk < 1:
        raise ExodusError("the file holds no element blocks")
    if "eb_prop1" in nc.variables:
        return [int(x) for x in np.asarray(nc.variables["eb_prop1"][:])]
    return list(range(1, nblk + 1))


def _connect(nc, b):
    """1-based connectivity of block index b (1-based) and its element type."""
    var = nc.variables.get("connect%d" % b)
    if var is None:
        raise ExodusError("element block %d has no connectivity (connect%d)" % (b, b))
    etype = getattr(var, "elem_type", b"")
    etype = etype.decode() if isinstance(etype, bytes) else str(etype)
    return np.asarray(var[:], dtype=np.int64), etype.upper()


def read_model(path):
    """(dim, model element centres M x dim, {variable name: length-M array})
    as ExodusModel holds them: elemental variables of block id 1 at time
    step 1 (none: an empty dict)."""
    nc = _open(path)
    try:
        dim, xyz = _coords(nc)
        conn, _ = _connect(nc, 1)  # the first block (ExodusModel.cpp:414-420)
        centres = xyz[conn - 1].mean(axis=1)
        variables = {}
        nvar = _dim(nc, "num_elem_var")
        if nvar > 0 and "name_elem_var" in nc.variables:
            names = _strings(nc.variables["name_elem_var"])
            ids = _block_ids(nc)
            if 1 not in ids:
                raise ExodusError("no element block with id 1 for the elemental variables")
            b = ids.index(1) + 1
            for i, name in enumerate(names[:nvar]):
                var = nc.variables.get("vals_elem_var%deb%d" % (i + 1, b))
                if var is None:
                    continue  # not defined on this block (elem_var_tab)
                variables[name] = np.array(var[0], dtype=np.float64)  # time step 1
        return dim, centres, variables
    finally:
        nc.close()


def read_mesh(path):
    """(dim, cells E x nv (0-based, reference vertex order), coords N x dim,
    side sets [(name, elements, sides)])."""
    nc = _open(path)
    try:
        dim, xyz = _coords(nc)
        blocks = []
        for b in range(1, len(_block_ids(nc)) + 1):
            conn, etype = _connect(nc, b)
            nv = conn.shape[1]
            if dim == 2 and nv != 4:
                raise ExodusError("block %d: %s elements; 2-D meshes must be QUAD4" % (b, etype or nv))
            if dim == 3 and nv != 8:
                raise ExodusError("block %d: %s elements; 3-D meshes must be HEX8" % (b, etype or nv))
            cells = conn - 1
            if dim == 3:
                cells = cells[:, [0, 3, 2, 1, 4, 5, 6, 7]]
            blocks.append(cells)
        cells = np.ascontiguousarray(np.concatenate(blocks).astype(np.int32))
        nss = _dim(nc, "num_side_sets")
        names = _strings(nc.variables["ss_names"]) if nss and "ss_names" in nc.variables else []
        side_sets = []
        for s in range(1, nss + 1):
            el = nc.variables.get("elem_ss%d" % s)
            sd = nc.variables.get("side_ss%d" % s)
            if el is None or sd is None:
                continue
            name = names[s - 1] if s - 1 < len(names) else ""
            side_sets.append((name, np.asarray(el[:], dtype=np.int64) - 1, np.asarray(sd[:], dtype=np.int64)))
        return dim, cells, xyz, side_sets
    finally:
        nc.close()
