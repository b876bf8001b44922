"""Vertex material from elemental model variables, the way the reference's
ExodusModel serves them to the elements.

Each element asks for every parameter at each of its vertices
(TensorQuad.cpp:243-251, Hexahedra.cpp:414-425); the model answers with the
elemental variable "<NAME>_<vertex>" of the model element whose centre is
nearest the element's centre (a k-d tree over model element centres,
ExodusModel.cpp:422-471), where the centre is the vertex mean
(TensorQuad.cpp:239, Hexahedra.cpp:157).  An isotropic "VP" request falls
back to "VPV" and an anisotropic "VPV" request to "VP"
(ExodusModel.cpp:440-456).  The physics ask for (Scalar.cpp:24,
Elastic2D.cpp:29-36, Elastic3D.cpp:33-40):

  acoustic         VP
  Elastic2D        RHO, C11, C13, C33, C55
  Elastic3D        RHO, VPV, VPH, VSV, VSH, ETA

`vertex_parameters` returns the E x nv x npar array that `sem.Mesh` takes;
`element_fluid` and `coupled_parameters` serve a mixed fluid/elastic model;
`exodus.read_model` supplies the model's element centres and elemental
variables from an Exodus file.
"""
import numpy as np
from scipy.spatial import cKDTree

PARAMETERS = {
    (2, 0): ("VP",),
    (3, 0): ("VP",),
    (2, 1): ("RHO", "C11", "C13", "C33", "C55"),
    (3, 1): ("RHO", "VPV", "VPH", "VSV", "VSH", "ETA"),
}


def element_centres(cells, coords):
    """Vertex mean of every element (cells: E x nv vertex ids)."""
    return np.asarray(coords, dtype=np.float64)[np.asarray(cells)].mean(axis=1)


def resolve_name(name, variables):
    """The elemental variable family a parameter request is served from."""
    if name + "_0" in variables:
        return name
    if name == "VP" and "VPV_0" in variables:
        return "VPV"
    if name == "VPV" and "VP_0" in variables:
        return "VP"
    raise KeyError("Requested parameter %s which is not stored as an elemental variable" % name)


def vertex_parameters(dim, physics, cells, coords, model_centres, variables):
    """E x nv x npar vertex material for sem.Mesh.

    model_centres: M x dim centres of the model's elements; variables:
    {"<NAME>_<v>": length-M array} for every parameter the physics needs and
    every vertex v."""
    if not variables:
        raise KeyError("You've tried to query a elemental parameter, but none are defined.")
    cells = np.asarray(cells)
    names = [resolve_name(n, variables) for n in PARAMETERS[(dim, physics)]]
    _, nearest = cKDTree(np.asarray(model_centres, dtype=np.float64)).query(
        element_centres(cells, coords))
    nv = cells.shape[1]
    out = np.empty((len(cells), nv, len(names)))
    for q, name in enumerate(names):
        for v in range(nv):
            out[:, v, q] = np.asarray(variables["%s_%d" % (name, v)], dtype=np.float64)[nearest]
    return out


def element_fluid(dim, cells, coords, model_centres, variables):
    """Per element, True where the model marks it fluid and False where it
    marks it elastic (ExodusModel.cpp:473-518): the elemental variable
    "fluid" of the model element nearest the element's centre, truncated to
    an integer, 1 fluid, 0 elastic; any other value, or no "fluid" variable,
    is an error.  A hexahedral mesh that mixes the two is refused, as the
    reference builds no 3-D fluid-solid coupled element (Element.cpp:60-112)."""
    if "fluid" not in variables:
        raise KeyError("Fluid flag not found in mesh. An elemental variable "
                       "'fluid' (1 fluid, 0 elastic) must mark every element.")
    cells = np.asarray(cells)
    _, nearest = cKDTree(np.asarray(model_centres, dtype=np.float64)).query(
        element_centres(cells, coords))
    flag = np.trunc(np.asarray(variables["fluid"], dtype=np.float64)[nearest])
    bad = (flag != 0) & (flag != 1)
    if bad.any():
        raise ValueError("element %d: unknown physics flag %g (0 elastic, 1 fluid)"
                         % (int(np.flatnonzero(bad)[0]), flag[bad][0]))
    fluid = flag == 1
    if dim == 3 and fluid.any() and not fluid.all():
        raise ValueError("a hexahedral mesh with both fluid and elastic elements "
                         "has no coupled element to build")
    return fluid


def coupled_parameters(cells, coords, model_centres, variables):
    """Everything a 2-D fluid-solid problem reads from the model: the fluid
    flags (element_fluid), the fluid cells' vertex VP (E_f x 4 x 1), the
    solid cells' vertex RHO, C11, C13, C33, C55 (E_s x 4 x 5) and the fluid
    cells' vertex RHO (E_f x 4), whose mean per cell is the density rho_0 the
    solid side of the interface reads (AcousticToElastic2D.cpp:33-41)."""
    cells = np.asarray(cells)
    fluid = element_fluid(2, cells, coords, model_centres, variables)
    par_f = vertex_parameters(2, 0, cells[fluid], coords, model_centres, variables)
    par_s = vertex_parameters(2, 1, cells[~fluid], coords, model_centres, variables)
    name = resolve_name("RHO", variables)  # the reference's error when absent
    _, nearest = cKDTree(np.asarray(model_centres, dtype=np.float64)).query(
        element_centres(cells[fluid], coords))
    rho = np.stack([np.asarray(variables["%s_%d" % (name, v)], dtype=np.float64)[nearest]
                    for v in range(cells.shape[1])], axis=1)
    return fluid, par_f, par_s, rho
