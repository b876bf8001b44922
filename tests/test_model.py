"""Elemental model variables to vertex material (ExodusModel.cpp:422-471
semantics): nearest model element centre, <NAME>_<vertex> variables, the
VP/VPV fallbacks and the physics' parameter lists."""
import numpy as np
import pytest

from salvus_amd import meshgen, model


def test_same_mesh_returns_the_variables():
    cells, coords = meshgen.box_hex(3, 2, 2, 30.0, 20.0, 20.0)
    rng = np.random.default_rng(1)
    E = len(cells)
    names = model.PARAMETERS[(3, 1)]
    truth = rng.uniform(1, 2, size=(E, 8, len(names)))
    variables = {"%s_%d" % (n, v): truth[:, v, q] for q, n in enumerate(names) for v in range(8)}
    centres = model.element_centres(cells, coords)
    perm = rng.permutation(E)  # model elements in another order
    vars_perm = {k: a[perm] for k, a in variables.items()}
    got = model.vertex_parameters(3, 1, cells, coords, centres[perm], vars_perm)
    np.testing.assert_array_equal(got, truth)


def test_nearest_coarse_element():
    cells, coords = meshgen.box_quad(8, 8, 80.0, 80.0)
    mcells, mcoords = meshgen.box_quad(3, 3, 80.0, 80.0)
    mcent = model.element_centres(mcells, mcoords)
    M = len(mcells)
    variables = {"VP_%d" % v: 1000.0 + 10.0 * np.arange(M) + v for v in range(4)}
    got = model.vertex_parameters(2, 0, cells, coords, mcent, variables)
    cent = model.element_centres(cells, coords)
    brute = np.argmin(((cent[:, None, :] - mcent[None, :, :]) ** 2).sum(-1), axis=1)
    for v in range(4):
        np.testing.assert_array_equal(got[:, v, 0], 1000.0 + 10.0 * brute + v)


def test_vp_vpv_fallbacks_and_errors():
    cells, coords = meshgen.box_quad(2, 2, 20.0, 20.0)
    cent = model.element_centres(cells, coords)
    only_vpv = {"VPV_%d" % v: np.full(4, 3000.0 + v) for v in range(4)}
    got = model.vertex_parameters(2, 0, cells, coords, cent, only_vpv)  # VP served by VPV
    assert got.shape == (4, 4, 1) and np.all(got[:, 2, 0] == 3002.0)
    assert model.resolve_name("VPV", {"VP_0": 1}) == "VP"
    with pytest.raises(KeyError, match="RHO"):  # the first elastic 2-D parameter
        model.vertex_parameters(2, 1, cells, coords, cent, only_vpv)
    with pytest.raises(KeyError, match="none are defined"):
        model.vertex_parameters(2, 0, cells, coords, cent, {})
