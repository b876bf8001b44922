"""Pins the oracle's P1 geometry (Jacobians, material interpolation, point
location) with the golden
values of the reference's test_QuadP1.cpp and test_HexP1.cpp."""
import numpy as np
import pytest

from oracle import pyoracle as po

QUAD_REF = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float)
HEX_REF = np.array([[-1, -1, -1], [-1, 1, -1], [1, 1, -1], [1, -1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


def test_quad_interpolation():  # test_QuadP1.cpp:12-34
    np.testing.assert_array_equal(po.interp((0, 0)), [0.25] * 4)
    for v, rs in enumerate(QUAD_REF):
        np.testing.assert_array_equal(po.interp(rs), np.eye(4)[v])


def test_hex_interpolation():  # test_HexP1.cpp:21-48
    np.testing.assert_array_equal(po.interp((0, 0, 0)), [0.125] * 8)
    for v, rst in enumerate(HEX_REF):
        np.testing.assert_array_equal(po.interp(rst), np.eye(8)[v])


def test_nodal_points():  # test_QuadP1.cpp:36-63, test_HexP1.cpp:50-137
    g = [-1, -0.654654, 0, 0.654654, 1]
    q = po.nodal_points(4, QUAD_REF)
    np.testing.assert_allclose(q[:, 0], np.tile(g, 5), atol=1e-6)
    np.testing.assert_allclose(q[:, 1], np.repeat(g, 5), atol=1e-6)
    h = po.nodal_points(4, HEX_REF)
    np.testing.assert_allclose(h[:, 0], np.tile(g, 25), atol=1e-6)
    np.testing.assert_allclose(h[:, 1], np.tile(np.repeat(g, 5), 5), atol=1e-6)
    np.testing.assert_allclose(h[:, 2], np.repeat(g, 25), atol=1e-6)


def test_check_hull():  # test_QuadP1.cpp:65-80, test_HexP1.cpp:139-165
    assert po.check_hull((0, 0), QUAD_REF) == 1
    assert po.check_hull((1, 1), QUAD_REF) == 1
    assert po.check_hull((-1.1, -1), QUAD_REF) == 0
    assert po.check_hull((0, 0), [[-2, -2], [2, 2], [2, -2], [-2, 2]]) == -1
    assert po.check_hull((0, 0), [[-1, -1], [-1, 1], [1, 1], [1, -1]]) == -1
    assert po.check_hull((0, 0, 0), HEX_REF) == 1
    assert po.check_hull((1, 1, 1), HEX_REF) == 1
    assert po.check_hull((-1.1, -1, -1), HEX_REF) == 0
    bad = [[-1, -1, -1], [1, -1, 1], [-1, 1, -1], [1, 1, -1], [1, -1, -1], [-1, -1, 1], [1, 1, 1],
           [-1, 1, 1]]
    assert po.check_hull((0, 0, 0), bad) == -1


def test_inverse_coordinate_transform():  # test_QuadP1.cpp:82-98, test_HexP1.cpp:167-180
    np.testing.assert_allclose(po.inverse_map((0, 0), 2 * QUAD_REF), [0, 0], atol=1e-6)
    np.testing.assert_allclose(po.inverse_map((2, 2), 2 * QUAD_REF), [1, 1], atol=1e-6)
    np.testing.assert_allclose(po.inverse_map((0, 0, 0), 2 * HEX_REF), [0, 0, 0], atol=1e-6)
    np.testing.assert_allclose(po.inverse_map((2, 2, 2), 2 * HEX_REF), [1, 1, 1], atol=1e-6)
    with pytest.raises(RuntimeError):
        po.inverse_map((100, 100), [[-2, -2], [2, 2], [2, -2], [-2, 2]])
    with pytest.raises(RuntimeError):
        po.inverse_map((100, 100, 100), [[-2, -3, -7], [2, -3, 7], [-2, 3, -7], [2, 3, -7],
                                         [2, -3, -7], [-2, -3, 7], [2, 3, 7], [-2, 3, 7]])


def test_inverse_jacobian_golden():  # test_QuadP1.cpp:100-116, test_HexP1.cpp:182-209
    det, inv = po.quad_invjac(0, 0, [[-1, -1], [2, -2], [3, 2], [-7, 2]])
    assert det == pytest.approx(5.375)
    np.testing.assert_allclose(inv, [[0.325581, 0.0465116], [0.232558, 0.604651]], rtol=1e-5)
    det, inv = po.hex_invjac(0, 0, 0, [[-2, -2, -2], [-3, 2, -2], [3, 3, -4], [2, -2, -2],
                                       [-1, -3, 2], [1, -2, 2], [2, 2, 2], [-2, 2, 2]])
    assert det == pytest.approx(10)
    np.testing.assert_allclose(inv, [[0.5, -0.05, 0.05], [0, 0.45, 0.05], [0, 0.05, 0.45]],
                               atol=1e-12)


def test_hex_jacobian_is_map_derivative():
    # Closed form: J rows are d(x,y,z)/d(r,s,t) of the trilinear map; compare
    # with central differences of the nodal-point map on a distorted hex.
    rng = np.random.default_rng(3)
    vtx = HEX_REF * 1.7 + rng.uniform(-0.2, 0.2, HEX_REF.shape)
    r, s, t = 0.3, -0.4, 0.1
    det, inv = po.hex_invjac(r, s, t, vtx)

    def x(rr, ss, tt):
        return po.interp((rr, ss, tt)) @ vtx
    h = 1e-6
    J = np.array([(x(r + h, s, t) - x(r - h, s, t)) / (2 * h),
                  (x(r, s + h, t) - x(r, s - h, t)) / (2 * h),
                  (x(r, s, t + h) - x(r, s, t - h)) / (2 * h)])
    np.testing.assert_allclose(inv, np.linalg.inv(J), rtol=1e-7)
    assert det == pytest.approx(np.linalg.det(J), rel=1e-8)


# a trapezoid: not a parallelogram, so the two quad Jacobians differ inside it
TRAPEZOID = np.array([[0.0, 0.0], [4.0, 0.0], [3.0, 2.5], [0.5, 2.0]])


def test_quad_consistent_jacobian_is_the_p1_derivative():
    """The consistent Jacobian is d x / d(r, s) of the P1 map (QuadP1.h:98-108
    weights times the vertices), checked by central differences; the
    reference's rows (QuadP1.h:20-27) agree with it on parallelograms and at
    the centre only."""
    def x_of(r, s):
        return po.interp((r, s)) @ TRAPEZOID

    h = 1e-6
    for r, s in ((0.3, -0.7), (-0.9, 0.8), (0.0, 0.0), (1.0, -1.0)):
        J = np.array([(x_of(r + h, s) - x_of(r - h, s)) / (2 * h),
                      (x_of(r, s + h) - x_of(r, s - h)) / (2 * h)])
        det, inv = po.quad_invjac_consistent(r, s, TRAPEZOID)
        assert det == pytest.approx(np.linalg.det(J), rel=1e-9)
        np.testing.assert_allclose(inv, np.linalg.inv(J), rtol=1e-8)
    det_c, inv_c = po.quad_invjac_consistent(0.0, 0.0, TRAPEZOID)
    det_r, inv_r = po.quad_invjac(0.0, 0.0, TRAPEZOID)
    assert det_c == det_r
    assert po.quad_invjac(0.5, -0.5, TRAPEZOID)[0] != pytest.approx(
        po.quad_invjac_consistent(0.5, -0.5, TRAPEZOID)[0], rel=1e-6)
    para = np.array([[0.0, 0.0], [4.0, 1.0], [5.0, 3.0], [1.0, 2.0]])
    for r, s in ((0.3, -0.7), (-0.9, 0.8)):
        dr_, ir_ = po.quad_invjac(r, s, para)
        dc_, ic_ = po.quad_invjac_consistent(r, s, para)
        assert dr_ == pytest.approx(dc_, rel=1e-15)
        np.testing.assert_allclose(ir_, ic_, rtol=1e-14)
