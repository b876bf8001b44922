"""Pins the oracle's GLL tables (points, weights, derivative matrix) against the reference's compiled
sympy tables, the closed form and exact polynomial identities."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po


@pytest.mark.parametrize("order", range(1, 11))
def test_literals_are_gll(order):
    xi, w = po.gll(order)
    xe, we = po.gll_exact(order)
    # The reference literals are 15-digit roundings of the GLL nodes/weights.
    np.testing.assert_allclose(xi, xe, atol=1e-15, rtol=0)
    np.testing.assert_allclose(w, we, atol=2e-14, rtol=0)


def test_order4_closed_form():
    # test_quadrature_points_weights.py:33-41
    xi, w = po.gll_exact(4)
    np.testing.assert_allclose(xi, [-1, -21 ** 0.5 / 7, 0, 21 ** 0.5 / 7, 1], atol=1e-15)
    np.testing.assert_allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], atol=1e-15)


@pytest.mark.parametrize("order", range(1, 11))
def test_literals_match_reference_tables(order):
    R = po.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    xi, w = po.gll(order)
    p, ww = np.zeros(order + 1), np.zeros(order + 1)
    getattr(R, "gll_coordinates_order%d_square" % order)(po.dp(p))
    getattr(R, "gll_weights_order%d_square" % order)(po.dp(ww))
    assert np.array_equal(p, xi) and np.array_equal(ww, w)


@pytest.mark.parametrize("order", range(2, 11))
def test_derivative_matrix_matches_reference(order):
    R = po.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    n = order + 1
    xi, _ = po.gll(order)
    f = getattr(R, "interpolate_eps_derivative_order%d_square" % order)
    f.argtypes = [C.c_double, C.c_double, po.D]
    Dr = np.zeros((n, n))
    for i in range(n):  # TensorQuad.cpp:116-140: grad.row(i) = test.col(0)
        t = np.zeros(n * n)
        f(xi[i], xi[0], po.dp(t))
        Dr[i] = t[:n]
    Dm = po.deriv_matrix(order)
    np.testing.assert_allclose(Dm, Dr, atol=1e-12 * np.abs(Dr).max(), rtol=0)


@pytest.mark.parametrize("order", range(1, 11))
def test_derivative_matrix_exact_on_polynomials(order):
    xi, _ = po.gll(order)
    Dm = po.deriv_matrix(order)
    for k in range(order + 1):
        np.testing.assert_allclose(Dm @ xi ** k, k * xi ** max(k - 1, 0) if k else 0 * xi,
                                   atol=1e-11)


@pytest.mark.parametrize("order", [3, 4, 5])
def test_tensor_lagrange_matches_reference(order):
    R = po.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    f = getattr(R, "interpolate_order%d_hex" % order)
    f.argtypes = [C.c_double] * 3 + [po.D]
    xi, _ = po.gll(order)
    rng = np.random.default_rng(1)
    for _ in range(5):
        r, s, t = rng.uniform(-1, 1, 3)
        ref = np.zeros((order + 1) ** 3)
        f(r, s, t, po.dp(ref))
        lr, ls, lt = po.lagrange(xi, r), po.lagrange(xi, s), po.lagrange(xi, t)
        mine = np.einsum("k,j,i->kji", lt, ls, lr).ravel()
        np.testing.assert_allclose(mine, ref, atol=1e-12)
