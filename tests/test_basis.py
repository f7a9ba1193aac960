"""Host Bernstein basis (bbdg/basis.py) for the line, triangle and tetrahedron, against
identities the mathematics fixes (src/utility.jl:8-147):

* partition of unity and linear precision (sum_i (i/N) B_i = lambda_0) on the line;
* the triangle basis on the edge lambda_2 = 0 is the line basis of that edge;
* Euler's identity for homogeneous polynomials, sum_m lambda_m dB/dlambda_m = N B, and
  central differences in barycentric space, for the derivative columns;
* the triangle derivative columns equal the tetrahedron's on the face lambda_3 = 0, which is
  how the reference defines them (src/utility.jl:140-147);
* exponents that do not sum to N are refused (the reference's DomainError).
"""
from __future__ import annotations

import numpy as np
import pytest

from bbdg import basis
from bbdg.mesh import multi_indices


@pytest.mark.parametrize("N", [1, 2, 5, 9, 14])
def test_line_basis(N):
    r = np.linspace(-1, 1, 23)
    V, D = basis.bernstein_basis_line(N, r)
    assert D is None and V.shape == (r.size, N + 1)
    np.testing.assert_allclose(V.sum(axis=1), 1.0, rtol=0, atol=1e-13)
    l0 = (r + 1) / 2
    np.testing.assert_allclose(V @ (np.arange(N + 1) / N), l0, rtol=0, atol=1e-13)
    # r = 1 is the vertex lambda_0 = 1: only the last basis function is non-zero there
    np.testing.assert_array_equal(V[-1], np.eye(N + 1)[N])
    np.testing.assert_array_equal(V[0], np.eye(N + 1)[0])


def tri_indices(N):
    return [(i, j, N - i - j) for j in range(N + 1) for i in range(N + 1 - j)]


@pytest.mark.parametrize("N", [1, 3, 6, 10])
def test_triangle_edge_is_the_line(N):
    t = np.linspace(0, 1, 17)
    V = basis.bernstein_basis_from_barycentric_tri(N, t, 1 - t, 0 * t)[0]
    Vl, _ = basis.bernstein_basis_from_barycentric_line(N, t, 1 - t)
    idx = tri_indices(N)
    for c, (i, j, k) in enumerate(idx):
        if k == 0:
            np.testing.assert_allclose(V[:, c], Vl[:, i], rtol=1e-14, atol=0)
        else:
            assert np.all(V[:, c] == 0)


@pytest.mark.parametrize("N", [1, 2, 4, 7])
def test_triangle_derivative_columns(N):
    rng = np.random.default_rng(N)
    lam = rng.uniform(0.1, 0.9, (3, 11))  # not on the plane sum = 1: the identities are homogeneous
    V, Vi, Vj, Vk = basis.bernstein_basis_from_barycentric_tri(N, *lam)
    np.testing.assert_allclose(lam[0][:, None] * Vi + lam[1][:, None] * Vj + lam[2][:, None] * Vk, N * V,
                               rtol=1e-12, atol=1e-14)
    h = 1e-5
    for m, Dm in enumerate((Vi, Vj, Vk)):
        lp, lm = lam.copy(), lam.copy()
        lp[m] += h
        lm[m] -= h
        fd = (basis.bernstein_basis_from_barycentric_tri(N, *lp)[0]
              - basis.bernstein_basis_from_barycentric_tri(N, *lm)[0]) / (2 * h)
        np.testing.assert_allclose(Dm, fd, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("N", [1, 3, 5])
def test_triangle_derivatives_are_the_tet_ones_on_a_face(N):
    rng = np.random.default_rng(10 + N)
    lam = rng.uniform(0, 1, (3, 9))
    lam /= lam.sum(axis=0)
    tri = basis.bernstein_basis_from_barycentric_tri(N, *lam)
    tet = basis.bernstein_basis_from_barycentric_tet(N, *lam, np.zeros(lam.shape[1]))
    mi = [tuple(int(v) for v in row) for row in multi_indices(N)]
    for c, (i, j, k) in enumerate(tri_indices(N)):
        ct = mi.index((i, j, k, 0))
        for m in range(3):
            np.testing.assert_allclose(tri[1 + m][:, c], tet[1 + m][:, ct], rtol=1e-14, atol=0)
        np.testing.assert_allclose(tri[0][:, c], tet[0][:, ct], rtol=1e-14, atol=0)


def test_tet_derivative_columns_euler():
    N = 4
    rng = np.random.default_rng(3)
    lam = rng.uniform(0.1, 0.9, (4, 7))
    V, *D = basis.bernstein_basis_from_barycentric_tet(N, *lam)
    np.testing.assert_allclose(sum(lam[m][:, None] * D[m] for m in range(4)), N * V, rtol=1e-12, atol=1e-14)


def test_derivative_exponents_must_sum_to_degree():
    lam = [np.ones(2)] * 4
    with pytest.raises(ValueError):
        basis.evaluate_bernstein_derivatives(3, lam, (1, 1, 1, 1))
    d = basis.evaluate_bernstein_derivatives(4, lam, (1, 1, 1, 1))
    assert all(np.all(v == 24.0) for v in d)  # 4!/(0! 1! 1! 1!) at lambda = 1
