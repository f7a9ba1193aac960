"""Pins of the CPU oracle against the mathematics (exact rationals), and against the
reference's own test identities (test/runtests.jl)."""
from math import comb

import numpy as np
import pytest

from exact import Fr, bary, domain_points, lift_exact, multi2, multi3, psi, solve


def np3(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


def nfp(N):
    return (N + 1) * (N + 2) // 2


# ------------------------------------------------------------------ tables (runtests.jl:51-80)

@pytest.mark.parametrize("N", range(1, 21))
def test_ordering_tables(orc, N):
    tri, tet = orc.tri_offsets(N), orc.tet_offsets(N)
    L = orc.lib()
    idx = 1
    for j in range(N + 1):
        for i in range(N + 1 - j):
            assert L.orc_ij_to_linear(i, j, tri) == idx
            idx += 1
    idx = 1
    for k in range(N + 1):
        for j in range(N + 1 - k):
            for i in range(N + 1 - j - k):
                assert L.orc_ijk_to_linear(i, j, k, tri, tet) == idx
                idx += 1
    # the whole 21-entry tuples: closed forms of the reference loops (src/ordering.jl:8-26)
    assert list(tri) == [sum(N + 2 - i for i in range(1, m + 1)) for m in range(21)]
    assert list(tet) == [sum(int((N + 2 - i) * (N + 3 - i) / 2) for i in range(1, m + 1)) for m in range(21)]
    assert [tuple(r) for r in orc.linear_to_ijkl(N)] == multi3(N)


@pytest.mark.parametrize("N", range(1, 21))
def test_l_j_bit_exact(orc, N):
    lj = orc.l_j(N)
    for j in range(1, 21):
        exact = Fr((-1) ** j * comb(N, j), 1 + j) if j <= N else Fr(0)
        assert lj[j - 1] == float(exact)  # float(Fraction) is correctly rounded


# ------------------------------------------------------------------ derivative

@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_derivative_is_exact_derivative(orc, N):
    """D^r x are the degree-N Bernstein coefficients of d/dr sum x_a psi_a:
    d/dr = (d/dl0 - d/dl3)/2 and d psi_a / d l_m = N psi^{N-1}_{a - e_m}."""
    rng = np.random.default_rng(N)
    x = [int(v) for v in rng.integers(-9, 10, np3(N))]
    idx = multi3(N)
    pts = domain_points(N)
    for axis in range(3):
        D = orc.deriv_dense(axis, N)
        Dx = [sum(Fr(D[r, c]) * x[c] for c in range(len(idx))) for r in range(len(idx))]
        for lam in pts:
            lhs = sum(Dx[r] * psi(idx[r], lam) for r in range(len(idx)))
            rhs = Fr(0)
            for c, a in enumerate(idx):
                for m, sgn in ((axis, 1), (3, -1)):
                    if a[m] > 0:
                        b = list(a)
                        b[m] -= 1
                        rhs += Fr(sgn * N, 2) * x[c] * psi(tuple(b), lam)
            assert lhs == rhs


@pytest.mark.parametrize("N", range(1, 13))
def test_stencil_matches_dense(orc, N):
    """runtests.jl:7-29"""
    rng = np.random.default_rng(N)
    x = rng.uniform(-1, 1, (3, np3(N)))
    for axis in range(3):
        D = orc.deriv_dense(axis, N) if N <= 8 else None
        got = orc.deriv_mul(axis, N, x)
        if D is not None:
            np.testing.assert_allclose(got, x @ D.T, rtol=1e-13, atol=1e-13)
        # at most 7 non-zeros per row of D = (D^axis - D^3)/2
        if D is not None:
            assert (np.count_nonzero(D, axis=1) <= 7).all()


def test_derivative_of_linear_is_constant(orc):
    """r itself: its coefficients in degree N are the domain-point values r_a; d/dr -> 1"""
    for N in (1, 5, 9, 15):
        ijkl = orc.linear_to_ijkl(N)
        r = 2.0 * ijkl[:, 0] / N - 1.0  # r = 2 l0 - 1, l0 = i/N at domain points (linear precision)
        for axis, expect in ((0, 1.0), (1, 0.0), (2, 0.0)):
            np.testing.assert_allclose(orc.deriv_mul(axis, N, r[None])[0], expect, atol=1e-12)


# ------------------------------------------------------------------ elevation / reduction / L0

@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_elevation_matrix_exact(orc, N):
    """E = V_N^{-1} V_{N-1} at 2-D domain points (the identity that defines elevation);
    the oracle's entries are the correctly rounded exact values."""
    pts = [(Fr(a, N), Fr(b, N), Fr(c, N)) for (a, b, c) in multi2(N)]
    VN = [[psi(m, p) for m in multi2(N)] for p in pts]
    VNm = [[psi(m, p) for m in multi2(N - 1)] for p in pts]
    E = solve(VN, VNm)
    got = orc.elev_dense(N)
    assert got.shape == (nfp(N), nfp(N - 1))
    for r in range(nfp(N)):
        for c in range(nfp(N - 1)):
            assert got[r, c] == float(E[r][c])


@pytest.mark.parametrize("N", range(1, 21))
def test_elevation_and_reduction_multiply(orc, N):
    """runtests.jl:82-90 and reduction = E^T"""
    rng = np.random.default_rng(N)
    E = orc.elev_dense(N + 1)
    x = rng.uniform(-1, 1, nfp(N))
    np.testing.assert_allclose(orc.elevation_multiply(N, x), E @ x, rtol=1e-13, atol=1e-14)
    y = rng.uniform(-1, 1, nfp(N + 1))
    np.testing.assert_allclose(orc.reduction_multiply(N + 1, y), E.T @ y, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("N", range(1, 10))
def test_face_lift_vs_naive_stack(orc, N):
    """runtests.jl:31-48: [L0; l_1 E_1^T L0; ...] with L0 = (N+1)^2/2 E^T E (runtests.jl:92-106)"""
    E1 = orc.elev_dense(N + 1)
    L0 = (N + 1) ** 2 / 2 * E1.T @ E1
    blocks, E = [L0], np.eye(nfp(N))
    for i in range(1, N + 1):
        E = E @ orc.elev_dense(N + 1 - i)
        blocks.append((-1) ** i * comb(N, i) / (1 + i) * E.T @ L0)
    Lf = np.vstack(blocks)
    rng = np.random.default_rng(N)
    x = rng.uniform(-1, 1, (4, nfp(N)))
    np.testing.assert_allclose(orc.face_lift(N, x), x @ Lf.T, rtol=1e-12, atol=1e-12)
    assert np.count_nonzero(np.abs(L0) > 1e-14, axis=1).max() <= 7


@pytest.mark.parametrize("N", [1, 2, 3])
def test_lift_equals_exact_mass_matrix_lift(orc, N):
    """M^{-1}[M^f1 .. M^f4] from exact Dirichlet integrals, faces ordered as the gather
    orders them (replaces get_bernstein_lift, src/utility.jl:157-180)."""
    L = lift_exact(N)
    got = orc.lift(N, np.eye(4 * nfp(N)))  # columns of the lift
    for r in range(np3(N)):
        for c in range(4 * nfp(N)):
            assert abs(got[c, r] - float(L[r][c])) <= 1e-13 * max(1.0, abs(float(L[r][c])))


@pytest.mark.parametrize("N", [4, 5, 6, 7])
def test_lift_equals_mass_matrix_lift_float(orc, N):
    from exact import face_mass, mass

    M = np.array([[float(v) for v in r] for r in mass(N)])
    Mf = np.hstack([np.array([[float(v) for v in r] for r in face_mass(N, f)]) for f in range(4)])
    L = np.linalg.solve(M, Mf)
    got = orc.lift(N, np.eye(4 * nfp(N))).T
    np.testing.assert_allclose(got, L, rtol=1e-8, atol=1e-8 * np.abs(L).max())


def test_fmask_is_the_face_trace(orc):
    """the coefficients whose multi-index has a zero in the face's barycentric are exactly
    the trace; their face 2-D order matches the lift's face permutations"""
    from exact import FACE_ORDER, FACE_ZERO

    for N in (1, 4, 9):
        fm = orc.fmask(N)
        ijkl = orc.linear_to_ijkl(N)
        for f in range(4):
            got = [tuple(ijkl[v]) for v in fm[f]]
            assert all(a[FACE_ZERO[f]] == 0 for a in got)
            assert [tuple(a[m] for m in FACE_ORDER[f]) for a in got] == multi2(N)


def test_lift_of_trace_of_constant(orc):
    """the constant 1 has all coefficients 1 in every degree; sum of the four face lifts
    applied to traces of a polynomial p equals M^{-1} int_dT p psi (spot check of
    linearity + positivity of L0)"""
    for N in (2, 6):
        x = np.ones((1, 4 * nfp(N)))
        y = orc.lift(N, x)
        assert np.isfinite(y).all()
        # L0 of a constant is (N+1)^2/2 * (E^T 1) = (N+1)^2/2 * (N+2)/(N+1) ... positive
        assert (orc.face_lift(N, np.ones((1, nfp(N))))[0, :nfp(N)] > 0).all()


@pytest.mark.parametrize("N", [2, 5, 11, 17])
def test_face_lift_long_double(orc, N):
    """the extended-precision face lift agrees with the FP64 one to rounding, and (N=2) with
    the exact rational naive stack"""
    rng = np.random.default_rng(N)
    x = rng.standard_normal((8, nfp(N)))
    ld = orc.face_lift_ld(N, x)
    np.testing.assert_allclose(ld, orc.face_lift(N, x), rtol=0, atol=1e-13 * np.abs(ld).max())
    if N == 2:
        pts = [(Fr(a, N + 1), Fr(b, N + 1), Fr(c, N + 1)) for (a, b, c) in multi2(N + 1)]
        E = solve([[psi(m, p) for m in multi2(N + 1)] for p in pts], [[psi(m, p) for m in multi2(N)] for p in pts])
        Et = [list(r) for r in zip(*E)]
        L0 = [[Fr((N + 1) ** 2, 2) * sum(Et[r][k] * E[k][c] for k in range(len(E))) for c in range(nfp(N))]
              for r in range(nfp(N))]
        xq = [Fr(v) for v in x[0]]
        slab = [sum(L0[r][c] * xq[c] for c in range(nfp(N))) for r in range(nfp(N))]
        exact = list(slab)
        cur = slab
        for j in range(1, N + 1):
            d = N - j  # reduce degree d+1 -> d
            pts = [(Fr(a, d + 1), Fr(b, d + 1), Fr(c, d + 1)) for (a, b, c) in multi2(d + 1)]
            Ed = solve([[psi(m, p) for m in multi2(d + 1)] for p in pts], [[psi(m, p) for m in multi2(d)] for p in pts])
            cur = [sum(Ed[k][r] * cur[k] for k in range(len(cur))) for r in range(nfp(d))]
            lj = Fr((-1) ** j * comb(N, j), 1 + j)
            exact += [lj * v for v in cur]
        np.testing.assert_allclose(ld[0], [float(v) for v in exact], rtol=1e-15, atol=1e-15)
