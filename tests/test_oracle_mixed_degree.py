"""Pins of the degree-transfer tables and of the oracle's mixed-degree RHS (elements that
carry their own degree; no counterpart in the reference, built from its pieces)."""
from fractions import Fraction as Fr

import numpy as np
import pytest


def nfp(n):
    return (n + 1) * (n + 2) // 2


def np3(n):
    return (n + 1) * (n + 2) * (n + 3) // 6


# ------------------------------------------------------------------ transfer tables

@pytest.mark.parametrize("N,Nn", [(1, 2), (2, 1), (3, 5), (5, 3), (4, 7), (6, 4)])
def test_transfer_tables_are_the_rounded_exact_values(N, Nn):
    from bbdg import transfer as TR

    exact = TR.transfer_exact(N, Nn)
    got = TR.transfer(N, Nn)
    assert got.shape == (nfp(N), nfp(Nn))
    assert all(got[r, c] == float(exact[r][c]) for r in range(nfp(N)) for c in range(nfp(Nn)))


def test_elevation_tables_match_the_reference_elevation(orc):
    """elevation N -> N+k = k applications of elevation_multiply! (src/BernsteinLift.jl:100-125)"""
    from bbdg import transfer as TR

    rng = np.random.default_rng(3)
    for n in range(1, 10):
        for N in range(n + 1, 11):
            x = rng.standard_normal(nfp(n))
            y = x
            for d in range(n, N):
                y = orc.elevation_multiply(d, y)
            np.testing.assert_allclose(TR.transfer(N, n) @ x, y, rtol=1e-13, atol=1e-13)


def test_projection_properties():
    """P E = I (projection of an elevated polynomial returns it), M_N P = E^T M_Nn (the L2
    normal equations), and rows of P and E sum to 1 (constants are preserved)"""
    from bbdg import transfer as TR

    def mass(n):
        return np.array([[float(v) for v in r] for r in TR.mass_exact(n)])

    for N in range(1, 10):
        for Nn in range(N + 1, 11):
            P, E = TR.transfer(N, Nn), TR.transfer(Nn, N)
            np.testing.assert_allclose(P @ E, np.eye(nfp(N)), atol=1e-12)
            lhs, rhs = mass(N) @ P, E.T @ mass(Nn)
            np.testing.assert_allclose(lhs, rhs, atol=1e-12 * np.abs(rhs).max())
            np.testing.assert_allclose(P.sum(1), 1.0, atol=1e-12)
            np.testing.assert_allclose(E.sum(1), 1.0, atol=1e-14)


def test_projection_exact_small():
    """degree 2 -> 1 by hand: the L2 projection of B^2_(2,0,0) = l0^2 onto the linear
    Bernstein basis (l0, l1, l2) solves M1 c = b with the Dirichlet integrals
    int l0^a l1^b l2^c = a! b! c! / (a+b+c+2)! (triangle of area 1/2)"""
    import sympy

    from bbdg import transfer as TR

    M1 = sympy.Matrix(3, 3, lambda i, j: sympy.Rational(2 if i == j else 1, 24))
    b = sympy.Matrix([sympy.Rational(6, 120), sympy.Rational(2, 120), sympy.Rational(2, 120)])
    c = M1.solve(b)
    P = TR.project_exact(2, 1)
    col = TR.multi2(2).index((2, 0, 0))
    for r, t in enumerate(TR.multi2(1)):
        assert P[r][col] == Fr(str(c[[(1, 0, 0), (0, 1, 0), (0, 0, 1)].index(t)]))


# ------------------------------------------------------------------ mixed-degree RHS

def make_batch(ns, degs, seed=0, periodic=True, F=4):
    from bbdg import mesh
    from bbdg.batch import Batch

    rng = np.random.default_rng(seed)
    meshes = [mesh.kuhn_mesh(n, 0.1, seed + i, periodic=periodic) for i, n in enumerate(ns)]
    degrees = [rng.integers(degs[0], degs[1] + 1, m.K) for m in meshes]
    return Batch(meshes, degrees, [0.5 * m.h / degs[1] ** 2 for m in meshes], F=F)


def test_md_uniform_degree_equals_rhs(orc):
    """one instance whose elements all have degree N: the mixed-degree RHS is the RHS"""
    from bbdg import basis, mesh
    from bbdg.batch import Batch

    N = 4
    m = mesh.kuhn_mesh(2, 0.1, 1, periodic=True)
    b = Batch([m], [np.full(m.K, N)], [0.01])
    fn = lambda x, y, z: np.stack([np.sin(2 * np.pi * x), np.cos(2 * np.pi * y), 0 * x + 0.3, np.sin(2 * np.pi * z)])
    q = b.interpolate([fn])
    ba = orc.BatchArrays(b, [1, 1, 1, 1])
    got = b.from_node_major(orc.rhs_md(0, b.to_node_major(q), ba))
    # the same state in the single-degree oracle (element order of the mesh)
    qk = np.ascontiguousarray(basis.interpolate(m, N, fn))
    ref = orc.rhs(0, N, qk, orc.MeshArrays(m, N, [1, 1, 1, 1]))  # (K, Np, F)
    ref_dev = np.concatenate([ref[b.order[e]].T.reshape(-1) for e in range(b.K)])
    np.testing.assert_allclose(got, ref_dev, rtol=0, atol=1e-13 * np.abs(ref_dev).max())


def test_md_constant_state_is_steady(orc):
    b = make_batch([2, 3], (1, 6), seed=2)
    q = b.interpolate([lambda x, y, z: np.stack([1 + 0 * x, 0.5 + 0 * x, -0.25 + 0 * x, 2 + 0 * x])] * 2)
    du = orc.rhs_md(0, b.to_node_major(q), orc.BatchArrays(b, [1, 1, 1, 1]))
    assert np.abs(du).max() < 1e-11


def test_md_linear_fields_exact(orc):
    """linear fields are exact at every degree and the traces match across every face of
    any degree pair, so the RHS is the exact -div u / -grad p (walls with uP = uM)"""
    b = make_batch([3], (1, 6), seed=4, periodic=False)
    g = (0.3, -0.2, 0.5)
    fn = lambda x, y, z: np.stack([g[0] * x + g[1] * y + g[2] * z + 0.1, 0.2 * y + 0.1 * z, -0.4 * x + 0.3 * z,
                                   0.7 * z - 0.2 * x])
    q = b.interpolate([fn])
    du = b.from_node_major(orc.rhs_md(0, b.to_node_major(q), orc.BatchArrays(b, [1, 1, 1, 1])))
    want = (-(0.0 + 0.0 + 0.7), -g[0], -g[1], -g[2])  # u_x + v_y + w_z = 0 + 0 + 0.7
    for N, s, c in b.groups():
        v = b.field_view(du, N)
        for f in range(4):
            np.testing.assert_allclose(v[:, f, :], want[f], atol=1e-11)


def test_md_conserves_the_integral_of_p(orc):
    """the upwind flux of both sides of a mixed-degree face integrates to the same value
    (the projection preserves integrals against constants), so sum_e int_e p_t = 0"""
    b = make_batch([2, 3], (1, 7), seed=6)
    rng = np.random.default_rng(1)
    fns = [lambda x, y, z, c=rng.uniform(0.3, 0.7, 3): np.stack(
        [np.exp(-20 * ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)), np.sin(2 * np.pi * y),
         np.cos(2 * np.pi * x), np.sin(2 * np.pi * (x + z))]) for _ in range(2)]
    q = b.interpolate(fns)
    du = b.from_node_major(orc.rhs_md(0, b.to_node_major(q), orc.BatchArrays(b, [1, 1, 1, 1])))
    total, scale = 0.0, 0.0
    for N, s, c in b.groups():
        v = b.field_view(du, N)
        ints = b.J[s:s + c] * (4 / 3) * v[:, 0, :].mean(axis=1)  # int of a Bernstein expansion
        total += ints.sum()
        scale += np.abs(ints).sum()
    assert abs(total) < 1e-12 * scale, (total, scale)


def test_md_step_matches_rhs_composition(orc):
    """one mixed-degree LSRK45 step equals the five stages composed from rhs_md"""
    b = make_batch([2], (2, 5), seed=8)
    q = b.interpolate([lambda x, y, z: np.stack([np.sin(2 * np.pi * x), 0 * x, 0 * x, np.cos(2 * np.pi * y)])])
    ba = orc.BatchArrays(b, [1, 1, 1, 1])
    qn = b.to_node_major(q)
    got = qn.copy()
    orc.lsrk_step_md(0, got, ba)
    a, bb = orc.lsrk_coeffs()
    dte = np.repeat(b.dte, [4 * np3(int(n)) for n in b.deg])
    res = np.zeros_like(qn)
    ref = qn.copy()
    for st in range(5):
        du = orc.rhs_md(0, ref, ba)
        res = a[st] * res + dte * du
        ref = ref + bb[st] * res
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-15 * np.abs(ref).max())
