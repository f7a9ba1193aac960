"""Bernstein basis evaluation on the host (set-up and output only, FP64).

Restates src/utility.jl:8-147: barycentric coordinates of the reference simplices,
Bernstein Vandermonde matrices and their barycentric derivative columns.  These are used
to turn point values into Bernstein coefficients (initial conditions, as in
examples/3d_wave.jl:111-113) and back (error measurement, examples/3d_wave.jl:137).
"""
from __future__ import annotations

from math import factorial

import numpy as np

from .mesh import multi_indices


def cartesian_to_barycentric_tet(r, s, t):
    """src/utility.jl:23-29"""
    r, s, t = (np.asarray(v, dtype=np.float64) for v in (r, s, t))
    return (1 + r) / 2, (1 + s) / 2, (1 + t) / 2, -(1 + r + s + t) / 2


def cartesian_to_barycentric_tri(r, s):
    """src/utility.jl:15-21"""
    r, s = (np.asarray(v, dtype=np.float64) for v in (r, s))
    return (r + 1) / 2, (s + 1) / 2, -(r + s) / 2


def bernstein_basis_from_barycentric_tet(N, l0, l1, l2, l3):
    """src/utility.jl:97-107: V[m, n] = psi_n(point m); columns in (k, j, i) order.
    Returns V, Vi, Vj, Vk, Vl (barycentric derivative columns, src/utility.jl:109-138)."""
    l0, l1, l2, l3 = (np.asarray(v, dtype=np.float64) for v in (l0, l1, l2, l3))
    mi = multi_indices(N)
    fN = factorial(N)
    V = np.empty((l0.size, mi.shape[0]))
    D = [np.zeros((l0.size, mi.shape[0])) for _ in range(4)]
    lam = [l0, l1, l2, l3]
    for c, (i, j, k, l) in enumerate(mi):
        a = (i, j, k, l)
        V[:, c] = fN / (factorial(i) * factorial(j) * factorial(k) * factorial(l)) * l0 ** i * l1 ** j * l2 ** k * l3 ** l
        for m in range(4):
            if a[m] > 0:
                b = list(a)
                b[m] -= 1
                coef = fN / (factorial(b[0]) * factorial(b[1]) * factorial(b[2]) * factorial(b[3]))
                D[m][:, c] = coef * lam[0] ** b[0] * lam[1] ** b[1] * lam[2] ** b[2] * lam[3] ** b[3]
    return (V, *D)


def bernstein_basis_tet(N, r, s, t):
    """src/utility.jl:62-63"""
    return bernstein_basis_from_barycentric_tet(N, *cartesian_to_barycentric_tet(r, s, t))


def evaluate_bernstein_derivatives(N, lam, a):
    """src/utility.jl:109-138 (Tet) and :140-147 (Tri, the Tet formula with the fourth
    coordinate 0 and exponent 0): the partial derivatives of the Bernstein polynomial of
    multi-index a (3 or 4 exponents summing to N) with respect to each barycentric
    coordinate, d/dlam_m = N!/(a - e_m)! prod lam^(a - e_m) (zero where a_m = 0), as a list
    of arrays over the points.  A multi-index whose exponents do not sum to N raises
    ValueError (the reference's DomainError)."""
    a = tuple(int(v) for v in a)
    if sum(a) != N or len(a) != len(lam) or min(a) < 0:
        raise ValueError(f"barycentric exponents {a} do not sum to the degree {N}")
    fN = factorial(N)
    out = []
    for m in range(len(a)):
        if a[m] == 0:
            out.append(np.zeros_like(lam[0]))
            continue
        b = list(a)
        b[m] -= 1
        coef = fN / float(np.prod([factorial(v) for v in b]))
        val = np.full_like(lam[0], coef)
        for c, e in zip(lam, b):
            val = val * c ** e
        out.append(val)
    return out


def cartesian_to_barycentric_line(r):
    """src/utility.jl:8-13: (r + 1)/2, (1 - r)/2"""
    r = np.asarray(r, dtype=np.float64)
    return (r + 1) / 2, (1 - r) / 2


def bernstein_basis_from_barycentric_line(N, l0, l1):
    """src/utility.jl:83-86: V[m, i] = C(N, i) l0^i l1^(N - i), i = 0..N; no derivative
    columns (the reference returns nothing for them)"""
    l0, l1 = (np.asarray(v, dtype=np.float64) for v in (l0, l1))
    fN = factorial(N)
    V = np.stack([fN / (factorial(i) * factorial(N - i)) * l0 ** i * l1 ** (N - i) for i in range(N + 1)], axis=1)
    return V, None


def bernstein_basis_line(N, r):
    """src/utility.jl:56-58"""
    return bernstein_basis_from_barycentric_line(N, *cartesian_to_barycentric_line(r))


def bernstein_basis_from_barycentric_tri(N, l0, l1, l2):
    """src/utility.jl:88-95: V[m, n] = psi_n(point m), columns (j outer, i inner), and the
    barycentric derivative columns Vi, Vj, Vk (src/utility.jl:140-147)."""
    l0, l1, l2 = (np.asarray(v, dtype=np.float64) for v in (l0, l1, l2))
    fN = factorial(N)
    cols = []
    D = [[], [], []]
    for j in range(N + 1):
        for i in range(N + 1 - j):
            k = N - i - j
            cols.append(fN / (factorial(i) * factorial(j) * factorial(k)) * l0 ** i * l1 ** j * l2 ** k)
            for m, d in enumerate(evaluate_bernstein_derivatives(N, (l0, l1, l2), (i, j, k))):
                D[m].append(d)
    return (np.stack(cols, axis=1), *(np.stack(d, axis=1) for d in D))


def bernstein_basis_tri(N, r, s):
    """src/utility.jl:60-61"""
    return bernstein_basis_from_barycentric_tri(N, *cartesian_to_barycentric_tri(r, s))


def equispaced_ref_points(N):
    """Domain points of the reference tet (rst) in the linear order of the basis."""
    from .mesh import REF_VERTS

    mi = multi_indices(N).astype(np.float64) / max(N, 1)
    return mi @ REF_VERTS


def interpolation_matrix(N):
    """inv(V) for Bernstein interpolation at the equispaced domain points (used in place
    of StartUpDG's nodes, examples/3d_wave.jl:111-113)."""
    p = equispaced_ref_points(N)
    V = bernstein_basis_tet(N, p[:, 0], p[:, 1], p[:, 2])[0]
    return np.linalg.inv(V), V


def interpolate(mesh, N, fn):
    """Bernstein coefficients (K, Np, F) of the interpolant of fn(x, y, z) -> (F, ...) at the
    equispaced domain points of every element."""
    from .mesh import domain_points

    Vinv, _ = interpolation_matrix(N)
    X = domain_points(mesh, N)  # (K,Np,3)
    vals = np.asarray(fn(X[..., 0], X[..., 1], X[..., 2]), dtype=np.float64)  # (F,K,Np)
    return np.ascontiguousarray(np.matmul(vals, Vinv.T).transpose(1, 2, 0))


def evaluate(N, coeffs):
    """Point values at the equispaced domain points: (K,Np,F) -> (K,Np,F)."""
    _, V = interpolation_matrix(N)
    return np.matmul(V, coeffs)
