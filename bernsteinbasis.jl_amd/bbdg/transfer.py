"""Face-trace transfer between polynomial degrees (meshes whose elements carry their own
degree).

A face shared by elements of degrees N and Nn carries two traces of different degree.  The
element of degree N brings its neighbour's trace to degree N before the usual upwind flux:

* Nn < N: degree elevation, the exact re-expansion of the degree-Nn Bernstein polynomial in
  the degree-N basis -- the product of single-step elevations, each the ElevationMatrix of
  src/ElevationMatrix.jl:11-24 (elevation_multiply!, src/BernsteinLift.jl:100-125);
* Nn > N: the L2 projection onto degree N on the face,
  P = M_N^{-1} E_{N->Nn}^T M_Nn, with M_n the Bernstein mass matrix of the triangle
  (M_ab = (n!/a!)(n!/b!)(a+b)!/(2n+2)! up to the common area factor) and E_{N->Nn} the
  elevation above.

Because the flux is linear in the jump and P E = I, this equals computing the flux on a
mortar of degree max(N, Nn) and projecting it to each side.  All entries are computed in
exact rational arithmetic (fractions) and rounded once to float64; the rounded tables for
degrees 1..10 are kept in transfer_tables.npz next to this file (regenerate with
``python -m bbdg.transfer``).
"""
from __future__ import annotations

import os
from fractions import Fraction as Fr
from functools import lru_cache
from math import factorial

import numpy as np

NMAX = 10
HERE = os.path.dirname(os.path.abspath(__file__))
NPZ = os.path.join(HERE, "transfer_tables.npz")


def nfp(n):
    return (n + 1) * (n + 2) // 2


def multi2(n):
    """face multi-indices (a, b, c) in the reference's 2-D order (b outer, a inner)"""
    return [(a, b, n - a - b) for b in range(n + 1) for a in range(n + 1 - b)]


def _matmul(A, B):
    Bt = list(zip(*B))
    return [[sum(a * b for a, b in zip(row, col)) for col in Bt] for row in A]


@lru_cache(None)
def elevate1_exact(n):
    """degree n -> n+1 (nfp(n+1) x nfp(n)): c'(a,b,c) = (a c(a-1,b,c) + b c(a,b-1,c) +
    c c(a,b,c-1)) / (n+1), the weights of elevation_multiply!"""
    lo, hi = multi2(n), multi2(n + 1)
    pos = {t: i for i, t in enumerate(lo)}
    E = [[Fr(0)] * len(lo) for _ in hi]
    for r, t in enumerate(hi):
        for m in range(3):
            s = list(t)
            s[m] -= 1
            if s[m] >= 0:
                E[r][pos[tuple(s)]] += Fr(t[m], n + 1)
    return tuple(tuple(r) for r in E)


@lru_cache(None)
def elevate_exact(n, N):
    """degree n -> N (nfp(N) x nfp(n)), n <= N"""
    E = [[Fr(int(i == j)) for j in range(nfp(n))] for i in range(nfp(n))]
    for d in range(n, N):
        E = _matmul([list(r) for r in elevate1_exact(d)], E)
    return tuple(tuple(r) for r in E)


@lru_cache(None)
def mass_exact(n):
    idx = multi2(n)
    M = []
    for a in idx:
        row = []
        for b in idx:
            v = Fr(factorial(n) ** 2, factorial(2 * n + 2))
            for m in range(3):
                v *= Fr(factorial(a[m] + b[m]), factorial(a[m]) * factorial(b[m]))
            row.append(v)
        M.append(row)
    return tuple(tuple(r) for r in M)


@lru_cache(None)
def mass_inverse_exact(n):
    M = [list(r) for r in mass_exact(n)]
    k = len(M)
    A = [row[:] + [Fr(int(i == j)) for j in range(k)] for i, row in enumerate(M)]
    for c in range(k):
        p = next(r for r in range(c, k) if A[r][c] != 0)
        A[c], A[p] = A[p], A[c]
        pv = A[c][c]
        A[c] = [v / pv for v in A[c]]
        for r in range(k):
            if r != c and A[r][c] != 0:
                f = A[r][c]
                A[r] = [x - f * y for x, y in zip(A[r], A[c])]
    return tuple(tuple(row[k:]) for row in A)


def project_exact(Nn, N):
    """L2 projection degree Nn -> N < Nn (nfp(N) x nfp(Nn))"""
    E = elevate_exact(N, Nn)
    Et = [list(r) for r in zip(*E)]
    return _matmul([list(r) for r in mass_inverse_exact(N)], _matmul(Et, [list(r) for r in mass_exact(Nn)]))


def transfer_exact(N, Nn):
    """the matrix that brings a degree-Nn face trace to degree N"""
    if Nn < N:
        return [list(r) for r in elevate_exact(Nn, N)]
    if Nn > N:
        return project_exact(Nn, N)
    return [[Fr(int(i == j)) for j in range(nfp(N))] for i in range(nfp(N))]


def generate(nmax=NMAX, path=NPZ):
    out = {}
    for N in range(1, nmax + 1):
        for Nn in range(1, nmax + 1):
            if N != Nn:
                out[f"T_{N}_{Nn}"] = np.array([[float(v) for v in r] for r in transfer_exact(N, Nn)])
    np.savez_compressed(path, **out)
    return out


@lru_cache(None)
def tables():
    if not os.path.exists(NPZ):
        raise RuntimeError(f"{NPZ} missing; regenerate with python -m bbdg.transfer")
    with np.load(NPZ, allow_pickle=False) as z:
        return {tuple(int(v) for v in k.split("_")[1:]): z[k] for k in z.files}


def transfer(N, Nn):
    """float64 matrix (nfp(N) x nfp(Nn)) bringing a degree-Nn trace to degree N"""
    if N == Nn:
        return np.eye(nfp(N))
    return tables()[(N, Nn)]


def pack(dim=NMAX + 1):
    """all tables in one float64 array with offsets off[N][Nn] (a dim x dim array; -1:
    identity or absent), for the C ABI and the oracle"""
    off = -np.ones((dim, dim), np.int64)
    parts, pos = [], 0
    for N in range(1, min(dim - 1, NMAX) + 1):
        for Nn in range(1, min(dim - 1, NMAX) + 1):
            if N != Nn:
                T = np.ascontiguousarray(transfer(N, Nn))
                off[N, Nn] = pos
                parts.append(T.reshape(-1))
                pos += T.size
    return np.concatenate(parts), off


if __name__ == "__main__":
    import time

    t = time.time()
    generate()
    print(f"wrote {NPZ} in {time.time() - t:.1f}s")
