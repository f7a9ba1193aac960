"""Exact rational closed forms used to pin the oracle (independent of its code).

* Bernstein basis values psi_alpha = N!/alpha! lambda^alpha at rational points
* Dirichlet integrals  int_T lambda^g dV = g! 3! |T| / (|g|+3)!,  int_f lambda^g dS = g! 2! |f| / (|g|+2)!
  on the reference tet (|T| = 4/3) and its faces (reference weight |f| = 2 for every face)
"""
from fractions import Fraction as Fr
from itertools import product
from math import factorial


def multi3(N):
    return [(i, j, k, N - i - j - k) for k in range(N + 1) for j in range(N + 1 - k) for i in range(N + 1 - j - k)]


def multi2(N):
    return [(a, b, N - a - b) for b in range(N + 1) for a in range(N + 1 - b)]


def mfac(t):
    r = 1
    for v in t:
        r *= factorial(v)
    return r


def psi(alpha, lam):
    """N!/alpha! prod lam^alpha (exact)"""
    N = sum(alpha)
    v = Fr(factorial(N), mfac(alpha))
    for a, l in zip(alpha, lam):
        v *= Fr(l) ** a
    return v


def bary(r, s, t):
    """docs/src/notation.md:14-18"""
    r, s, t = Fr(r), Fr(s), Fr(t)
    return ((1 + r) / 2, (1 + s) / 2, (1 + t) / 2, -(1 + r + s + t) / 2)


def domain_points(N):
    """barycentric domain points alpha/N (unisolvent for degree N)"""
    return [tuple(Fr(a, N) for a in al) for al in multi3(N)]


def mass(N):
    """M[a][b] = int_T psi_a psi_b,  |T| = 4/3"""
    idx = multi3(N)
    c = Fr(factorial(3) * 4, 3 * factorial(2 * N + 3))
    return [[Fr(factorial(N) ** 2, mfac(a) * mfac(b)) * mfac([x + y for x, y in zip(a, b)]) * c for b in idx]
            for a in idx]


# face f (0-based f1..f4): the barycentric that vanishes and the order of the remaining three
FACE_ZERO = [1, 3, 0, 2]
FACE_ORDER = [(0, 2, 3), (1, 2, 0), (1, 2, 3), (0, 1, 3)]


def face_mass(N, f):
    """M^f[alpha][m] = int_f psi_alpha phi_m, faces ordered by FACE_ORDER, |f| = 2"""
    idx = multi3(N)
    c = Fr(factorial(2) * 2, factorial(2 * N + 2))
    out = []
    for a in idx:
        row = []
        for abc in multi2(N):
            if a[FACE_ZERO[f]] != 0:
                row.append(Fr(0))
                continue
            ap = tuple(a[m] for m in FACE_ORDER[f])
            row.append(Fr(factorial(N) ** 2, mfac(ap) * mfac(abc)) * mfac([x + y for x, y in zip(ap, abc)]) * c)
        out.append(row)
    return out


def solve(A, B):
    """exact A X = B by Gauss-Jordan on Fractions"""
    n = len(A)
    m = len(B[0])
    M = [list(A[i]) + list(B[i]) for i in range(n)]
    for col in range(n):
        piv = next(r for r in range(col, n) if M[r][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        p = M[col][col]
        M[col] = [v / p for v in M[col]]
        for r in range(n):
            if r != col and M[r][col] != 0:
                fct = M[r][col]
                M[r] = [x - fct * y for x, y in zip(M[r], M[col])]
    return [row[n:n + m] for row in M]


def lift_exact(N):
    """Np x 4Nfp exact lift M^{-1} [M^f1 .. M^f4] (docs/src/lift.md:5-23)"""
    Mf = [face_mass(N, f) for f in range(4)]
    B = [sum((Mf[f][r] for f in range(4)), []) for r in range(len(Mf[0]))]
    return solve(mass(N), B)


__all__ = ["Fr", "product", "multi2", "multi3", "psi", "bary", "domain_points", "mass", "face_mass", "solve",
           "lift_exact", "FACE_ORDER", "FACE_ZERO"]
