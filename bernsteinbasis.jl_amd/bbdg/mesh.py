"""Affine tetrahedral meshes of the unit cube, in the data layout the kernels read.

The reference builds its meshes through StartUpDG (``uniform_mesh`` + ``MeshData``,
examples/3d_wave.jl:89-97), which is not part of the reference and not importable here.
This module restates what the right-hand sides consume from ``MeshData``:

* ``rxJ ... tzJ`` and ``J``        (examples/3d_wave.jl:76-78, :83)
* ``nxJ, nyJ, nzJ``                (examples/3d_wave.jl:46-53)
* ``mapP``                         (examples/3d_wave.jl:45)

Conventions (the reference tetrahedron of docs/src/notation.md:9-18):

* barycentrics l0=(1+r)/2, l1=(1+s)/2, l2=(1+t)/2, l3=-(1+r+s+t)/2; tet-local vertex m is
  where l_m = 1, so v0=(1,-1,-1), v1=(-1,1,-1), v2=(-1,-1,1), v3=(-1,-1,-1).
* faces in StartUpDG order (src/utility.jl:171-174): f1 s=-1 (l1=0), f2 r+s+t=-1 (l3=0),
  f3 r=-1 (l0=0), f4 t=-1 (l2=0).  Reference normals nrJ: (0,-1,0), (1,1,1), (-1,0,0),
  (0,0,-1); physical ``nJ = J G^{-T} nrJ`` (Nanson), so |nJ| = face area / 2.
* face-local vertex order (the 2-D Bernstein index (a,b,c) of a face node weights these
  vertices), fixed by the face permutations of src/BernsteinLift.jl:207-217:
  f1 [0,2,3], f2 [1,2,0], f3 [1,2,3], f4 [0,1,3].

Compact connectivity for the fused kernels: per face, the neighbour element, its face
and the permutation p (one of 6) with ``nbr_face_vertex[m] = my_face_vertex[p[m]]``;
boundary faces point to themselves and carry a boundary code.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

FACE_VERTS = np.array([[0, 2, 3], [1, 2, 0], [1, 2, 3], [0, 1, 3]], dtype=np.int64)
REF_NORMALS = np.array([[0.0, -1.0, 0.0], [1.0, 1.0, 1.0], [-1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])
REF_VERTS = np.array([[1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0], [-1.0, -1.0, -1.0]])
PERMS = np.array(list(itertools.permutations(range(3))), dtype=np.int64)  # 6 x 3, lexicographic

BC_INTERIOR = 0
BC_WALL = 1  # mirror state: uP = sign .* uM (PEC for Maxwell)


@dataclass
class TetMesh:
    """Geometry + connectivity of K affine tets.

    vx      : (K,4,3) physical vertex coordinates (tet-local vertex order)
    geo     : (K,9)   rxJ sxJ txJ ryJ syJ tyJ rzJ szJ tzJ
    J       : (K,)    Jacobian determinant of the reference->physical map
    nJ      : (K,4,3) scaled outward normals per face (nxJ, nyJ, nzJ)
    nbr     : (K,4)   neighbour element across each face (self on boundary faces)
    nbr_face: (K,4)   neighbour's face index (0..3)
    perm    : (K,4)   face-vertex permutation index into PERMS
    bc      : (K,4)   BC_INTERIOR or BC_WALL
    h       : cube edge length
    """

    vx: np.ndarray
    geo: np.ndarray
    J: np.ndarray
    nJ: np.ndarray
    nbr: np.ndarray
    nbr_face: np.ndarray
    perm: np.ndarray
    bc: np.ndarray
    h: float
    periodic: bool = True
    meta: dict = field(default_factory=dict)

    @property
    def K(self) -> int:
        return int(self.J.shape[0])


def geometric_factors(vx: np.ndarray):
    """rxJ..tzJ, J and nJ of affine tets with tet-local vertices vx (K,4,3)."""
    X0, X1, X2, X3 = (vx[:, m, :] for m in range(4))
    # columns dx/dr, dx/ds, dx/dt  (x = X3 + (X0-X3)(1+r)/2 + (X1-X3)(1+s)/2 + (X2-X3)(1+t)/2)
    G = np.stack([(X0 - X3) / 2, (X1 - X3) / 2, (X2 - X3) / 2], axis=2)  # (K,3,3) G[:,x,r]
    J = np.linalg.det(G)
    # J * G^{-1} = adjugate: rows (r,s,t), columns (x,y,z)
    adj = np.empty_like(G)
    a = G
    adj[:, 0, 0] = a[:, 1, 1] * a[:, 2, 2] - a[:, 1, 2] * a[:, 2, 1]
    adj[:, 0, 1] = a[:, 0, 2] * a[:, 2, 1] - a[:, 0, 1] * a[:, 2, 2]
    adj[:, 0, 2] = a[:, 0, 1] * a[:, 1, 2] - a[:, 0, 2] * a[:, 1, 1]
    adj[:, 1, 0] = a[:, 1, 2] * a[:, 2, 0] - a[:, 1, 0] * a[:, 2, 2]
    adj[:, 1, 1] = a[:, 0, 0] * a[:, 2, 2] - a[:, 0, 2] * a[:, 2, 0]
    adj[:, 1, 2] = a[:, 0, 2] * a[:, 1, 0] - a[:, 0, 0] * a[:, 1, 2]
    adj[:, 2, 0] = a[:, 1, 0] * a[:, 2, 1] - a[:, 1, 1] * a[:, 2, 0]
    adj[:, 2, 1] = a[:, 0, 1] * a[:, 2, 0] - a[:, 0, 0] * a[:, 2, 1]
    adj[:, 2, 2] = a[:, 0, 0] * a[:, 1, 1] - a[:, 0, 1] * a[:, 1, 0]
    # geo order rxJ sxJ txJ ryJ syJ tyJ rzJ szJ tzJ : (r,s,t) fastest, then (x,y,z)
    geo = np.ascontiguousarray(np.transpose(adj, (0, 2, 1)).reshape(-1, 9))
    # nJ_f = (J G^{-T}) nrJ_f = adj^T nrJ_f
    nJ = np.einsum("kij,fi->kfj", adj, REF_NORMALS)
    return geo, J, nJ


def morton3(x, y, z):
    """interleave the bits of three non-negative integers (x lowest)"""
    code = np.zeros_like(np.asarray(x, dtype=np.int64))
    for bit in range(21):
        for d, v in enumerate((x, y, z)):
            code |= ((np.asarray(v, dtype=np.int64) >> bit) & 1) << (3 * bit + d)
    return code


def kuhn_mesh(n: int, jitter: float = 0.1, seed: int | None = 0, periodic: bool = True,
              order: str = "morton") -> TetMesh:
    """Unit cube, n^3 cubes, each Kuhn-split into 6 tets (K = 6 n^3).

    Interior lattice vertices (all coordinates strictly inside (0,n)) are moved by
    i.i.d. U[-jitter*h, jitter*h] per coordinate from ``numpy.random.default_rng(seed)``
    drawn as one ((n-1)^3, 3) array in lattice order (x fastest).  Boundary vertices stay on
    the lattice, so periodic images match exactly.

    Element order: cubes along a Morton (Z-order) curve of (cx, cy, cz) -- ``order="morton"``,
    so that face neighbours in every direction have nearby element numbers -- or
    lexicographically (cx fastest, ``order="lex"``); within a cube, the 6 permutations of
    (x,y,z) in lexicographic order.  The numbering changes no geometry.
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    h = 1.0 / n
    ncube = n ** 3
    K = 6 * ncube
    c = np.arange(ncube)
    cx, cy, cz = c % n, (c // n) % n, c // (n * n)
    if order == "morton":
        perm_c = np.argsort(morton3(cx, cy, cz), kind="stable")
        cx, cy, cz = cx[perm_c], cy[perm_c], cz[perm_c]
    elif order != "lex":
        raise ValueError(f"unknown element order {order!r}")
    corner = np.stack([cx, cy, cz], axis=1)  # (ncube,3)
    lat = np.empty((K, 4, 3), dtype=np.int64)
    for t, perm in enumerate(itertools.permutations(range(3))):
        e0 = np.zeros(3, dtype=np.int64)
        e0[perm[0]] = 1
        e1 = e0.copy()
        e1[perm[1]] = 1
        P = [np.zeros(3, np.int64), e0, e1, np.ones(3, np.int64)]
        for m in range(4):
            lat[t::6, m, :] = corner + P[m]
    # jitter of interior lattice vertices
    disp = np.zeros((n + 1, n + 1, n + 1, 3))
    if n >= 2 and jitter != 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter * h, jitter * h, size=((n - 1) ** 3, 3))
        # lattice order x fastest: index (i-1) + (n-1)*((j-1) + (n-1)*(k-1))
        disp[1:n, 1:n, 1:n, :] = d.reshape(n - 1, n - 1, n - 1, 3).transpose(2, 1, 0, 3)
    vx = lat * h + disp[lat[..., 0], lat[..., 1], lat[..., 2]]
    geo, J, nJ = geometric_factors(vx)
    flip = J < 0
    if np.any(flip):
        lat[flip] = lat[flip][:, [1, 0, 2, 3], :]
        vx[flip] = vx[flip][:, [1, 0, 2, 3], :]
        geo, J, nJ = geometric_factors(vx)
    nbr, nbr_face, perm, bc = _connect(lat, n, periodic)
    return TetMesh(vx=vx, geo=geo, J=J, nJ=nJ, nbr=nbr, nbr_face=nbr_face, perm=perm, bc=bc, h=h,
                   periodic=periodic, meta={"n": n, "jitter": jitter, "seed": seed, "order": order})


def _connect(lat: np.ndarray, n: int, periodic: bool):
    """Match faces by their (periodically shifted) lattice vertices."""
    K = lat.shape[0]
    fv = lat[:, FACE_VERTS, :]  # (K,4,3 verts,3 coords)
    if periodic:
        shift = n * np.floor_divide(fv.min(axis=2), n)  # (K,4,3)
        fv = fv - shift[:, :, None, :]
    # vertex keys as single integers
    base = n + 2
    vkey = (fv[..., 0] * base + fv[..., 1]) * base + fv[..., 2]  # (K,4,3)
    skey = np.sort(vkey, axis=2)
    fkey = (skey[..., 0] * base ** 3 + skey[..., 1]) * base ** 3 + skey[..., 2]  # (K,4)
    flat = fkey.reshape(-1)
    order = np.argsort(flat, kind="stable")
    sk = flat[order]
    nbr = np.repeat(np.arange(K)[:, None], 4, axis=1)
    nbr_face = np.tile(np.arange(4), (K, 1))
    perm = np.zeros((K, 4), dtype=np.int64)
    bc = np.full((K, 4), BC_WALL, dtype=np.int64)
    same = sk[1:] == sk[:-1]
    if np.any(same[1:] & same[:-1]):
        raise RuntimeError("a face is shared by more than two elements")
    idx = np.nonzero(same)[0]
    a, b = order[idx], order[idx + 1]
    ea, fa = a // 4, a % 4
    eb, fb = b // 4, b % 4
    nbr[ea, fa], nbr_face[ea, fa] = eb, fb
    nbr[eb, fb], nbr_face[eb, fb] = ea, fa
    bc[ea, fa] = BC_INTERIOR
    bc[eb, fb] = BC_INTERIOR
    # permutations: nbr_face_vertex[m] == my_face_vertex[p[m]]
    va, vb = vkey[ea, fa], vkey[eb, fb]  # (P,3)
    perm[ea, fa] = _perm_index(va, vb)
    perm[eb, fb] = _perm_index(vb, va)
    if periodic and np.any(bc != BC_INTERIOR):
        raise RuntimeError("periodic mesh has unmatched faces")
    return nbr, nbr_face, perm, bc


def _perm_index(mine: np.ndarray, theirs: np.ndarray) -> np.ndarray:
    """index p into PERMS with theirs[m] == mine[PERMS[p][m]]"""
    out = np.full(mine.shape[0], -1, dtype=np.int64)
    for p, pm in enumerate(PERMS):
        ok = np.all(mine[:, pm] == theirs, axis=1)
        out[ok] = p
    if np.any(out < 0):
        raise RuntimeError("face vertices do not match")
    return out


# ----------------------------------------------------------------- index helpers

def np_(N: int) -> int:
    return (N + 1) * (N + 2) * (N + 3) // 6


def nfp(N: int) -> int:
    return (N + 1) * (N + 2) // 2


def lin3(N: int, i, j, k):
    """0-based volume index (src/ordering.jl:33-35 minus one)."""
    i, j, k = np.asarray(i), np.asarray(j), np.asarray(k)
    tri = j * (2 * (N - k) + 3 - j) // 2
    tet = np_(N) - (N - k + 1) * (N - k + 2) * (N - k + 3) // 6
    return i + tri + tet


def face_node_abc(N: int) -> np.ndarray:
    """(Nfp,3) face-local Bernstein indices (a,b,c) in 2-D order (b outer, a inner)."""
    return np.array([(a, b, N - a - b) for b in range(N + 1) for a in range(N + 1 - b)], dtype=np.int64)


def fmask(N: int) -> np.ndarray:
    """(4,Nfp) volume index of each face node; plays the role of rd.Fmask."""
    abc = face_node_abc(N)
    a, b, c = abc[:, 0], abc[:, 1], abc[:, 2]
    zero = np.zeros_like(a)
    return np.stack([lin3(N, a, zero, b), lin3(N, c, a, b), lin3(N, zero, a, b), lin3(N, a, b, zero)])


def face_node_perm_table(N: int) -> np.ndarray:
    """(6,Nfp): face-node index on the neighbour side for each permutation."""
    abc = face_node_abc(N)
    out = np.empty((6, abc.shape[0]), dtype=np.int64)
    for p, pm in enumerate(PERMS):
        t = abc[:, pm]
        out[p] = t[:, 0] + t[:, 1] * (2 * N + 3 - t[:, 1]) // 2
    return out


def map_p(mesh: TetMesh, N: int) -> tuple[np.ndarray, np.ndarray]:
    """StartUpDG-style mapP (K,4Nfp) into the flattened [K*4Nfp] face-node array, and the
    per-face-node boundary code."""
    Nf = nfp(N)
    pt = face_node_perm_table(N)  # (6,Nf)
    K = mesh.K
    loc = pt[mesh.perm]  # (K,4,Nf)
    mp = (mesh.nbr[:, :, None] * 4 + mesh.nbr_face[:, :, None]) * Nf + loc
    self_idx = (np.arange(K)[:, None, None] * 4 + np.arange(4)[None, :, None]) * Nf + np.arange(Nf)[None, None, :]
    wall = mesh.bc[:, :, None] != BC_INTERIOR
    mp = np.where(wall, self_idx, mp)
    bcn = np.broadcast_to(mesh.bc[:, :, None], (K, 4, Nf))
    return np.ascontiguousarray(mp.reshape(K, 4 * Nf)).astype(np.int32), np.ascontiguousarray(bcn.reshape(K, 4 * Nf)).astype(np.int8)


def face_node_normals(mesh: TetMesh, N: int):
    """nxJ, nyJ, nzJ broadcast to (K,4Nfp) (constant per face of an affine tet)."""
    Nf = nfp(N)
    out = []
    for d in range(3):
        out.append(np.ascontiguousarray(np.repeat(mesh.nJ[:, :, d], Nf, axis=1)))
    return out


def multi_indices(N: int) -> np.ndarray:
    """(Np,4) (i,j,k,l) in the reference's linear order (src/ordering.jl:37-39)."""
    return np.array([(i, j, k, N - i - j - k) for k in range(N + 1) for j in range(N + 1 - k)
                     for i in range(N + 1 - j - k)], dtype=np.int64)


def domain_points(mesh: TetMesh, N: int) -> np.ndarray:
    """(K,Np,3) physical domain points x_alpha = sum_m alpha_m/N X_m."""
    mi = multi_indices(N).astype(np.float64) / N
    return np.einsum("am,kmx->kax", mi, mesh.vx)
