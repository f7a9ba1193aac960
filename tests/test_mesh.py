"""Invariants of the mesh builder (host set-up code): they pin geometric factors, normals
and the face matching that replace StartUpDG's MeshData."""
import numpy as np
import pytest


def np3(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


@pytest.mark.parametrize("n,periodic", [(1, True), (2, True), (3, False), (4, True)])
def test_geometry_invariants(n, periodic):
    from bbdg import mesh

    m = mesh.kuhn_mesh(n, 0.1, 7, periodic=periodic)
    assert m.K == 6 * n ** 3
    assert (m.J > 0).all()
    # volumes add up to the unit cube (reference volume 4/3)
    assert abs(m.J.sum() * 4 / 3 - 1.0) < 1e-12
    # closed surface: sum_f (reference face weight 2) nJ_f = 0
    assert np.abs((2 * m.nJ).sum(axis=1)).max() < 1e-14
    # |nJ| = physical face area / 2
    for f, fv in enumerate(mesh.FACE_VERTS):
        a = m.vx[:, fv[1]] - m.vx[:, fv[0]]
        b = m.vx[:, fv[2]] - m.vx[:, fv[0]]
        area = 0.5 * np.linalg.norm(np.cross(a, b), axis=1)
        np.testing.assert_allclose(np.linalg.norm(m.nJ[:, f], axis=1), area / 2, rtol=1e-12)
        # outward: normal points away from the opposite vertex
        opp = m.vx[:, [v for v in range(4) if v not in fv][0]]
        assert (np.einsum("kx,kx->k", m.nJ[:, f], m.vx[:, fv[0]] - opp) > 0).all()
    # J G^{-1} G = J I
    X0, X1, X2, X3 = (m.vx[:, i] for i in range(4))
    G = np.stack([(X0 - X3) / 2, (X1 - X3) / 2, (X2 - X3) / 2], axis=2)
    adj = m.geo.reshape(-1, 3, 3).transpose(0, 2, 1)  # rows r,s,t ; cols x,y,z
    np.testing.assert_allclose(adj @ G, m.J[:, None, None] * np.eye(3), atol=1e-14)


@pytest.mark.parametrize("n,periodic", [(2, True), (3, False), (3, True)])
@pytest.mark.parametrize("N", [1, 3, 6])
def test_face_matching(n, periodic, N):
    """matched face nodes sit at the same physical point (modulo the periodic shift), the
    neighbour relation is an involution and normals are opposite"""
    from bbdg import mesh

    m = mesh.kuhn_mesh(n, 0.1, 5, periodic=periodic)
    mp, bc = mesh.map_p(m, N)
    X = mesh.domain_points(m, N).reshape(-1, 3)
    fm = mesh.fmask(N).reshape(-1)
    K, Nf = m.K, mesh.nfp(N)
    pos = X.reshape(K, -1, 3)[:, fm, :].reshape(-1, 3)
    inner = bc.reshape(-1) == 0
    d = pos[mp.reshape(-1)] - pos
    if periodic:
        d = d - np.round(d)
    assert np.abs(d[inner]).max() < 1e-12
    assert (mp.reshape(-1)[mp.reshape(-1)] == np.arange(mp.size)).all()
    e = np.arange(K)[:, None]
    inter = m.bc == 0
    np.testing.assert_allclose(m.nJ[m.nbr, m.nbr_face][inter], -m.nJ[inter], atol=1e-14)
    if not periodic:
        # boundary faces lie on the cube boundary
        wall = pos[~inner]
        on = np.isclose(wall, 0).any(1) | np.isclose(wall, 1).any(1)
        assert on.all()
        assert (m.nbr[~inter] == np.broadcast_to(e, m.nbr.shape)[~inter]).all()


def test_fmask_matches_oracle(orc):
    from bbdg import mesh

    for N in (1, 5, 10):
        assert np.array_equal(mesh.fmask(N), orc.fmask(N))


def test_interpolation_reproduces_polynomials():
    from bbdg import basis, mesh

    m = mesh.kuhn_mesh(2, 0.1, 1)
    N = 4
    q = basis.interpolate(m, N, lambda x, y, z: np.stack([x * x * y - z ** 3, 1 + 0 * x]))
    X = mesh.domain_points(m, N)
    vals = basis.evaluate(N, q)
    np.testing.assert_allclose(vals[..., 0], X[..., 0] ** 2 * X[..., 1] - X[..., 2] ** 3, atol=1e-12)
    np.testing.assert_allclose(q[..., 1], 1.0, atol=1e-12)  # partition of unity


def test_morton_order_is_a_renumbering():
    """the Morton numbering permutes the same elements (same geometry, same volumes)"""
    from bbdg import mesh

    a = mesh.kuhn_mesh(8, 0.1, 3, order="lex")
    b = mesh.kuhn_mesh(8, 0.1, 3, order="morton")
    key = lambda m: np.round(m.vx.reshape(m.K, -1), 12)
    ka = {tuple(r) for r in key(a)}
    kb = {tuple(r) for r in key(b)}
    assert ka == kb
    np.testing.assert_allclose(np.sort(a.J), np.sort(b.J))
    # more face neighbours lie within a short window of element numbers
    win = lambda m: (np.abs(m.nbr - np.arange(m.K)[:, None])[m.bc == 0] < 200).mean()
    assert win(b) > win(a)
