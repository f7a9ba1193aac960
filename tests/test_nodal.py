"""Dense nodal comparator (extras/3d_nodal_wave.jl:57-73): host operators on the CPU, and
on the GPU the nodal RHS of V c against V times the oracle's Bernstein RHS of c."""
import numpy as np
import pytest


def np3(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


@pytest.mark.parametrize("N", [1, 2, 4, 6])
def test_nodal_derivatives_are_exact_on_polynomials(N):
    """d/dr of a degree-N polynomial sampled at the nodes, r s t the reference coordinates"""
    from bbdg import basis
    from bbdg.nodal import nodal_derivative_matrices

    Dr, Ds, Dt = nodal_derivative_matrices(N)
    p = basis.equispaced_ref_points(N)
    r, s, t = p[:, 0], p[:, 1], p[:, 2]
    if N == 1:
        f, fr, fs, ft = 2 * r - s + 3 * t, 2 + 0 * r, -1 + 0 * r, 3 + 0 * r
    else:  # degree N: r^(N-1) s + t^N + 0.3 r t
        f = r ** (N - 1) * s + t ** N + 0.3 * r * t
        fr = (N - 1) * r ** (N - 2) * s + 0.3 * t
        fs = r ** (N - 1)
        ft = N * t ** (N - 1) + 0.3 * r
    np.testing.assert_allclose(Dr @ f, fr, atol=1e-10)
    np.testing.assert_allclose(Ds @ f, fs, atol=1e-10)
    np.testing.assert_allclose(Dt @ f, ft, atol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("pde,N,periodic", [("wave", 4, True), ("maxwell", 3, False), ("wave", 7, True),
                                            ("maxwell", 6, True)])
def test_nodal_rhs_is_V_times_bernstein_rhs(ctx, orc, pde, N, periodic):
    from bbdg import basis, mesh
    from bbdg.nodal import NodalOperators, NodalRHS

    m = mesh.kuhn_mesh(2, 0.1, 4, periodic=periodic)
    F = 4 if pde == "wave" else 6
    rng = np.random.default_rng(N)
    k = rng.uniform(1, 2, (F, 3))
    c = basis.interpolate(m, N, lambda x, y, z: np.stack(
        [np.sin(2 * np.pi * (k[f, 0] * x + k[f, 1] * y) + k[f, 2] * z) for f in range(F)]))
    sign = [1.0] * 4 if pde == "wave" else [1, 1, 1, -1, -1, -1]
    ref_b = orc.rhs(0 if pde == "wave" else 1, N, c, orc.MeshArrays(m, N, sign), nthreads=4, flux_kind=1)
    ops = NodalOperators(N)
    V = ops.V
    qn = np.einsum("ab,kbf->kfa", V, c)                       # nodal values, [K][F][Np]
    rhs = NodalRHS(m, N, pde, ops=ops)
    du = rhs.rhs(qn)
    ref_n = np.einsum("ab,kbf->kfa", V, ref_b)
    err = np.linalg.norm(du - ref_n) / np.linalg.norm(ref_n)
    assert err <= 1e-10, err
    rhs.close()
