"""Pins of the oracle's RHS (examples/3d_wave.jl:36-84, 3d_maxwell.jl:36-93) and of the
LSRK45 integrator.  The reference has no tests for these: they are pinned by invariants,
exact solutions of the PDEs and the published integrator coefficients."""
from fractions import Fraction
from math import factorial

import numpy as np
import pytest


def np3(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


def _setup(n, periodic, N, pde, seed=0):
    import oracle as O
    from bbdg import mesh

    m = mesh.kuhn_mesh(n, 0.1, seed, periodic=periodic)
    sign = [1.0] * 4 if pde == 0 else [1, 1, 1, -1, -1, -1]
    return m, O.MeshArrays(m, N, sign)


@pytest.mark.parametrize("pde", [0, 1])
@pytest.mark.parametrize("kind", [0, 1])
def test_constant_state_is_steady(orc, pde, kind):
    N, F = 4, (4 if pde == 0 else 6)
    m, ma = _setup(2, True, N, pde)
    q = np.tile(np.linspace(0.5, 2.0, F), (m.K, np3(N), 1))
    du = orc.rhs(pde, N, q, ma, flux_kind=kind)
    assert np.abs(du).max() < 1e-9 * np.abs(1 / m.J).max() * 1e-3


def test_linear_fields_exact_gradients(orc):
    """continuous linear fields: no jumps, so du is the exact flux divergence.
    wave: p = x  ->  du/dt = -dp/dx = -1.  Maxwell: Ez = x -> dHy/dt = dEz/dx = 1
    (the reference's Maxwell RHS adds +sum geo*D f, examples/3d_maxwell.jl:84-87)."""
    from bbdg import basis

    N = 3
    for pde, F, fn, comp in ((0, 4, lambda x, y, z: np.stack([x, 0 * x, 0 * x, 0 * x]), 1),
                             (1, 6, lambda x, y, z: np.stack([0 * x] * 5 + [x]), 1)):
        m, ma = _setup(3, False, N, pde)
        ma.bc[:] = 1
        ma.bc_sign[:] = 1.0  # mirror of the own state: no jump on the walls either
        q = basis.interpolate(m, N, fn)
        du = orc.rhs(pde, N, q, ma, flux_kind=1)
        expect = np.zeros_like(du)
        expect[..., comp] = -1.0 if pde == 0 else 1.0
        np.testing.assert_allclose(du, expect, atol=1e-9)


def _run(orc, pde, N, n, T, fn, periodic, kind=1):
    from bbdg import basis, mesh

    m, ma = _setup(n, periodic, N, pde, seed=11)
    q = np.ascontiguousarray(basis.interpolate(m, N, lambda x, y, z: fn(x, y, z, 0.0)))
    nsteps = int(np.ceil(T / (0.5 * m.h / N ** 2)))
    dt = T / nsteps
    for _ in range(nsteps):
        orc.lsrk_step(pde, N, q, dt, ma, nthreads=8, flux_kind=kind)
    X = mesh.domain_points(m, N)
    ex = fn(X[..., 0], X[..., 1], X[..., 2], T)
    got = basis.evaluate(N, q)
    return np.abs(got - np.moveaxis(ex, 0, -1)).max()


def wave_mode(x, y, z, t):
    """exact periodic standing wave of p_t + div u = 0, u_t + grad p = 0 on the unit cube"""
    w = 2 * np.pi * np.sqrt(3)
    c = np.cos
    s = np.sin
    P = c(2 * np.pi * x) * c(2 * np.pi * y) * c(2 * np.pi * z)
    a = s(w * t) / np.sqrt(3)
    return np.stack([P * c(w * t), a * s(2 * np.pi * x) * c(2 * np.pi * y) * c(2 * np.pi * z),
                     a * c(2 * np.pi * x) * s(2 * np.pi * y) * c(2 * np.pi * z),
                     a * c(2 * np.pi * x) * c(2 * np.pi * y) * s(2 * np.pi * z)])


def maxwell_mode(x, y, z, t):
    """TM110 cavity mode with PEC walls (the analytic solution of examples/3d_maxwell.jl:144-150)"""
    w = np.pi * np.sqrt(2)
    z0 = 0 * x
    return np.stack([-np.sin(np.pi * x) * np.cos(np.pi * y) * np.sin(w * t) / np.sqrt(2),
                     np.cos(np.pi * x) * np.sin(np.pi * y) * np.sin(w * t) / np.sqrt(2), z0, z0, z0,
                     np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(w * t)])


def test_wave_converges_to_exact_mode(orc):
    """max error at the domain points after T = 0.05; asymptotic rate ~ N + 1/2"""
    N, T = 4, 0.05
    e1 = _run(orc, 0, N, 4, T, wave_mode, True)
    e2 = _run(orc, 0, N, 8, T, wave_mode, True)
    rate = np.log2(e1 / e2)
    assert e2 < 2e-3 and rate > 4.0, (e1, e2, rate)


def test_maxwell_pec_converges_to_exact_mode(orc):
    N, T = 4, 0.1
    e1 = _run(orc, 1, N, 2, T, maxwell_mode, False)
    e2 = _run(orc, 1, N, 4, T, maxwell_mode, False)
    rate = np.log2(e1 / e2)
    assert e2 < 2e-3 and rate > 4.0, (e1, e2, rate)


def test_upwind_energy_decays(orc):
    """sum_e J_e q_e^T M q_e is non-increasing for the (Jf-consistent) upwind flux"""
    from exact import mass

    N = 3
    M = np.array([[float(v) for v in r] for r in mass(N)])
    m, ma = _setup(2, True, N, 0)
    rng = np.random.default_rng(0)
    q = np.ascontiguousarray(rng.uniform(-1, 1, (m.K, np3(N), 4)))

    def energy(q):
        return float(np.einsum("k,kaf,ab,kbf->", m.J, q, M, q))

    e = [energy(q)]
    for _ in range(5):
        orc.lsrk_step(0, N, q, 0.5 * m.h / N ** 2, ma, flux_kind=1)
        e.append(energy(q))
    assert all(b <= a * (1 + 1e-12) for a, b in zip(e, e[1:])), e
    assert e[-1] < e[0]


@pytest.mark.parametrize("pde,F", [(0, 4), (1, 6)])
def test_flux_kinds_agree_on_unit_normals_and_scale_with_jf(orc, pde, F):
    """The reference's flux (kind 0, examples/3d_wave.jl:45-54, examples/3d_maxwell.jl:45-63)
    and the Jf-consistent one (kind 1) are the same upwind flux written for different
    normal scalings: with nJ = Jf n (n a unit normal), every term of kind 1 carries exactly
    one factor Jf, so flux_1(d, Jf n) = Jf flux_0(d, n), and the two coincide where
    |nJ| = 1.  A dropped, mis-signed or mis-scaled penalty term in either breaks this."""
    from types import SimpleNamespace

    N = 2
    nf = 6
    K = 40
    rng = np.random.default_rng(7 + pde)
    n = rng.normal(size=(K, 4, 3))
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    Jf = rng.uniform(0.3, 3.0, size=(K, 4, 1))
    uM = np.ascontiguousarray(rng.uniform(-1, 1, (K, 4 * nf, F)))
    # every face node's neighbour value is another random face node (interior faces)
    mapP = np.ascontiguousarray(rng.permutation(K * 4 * nf).reshape(K, 4 * nf).astype(np.int32))

    def arrays(nrm):
        rep = [np.ascontiguousarray(np.repeat(nrm[:, :, d], nf, axis=1)) for d in range(3)]
        return SimpleNamespace(mapP=mapP, bc=np.zeros((K, 4 * nf), np.int8), bc_sign=np.ones(F), nx=rep[0],
                               ny=rep[1], nz=rep[2])

    f0_unit = orc.flux(pde, N, uM, arrays(n), flux_kind=0)
    f1_unit = orc.flux(pde, N, uM, arrays(n), flux_kind=1)
    np.testing.assert_allclose(f1_unit, f0_unit, rtol=0, atol=1e-15 * np.abs(f0_unit).max())
    f1_scaled = orc.flux(pde, N, uM, arrays(n * Jf), flux_kind=1)
    jf_nodes = np.repeat(Jf[:, :, 0], nf, axis=1)[:, :, None]
    np.testing.assert_allclose(f1_scaled, jf_nodes * f0_unit, rtol=1e-14, atol=1e-15)
    # and the reference flux itself does not scale this way (its penalties carry Jf^0 and Jf^2)
    f0_scaled = orc.flux(pde, N, uM, arrays(n * Jf), flux_kind=0)
    assert np.abs(f0_scaled - jf_nodes * f0_unit).max() > 1e-3


def test_reference_flux_penalty_is_too_stiff_for_lsrk(orc):
    """The reference's flux (penalties with Jf^0 and Jf^2) blows up under LSRK45 with
    dt = 0.5 h / N^2 on the 8^3 x 6 mesh; the Jf-consistent upwind flux does not."""
    from bbdg import basis

    N = 3
    m, ma = _setup(8, True, N, 0)
    q0 = np.ascontiguousarray(basis.interpolate(m, N, lambda x, y, z: np.stack(
        [np.exp(-25 * ((x - .5) ** 2 + (y - .5) ** 2 + (z - .5) ** 2)), 0 * x, 0 * x, 0 * x])))
    growth = {}
    for kind in (0, 1):
        q = q0.copy()
        for _ in range(3):
            orc.lsrk_step(0, N, q, 0.5 * m.h / N ** 2, ma, nthreads=8, flux_kind=kind)
        growth[kind] = np.abs(q).max() / np.abs(q0).max()
    assert growth[0] > 1e3 and growth[1] < 1.01


def test_lsrk_stability_polynomial(orc):
    """R(z) of the 2N-storage scheme with the published coefficients matches exp(z) to
    fourth order (the coefficients are 13-digit rationals, so up to ~1e-12)."""
    a, b = orc.lsrk_coeffs()
    a = [Fraction(v) for v in a]
    b = [Fraction(v) for v in b]
    # y' = z y with y0 = 1: track polynomials in z as coefficient lists
    y = [Fraction(1)]
    res = [Fraction(0)]

    def add(p, q):
        n = max(len(p), len(q))
        return [(p[i] if i < len(p) else 0) + (q[i] if i < len(q) else 0) for i in range(n)]

    for s in range(5):
        zy = [Fraction(0)] + y
        res = add([a[s] * c for c in res], zy)
        y = add(y, [b[s] * c for c in res])
    for k in range(5):
        assert abs(float(y[k]) - 1 / factorial(k)) < 1e-12, (k, float(y[k]))
    assert abs(float(y[5]) - 1 / 120) > 1e-6  # fifth order is not matched (R has degree 5)


def test_lsrk_fourth_order(orc):
    """error of y' = -y over [0,1] falls ~16x per halving of dt"""
    a, b = orc.lsrk_coeffs()

    def run(n):
        y, r, dt = 1.0, 0.0, 1.0 / n
        for _ in range(n):
            for s in range(5):
                r = a[s] * r + dt * (-y)
                y = y + b[s] * r
        return abs(y - np.exp(-1.0))

    e1, e2 = run(10), run(20)
    assert 13 < e1 / e2 < 20
