"""GPU parity at the benchmarked sizes: one fused RHS of the Maxwell N=9 workload over all
K = 196608 elements (32^3 x 6 Kuhn cubes, jitter seed 3, PEC walls), and the N=7 derivative
and 4-face lift operators over K = 262144 elements (F = 4, geometric factors cycled from
the 32^3 x 6 seed-1 mesh), each against the oracle on the same seeded random inputs.  The
smaller parity tests leave the large-K indexing (element offsets beyond 2^24 values, the
XCD-aware block remap at these grid sizes) to this file."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def threads():
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else min(n, 16)


def worst_element(got, ref):
    """largest per-element relative error (an indexing mistake confined to a few elements
    shows here even when the global norm hides it)"""
    K = got.shape[0]
    d = np.abs(got - ref).reshape(K, -1).max(axis=1)
    s = np.abs(ref).reshape(K, -1).max(axis=1)
    return float((d / np.maximum(s, 1e-300)).max())


def test_fused_rhs_maxwell_n9_full_size(ctx):
    import bbdg
    import oracle as O
    from bbdg import mesh

    N, n = 9, 32
    m = mesh.kuhn_mesh(n, 0.1, 3, periodic=False)
    assert m.K == 196608
    rng = np.random.default_rng(33)
    q = np.ascontiguousarray(rng.uniform(-1, 1, (m.K, 220, 6)))
    ma = O.MeshArrays(m, N, [1, 1, 1, -1, -1, -1])
    print(f"full-size fused RHS: Maxwell N={N} K={m.K}", flush=True)
    ref = O.rhs(1, N, q, ma, nthreads=threads(), flux_kind=1)
    s = bbdg.Solver(m, N, "maxwell", flux="upwind")
    s.set_state(q)
    for prec, tol, etol in (("f64", TOL64, 1e-10), ("mixed", 2e-5, 1e-3)):
        s.autotune(prec)  # the tiling the benchmark runs
        got = s.rhs("fused", prec)
        assert rel(got, ref) <= tol, (prec, rel(got, ref))
        assert worst_element(got, ref) <= etol, (prec, worst_element(got, ref))
    s.close()


def test_operators_n7_full_size(ctx):
    import bbdg
    import oracle as O
    from bbdg import mesh

    N, F, K, NP, NFP = 7, 4, 262144, 120, 36
    m = mesh.kuhn_mesh(32, 0.1, 1, periodic=False)
    geo_h = np.ascontiguousarray(m.geo[np.arange(K) % m.K])
    rng = np.random.default_rng(2)
    x_h = rng.uniform(-1, 1, (K, F, NP))
    f_h = rng.uniform(-1, 1, (K, F, 4 * NFP))
    print(f"full-size operators: N={N} F={F} K={K}", flush=True)
    # oracle: the three reference-coordinate derivatives, combined with each element's factors
    d_ref = [O.deriv_mul(ax, N, x_h.reshape(-1, NP)).reshape(K, F, NP) for ax in range(3)]
    g = geo_h[:, None, None, :]
    grad_ref = [g[..., 3 * c] * d_ref[0] + g[..., 3 * c + 1] * d_ref[1] + g[..., 3 * c + 2] * d_ref[2]
                for c in range(3)]
    del d_ref
    lift_ref = O.lift(N, f_h.reshape(-1, 4 * NFP)).reshape(K, F, NP)
    geo = bbdg.DeviceArray.from_numpy(geo_h, ctx=ctx)
    for dt, tg, tl in ((np.float64, TOL64, TOL64), (np.float32, 2e-6, 1e-5)):
        x = bbdg.DeviceArray.from_numpy(x_h.astype(dt), ctx=ctx)
        xv = x.view((NP, K, F), (1, F * NP, NP))
        o = [bbdg.DeviceArray.empty((K, F, NP), dt, ctx=ctx) for _ in range(3)]
        ov = [a.view(xv.shape, xv.strides) for a in o]
        bbdg.grad_(ov[0], ov[1], ov[2], N, xv, geo)
        for c in range(3):
            got = o[c].numpy().astype(np.float64)
            assert rel(got, grad_ref[c]) <= tg, (dt, c, rel(got, grad_ref[c]))
            assert worst_element(got, grad_ref[c]) <= 100 * tg, (dt, c)
        for a in o + [x]:
            a.free()
        fd = bbdg.DeviceArray.from_numpy(f_h.astype(dt), ctx=ctx)
        fv = fd.view((4 * NFP, K, F), (1, F * 4 * NFP, 4 * NFP))
        lo = bbdg.DeviceArray.empty((K, F, NP), dt, ctx=ctx)
        lv = lo.view((NP, K, F), (1, F * NP, NP))
        bbdg.mul_(lv, bbdg.BernsteinLift(N, dt, "acc64"), fv)
        got = lo.numpy().astype(np.float64)
        assert rel(got, lift_ref) <= tl, (dt, rel(got, lift_ref))
        assert worst_element(got, lift_ref) <= 100 * tl, dt
        for a in (fd, lo):
            a.free()
    geo.free()
