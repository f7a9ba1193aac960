"""GPU parity of the solver paths (plain op-by-op FP64, fused FP64, fused mixed) against
the oracle's restatement of examples/3d_wave.jl / 3d_maxwell.jl + LSRK45."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def setup(pde, N, n=2, seed=0, periodic=None, flux="upwind"):
    import bbdg
    import oracle as O
    from bbdg import basis, mesh

    if periodic is None:
        periodic = pde == "wave"
    m = mesh.kuhn_mesh(n, 0.1, seed, periodic=periodic)
    if pde == "wave":
        fn = lambda x, y, z: np.stack([np.exp(-25 * ((x - .5) ** 2 + (y - .5) ** 2 + (z - .5) ** 2)),
                                       0.3 * np.sin(2 * np.pi * y), 0 * x, 0.1 * np.cos(2 * np.pi * x)])
        sign = [1.0] * 4
    else:
        fn = lambda x, y, z: np.stack([0.2 * np.sin(np.pi * z), 0 * x, 0.1 * x * y, 0 * x, 0.05 * z,
                                       np.sin(np.pi * x) * np.sin(np.pi * y)])
        sign = [1, 1, 1, -1, -1, -1]
    q0 = np.ascontiguousarray(basis.interpolate(m, N, fn))
    s = bbdg.Solver(m, N, pde, flux=flux)
    s.set_state(q0)
    ma = O.MeshArrays(m, N, sign)
    return m, q0, s, ma, O


@pytest.mark.parametrize("pde,N", [("wave", 1), ("wave", 3), ("wave", 5), ("wave", 7), ("wave", 9), ("wave", 10),
                                   ("maxwell", 2), ("maxwell", 5), ("maxwell", 9)])
@pytest.mark.parametrize("flux", ["upwind", "reference"])
def test_rhs_parity(ctx, pde, N, flux):
    m, q0, s, ma, O = setup(pde, N, flux=flux)
    p = 0 if pde == "wave" else 1
    ref = O.rhs(p, N, q0, ma, nthreads=4, flux_kind=1 if flux == "upwind" else 0)
    for path in ("plain", "fused"):
        got = s.rhs(path, "f64")
        assert rel(got, ref) <= TOL64, (path, rel(got, ref))
    mixed = s.rhs("fused", "mixed")
    assert rel(mixed, ref) <= 2e-5, rel(mixed, ref)


@pytest.mark.parametrize("pde,N", [("wave", 5), ("wave", 7), ("maxwell", 9)])
def test_step_parity(ctx, pde, N):
    m, q0, s, ma, O = setup(pde, N)
    p = 0 if pde == "wave" else 1
    dt = 0.5 * m.h / N ** 2
    qr = q0.copy()
    for _ in range(3):
        O.lsrk_step(p, N, qr, dt, ma, nthreads=4, flux_kind=1)
    for path, prec, tol in (("plain", "f64", TOL64), ("fused", "f64", TOL64), ("fused", "mixed", 1e-5)):
        s.set_state(q0)
        s.step(dt, 3, path, prec)
        got = s.get_state()
        assert rel(got, qr) <= tol, (path, prec, rel(got, qr))


def test_graph_replay_matches_eager(ctx):
    m, q0, s, ma, O = setup("wave", 7, n=3)
    dt = 0.5 * m.h / 49
    s.step(dt, 6, "fused", "f64")
    eager = s.get_state()
    for nsteps in (3, 2):  # odd and even stage counts per graph
        s.set_state(q0)
        s.capture(dt, nsteps, "fused", "f64")
        for _ in range(6 // nsteps):
            s.launch()
            s.get_state()  # eager host work between replays
        assert np.array_equal(s.get_state(), eager)
    s.set_state(q0)
    s.capture(dt, 2, "plain", "f64")
    for _ in range(3):
        s.launch()
    assert rel(s.get_state(), eager) <= TOL64


def test_autotune_and_tiling_invariance(ctx):
    """results do not depend on elements per workgroup (bit for bit) nor on the lift
    organisation (same sums in another order)"""
    m, q0, s, ma, O = setup("wave", 5, n=3)
    ref = O.rhs(0, 5, q0, ma, nthreads=4, flux_kind=1)
    outs = {}
    import bbdg

    for lift in ("gather", "pingpong", "wave", "wave_gt", "wave1"):
        for eb in (1, 2, 4):
            try:
                s.set_tuning("f64", s.tiling_code(eb, lift))
            except bbdg.BBError:
                continue  # not compiled (the wave organisation has EB = 1, 2 only)
            outs[(lift, eb)] = s.rhs("fused", "f64")
            assert rel(outs[(lift, eb)], ref) <= TOL64
        for eb in (2, 4):
            if (lift, eb) in outs:
                assert np.array_equal(outs[(lift, eb)], outs[(lift, 1)])
    # the wave organisations compute every chain item exactly as the gather one
    assert np.array_equal(outs[("wave", 1)], outs[("gather", 1)])
    assert np.array_equal(outs[("wave_gt", 1)], outs[("gather", 1)])
    assert np.array_equal(outs[("wave1", 1)], outs[("gather", 1)])
    t = s.autotune("f64")
    assert t in [s.tiling_code(e, lm) for e in (1, 2, 4) for lm in ("gather", "pingpong", "wave", "wave_gt", "wave1")]
    tm = s.autotune("mixed")
    assert tm in [s.tiling_code(e, lm) for e in (1, 2, 4)
                  for lm in ("gather", "pingpong", "gather_mfma", "wave", "wave_gt", "wave1")]
    assert rel(s.rhs("fused", "f64"), ref) <= TOL64
    for prec in ("f64", "mixed"):
        s.set_tuning(prec, s.tiling_code(1, "pingpong"))
    for path, prec, tol in (("fused", "f64", TOL64), ("fused", "mixed", 2e-5)):
        assert rel(s.rhs(path, prec), ref) <= tol


@pytest.mark.parametrize("pde,N,n", [("wave", 2, 3), ("wave", 3, 3), ("wave", 7, 2), ("wave", 10, 2),
                                     ("maxwell", 3, 3), ("maxwell", 5, 3), ("maxwell", 9, 2), ("maxwell", 8, 2)])
def test_wave_owned_chains(ctx, pde, N, n):
    """chains owned by single waves (no workgroup barrier between flux, L0 and the chain),
    index tables in LDS or in global memory: FP64 bit-identical to the gather organisation, mixed within its tolerance, for every
    compiled elements-per-workgroup count (partial last workgroups included), RHS and two
    steps"""
    # N = 2: the flux has its own buffer; N = 3: each chain's flux fills exactly the rest of
    # its chain region (2 Nfp = Np); N >= 5: the flux sits inside the chain region
    import bbdg

    m, q0, s, ma, O = setup(pde, N, n=n)
    p = 0 if pde == "wave" else 1
    ref = O.rhs(p, N, q0, ma, nthreads=4, flux_kind=1)
    dt = 0.5 * m.h / N ** 2
    ran = 0
    for prec in ("f64", "mixed"):
        for eb in (1, 2):
            for org in ("wave", "wave_gt", "wave1"):
                try:
                    s.set_tuning(prec, s.tiling_code(eb, org))
                except bbdg.BBError:
                    continue  # does not fit in LDS
                ran += 1
                got = s.rhs("fused", prec)
                assert rel(got, ref) <= (TOL64 if prec == "f64" else 2e-5)
                s.set_state(q0)
                s.step(dt, 2, "fused", prec)
                a = s.get_state()
                s.set_tuning(prec, s.tiling_code(eb, "gather"))
                s.set_state(q0)
                s.step(dt, 2, "fused", prec)
                b = s.get_state()
                s.set_state(q0)
                if prec == "f64":
                    assert np.array_equal(a, b)
                    assert np.array_equal(got, s.rhs("fused", prec))
                else:
                    assert rel(a, b) <= 1e-6
    assert ran >= 1


@pytest.mark.parametrize("pde,N,n", [("wave", 1, 2), ("wave", 4, 3), ("wave", 7, 2), ("wave", 10, 2),
                                     ("maxwell", 3, 3), ("maxwell", 6, 2), ("maxwell", 9, 2)])
def test_mfma_l0_tilings(ctx, pde, N, n):
    """L0 on the matrix cores (mixed precision): every compiled elements-per-workgroup
    count, partial last workgroups (K = 6 n^3 is not a multiple of 4), against the oracle
    and against the same tiling with L0 on the vector units, for the RHS and one step"""
    import bbdg

    m, q0, s, ma, O = setup(pde, N, n=n)
    p = 0 if pde == "wave" else 1
    ref = O.rhs(p, N, q0, ma, nthreads=4, flux_kind=1)
    dt = 0.5 * m.h / N ** 2
    ran = 0
    for eb in (1, 2, 4):
        try:
            s.set_tuning("mixed", s.tiling_code(eb, "gather_mfma"))
        except bbdg.BBError:
            continue
        ran += 1
        got = s.rhs("fused", "mixed")
        assert rel(got, ref) <= 2e-5, (eb, rel(got, ref))
        s.set_tuning("mixed", s.tiling_code(eb, "gather"))
        valu = s.rhs("fused", "mixed")
        assert rel(got, valu) <= 2e-6, (eb, rel(got, valu))
        s.set_tuning("mixed", s.tiling_code(eb, "gather_mfma"))
        s.set_state(q0)
        s.step(dt, 1, "fused", "mixed")
        a = s.get_state()
        s.set_tuning("mixed", s.tiling_code(eb, "gather"))
        s.set_state(q0)
        s.step(dt, 1, "fused", "mixed")
        assert rel(a, s.get_state()) <= 1e-6
        s.set_state(q0)
    assert ran >= 1


def test_config1_every_step(ctx):
    """3-D acoustic wave, N=5, 8^3 x 6 jittered Kuhn mesh (seed 0), periodic,
    p = exp(-25|x-0.5|^2), LSRK45 dt = 0.5 h / N^2, 50 steps; fields compared after every
    step against the oracle (FP64 fused path)."""
    import bbdg
    import oracle as O
    from bbdg import basis, mesh

    N = 5
    m = mesh.kuhn_mesh(8, 0.1, 0, periodic=True)
    q0 = np.ascontiguousarray(basis.interpolate(m, N, lambda x, y, z: np.stack(
        [np.exp(-25 * ((x - .5) ** 2 + (y - .5) ** 2 + (z - .5) ** 2)), 0 * x, 0 * x, 0 * x])))
    s = bbdg.Solver(m, N, "wave")
    s.set_state(q0)
    ma = O.MeshArrays(m, N, [1, 1, 1, 1])
    dt = 0.5 * m.h / N ** 2
    qr = q0.copy()
    worst = 0.0
    for step in range(50):
        O.lsrk_step(0, N, qr, dt, ma, nthreads=8, flux_kind=1)
        s.step(dt, 1, "fused", "f64")
        e = rel(s.get_state(), qr)
        worst = max(worst, e)
        assert e <= TOL64, (step, e)
    print(f"config 1: worst per-step relative L2 difference {worst:.3e}")


@pytest.mark.parametrize("pde,N,periodic", [("wave", 3, True), ("wave", 9, False), ("maxwell", 5, True),
                                            ("maxwell", 9, False), ("maxwell", 10, False)])
def test_random_state_parity(ctx, pde, N, periodic):
    """every field non-zero (smooth initial states can hide a field mix-up behind zeros)"""
    import bbdg
    import oracle as O
    from bbdg import mesh

    p, F = (0, 4) if pde == "wave" else (1, 6)
    m = mesh.kuhn_mesh(2, 0.1, 4, periodic=periodic)
    sign = [1.0] * F if p == 0 else [1, 1, 1, -1, -1, -1]
    ma = O.MeshArrays(m, N, sign)
    s = bbdg.Solver(m, N, pde, bc_sign=sign)
    q = np.ascontiguousarray(np.random.default_rng(N).uniform(-1, 1, (m.K, s.Np, F)))
    s.set_state(q)
    ref = O.rhs(p, N, q, ma, nthreads=4, flux_kind=1)
    for path in ("plain", "fused"):
        got = s.rhs(path, "f64")
        for f in range(F):
            assert rel(got[..., f], ref[..., f]) <= TOL64, (path, f, rel(got[..., f], ref[..., f]))
    assert rel(s.rhs("fused", "mixed"), ref) <= 2e-5
    dt = 0.5 * m.h / N ** 2
    qr = q.copy()
    O.lsrk_step(p, N, qr, dt, ma, nthreads=4, flux_kind=1)
    for path in ("plain", "fused"):
        s.set_state(q)
        s.step(dt, 1, path, "f64")
        assert rel(s.get_state(), qr) <= TOL64, path
