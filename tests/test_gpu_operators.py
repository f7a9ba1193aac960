"""GPU parity of the operator kernels against the oracle (FP64: <= 1e-12 relative,
FP32: <= 2e-6 relative in the norm of each output), through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def np3(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


def nfp(N):
    return (N + 1) * (N + 2) // 2


def to_dev_cols(x_kfn, layout, dtype):
    """x (K, F, n) host -> Julia-shaped (n, K, F) device view in the requested layout."""
    import bbdg

    K, F, n = x_kfn.shape
    if layout == "aos":  # Julia column-major SVector{F}: [K][n][F]
        d = bbdg.DeviceArray.from_numpy(np.ascontiguousarray(x_kfn.transpose(0, 2, 1), dtype=dtype))
        return d, d.view((n, K, F), (F, n * F, 1))
    d = bbdg.DeviceArray.from_numpy(np.ascontiguousarray(x_kfn, dtype=dtype))  # [K][F][n]
    return d, d.view((n, K, F), (1, F * n, n))


def from_dev_cols(v):
    """Julia-shaped (n, K, F) device view -> (K, F, n) host"""
    return v.numpy().transpose(1, 2, 0)


def empty_cols(n, K, F, layout, dtype):
    import bbdg

    if layout == "aos":
        d = bbdg.DeviceArray.zeros((K, n, F), dtype)
        return d, d.view((n, K, F), (F, n * F, 1))
    d = bbdg.DeviceArray.zeros((K, F, n), dtype)
    return d, d.view((n, K, F), (1, F * n, n))


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7, 9, 12, 20])
@pytest.mark.parametrize("layout", ["aos", "soa"])
def test_deriv_fp64(ctx, orc, N, layout):
    import bbdg

    rng = np.random.default_rng(100 + N)
    K, F = 7, 3
    x = rng.uniform(-1, 1, (K, F, np3(N)))
    _, xv = to_dev_cols(x, layout, np.float64)
    for axis, D in enumerate([bbdg.BernsteinDerivativeMatrix_3D_r(N), bbdg.BernsteinDerivativeMatrix_3D_s(N),
                              bbdg.BernsteinDerivativeMatrix_3D_t(N)]):
        _, ov = empty_cols(np3(N), K, F, layout, np.float64)
        bbdg.mul_(ov, D, xv)
        got = from_dev_cols(ov)
        ref = orc.deriv_mul(axis, N, x.reshape(-1, np3(N))).reshape(K, F, -1)
        assert rel(got, ref) <= TOL64, (N, axis, layout, rel(got, ref))


@pytest.mark.parametrize("N", [1, 4, 7, 9])
def test_deriv_vector_and_fp32(ctx, orc, N):
    import bbdg

    rng = np.random.default_rng(7)
    n = np3(N)
    x = rng.uniform(-1, 1, n)
    D = bbdg.BernsteinDerivativeMatrix_3D_t(N)
    xd = bbdg.DeviceArray.from_numpy(x)
    od = bbdg.DeviceArray.zeros((n,))
    bbdg.mul_(od, D, xd)
    assert rel(od.numpy(), orc.deriv_mul(2, N, x[None])[0]) <= TOL64
    x32 = bbdg.DeviceArray.from_numpy(x.astype(np.float32))
    o32 = bbdg.DeviceArray.zeros((n,), np.float32)
    bbdg.mul_(o32, D, x32)
    assert rel(o32.numpy(), orc.deriv_mul(2, N, x.astype(np.float32).astype(np.float64)[None])[0]) <= 2e-6


@pytest.mark.parametrize("N", [1, 3, 7, 9])
@pytest.mark.parametrize("dtype,tol", [(np.float64, TOL64), (np.float32, 2e-6)])
def test_grad(ctx, orc, N, dtype, tol):
    import bbdg
    from bbdg import mesh

    m = mesh.kuhn_mesh(2, 0.1, 1)
    K, F = m.K, 4
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (K, F, np3(N)))
    _, xv = to_dev_cols(x, "soa", dtype)
    outs = [empty_cols(np3(N), K, F, "soa", dtype) for _ in range(3)]
    geo = bbdg.DeviceArray.from_numpy(m.geo)
    bbdg.grad_(outs[0][1], outs[1][1], outs[2][1], N, xv, geo)
    xr = x.astype(dtype).astype(np.float64).reshape(-1, np3(N))
    d = [orc.deriv_mul(a, N, xr).reshape(K, F, -1) for a in range(3)]
    g = m.geo[:, :, None, None]
    for c in range(3):
        ref = g[:, 3 * c + 0] * d[0] + g[:, 3 * c + 1] * d[1] + g[:, 3 * c + 2] * d[2]
        assert rel(from_dev_cols(outs[c][1]), ref) <= tol, (N, c, rel(from_dev_cols(outs[c][1]), ref))


@pytest.mark.parametrize("N", [1, 4, 7, 8])
@pytest.mark.parametrize("F", [1, 3, 4])
@pytest.mark.parametrize("layout", ["aos", "soa", "padded"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, TOL64), (np.float32, 2e-6)])
def test_grad_many_blocks(ctx, orc, N, F, layout, dtype, tol):
    """enough elements for many blocks and a partial last block, odd node counts (scalar
    loads), both compact layouts, and a padded element stride (the strided kernel)"""
    import bbdg

    rng = np.random.default_rng(N * 10 + F)
    K, n = 1003, np3(N)
    x = rng.uniform(-1, 1, (K, F, n))
    geo_h = rng.uniform(-2, 2, (K, 9))
    geo = bbdg.DeviceArray.from_numpy(geo_h)
    if layout == "padded":
        pad = np.zeros((K, F, n + 3), dtype)
        pad[:, :, :n] = x
        xd = bbdg.DeviceArray.from_numpy(pad)
        xv = xd.view((n, K, F), (1, F * (n + 3), n + 3))
        outs = []
        for _ in range(3):
            d = bbdg.DeviceArray.zeros((K, F, n + 3), dtype)
            outs.append((d, d.view(xv.shape, xv.strides)))
    else:
        _, xv = to_dev_cols(x, layout, dtype)
        outs = [empty_cols(n, K, F, layout, dtype) for _ in range(3)]
    bbdg.grad_(outs[0][1], outs[1][1], outs[2][1], N, xv, geo)
    xr = x.astype(dtype).astype(np.float64).reshape(-1, n)
    d = [orc.deriv_mul(a, N, xr).reshape(K, F, -1) for a in range(3)]
    g = geo_h[:, :, None, None]
    for c in range(3):
        ref = g[:, 3 * c + 0] * d[0] + g[:, 3 * c + 1] * d[1] + g[:, 3 * c + 2] * d[2]
        got = from_dev_cols(outs[c][1])
        assert rel(got, ref) <= tol, (N, F, layout, c, rel(got, ref))
    D = bbdg.BernsteinDerivativeMatrix_3D_s(N)
    bbdg.mul_(outs[0][1], D, xv)
    assert rel(from_dev_cols(outs[0][1]), d[1]) <= tol


@pytest.mark.parametrize("N", list(range(1, 13)) + [15, 17, 20])
def test_lift_fp64(ctx, orc, N):
    import bbdg

    rng = np.random.default_rng(300 + N)
    K, F = 5, 3
    x = rng.uniform(-1, 1, (K, F, 4 * nfp(N)))
    for layout in ("aos", "soa"):
        _, xv = to_dev_cols(x, layout, np.float64)
        _, ov = empty_cols(np3(N), K, F, layout, np.float64)
        bbdg.mul_(ov, bbdg.BernsteinLift(N), xv)
        ref = orc.lift(N, x.reshape(-1, 4 * nfp(N))).reshape(K, F, -1)
        assert rel(from_dev_cols(ov), ref) <= TOL64, (N, layout, rel(from_dev_cols(ov), ref))


@pytest.mark.parametrize("N", [1, 3, 7, 9, 12])
@pytest.mark.parametrize("F", [1, 4, 5, 6])
def test_lift_field_counts(ctx, orc, N, F):
    """field counts around the four-field batches of the one-wave-per-element kernel (F=5
    leaves a remainder of one), in both compact layouts and all precision modes"""
    import bbdg

    rng = np.random.default_rng(500 + 10 * N + F)
    K = 37
    x = rng.uniform(-1, 1, (K, F, 4 * nfp(N)))
    ref = orc.lift(N, x.reshape(-1, 4 * nfp(N))).reshape(K, F, -1)
    ref32 = orc.lift(N, x.astype(np.float32).astype(np.float64).reshape(-1, 4 * nfp(N))).reshape(K, F, -1)
    for layout in ("soa", "aos"):
        _, xv = to_dev_cols(x, layout, np.float64)
        _, ov = empty_cols(np3(N), K, F, layout, np.float64)
        bbdg.mul_(ov, bbdg.BernsteinLift(N), xv)
        assert rel(from_dev_cols(ov), ref) <= TOL64, (layout, rel(from_dev_cols(ov), ref))
        for mode in ("plain", "acc64", "acc64_mfma"):
            _, xv = to_dev_cols(x, layout, np.float32)
            _, ov = empty_cols(np3(N), K, F, layout, np.float32)
            bbdg.mul_(ov, bbdg.BernsteinLift(N, np.float32, mode), xv)
            assert rel(from_dev_cols(ov), ref32) <= 1e-5, (layout, mode, rel(from_dev_cols(ov), ref32))


@pytest.mark.parametrize("N", [2, 5, 7, 9])
@pytest.mark.parametrize("mode", ["plain", "acc64", "acc64_mfma"])
def test_lift_fp32(ctx, orc, N, mode):
    import bbdg

    rng = np.random.default_rng(N)
    K, F = 9, 2
    x = rng.uniform(-1, 1, (K, F, 4 * nfp(N)))
    _, xv = to_dev_cols(x, "soa", np.float32)
    _, ov = empty_cols(np3(N), K, F, "soa", np.float32)
    bbdg.mul_(ov, bbdg.BernsteinLift(N, np.float32, mode), xv)
    ref = orc.lift(N, x.astype(np.float32).astype(np.float64).reshape(-1, 4 * nfp(N))).reshape(K, F, -1)
    assert rel(from_dev_cols(ov), ref) <= 1e-5


@pytest.mark.parametrize("N", [1, 2, 5, 8, 11, 14, 17, 20])
def test_face_lift(ctx, orc, N):
    import bbdg

    rng = np.random.default_rng(400 + N)
    K = 6
    x = rng.standard_normal((K, 1, nfp(N)))
    ref = orc.face_lift(N, x.reshape(-1, nfp(N))).reshape(K, 1, -1)
    L = bbdg.BernsteinLift(N)
    _, xv = to_dev_cols(x, "soa", np.float64)
    _, ov = empty_cols(np3(N), K, 1, "soa", np.float64)
    bbdg.face_lift_multiply_(ov, L, xv)
    assert rel(from_dev_cols(ov), ref) <= TOL64
    # FP32 input, FP64 output: the chain precision variants
    ref32 = orc.face_lift(N, x.astype(np.float32).astype(np.float64).reshape(-1, nfp(N))).reshape(K, 1, -1)
    _, x32 = to_dev_cols(x, "soa", np.float32)
    errs = {}
    for mode in ("plain", "acc64", "twosum"):
        _, ov = empty_cols(np3(N), K, 1, "soa", np.float64)
        bbdg.face_lift_multiply_(ov, L, x32, mode=mode)
        errs[mode] = rel(from_dev_cols(ov), ref32)
    assert errs["plain"] <= 1e-5
    assert errs["acc64"] <= 1e-5 and errs["twosum"] <= 1e-5


@pytest.mark.parametrize("N", [8, 11, 14, 17])
def test_face_lift_fp64_conditioning(ctx, orc, N):
    """the FP64 single-face lift against the long-double restatement, on the inputs of the
    lift conditioning study: N(0,1) face coefficients (seed 4) and the same scaled per
    column by 10^U[-3,3]; the relative L2 error stays at FP64 rounding level"""
    import bbdg

    rng = np.random.default_rng(4)
    K, nf, n = 512, nfp(N), np3(N)
    base = rng.standard_normal((K, nf))
    scale = 10.0 ** rng.uniform(-3, 3, (K, 1))
    L = bbdg.BernsteinLift(N)
    for x in (base, base * scale):
        xd = bbdg.DeviceArray.from_numpy(x)
        od = bbdg.DeviceArray.zeros((K, n))
        bbdg.face_lift_multiply_(od.view((n, K), (1, n)), L, xd.view((nf, K), (1, nf)))
        ld = orc.face_lift_ld(N, x)
        err = rel(od.numpy(), ld)
        print(f"N={N} fp64 vs long double: {err:.3e}")
        assert err <= 1e-14, (N, err)


@pytest.mark.parametrize("N", [1, 2, 4, 7, 10, 20])
def test_elevation(ctx, orc, N):
    import bbdg

    rng = np.random.default_rng(N)
    K, F = 4, 2
    x = rng.uniform(-1, 1, (K, F, nfp(N - 1)))
    _, xv = to_dev_cols(x, "aos", np.float64)
    _, ov = empty_cols(nfp(N), K, F, "aos", np.float64)
    bbdg.mul_(ov, bbdg.ElevationMatrix(N), xv)
    ref = np.stack([[orc.elevation_multiply(N - 1, x[k, f]) for f in range(F)] for k in range(K)])
    assert rel(from_dev_cols(ov), ref) <= TOL64


def test_face_gather(ctx, orc):
    import bbdg

    N, K, F = 6, 5, 4
    rng = np.random.default_rng(0)
    q = rng.uniform(-1, 1, (K, F, np3(N)))
    _, qv = to_dev_cols(q, "aos", np.float64)
    _, uv = empty_cols(4 * nfp(N), K, F, "aos", np.float64)
    bbdg.face_gather_(uv, N, qv)
    fm = orc.fmask(N).reshape(-1)
    assert np.array_equal(from_dev_cols(uv), q[:, :, fm])


@pytest.mark.parametrize("pde", ["wave", "maxwell"])
@pytest.mark.parametrize("flux_kind", ["reference", "upwind"])
def test_interface_flux(ctx, orc, pde, flux_kind):
    import bbdg
    from bbdg import mesh

    N = 4
    F = 4 if pde == "wave" else 6
    periodic = pde == "wave"
    m = mesh.kuhn_mesh(2, 0.1, 3, periodic=periodic)
    sign = [1.0] * F if pde == "wave" else [1, 1, 1, -1, -1, -1]
    ma = orc.MeshArrays(m, N, sign)
    rng = np.random.default_rng(5)
    q = rng.uniform(-1, 1, (m.K, np3(N), F))
    uM = orc.face_gather(N, q)
    d_uM = bbdg.DeviceArray.from_numpy(uM)
    uMv = d_uM.view((4 * nfp(N), m.K, F), (F, 4 * nfp(N) * F, 1))
    d_fl = bbdg.DeviceArray.zeros(uM.shape)
    flv = d_fl.view(uMv.shape, uMv.strides)
    dev = {k: bbdg.DeviceArray.from_numpy(v) for k, v in
           dict(mapP=ma.mapP, nx=ma.nx, ny=ma.ny, nz=ma.nz, bc=ma.bc).items()}
    bbdg.interface_flux_(flv, pde, N, uMv, dev["mapP"], dev["nx"], dev["ny"], dev["nz"], bc=dev["bc"], bc_sign=sign,
                         flux_kind=flux_kind)
    ref = orc.flux(0 if pde == "wave" else 1, N, uM, ma, flux_kind=0 if flux_kind == "reference" else 1)
    assert rel(d_fl.numpy(), ref) <= TOL64


def test_errors_are_reported(ctx):
    import bbdg

    D = bbdg.BernsteinDerivativeMatrix_3D_r(3)
    x = bbdg.DeviceArray.zeros((20, 2))
    o = bbdg.DeviceArray.zeros((21, 2))
    with pytest.raises(ValueError):
        bbdg.mul_(o, D, x)
    lay = bbdg._lib.bb_layout(20, 1, 0)
    rc = bbdg.lib().bb_deriv_apply(ctx.h, 0, 21, 1, 2, 0, x.ptr, lay, x.ptr, lay)
    assert rc == bbdg._lib.BB_ERR_ARG and b"outside" in bbdg.lib().bb_last_error(ctx.h)


def test_empty_batch_is_a_no_op(ctx):
    """K = 0 columns (an empty batch): every operator entry point returns BB_OK and writes
    nothing (the output keeps its sentinel values)"""
    import bbdg
    from bbdg import _lib as B

    lib = bbdg.lib()
    N, F = 4, 2
    n, nf = np3(N), nfp(N)
    x = bbdg.DeviceArray.from_numpy(np.ones(4 * F * max(n, 4 * nf)))
    geo = bbdg.DeviceArray.from_numpy(np.ones(9))
    sentinel = np.full(F * n, 7.0)
    outs = [bbdg.DeviceArray.from_numpy(sentinel.copy()) for _ in range(3)]
    vol = B.bb_layout(F * n, 1, n)
    face4 = B.bb_layout(F * 4 * nf, 1, 4 * nf)
    face1 = B.bb_layout(F * nf, 1, nf)
    for dt in (B.BB_F64, B.BB_F32):
        calls = {
            "deriv": lambda: lib.bb_deriv_apply(ctx.h, 0, N, F, 0, dt, x.ptr, vol, outs[0].ptr, vol),
            "grad": lambda: lib.bb_grad_apply(ctx.h, N, F, 0, dt, x.ptr, vol, geo.ptr, outs[0].ptr, outs[1].ptr,
                                              outs[2].ptr, vol),
            "lift": lambda: lib.bb_lift_apply(ctx.h, N, F, 0, dt, 0, x.ptr, face4, outs[0].ptr, vol),
            "face_lift": lambda: lib.bb_face_lift_apply(ctx.h, N, F, 0, dt, dt, 0, x.ptr, face1, outs[0].ptr, vol),
            "elevate": lambda: lib.bb_elevate_apply(ctx.h, N, F, 0, dt, x.ptr, face1, outs[0].ptr, face1),
            "gather": lambda: lib.bb_face_gather(ctx.h, N, F, 0, dt, x.ptr, vol, outs[0].ptr, face4),
        }
        for name, call in calls.items():
            rc = call()
            assert rc == B.BB_OK, (name, dt, rc, lib.bb_last_error(ctx.h))
    ctx.sync()
    for o in outs:
        assert np.array_equal(o.numpy(), sentinel)
