"""Failure detection through the C ABI: NaN / Inf in an input is refused with
BB_ERR_NONFINITE, and the device scan finds non-finite values in arrays and in the state of
a diverged run.  (The reference checks nothing; these are the build's own checks.)"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 255, 257, 70001, 3_000_001])
def test_device_scan_counts(ctx, dtype, n):
    import bbdg

    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(dtype)
    assert bbdg.DeviceArray.from_numpy(x, ctx=ctx).count_nonfinite() == 0
    k = min(n, 37)
    pos = rng.choice(n, k, replace=False)
    x[pos] = rng.choice(np.array([np.nan, np.inf, -np.inf], dtype), k)
    if n > 2:
        x[-1] = np.nan  # the last value of a ragged tail
    d = bbdg.DeviceArray.from_numpy(x, ctx=ctx)
    assert d.count_nonfinite() == int((~np.isfinite(x)).sum())
    d.free()


def test_scan_status_codes(ctx):
    import bbdg
    from bbdg import _lib as B

    x = np.array([1.0, np.inf, 2.0, np.nan])
    d = bbdg.DeviceArray.from_numpy(x, ctx=ctx)
    n = C.c_int64(-1)
    L = B.lib()
    assert L.bb_count_nonfinite(ctx.h, B.BB_F64, C.c_void_p(d.ptr), 4, C.byref(n)) == B.BB_ERR_NONFINITE
    assert n.value == 2 and "2 of 4" in L.bb_last_error(ctx.h).decode()
    assert L.bb_count_nonfinite(ctx.h, B.BB_F64, C.c_void_p(d.ptr), 1, C.byref(n)) == B.BB_OK and n.value == 0
    assert L.bb_count_nonfinite(ctx.h, B.BB_F64, C.c_void_p(d.ptr), 0, C.byref(n)) == B.BB_OK and n.value == 0
    assert L.bb_count_nonfinite(ctx.h, 9, C.c_void_p(d.ptr), 4, C.byref(n)) == B.BB_ERR_ARG
    assert L.bb_count_nonfinite(ctx.h, B.BB_F64, C.c_void_p(d.ptr), -1, C.byref(n)) == B.BB_ERR_ARG
    assert L.bb_count_nonfinite(ctx.h, B.BB_F64, None, 4, C.byref(n)) == B.BB_ERR_NULL
    d.free()


def _solver(ctx, N=3, n=2):
    import bbdg
    from bbdg import basis, mesh

    m = mesh.kuhn_mesh(n, 0.1, 0, periodic=True)
    q0 = basis.interpolate(m, N, lambda x, y, z: np.stack(
        [np.exp(-25 * ((x - .5) ** 2 + (y - .5) ** 2 + (z - .5) ** 2)), 0 * x, 0 * x, 0 * x]))
    s = bbdg.Solver(m, N, "wave", ctx=ctx)
    s.set_state(q0)
    return m, q0, s


def test_solver_refuses_nonfinite_state(ctx):
    from bbdg import _lib as B

    m, q0, s = _solver(ctx)
    try:
        bad = q0.copy()
        bad[5, 2, 0] = np.nan
        with pytest.raises(B.BBError) as e:
            s.set_state(bad)
        assert e.value.code == B.BB_ERR_NONFINITE
        # the refused state was not copied
        np.testing.assert_array_equal(s.get_state(), q0)
        assert s.check_state() == 0
    finally:
        s.close()


def test_solver_state_scan_finds_injected_and_diverged_values(ctx):
    import bbdg
    from bbdg import _lib as B

    m, q0, s = _solver(ctx)
    try:
        p = C.c_void_p()
        ctx.check(B.lib().bb_solver_state_ptr(s.h, C.byref(p)))
        one = np.array([np.inf])
        ctx.check(B.lib().bb_memcpy_h2d(ctx.h, C.c_void_p(p.value + 8 * 17), one.ctypes.data_as(C.c_void_p), 8))
        assert s.check_state() == 1
        n = C.c_int64()
        assert B.lib().bb_solver_check_state(s.h, C.byref(n)) == B.BB_ERR_NONFINITE and n.value == 1
        s.set_state(q0)
        # a step 200x beyond the stable one diverges: the scan reports it, the healthy run does not
        dt = 0.5 * m.h / 3 ** 2
        s.step(dt, 20, "fused", "f64")
        assert s.check_state() == 0
        s.set_state(q0)
        s.step(200 * dt, 200, "fused", "f64")
        assert s.check_state() > 0
    finally:
        s.close()


def test_batch_refuses_and_scans(ctx):
    from bbdg import _lib as B
    from bbdg import mesh
    from bbdg.batch import Batch, BatchSolver

    meshes = [mesh.kuhn_mesh(2, 0.1, 7, periodic=True)]
    rng = np.random.default_rng(7)
    b = Batch(meshes, [rng.integers(2, 5, meshes[0].K)], [0.5 * meshes[0].h / 16])
    q = b.interpolate([lambda x, y, z: np.stack([np.exp(-25 * ((x - .5) ** 2 + (y - .5) ** 2)), 0 * x, 0 * x,
                                                 0 * x])])
    s = BatchSolver(b)
    try:
        s.set_state(q)
        assert s.check_state() == 0
        bad = q.copy()
        bad[-1] = -np.inf
        with pytest.raises(B.BBError) as e:
            s.set_state(bad)
        assert e.value.code == B.BB_ERR_NONFINITE
        assert s.check_state() == 0
    finally:
        s.close()
