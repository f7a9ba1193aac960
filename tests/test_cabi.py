"""The C-ABI library on the CPU: it loads, exports every function include/bbdg.h declares,
and its host-side tables and dense entries agree bit for bit with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "bbdg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bb_\w+)\s*\(", src)))


def test_exports_every_declared_symbol():
    import bbdg

    L = bbdg.lib()
    names = declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding knows every declared signature
    from bbdg import _lib

    assert set(names) == set(_lib.SIGS), set(names) ^ set(_lib.SIGS)


@pytest.mark.parametrize("N", range(1, 21))
def test_host_tables_bit_exact(orc, N):
    import bbdg

    assert bbdg.tri_offsets(N) == tuple(orc.tri_offsets(N))
    assert bbdg.tet_offsets(N) == tuple(orc.tet_offsets(N))
    assert bbdg.l_j(N) == tuple(orc.l_j(N))
    assert bbdg.linear_to_ijkl_lookup(N) == [tuple(r) for r in orc.linear_to_ijkl(N)]
    assert np.array_equal(bbdg.fmask(N), orc.fmask(N))
    s2m, m2s = bbdg.bernstein_2d_scalar_multiindex_lookup(N)
    assert s2m == [(i, j, N - i - j) for j in range(N + 1) for i in range(N + 1 - j)]
    assert all(m2s[m] == k + 1 for k, m in enumerate(s2m))
    t = bbdg.tri_offsets(N)
    assert bbdg.ij_to_linear(1, 1, t) == orc.lib().orc_ij_to_linear(1, 1, orc.tri_offsets(N))


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7])
def test_dense_entries_bit_exact(orc, N):
    import bbdg

    for axis, D in enumerate([bbdg.BernsteinDerivativeMatrix_3D_r(N), bbdg.BernsteinDerivativeMatrix_3D_s(N),
                              bbdg.BernsteinDerivativeMatrix_3D_t(N)]):
        assert D.size == (orc.Np(N), orc.Np(N))
        assert np.array_equal(D.dense(), orc.deriv_dense(axis, N))
    E = bbdg.ElevationMatrix(N)
    assert np.array_equal(E.dense(), orc.elev_dense(N))


def test_argument_errors_need_no_gpu():
    import bbdg
    from bbdg import _lib as B

    L = bbdg.lib()
    tri = np.zeros(21, np.int64)
    assert L.bb_tables(0, tri.ctypes.data_as(C.c_void_p), None, None) == B.BB_ERR_ARG
    assert L.bb_tables(21, tri.ctypes.data_as(C.c_void_p), None, None) == B.BB_ERR_ARG
    v = C.c_double()
    assert L.bb_deriv_entry(3, 4, 0, 0, C.byref(v)) == B.BB_ERR_ARG
    assert L.bb_deriv_entry(0, 4, 35, 0, C.byref(v)) == B.BB_ERR_ARG
    assert L.bb_deriv_entry(0, 4, 0, 0, None) == B.BB_ERR_NULL
    assert L.bb_elevation_entry(4, 15, 0, C.byref(v)) == B.BB_ERR_ARG
    lay = B.bb_layout(1, 1, 1)
    assert L.bb_deriv_apply(None, 0, 3, 1, 1, 0, None, lay, None, lay) == B.BB_ERR_NULL
    with pytest.raises(IndexError):
        bbdg.BernsteinDerivativeMatrix_3D_r(2)[10, 0]
    with pytest.raises(ValueError):
        bbdg.BernsteinLift(21)


def test_uploads_keep_integer_dtypes():
    """index tables (mapP int32, bc int8) must reach the device with their own dtype"""
    from bbdg.device import host_buffer

    for dt in (np.int32, np.int8, np.float32, np.float64):
        a = np.arange(12, dtype=dt).reshape(3, 4)
        buf, dtype, strides, shape = host_buffer(a)
        assert dtype == np.dtype(dt) and buf.dtype == np.dtype(dt) and strides == (4, 1) and shape == (3, 4)
        buf, dtype, strides, shape = host_buffer(np.asfortranarray(a))
        assert strides == (1, 3) and np.array_equal(buf, a.T.reshape(-1))
    with pytest.raises(TypeError):
        host_buffer(np.zeros(3, np.complex128))
