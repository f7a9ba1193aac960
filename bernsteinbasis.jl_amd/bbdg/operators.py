"""The reference's operator API (src/BernsteinBasis.jl:19-23) over the C ABI.

Julia                                            here
-----                                            ----
BernsteinDerivativeMatrix_3D_{r,s,t}(N)          same names; ``size``, ``D[m, n]`` (0-based)
mul!(out, D, x)                                  mul_(out, D, x)
BernsteinLift{T}(N), MultithreadedBernsteinLift  BernsteinLift(N, dtype), MultithreadedBernsteinLift(N, nthreads)
mul!(out, L, x), threaded_mul!(out, L, x, tid)   mul_(out, L, x), threaded_mul_(out, L, x, tid)
face_lift_multiply!(out, L, x)                   face_lift_multiply_(out, L, x)
ElevationMatrix{N}()                             ElevationMatrix(N); ``E[m, n]``, mul_(out, E, x)
tri_offsets, tet_offsets, l_j, ...               same names (bit-exact tables)

Arrays are ``DeviceArray``s shaped like the Julia arrays: a vector (n,), a matrix of
columns (n, ncols), or a matrix of ``SVector{F}`` columns (n, ncols, F).  Any strides are
accepted, so Julia's column-major SVector layout (strides (F, n F, 1)) works without a copy.
Errors raise ``BBError`` (the reference does no checking and has undefined behaviour on a
size mismatch).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from .device import DeviceArray

# ----------------------------------------------------------------------- tables


def _tables(N):
    tri = np.zeros(21, np.int64)
    tet = np.zeros(21, np.int64)
    lj = np.zeros(20)
    rc = B.lib().bb_tables(N, tri.ctypes.data_as(C.c_void_p), tet.ctypes.data_as(C.c_void_p),
                           lj.ctypes.data_as(C.c_void_p))
    if rc:
        raise B.BBError(rc, f"N={N}")
    return tri, tet, lj


def tri_offsets(N):
    """src/ordering.jl:8-16"""
    return tuple(int(v) for v in _tables(N)[0])


def tet_offsets(N):
    """src/ordering.jl:18-26"""
    return tuple(int(v) for v in _tables(N)[1])


def l_j(N):
    """src/constants.jl:10-12"""
    return tuple(float(v) for v in _tables(N)[2])


def ij_to_linear(i, j, tri_offset):
    """src/ordering.jl:6 (1-based result, like the reference)"""
    return i + tri_offset[j] + 1


def ijk_to_linear(i, j, k, tri_offset, tet_offset):
    """src/ordering.jl:33-35 (1-based result)"""
    return i + tri_offset[j] + 1 + tet_offset[k] - j * k


def linear_to_ijkl_lookup(N):
    """src/ordering.jl:37-39"""
    n = (N + 1) * (N + 2) * (N + 3) // 6
    out = np.zeros(4 * n, np.int32)
    rc = B.lib().bb_linear_to_ijkl(N, out.ctypes.data_as(C.c_void_p))
    if rc:
        raise B.BBError(rc, f"N={N}")
    return [tuple(int(v) for v in r) for r in out.reshape(-1, 4)]


def bernstein_2d_scalar_multiindex_lookup(N):
    """src/ordering.jl:47-51 (1-based scalar indices in the dict, like the reference)"""
    n = (N + 1) * (N + 2) // 2
    out = np.zeros(3 * n, np.int32)
    rc = B.lib().bb_linear_to_ijk_2d(N, out.ctypes.data_as(C.c_void_p))
    if rc:
        raise B.BBError(rc, f"N={N}")
    s2m = [tuple(int(v) for v in r) for r in out.reshape(-1, 3)]
    return s2m, {m: i + 1 for i, m in enumerate(s2m)}


def fmask(N):
    """Face-trace index table (4, Nfp), 0-based (plays the role of rd.Fmask)."""
    n = (N + 1) * (N + 2) // 2
    out = np.zeros(4 * n, np.int32)
    rc = B.lib().bb_fmask(N, out.ctypes.data_as(C.c_void_p))
    if rc:
        raise B.BBError(rc, f"N={N}")
    return out.reshape(4, n)


# ----------------------------------------------------------------------- operator types

def _np3(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


def _nfp(N):
    return (N + 1) * (N + 2) // 2


class _Deriv:
    axis = 0

    def __init__(self, N):
        if not 1 <= N <= 20:
            raise ValueError("Supports 1 <= N <= 20")
        self.N = N
        self.tri_offsets = tri_offsets(N)
        self.tet_offsets = tet_offsets(N)

    @property
    def size(self):
        n = _np3(self.N)
        return (n, n)

    def __getitem__(self, mn):
        m, n = mn
        v = C.c_double()
        rc = B.lib().bb_deriv_entry(self.axis, self.N, int(m), int(n), C.byref(v))
        if rc:
            raise IndexError(f"({m},{n}) outside {self.size}")
        return v.value

    def dense(self):
        n = _np3(self.N)
        return np.array([[self[i, j] for j in range(n)] for i in range(n)])


class BernsteinDerivativeMatrix_3D_r(_Deriv):
    """src/BernsteinDerivativeMatrix_3D.jl:9-14"""
    axis = 0


class BernsteinDerivativeMatrix_3D_s(_Deriv):
    """src/BernsteinDerivativeMatrix_3D.jl:24-29"""
    axis = 1


class BernsteinDerivativeMatrix_3D_t(_Deriv):
    """src/BernsteinDerivativeMatrix_3D.jl:39-44"""
    axis = 2


class BernsteinLift:
    """src/BernsteinLift.jl:12-27.  ``dtype`` float64 or float32; ``mode`` selects the FP32
    arithmetic ('plain'; 'acc64': FP32 chain with FP64 weighted accumulation; 'acc64_mfma': as
    'acc64' with L0 on the matrix cores where that kernel exists).  No
    host scratch: the scratch vectors of the reference live in LDS."""

    def __init__(self, N, dtype=np.float64, mode="acc64"):
        if not 1 <= N <= 20:
            raise ValueError("Supports 1 <= N <= 20")
        self.N = N
        self.dtype = np.dtype(dtype)
        self.mode = {"plain": B.BB_LIFT_PLAIN, "acc64": B.BB_LIFT_ACC64, "twosum": B.BB_LIFT_TWOSUM,
                     "acc64_mfma": B.BB_LIFT_ACC64_MFMA}[mode]

    @property
    def size(self):
        return (_np3(self.N), 4 * _nfp(self.N))


class MultithreadedBernsteinLift(BernsteinLift):
    """src/BernsteinLift.jl:40-56.  On the GPU every column is its own work item, so the
    per-thread scratch columns of the reference are not needed; ``threads`` is kept for
    signature compatibility."""

    def __init__(self, N, threads=1, dtype=np.float64, mode="acc64"):
        super().__init__(N, dtype, mode)
        self.threads = threads


class ElevationMatrix:
    """src/ElevationMatrix.jl:11-24: degree N-1 -> N on a triangle."""

    def __init__(self, N):
        if not 1 <= N <= 20:
            raise ValueError("Supports 1 <= N <= 20")
        self.N = N

    @property
    def size(self):
        return (_nfp(self.N), _nfp(self.N - 1))

    def __getitem__(self, mn):
        m, n = mn
        v = C.c_double()
        rc = B.lib().bb_elevation_entry(self.N, int(m), int(n), C.byref(v))
        if rc:
            raise IndexError(f"({m},{n}) outside {self.size}")
        return v.value

    def dense(self):
        r, c = self.size
        return np.array([[self[i, j] for j in range(c)] for i in range(r)])


# ----------------------------------------------------------------------- mul!

def _cols(a: DeviceArray, nrows: int, what: str):
    """(layout, ncols, F) of a Julia-shaped array of columns of length nrows."""
    if a.shape[0] != nrows:
        raise ValueError(f"{what}: first dimension {a.shape[0]} != {nrows}")
    if a.ndim == 1:
        return B.bb_layout(0, a.strides[0], 0), 1, 1
    if a.ndim == 2:
        return B.bb_layout(a.strides[1], a.strides[0], 0), a.shape[1], 1
    if a.ndim == 3:
        return B.bb_layout(a.strides[1], a.strides[0], a.strides[2]), a.shape[1], a.shape[2]
    raise ValueError(f"{what}: expected 1-3 dimensions")


def _same(x: DeviceArray, out: DeviceArray, kx, Fx, ko, Fo):
    if (kx, Fx) != (ko, Fo):
        raise ValueError(f"column count mismatch: x has {kx}x{Fx}, out has {ko}x{Fo}")
    if x.dtype != out.dtype:
        raise ValueError("x and out must have the same dtype")


def mul_(out: DeviceArray, op, x: DeviceArray):
    """mul!(out, op, x): derivative (src/BernsteinDerivativeMatrix_3D.jl:134-231), lift
    (src/BernsteinLift.jl:191-269) or elevation (src/ElevationMatrix.jl)."""
    ctx = x.ctx
    L = B.lib()
    if isinstance(op, _Deriv):
        n = _np3(op.N)
        lx, kx, Fx = _cols(x, n, "x")
        lo, ko, Fo = _cols(out, n, "out")
        _same(x, out, kx, Fx, ko, Fo)
        ctx.check(L.bb_deriv_apply(ctx.h, op.axis, op.N, Fx, kx, x.code, C.c_void_p(x.ptr), lx, C.c_void_p(out.ptr), lo))
    elif isinstance(op, BernsteinLift):
        lx, kx, Fx = _cols(x, 4 * _nfp(op.N), "x")
        lo, ko, Fo = _cols(out, _np3(op.N), "out")
        _same(x, out, kx, Fx, ko, Fo)
        ctx.check(L.bb_lift_apply(ctx.h, op.N, Fx, kx, x.code, op.mode, C.c_void_p(x.ptr), lx, C.c_void_p(out.ptr), lo))
    elif isinstance(op, ElevationMatrix):
        lx, kx, Fx = _cols(x, _nfp(op.N - 1), "x")
        lo, ko, Fo = _cols(out, _nfp(op.N), "out")
        _same(x, out, kx, Fx, ko, Fo)
        ctx.check(L.bb_elevate_apply(ctx.h, op.N, Fx, kx, x.code, C.c_void_p(x.ptr), lx, C.c_void_p(out.ptr), lo))
    else:
        raise TypeError(f"mul_ does not know {type(op).__name__}")
    return out


def threaded_mul_(out: DeviceArray, op: MultithreadedBernsteinLift, x: DeviceArray, thread: int = 1):
    """threaded_mul!(out, L, x, thread), src/BernsteinLift.jl:233-262"""
    return mul_(out, op, x)


def face_lift_multiply_(out: DeviceArray, op: BernsteinLift, x: DeviceArray, out_dtype=None, mode=None):
    """face_lift_multiply!(out, L, x), src/BernsteinLift.jl:139-179 (one face, Nfp -> Np)."""
    ctx = x.ctx
    lx, kx, Fx = _cols(x, _nfp(op.N), "x")
    lo, ko, Fo = _cols(out, _np3(op.N), "out")
    if (kx, Fx) != (ko, Fo):
        raise ValueError("column count mismatch")
    m = op.mode if mode is None else {"plain": 0, "acc64": 1, "twosum": 2}[mode]
    ctx.check(B.lib().bb_face_lift_apply(ctx.h, op.N, Fx, kx, x.code, out.code, m, C.c_void_p(x.ptr), lx,
                                         C.c_void_p(out.ptr), lo))
    return out


def grad_(dx: DeviceArray, dy: DeviceArray, dz: DeviceArray, N: int, x: DeviceArray, geo: DeviceArray):
    """Physical gradient: the three derivative matvecs with per-element rxJ..tzJ
    (examples/3d_wave.jl:63-71, :75-78).  geo: (K, 9) float64 device array."""
    ctx = x.ctx
    n = _np3(N)
    lx, kx, Fx = _cols(x, n, "x")
    lo, ko, Fo = _cols(dx, n, "dx")
    for d in (dy, dz):
        if d.shape != dx.shape or d.strides != dx.strides or d.dtype != dx.dtype:
            raise ValueError("dx, dy, dz must share shape, strides and dtype")
    _same(x, dx, kx, Fx, ko, Fo)
    if geo.shape != (kx, 9) or geo.dtype != np.float64:
        raise ValueError("geo must be (K, 9) float64")
    ctx.check(B.lib().bb_grad_apply(ctx.h, N, Fx, kx, x.code, C.c_void_p(x.ptr), lx, C.c_void_p(geo.ptr),
                                    C.c_void_p(dx.ptr), C.c_void_p(dy.ptr), C.c_void_p(dz.ptr), lo))
    return dx, dy, dz


def face_gather_(uM: DeviceArray, N: int, q: DeviceArray):
    """uM .= q[Fmask, :] (examples/3d_wave.jl:41)"""
    ctx = q.ctx
    lq, kq, Fq = _cols(q, _np3(N), "q")
    lm, km, Fm = _cols(uM, 4 * _nfp(N), "uM")
    _same(q, uM, kq, Fq, km, Fm)
    ctx.check(B.lib().bb_face_gather(ctx.h, N, Fq, kq, q.code, C.c_void_p(q.ptr), lq, C.c_void_p(uM.ptr), lm))
    return uM


def interface_flux_(flux: DeviceArray, pde: str, N: int, uM: DeviceArray, mapP: DeviceArray, nx: DeviceArray,
                    ny: DeviceArray, nz: DeviceArray, bc: DeviceArray | None = None, bc_sign=None,
                    flux_kind: str = "reference"):
    """The interface-flux loop of examples/3d_wave.jl:43-56 / 3d_maxwell.jl:43-65 (FP64)."""
    ctx = uM.ctx
    p = {"wave": B.BB_PDE_WAVE, "maxwell": B.BB_PDE_MAXWELL}[pde]
    F = 4 if p == B.BB_PDE_WAVE else 6
    lm, km, Fm = _cols(uM, 4 * _nfp(N), "uM")
    lf, kf, Ff = _cols(flux, 4 * _nfp(N), "flux")
    if Fm != F or (km, Fm) != (kf, Ff):
        raise ValueError("uM/flux shape")
    if uM.dtype != np.float64 or flux.dtype != np.float64:
        raise TypeError("the interface flux is FP64")
    M = 4 * _nfp(N)
    for name, arr, dt in (("mapP", mapP, np.int32), ("nx", nx, np.float64), ("ny", ny, np.float64),
                          ("nz", nz, np.float64)) + ((("bc", bc, np.int8),) if bc is not None else ()):
        if arr.dtype != np.dtype(dt) or arr.shape != (km, M) or arr.strides != (M, 1):
            raise TypeError(f"{name} must be a C-contiguous ({km}, {M}) {np.dtype(dt).name} array")
    sign = None
    if bc is not None:
        sign = np.ascontiguousarray(np.asarray(bc_sign if bc_sign is not None else np.ones(F), np.float64))
    k = {"reference": B.BB_FLUX_REFERENCE, "upwind": B.BB_FLUX_UPWIND}[flux_kind]
    ctx.check(B.lib().bb_interface_flux(
        ctx.h, p, k, N, km, C.c_void_p(uM.ptr), lm, C.c_void_p(mapP.ptr),
        C.c_void_p(bc.ptr) if bc is not None else None, sign.ctypes.data_as(C.c_void_p) if sign is not None else None,
        C.c_void_p(nx.ptr), C.c_void_p(ny.ptr), C.c_void_p(nz.ptr), C.c_void_p(flux.ptr), lf))
    return flux
