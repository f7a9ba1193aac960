"""CPU oracle: a plain-C Float64 restatement of the reference's operators (bbref.c).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this module, and only as the checker
(or as the timed CPU baseline).  The product library never calls it.

``build()`` compiles ``bbref.c`` with gcc into ``oracle/_build/liboracle.so``; ``lib()``
loads it (building first only when nothing is built).  The checker is built without
floating-point contraction (-ffp-contract=off), so every product and sum rounds as written.

``baseline_lib()`` is the same source built the way a CPU baseline of the reference's
threaded path is built: -O3 -march=native (contraction allowed), OpenMP.  -march=native
targets the machine that compiles it, so this variant is compiled on first use on the
machine that runs it, into a file keyed by that machine's CPU model and the flags.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(HERE, "_build")
SO = os.path.join(BUILD_DIR, "liboracle.so")
SRC = os.path.join(HERE, "bbref.c")
CHECK_FLAGS = ["-O2", "-ffp-contract=off"]
# (ISO C mode turns contraction off by default; the baseline asks for it explicitly)
BASELINE_FLAGS = ["-O3", "-march=native", "-ffp-contract=fast"]

_lib = None
_baseline = None


def _compile(out: str, flags: list, extra: list | None = None) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    tmp = f"{out}.tmp{os.getpid()}"
    cmd = ["gcc", *flags, "-fPIC", "-shared", "-fopenmp", "-std=c11", *(extra or []), "-o", tmp, SRC, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build() -> str:
    return _compile(SO, CHECK_FLAGS)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def baseline_path() -> str:
    key = hashlib.sha1((cpu_model() + " ".join(BASELINE_FLAGS)).encode()).hexdigest()[:12]
    return os.path.join(BUILD_DIR, f"liboracle_baseline_{key}.so")


def baseline_lib():
    """the oracle compiled -O3 -march=native for this machine (timed CPU baseline only)"""
    global _baseline
    if _baseline is None:
        path = baseline_path()
        if not os.path.exists(path):
            _compile(path, BASELINE_FLAGS)
        _baseline = _bind(C.CDLL(path))
    return _baseline


def build_sanitized(out: str) -> str:
    """the checker built with the address and undefined-behaviour sanitizers (host tests)"""
    return _compile(out, CHECK_FLAGS + ["-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                                        "-fno-omit-frame-pointer"])


def lib():
    """the checker build; BB_ORACLE_SO names another build of bbref.c to load instead (the
    sanitizer run of the oracle tests)"""
    global _lib
    if _lib is None:
        path = os.environ.get("BB_ORACLE_SO") or SO
        if path == SO and not os.path.exists(SO):
            build()
        _lib = _bind(C.CDLL(path))
    return _lib


def _bind(L):
    i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.orc_np.argtypes = [C.c_int]
    L.orc_nfp.argtypes = [C.c_int]
    L.orc_tri_offsets.argtypes = [C.c_int, i64p]
    L.orc_tet_offsets.argtypes = [C.c_int, i64p]
    L.orc_ij_to_linear.argtypes = [C.c_int, C.c_int, i64p]
    L.orc_ij_to_linear.restype = C.c_int64
    L.orc_ijk_to_linear.argtypes = [C.c_int, C.c_int, C.c_int, i64p, i64p]
    L.orc_ijk_to_linear.restype = C.c_int64
    L.orc_linear_to_ijkl.argtypes = [C.c_int, i32p]
    L.orc_linear_to_ijk_2d.argtypes = [C.c_int, i32p]
    L.orc_l_j.argtypes = [C.c_int, dp]
    L.orc_deriv_entry.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_deriv_entry.restype = C.c_double
    L.orc_deriv_mul.argtypes = [C.c_int, C.c_int, dp, dp]
    L.orc_deriv_mul_cols.argtypes = [C.c_int, C.c_int, C.c_int64, dp, dp]
    L.orc_elev_entry.argtypes = [C.c_int, C.c_int, C.c_int]
    L.orc_elev_entry.restype = C.c_double
    L.orc_elevation_multiply.argtypes = [C.c_int, dp, dp]
    L.orc_reduction_multiply.argtypes = [C.c_int, dp, dp]
    L.orc_face_lift.argtypes = [C.c_int, dp, dp, dp, dp]
    L.orc_lift.argtypes = [C.c_int, dp, dp, dp]
    L.orc_lift_scratch_len.argtypes = [C.c_int]
    L.orc_lift_cols.argtypes = [C.c_int, C.c_int64, dp, dp]
    L.orc_face_lift_cols.argtypes = [C.c_int, C.c_int64, dp, dp]
    L.orc_fmask.argtypes = [C.c_int, i32p]
    L.orc_face_gather.argtypes = [C.c_int, C.c_int, C.c_int64, dp, dp]
    i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
    L.orc_rhs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, dp, i32p, i8p, dp, dp, dp, dp, dp, dp, dp, C.c_int]
    L.orc_lsrk_step.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, dp, dp, C.c_double, i32p, i8p, dp, dp, dp, dp,
                                dp, dp, C.c_int]
    L.orc_lsrk_coeffs.argtypes = [dp, dp]
    L.orc_face_lift_ld_cols.argtypes = [C.c_int, C.c_int64, dp, dp]
    L.orc_flux.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, dp, i32p, i8p, dp, dp, dp, dp, dp]
    L.orc_rhs_md.argtypes = [C.c_int, C.c_int, C.c_int64, i32p, i64p, dp, i32p, i32p, i32p, i32p, dp, dp, dp, dp,
                             dp, i64p, dp, C.c_int]
    L.orc_lsrk_step_md.argtypes = [C.c_int, C.c_int, C.c_int64, i32p, i64p, dp, dp, dp, i32p, i32p, i32p, i32p,
                                   dp, dp, dp, dp, dp, i64p, C.c_int]
    return L


# ----------------------------------------------------------------- thin wrappers

def Np(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


def Nfp(N):
    return (N + 1) * (N + 2) // 2


def tri_offsets(N):
    o = np.zeros(21, np.int64)
    lib().orc_tri_offsets(N, o)
    return o


def tet_offsets(N):
    o = np.zeros(21, np.int64)
    lib().orc_tet_offsets(N, o)
    return o


def l_j(N):
    o = np.zeros(20)
    lib().orc_l_j(N, o)
    return o


def linear_to_ijkl(N):
    o = np.zeros(4 * Np(N), np.int32)
    lib().orc_linear_to_ijkl(N, o)
    return o.reshape(-1, 4)


def fmask(N):
    o = np.zeros(4 * Nfp(N), np.int32)
    lib().orc_fmask(N, o)
    return o.reshape(4, -1)


def deriv_dense(axis, N):
    n = Np(N)
    L = lib()
    return np.array([[L.orc_deriv_entry(axis, N, m, k) for k in range(n)] for m in range(n)])


def deriv_mul(axis, N, x):
    """x: (ncols, Np) C-order == Julia Np x ncols column-major"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    ncols = x.size // Np(N)
    lib().orc_deriv_mul_cols(axis, N, ncols, x.reshape(-1), out.reshape(-1))
    return out


def elev_dense(N):
    L = lib()
    return np.array([[L.orc_elev_entry(N, m, k) for k in range(Nfp(N - 1))] for m in range(Nfp(N))])


def elevation_multiply(N, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(Nfp(N + 1))
    lib().orc_elevation_multiply(N, x, out)
    return out


def reduction_multiply(N, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(Nfp(N - 1))
    lib().orc_reduction_multiply(N, x, out)
    return out


def face_lift(N, x):
    """x: (ncols, Nfp) -> (ncols, Np)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    ncols = x.size // Nfp(N)
    out = np.zeros((ncols, Np(N)))
    lib().orc_face_lift_cols(N, ncols, x.reshape(-1), out.reshape(-1))
    return out


def face_lift_ld(N, x):
    """face lift evaluated in long double, rounded to double at the end: (ncols, Nfp) -> (ncols, Np)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    ncols = x.size // Nfp(N)
    out = np.zeros((ncols, Np(N)))
    lib().orc_face_lift_ld_cols(N, ncols, x.reshape(-1), out.reshape(-1))
    return out


def lift(N, x):
    """x: (ncols, 4Nfp) -> (ncols, Np)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    ncols = x.size // (4 * Nfp(N))
    out = np.zeros((ncols, Np(N)))
    lib().orc_lift_cols(N, ncols, x.reshape(-1), out.reshape(-1))
    return out


def lsrk_coeffs():
    a, b = np.zeros(5), np.zeros(5)
    lib().orc_lsrk_coeffs(a, b)
    return a, b


class MeshArrays:
    """StartUpDG-style per-face-node arrays for one mesh at degree N (oracle input)."""

    def __init__(self, mesh, N, bc_sign):
        from bbdg import mesh as M  # host setup code, not the compute path

        self.N = N
        self.K = mesh.K
        self.mapP, self.bc = M.map_p(mesh, N)
        self.nx, self.ny, self.nz = M.face_node_normals(mesh, N)
        self.geo = np.ascontiguousarray(mesh.geo)
        self.J = np.ascontiguousarray(mesh.J)
        self.bc_sign = np.ascontiguousarray(np.asarray(bc_sign, dtype=np.float64))


def rhs(pde, N, q_aos, ma: MeshArrays, nthreads=1, flux_kind=1, L=None):
    """q_aos: (K, Np, F) -> du (K, Np, F)  (examples/3d_wave.jl:36-84 / 3d_maxwell.jl:36-93).
    L: the loaded oracle to run (default the checker build; baseline_lib() for timing)"""
    q = np.ascontiguousarray(q_aos, dtype=np.float64)
    du = np.zeros_like(q)
    (L or lib()).orc_rhs(pde, flux_kind, N, ma.K, q.reshape(-1), ma.mapP.reshape(-1), ma.bc.reshape(-1), ma.bc_sign,
                  ma.nx.reshape(-1), ma.ny.reshape(-1), ma.nz.reshape(-1), ma.geo.reshape(-1), ma.J, du.reshape(-1),
                  nthreads)
    return du


def face_gather(N, q_aos):
    """uM = q[Fmask, :] on (K, Np, F) -> (K, 4Nfp, F)"""
    q = np.ascontiguousarray(q_aos, dtype=np.float64)
    K, _, F = q.shape
    uM = np.zeros((K, 4 * Nfp(N), F))
    lib().orc_face_gather(N, F, K, q.reshape(-1), uM.reshape(-1))
    return uM


def flux(pde, N, uM, ma: MeshArrays, flux_kind=1):
    """interface flux on (K, 4Nfp, F)"""
    uM = np.ascontiguousarray(uM, dtype=np.float64)
    out = np.zeros_like(uM)
    lib().orc_flux(pde, flux_kind, N, uM.shape[0], uM.reshape(-1), ma.mapP.reshape(-1), ma.bc.reshape(-1),
                   ma.bc_sign, ma.nx.reshape(-1), ma.ny.reshape(-1), ma.nz.reshape(-1), out.reshape(-1))
    return out


def lsrk_step(pde, N, q_aos, dt, ma: MeshArrays, nthreads=1, flux_kind=1, L=None):
    """in-place LSRK45 step on q_aos (K, Np, F) float64 C-contiguous (L as for rhs)"""
    assert q_aos.flags.c_contiguous and q_aos.dtype == np.float64
    res = np.zeros_like(q_aos)
    (L or lib()).orc_lsrk_step(pde, flux_kind, N, ma.K, q_aos.reshape(-1), res.reshape(-1), dt, ma.mapP.reshape(-1),
                        ma.bc.reshape(-1), ma.bc_sign, ma.nx.reshape(-1), ma.ny.reshape(-1), ma.nz.reshape(-1),
                        ma.geo.reshape(-1), ma.J, nthreads)
    return q_aos


class BatchArrays:
    """Oracle input for elements of mixed degree: a bbdg.batch.Batch (host set-up data,
    device element order) plus the degree-transfer tables, offsets in a 21 x 21 table."""

    def __init__(self, batch, bc_sign):
        from bbdg import transfer as TR  # host set-up tables (exact rationals, rounded once)

        self.b = batch
        self.K = batch.K
        self.deg = np.ascontiguousarray(batch.deg.astype(np.int32))
        self.qoff = np.ascontiguousarray(batch.qoff.astype(np.int64))
        self.nbr = np.ascontiguousarray(batch.nbr.astype(np.int32).reshape(-1))
        self.nbf = np.ascontiguousarray(batch.nbr_face.astype(np.int32).reshape(-1))
        self.perm = np.ascontiguousarray(batch.perm.astype(np.int32).reshape(-1))
        self.bc = np.ascontiguousarray(batch.bc.astype(np.int32).reshape(-1))
        self.nJ = np.ascontiguousarray(batch.nJ.reshape(-1))
        self.geo = np.ascontiguousarray(batch.geo.reshape(-1))
        self.J = np.ascontiguousarray(batch.J)
        self.dte = np.ascontiguousarray(batch.dte)
        T, off = TR.pack(21)
        self.T = np.ascontiguousarray(T)
        self.toff = np.ascontiguousarray(off.reshape(-1))
        self.bc_sign = np.ascontiguousarray(np.asarray(bc_sign, dtype=np.float64))


def rhs_md(pde, q_node_major, ba: BatchArrays, nthreads=1, flux_kind=1):
    """mixed-degree RHS on a per-element [Np][F] state (Batch.to_node_major layout)"""
    q = np.ascontiguousarray(q_node_major, dtype=np.float64)
    du = np.zeros_like(q)
    lib().orc_rhs_md(pde, flux_kind, ba.K, ba.deg, ba.qoff, q, ba.nbr, ba.nbf, ba.perm, ba.bc, ba.bc_sign, ba.nJ,
                     ba.geo, ba.J, ba.T, ba.toff, du, nthreads)
    return du


def lsrk_step_md(pde, q_node_major, ba: BatchArrays, dt_scale=1.0, nthreads=1, flux_kind=1):
    """one LSRK45 step in place, element e with step dt_scale * dt_e"""
    res = np.zeros_like(q_node_major)
    dt = np.ascontiguousarray(ba.dte * dt_scale)
    lib().orc_lsrk_step_md(pde, flux_kind, ba.K, ba.deg, ba.qoff, q_node_major, res, dt, ba.nbr, ba.nbf, ba.perm,
                           ba.bc, ba.bc_sign, ba.nJ, ba.geo, ba.J, ba.T, ba.toff, nthreads)
