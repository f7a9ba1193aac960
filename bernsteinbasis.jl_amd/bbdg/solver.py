"""Device-resident wave / Maxwell DG solver over the C ABI.

Replaces the rhs_matvec! + ODEProblem pair of examples/3d_wave.jl:36-134 and
examples/3d_maxwell.jl:36-138.  Two right-hand-side paths:

* ``path="plain"``: one kernel per operator in the reference's order (face gather,
  flux, fx/fy/fz, nine derivative matvecs, lift, combine + /J), FP64;
* ``path="fused"``: one kernel per low-storage RK stage, ``prec="f64"`` or ``"mixed"``.

Host state is exchanged in the reference's layout (K, Np, F) (Julia's column-major
Np x K matrix of SVector{F}); on the device it is (K, F, Np).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from .device import Context, default_context
from .mesh import TetMesh, nfp, np_

PDES = {"wave": B.BB_PDE_WAVE, "maxwell": B.BB_PDE_MAXWELL}
FLUX = {"reference": B.BB_FLUX_REFERENCE, "upwind": B.BB_FLUX_UPWIND}
PATH = {"plain": B.BB_PATH_PLAIN, "fused": B.BB_PATH_FUSED}
PREC = {"f64": B.BB_PREC_F64, "mixed": B.BB_PREC_MIXED}
NFIELDS = {"wave": 4, "maxwell": 6}
PEC_SIGN = (1.0, 1.0, 1.0, -1.0, -1.0, -1.0)  # Maxwell PEC wall: H+ = H-, E+ = -E-


def _p(a, dt):
    return np.ascontiguousarray(a, dtype=dt).ctypes.data_as(C.c_void_p)


class Solver:
    def __init__(self, mesh: TetMesh, N: int, pde: str = "wave", flux: str = "upwind", bc_sign=None,
                 ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.mesh, self.N, self.pde = mesh, N, pde
        self.F = NFIELDS[pde]
        self.Np, self.Nfp, self.K = np_(N), nfp(N), mesh.K
        if bc_sign is None:
            bc_sign = PEC_SIGN if pde == "maxwell" else (1.0,) * self.F
        self._keep = [np.ascontiguousarray(mesh.geo, np.float64), np.ascontiguousarray(mesh.J, np.float64),
                      np.ascontiguousarray(mesh.nJ, np.float64), np.ascontiguousarray(mesh.nbr, np.int32),
                      np.ascontiguousarray(mesh.nbr_face, np.int8), np.ascontiguousarray(mesh.perm, np.int8),
                      np.ascontiguousarray(mesh.bc, np.int8), np.ascontiguousarray(bc_sign, np.float64)]
        h = C.c_void_p()
        self.ctx.check(B.lib().bb_solver_create(self.ctx.h, PDES[pde], FLUX[flux], N, self.K,
                                                *[a.ctypes.data_as(C.c_void_p) for a in self._keep], C.byref(h)))
        self.h = h
        B.track(self)

    def close(self):
        if getattr(self, "h", None):
            B.lib().bb_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state (host layout (K, Np, F))
    def set_state(self, q):
        q = np.asarray(q, np.float64)
        if q.shape != (self.K, self.Np, self.F):
            raise ValueError(f"state must be {(self.K, self.Np, self.F)}")
        soa = np.ascontiguousarray(q.transpose(0, 2, 1))
        self.ctx.check(B.lib().bb_solver_set_state(self.h, soa.ctypes.data_as(C.c_void_p)))

    def get_state(self):
        soa = np.empty((self.K, self.F, self.Np))
        self.ctx.check(B.lib().bb_solver_get_state(self.h, soa.ctypes.data_as(C.c_void_p)))
        return np.ascontiguousarray(soa.transpose(0, 2, 1))

    def check_state(self):
        """number of NaN / Inf values in the device state (0 while the run is healthy)"""
        n = C.c_int64()
        rc = B.lib().bb_solver_check_state(self.h, C.byref(n))
        if rc not in (B.BB_OK, B.BB_ERR_NONFINITE):
            self.ctx.check(rc)
        return n.value

    def rhs(self, path="fused", prec="f64"):
        soa = np.empty((self.K, self.F, self.Np))
        self.ctx.check(B.lib().bb_solver_rhs(self.h, PATH[path], PREC[prec], soa.ctypes.data_as(C.c_void_p)))
        return np.ascontiguousarray(soa.transpose(0, 2, 1))

    def step(self, dt, nsteps=1, path="fused", prec="f64"):
        self.ctx.check(B.lib().bb_solver_step(self.h, PATH[path], PREC[prec], float(dt), int(nsteps)))

    def capture(self, dt, nsteps, path="fused", prec="f64"):
        self.ctx.check(B.lib().bb_solver_graph_capture(self.h, PATH[path], PREC[prec], float(dt), int(nsteps)))

    def launch(self):
        self.ctx.check(B.lib().bb_solver_graph_launch(self.h))

    def autotune(self, prec="f64"):
        """time every compiled tiling of the fused kernel; returns the chosen tiling code
        (elements per workgroup + 16 x lift organisation)"""
        e = C.c_int()
        self.ctx.check(B.lib().bb_solver_autotune(self.h, PREC[prec], C.byref(e)))
        return e.value

    def set_tuning(self, prec, tiling):
        self.ctx.check(B.lib().bb_solver_set_tuning(self.h, PREC[prec], int(tiling)))

    @staticmethod
    def tiling_code(elems_per_block, lift="gather"):
        return int(elems_per_block) + 16 * {"gather": 0, "pingpong": 1, "gather_mfma": 2, "wave": 3,
                                            "wave_gt": 4, "wave1": 5}[lift]

    def stage_bytes(self, prec="f64"):
        b, d = C.c_double(), C.c_double()
        self.ctx.check(B.lib().bb_solver_stage_bytes(self.h, PREC[prec], C.byref(b), C.byref(d)))
        return b.value, d.value

    @property
    def dofs(self):
        return self.K * self.Np * self.F
