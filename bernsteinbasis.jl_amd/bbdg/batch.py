"""Many independent tet meshes whose elements carry their own polynomial degree, stepped
together (host set-up; the compute path is bb_batch_* in the C ABI).

Layout: all elements of all instances are numbered instance after instance ("global"
order); the device order groups them by degree -- group N holds every element of degree
N in global order, groups in increasing N -- and the state of group N is [K_N][F][Np(N)]
FP64, the groups concatenated.  Face neighbours are indices in the device order.  A face
between elements of different degree is handled by bringing the neighbour's trace to
the element's degree (bbdg.transfer).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as B
from . import transfer as TR
from .basis import interpolation_matrix
from .device import Context, default_context
from .mesh import BC_INTERIOR, multi_indices

PDE = {"wave": B.BB_PDE_WAVE, "maxwell": B.BB_PDE_MAXWELL}
FLUX = {"reference": B.BB_FLUX_REFERENCE, "upwind": B.BB_FLUX_UPWIND}
PREC = {"f64": B.BB_PREC_F64, "mixed": B.BB_PREC_MIXED}


def np_(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


@dataclass
class Batch:
    """meshes: list of TetMesh; degrees: per instance an int array (K_i,) of element
    degrees; dt: per instance step size; F: fields per node"""

    meshes: list
    degrees: list
    dt: list
    F: int = 4

    def __post_init__(self):
        self.inst_off = np.concatenate([[0], np.cumsum([m.K for m in self.meshes])]).astype(np.int64)
        deg_g = np.concatenate([np.asarray(d, np.int64) for d in self.degrees])
        self.K = int(deg_g.size)
        self.order = np.argsort(deg_g, kind="stable")          # device position -> global index
        self.pos = np.empty(self.K, np.int64)
        self.pos[self.order] = np.arange(self.K)                # global index -> device position
        self.deg = deg_g[self.order]                            # device order
        self.Ns = sorted(set(int(v) for v in self.deg))
        self.inst = np.repeat(np.arange(len(self.meshes)), [m.K for m in self.meshes])[self.order]
        npe = np.array([np_(int(n)) for n in self.deg]) * self.F
        self.qoff = np.concatenate([[0], np.cumsum(npe)[:-1]]).astype(np.int64)  # per element, device order
        self.words = int(npe.sum())
        self.dofs = self.words
        cat = lambda name: np.concatenate([getattr(m, name) for m in self.meshes])[self.order]
        self.geo = np.ascontiguousarray(cat("geo"))
        self.J = np.ascontiguousarray(cat("J"))
        self.nJ = np.ascontiguousarray(cat("nJ"))
        nbr = np.concatenate([m.nbr + off for m, off in zip(self.meshes, self.inst_off[:-1])])[self.order]
        self.nbr = np.ascontiguousarray(self.pos[nbr]).astype(np.int32)
        self.nbr_face = np.ascontiguousarray(cat("nbr_face")).astype(np.int32)
        self.perm = np.ascontiguousarray(cat("perm")).astype(np.int32)
        self.bc = np.ascontiguousarray((cat("bc") != BC_INTERIOR).astype(np.int32))
        self.dte = np.ascontiguousarray(np.asarray(self.dt, np.float64)[self.inst])
        self.vx = np.concatenate([m.vx for m in self.meshes])[self.order]

    def groups(self):
        """(N, first device position, count) per degree present"""
        out = []
        for N in self.Ns:
            idx = np.nonzero(self.deg == N)[0]
            out.append((N, int(idx[0]), int(idx.size)))
        return out

    def interpolate(self, fn_per_instance):
        """device-layout state from fn_i(x, y, z) -> (F, ...) per instance"""
        q = np.zeros(self.words)
        for N, s, c in self.groups():
            Vinv, _ = interpolation_matrix(N)
            mi = multi_indices(N).astype(np.float64) / N
            X = np.einsum("am,kmx->kax", mi, self.vx[s:s + c])
            vals = np.empty((self.F, c, np_(N)))
            inst = self.inst[s:s + c]
            for i in np.unique(inst):
                sel = inst == i
                vals[:, sel] = np.asarray(fn_per_instance[i](X[sel, :, 0], X[sel, :, 1], X[sel, :, 2]))
            coef = np.matmul(vals, Vinv.T)  # (F, c, Np)
            q[self.qoff[s]:self.qoff[s] + c * self.F * np_(N)] = coef.transpose(1, 0, 2).reshape(-1)
        return q

    def to_node_major(self, q):
        """device layout ([F][Np] per element) -> per element [Np][F] (the oracle's layout)"""
        out = np.empty_like(q)
        for N, s, c in self.groups():
            a, w = self.qoff[s], c * self.F * np_(N)
            out[a:a + w] = q[a:a + w].reshape(c, self.F, np_(N)).transpose(0, 2, 1).reshape(-1)
        return out

    def from_node_major(self, q):
        out = np.empty_like(q)
        for N, s, c in self.groups():
            a, w = self.qoff[s], c * self.F * np_(N)
            out[a:a + w] = q[a:a + w].reshape(c, np_(N), self.F).transpose(0, 2, 1).reshape(-1)
        return out

    def field_view(self, q, N):
        """(K_N, F, Np) view of group N of a device-layout state"""
        for n, s, c in self.groups():
            if n == N:
                return q[self.qoff[s]:self.qoff[s] + c * self.F * np_(N)].reshape(c, self.F, np_(N))
        raise KeyError(N)


class BatchSolver:
    """Device-resident batch (bb_batch_* of the C ABI): per stage, a transfer pass for the
    faces whose neighbour has another degree, then one fused stage kernel per degree group, the
    groups on four streams; hipGraph-captured steps, per-group autotuned tilings."""

    def __init__(self, batch: Batch, pde="wave", flux="upwind", bc_sign=None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.batch = batch
        F = 4 if pde == "wave" else 6
        if F != batch.F:
            raise ValueError("field count of the batch does not match the PDE")
        T, off = TR.pack(11)
        self._T = np.ascontiguousarray(T)
        self._off = np.ascontiguousarray(off.reshape(-1))
        sign = np.ones(F) if bc_sign is None else np.asarray(bc_sign, np.float64)
        h = C.c_void_p()
        p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        deg = np.ascontiguousarray(batch.deg.astype(np.int32))
        self.ctx.check(B.lib().bb_batch_create(
            self.ctx.h, PDE[pde], FLUX[flux], batch.K, p(deg, C.c_int32), p(batch.geo, C.c_double),
            p(batch.J, C.c_double), p(batch.nJ, C.c_double), p(batch.nbr, C.c_int32), p(batch.nbr_face, C.c_int32),
            p(batch.perm, C.c_int32), p(batch.bc, C.c_int32), p(np.ascontiguousarray(sign), C.c_double),
            p(batch.dte, C.c_double), p(self._T, C.c_double), p(self._off, C.c_int64), self._T.size, C.byref(h)))
        self.h = h
        B.track(self)

    def close(self):
        if self.h:
            B.lib().bb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.size != self.batch.words:
            raise ValueError("state size")
        self.ctx.check(B.lib().bb_batch_set_state(self.h, q.ctypes.data_as(C.POINTER(C.c_double))))

    def get_state(self):
        q = np.empty(self.batch.words)
        self.ctx.check(B.lib().bb_batch_get_state(self.h, q.ctypes.data_as(C.POINTER(C.c_double))))
        return q

    def check_state(self):
        """number of NaN / Inf values in the device state of every degree group"""
        n = C.c_int64()
        rc = B.lib().bb_batch_check_state(self.h, C.byref(n))
        if rc not in (B.BB_OK, B.BB_ERR_NONFINITE):
            self.ctx.check(rc)
        return n.value

    def rhs(self, prec="f64"):
        du = np.empty(self.batch.words)
        self.ctx.check(B.lib().bb_batch_rhs(self.h, PREC[prec], du.ctypes.data_as(C.POINTER(C.c_double))))
        return du

    def step(self, nsteps=1, prec="f64", dt_scale=1.0):
        self.ctx.check(B.lib().bb_batch_step(self.h, PREC[prec], float(dt_scale), int(nsteps)))

    def capture(self, nsteps=1, prec="f64", dt_scale=1.0):
        self.ctx.check(B.lib().bb_batch_graph_capture(self.h, PREC[prec], float(dt_scale), int(nsteps)))

    def launch(self):
        self.ctx.check(B.lib().bb_batch_graph_launch(self.h))

    def autotune(self, prec="f64"):
        t = (C.c_int * 11)()
        self.ctx.check(B.lib().bb_batch_autotune(self.h, PREC[prec], t))
        return {N: int(t[N]) for N in self.batch.Ns}

    def set_tuning(self, prec, N, tiling):
        self.ctx.check(B.lib().bb_batch_set_tuning(self.h, PREC[prec], int(N), int(tiling)))
