"""Dense nodal comparator: the RHS of extras/3d_nodal_wave.jl:57-73 on the GPU.

The reference's comparison solver keeps nodal values and applies dense matrices: Dr, Ds,
Dt (Np x Np) to fx.(u), fy.(u), fz.(u) and a dense LIFT (Np x 4Nfp) to the interface flux
-- O(Np^2) per element and field where the Bernstein operators are O(Np).  Here the nodes
are the equispaced domain points in the Bernstein linear order (the points the initial
conditions are interpolated at), so that with V the Bernstein Vandermonde matrix at these
nodes and V_f its face counterpart:

    Dr_nodal = V Dr V^{-1},   LIFT_nodal = V LIFT [V_f^{-1}, .., V_f^{-1}] (4 face blocks),

and the nodal RHS of V c equals V times the Bernstein RHS of c.  The dense products run as
library GEMMs (bb_gemm) over all columns of all elements; the pointwise pieces are the plain
path's kernels (face gather, interface flux, flux directions, geometric combine).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from .basis import bernstein_basis_from_barycentric_tri, interpolation_matrix
from .device import Context, DeviceArray, default_context
from .mesh import BC_INTERIOR, face_node_abc, face_node_normals, map_p, nfp, np_
from .operators import (BernsteinDerivativeMatrix_3D_r, BernsteinDerivativeMatrix_3D_s,
                        BernsteinDerivativeMatrix_3D_t, BernsteinLift, mul_)

PDE = {"wave": B.BB_PDE_WAVE, "maxwell": B.BB_PDE_MAXWELL}
FLUX = {"reference": B.BB_FLUX_REFERENCE, "upwind": B.BB_FLUX_UPWIND}


def nodal_derivative_matrices(N: int):
    """Dr, Ds, Dt of the Lagrange basis at the equispaced domain points (host, FP64):
    V D V^{-1} with the Bernstein derivative matrices of bb_deriv_entry"""
    Vinv, V = interpolation_matrix(N)
    D = [cls(N).dense() for cls in (BernsteinDerivativeMatrix_3D_r, BernsteinDerivativeMatrix_3D_s,
                                     BernsteinDerivativeMatrix_3D_t)]
    return tuple(V @ d @ Vinv for d in D)


class NodalOperators:
    """Dense nodal Dr, Ds, Dt (Np x Np) and LIFT (Np x 4 Nfp), FP64, and V / V^{-1}."""

    def __init__(self, N: int, ctx: Context | None = None):
        ctx = ctx or default_context()
        self.N = N
        Vinv, V = interpolation_matrix(N)
        self.V, self.Vinv = V, Vinv
        self.Dr, self.Ds, self.Dt = nodal_derivative_matrices(N)
        # the Bernstein lift as a matrix: the GPU lift of the identity columns
        n, m = np_(N), 4 * nfp(N)
        x = DeviceArray.from_numpy(np.eye(m), ctx=ctx)                  # column c = e_c
        out = DeviceArray.zeros((m, n), ctx=ctx)
        mul_(out.view((n, m), (1, n)), BernsteinLift(N), x.view((m, m), (1, m)))
        LB = out.numpy().T                                              # (Np, 4 Nfp)
        x.free()
        out.free()
        abc = face_node_abc(N).astype(np.float64) / N
        Vf = bernstein_basis_from_barycentric_tri(N, abc[:, 0], abc[:, 1], abc[:, 2])[0]
        Vf_inv = np.linalg.inv(Vf)
        self.LIFT = V @ LB @ np.kron(np.eye(4), Vf_inv)

    def flops_per_rhs(self, K: int, F: int) -> float:
        n, m = np_(self.N), 4 * nfp(self.N)
        return 2.0 * K * F * (9 * n * n + n * m)


class NodalRHS:
    """Device-resident mesh arrays and scratch for the dense nodal RHS of one mesh at one
    degree (FP64, state layout [K][F][Np] of nodal values)."""

    def __init__(self, mesh, N: int, pde: str = "wave", flux: str = "upwind", bc_sign=None,
                 ops: NodalOperators | None = None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.mesh, self.N, self.pde = mesh, N, pde
        self.F = 4 if pde == "wave" else 6
        self.K = mesh.K
        self.ops = ops or NodalOperators(N, self.ctx)
        n, m, K, F = np_(N), 4 * nfp(N), mesh.K, self.F
        self.n, self.m = n, m
        if bc_sign is None:
            bc_sign = [1.0] * 4 if pde == "wave" else [1, 1, 1, -1, -1, -1]
        self.sign = np.ascontiguousarray(np.asarray(bc_sign, np.float64))
        self.flux_kind = FLUX[flux]
        mp, bc = map_p(mesh, N)
        nx, ny, nz = face_node_normals(mesh, N)
        dev = lambda a: DeviceArray.from_numpy(np.ascontiguousarray(a), ctx=self.ctx)
        self.mapP, self.bc = dev(mp), dev(bc)
        self.nx, self.ny, self.nz = dev(nx), dev(ny), dev(nz)
        self.geo, self.J = dev(mesh.geo), dev(mesh.J)
        self.has_wall = bool((mesh.bc != BC_INTERIOR).any())
        # column-major operator matrices for the GEMMs
        self.Dd = [dev(np.asfortranarray(d).reshape(-1, order="F")) for d in (self.ops.Dr, self.ops.Ds, self.ops.Dt)]
        self.Ld = dev(np.asfortranarray(self.ops.LIFT).reshape(-1, order="F"))
        z = lambda shape: DeviceArray.zeros(shape, ctx=self.ctx)
        self.uM, self.flux = z((K, F, m)), z((K, F, m))
        self.fu = [z((K, F, n)) for _ in range(3)]
        self.d9 = [z((K, F, n)) for _ in range(9)]
        self.du = z((K, F, n))

    def lift_device(self):
        """du = LIFT_nodal flux: one Np x 4Nfp GEMM over all K*F columns"""
        n, m = self.n, self.m
        self.ctx.check(B.lib().bb_gemm(self.ctx.h, B.BB_F64, n, self.K * self.F, m, 1.0, C.c_void_p(self.Ld.ptr),
                                       n, C.c_void_p(self.flux.ptr), m, 0.0, C.c_void_p(self.du.ptr), n))

    def derivatives_device(self):
        """the nine reference derivatives D_a fu[c]: nine Np x Np GEMMs over all K*F columns"""
        n = self.n
        for c in range(3):
            for ax in range(3):
                self.ctx.check(B.lib().bb_gemm(self.ctx.h, B.BB_F64, n, self.K * self.F, n, 1.0,
                                               C.c_void_p(self.Dd[ax].ptr), n, C.c_void_p(self.fu[c].ptr), n, 0.0,
                                               C.c_void_p(self.d9[3 * c + ax].ptr), n))

    def surface_device(self, q: DeviceArray):
        """du = LIFT_nodal (interface flux of q): face gather, flux, lift GEMM"""
        L, h = B.lib(), self.ctx.h
        K, F, m, N = self.K, self.F, self.m, self.N
        lv = B.bb_layout(F * self.n, 1, self.n)
        lf = B.bb_layout(F * m, 1, m)
        vp = lambda a: C.c_void_p(a.ptr)
        self.ctx.check(L.bb_face_gather(h, N, F, K, B.BB_F64, vp(q), lv, vp(self.uM), lf))
        self.ctx.check(L.bb_interface_flux(
            h, PDE[self.pde], self.flux_kind, N, K, vp(self.uM), lf, vp(self.mapP),
            vp(self.bc) if self.has_wall else None,
            self.sign.ctypes.data_as(C.c_void_p) if self.has_wall else None,
            vp(self.nx), vp(self.ny), vp(self.nz), vp(self.flux), lf))
        self.lift_device()

    def volume_device(self, q: DeviceArray):
        """flux directions of q, then the nine derivative GEMMs"""
        vp = lambda a: C.c_void_p(a.ptr)
        self.ctx.check(B.lib().bb_pde_flux_dirs(self.ctx.h, PDE[self.pde], self.N, self.K, vp(q), vp(self.fu[0]),
                                                vp(self.fu[1]), vp(self.fu[2])))
        self.derivatives_device()

    def rhs_device(self, q: DeviceArray):
        """du (device, [K][F][Np]) = nodal RHS of the nodal state q (device, same layout)"""
        self.surface_device(q)
        self.volume_device(q)
        ptrs = (C.c_void_p * 9)(*[a.ptr for a in self.d9])
        self.ctx.check(B.lib().bb_rhs_combine(self.ctx.h, PDE[self.pde], self.N, self.K, C.c_void_p(self.du.ptr),
                                              ptrs, C.c_void_p(self.geo.ptr), C.c_void_p(self.J.ptr)))
        return self.du

    def rhs(self, q_host_kfn):
        """host convenience: q (K, F, Np) nodal values -> du (K, F, Np)"""
        q = DeviceArray.from_numpy(np.ascontiguousarray(q_host_kfn, dtype=np.float64), ctx=self.ctx)
        du = self.rhs_device(q).numpy().copy()
        q.free()
        return du

    def close(self):
        for a in [self.mapP, self.bc, self.nx, self.ny, self.nz, self.geo, self.J, *self.Dd, self.Ld, self.uM,
                  self.flux, *self.fu, *self.d9, self.du]:
            a.free()
