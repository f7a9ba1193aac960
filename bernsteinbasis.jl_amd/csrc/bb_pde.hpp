// bb_pde.hpp -- the two first-order systems of the reference examples, as device code.
//
// wave    (examples/3d_wave.jl:20-33):   q = (p,u,v,w), fx = (u,p,0,0), fy = (v,0,p,0),
//                                        fz = (w,0,0,p); du = LIFT*flux - sum geo*D f
// Maxwell (examples/3d_maxwell.jl:20-33): q = (Hx,Hy,Hz,Ex,Ey,Ez), fx = (0,Ez,-Ey,0,-Hz,Hy),
//                                        fy = (-Ez,0,Ex,Hz,0,-Hx), fz = (Ey,-Ex,0,-Hy,Hx,0);
//                                        du = LIFT*flux + sum geo*D f
// Both then divide by J (examples/3d_wave.jl:83, 3d_maxwell.jl:92).
#pragma once
#include "bb_common.hpp"

namespace bb {

template <int PDE> struct Pde;

template <> struct Pde<BB_PDE_WAVE> {
    static constexpr int F = 4;
    static constexpr bool plus = false;  // du -= sum geo * D f
    // flux vector fx/fy/fz (dir 0/1/2) applied to a state; call with a constant dir
    template <typename T> __device__ static void fdir(int dir, const T *q, T *o) {
        o[0] = dir == 0 ? q[1] : dir == 1 ? q[2] : q[3];
        o[1] = dir == 0 ? q[0] : T(0);
        o[2] = dir == 1 ? q[0] : T(0);
        o[3] = dir == 2 ? q[0] : T(0);
    }
    // volume term: add the contribution of field f's physical gradient (gx,gy,gz)
    // to vol = -(d/dx fx + d/dy fy + d/dz fz) (the sign of the wave RHS folded in)
    template <typename T> __device__ static void vol_add(int f, T gx, T gy, T gz, T *vol) {
        if (f == 0) { vol[1] -= gx; vol[2] -= gy; vol[3] -= gz; }
        else if (f == 1) vol[0] -= gx;
        else if (f == 2) vol[0] -= gy;
        else vol[0] -= gz;
    }
    // interface flux, examples/3d_wave.jl:45-54 (kind 0 verbatim; kind 1: every term
    // carries Jf^1, i.e. d_p -> Jf d_p and (n.d) -> (n.d)/Jf in the penalty).  Jf = |nJ| and
    // invJf = 1/Jf come from the caller (per face of an affine element).
    // (sums of products are written as explicit fma, so that every kernel that inlines this
    // rounds the same way whatever the compiler's contraction choices)
    template <typename T>
    __device__ static void flux(int kind, const T *d, T nx, T ny, T nz, T Jf, T invJf, T *o) {
        const T ndotdU = fma(nz, d[3], fma(ny, d[2], nx * d[1]));
        if (kind == BB_FLUX_REFERENCE) {
            o[0] = T(0.5) * (d[0] - ndotdU);
            o[1] = T(0.5) * ((ndotdU - d[0]) * nx);
            o[2] = T(0.5) * ((ndotdU - d[0]) * ny);
            o[3] = T(0.5) * ((ndotdU - d[0]) * nz);
        } else {
            T un = ndotdU * invJf;
            o[0] = T(0.5) * fma(Jf, d[0], -ndotdU);
            o[1] = T(0.5) * ((un - d[0]) * nx);
            o[2] = T(0.5) * ((un - d[0]) * ny);
            o[3] = T(0.5) * ((un - d[0]) * nz);
        }
    }
};

template <> struct Pde<BB_PDE_MAXWELL> {
    static constexpr int F = 6;
    static constexpr bool plus = true;  // du += sum geo * D f
    template <typename T> __device__ static void fdir(int dir, const T *q, T *o) {
        if (dir == 0) { o[0] = T(0); o[1] = q[5]; o[2] = -q[4]; o[3] = T(0); o[4] = -q[2]; o[5] = q[1]; }
        else if (dir == 1) { o[0] = -q[5]; o[1] = T(0); o[2] = q[3]; o[3] = q[2]; o[4] = T(0); o[5] = -q[0]; }
        else { o[0] = q[4]; o[1] = -q[3]; o[2] = T(0); o[3] = -q[1]; o[4] = q[0]; o[5] = T(0); }
    }
    // vol = +(d/dx fx + d/dy fy + d/dz fz), regrouped by source field
    template <typename T> __device__ static void vol_add(int f, T gx, T gy, T gz, T *vol) {
        switch (f) {
        case 0: vol[4] += gz; vol[5] -= gy; break;  // Hx
        case 1: vol[3] -= gz; vol[5] += gx; break;  // Hy
        case 2: vol[3] += gy; vol[4] -= gx; break;  // Hz
        case 3: vol[1] -= gz; vol[2] += gy; break;  // Ex
        case 4: vol[0] += gz; vol[2] -= gx; break;  // Ey
        default: vol[0] -= gy; vol[1] += gx; break; // Ez
        }
    }
    // examples/3d_maxwell.jl:45-63 (kind 1: penalties d -> Jf d, (n.d)n -> (n.d)n/Jf)
    // 1/2 (fx(d) nx + fy(d) ny + fz(d) nz + d - pen): the two non-zero entries of
    // fx nx + fy ny + fz nz per field (n x E and -n x H), as explicit fma; the factor 1/2 is
    // applied last (exact)
    template <typename T>
    __device__ static void flux(int kind, const T *d, T nx, T ny, T nz, T Jf, T invJf, T *o) {
        T ndH = fma(nz, d[2], fma(ny, d[1], nx * d[0]));
        T ndE = fma(nz, d[5], fma(ny, d[4], nx * d[3]));
        if (kind != BB_FLUX_REFERENCE) {
            ndH = ndH * invJf;
            ndE = ndE * invJf;
        }
        const T cr[6] = {fma(d[4], nz, -d[5] * ny), fma(-d[3], nz, d[5] * nx), fma(d[3], ny, -d[4] * nx),
                         fma(-d[1], nz, d[2] * ny), fma(d[0], nz, -d[2] * nx), fma(-d[0], ny, d[1] * nx)};
        const T pen[6] = {ndH * nx, ndH * ny, ndH * nz, ndE * nx, ndE * ny, ndE * nz};
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const T jd = kind != BB_FLUX_REFERENCE ? Jf * d[f] : d[f];
            o[f] = T(0.5) * ((cr[f] + jd) - pen[f]);
        }
    }
};

}  // namespace bb
