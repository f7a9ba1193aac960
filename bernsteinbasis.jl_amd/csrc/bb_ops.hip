// bb_ops.hip -- derivative, elevation, trace-gather and interface-flux operators.
//
// Derivative kernels stage a block of whole columns through LDS with coalesced loads
// (the block's elements are contiguous in both the reference's [K][Np][F] layout and the
// solver's [K][F][Np] layout), then every thread evaluates the <= 13-point barycentric
// stencil of its node from LDS.  They are HBM-bound: per column Np values in, Np (or
// 3 Np) values out.
#include "bb_dispatch.hpp"
#include "bb_pde.hpp"
#include "bb_stencil.hpp"

namespace bb {

constexpr int kDerivThreads = 256;
constexpr int kDerivTarget = 2048;  // staged values per block

// split a block-local work item t into (element, field, node): the fastest index is the
// one with the smaller stride in memory, so consecutive threads touch consecutive words.
__device__ __forceinline__ void split3(int t, bool field_fast, int FB, int NP, int &e, int &f, int &n) {
    if (field_fast) {
        f = t % FB;
        int r = t / FB;
        n = r % NP;
        e = r / NP;
    } else {
        n = t % NP;
        int r = t / NP;
        f = r % FB;
        e = r / FB;
    }
}

template <int N, typename T, bool GRAD>
__global__ __launch_bounds__(kDerivThreads) void k_deriv(const T *__restrict__ x, Lay lx, T *__restrict__ o0,
                                                          T *__restrict__ o1, T *__restrict__ o2, Lay lo,
                                                          const double *__restrict__ geo, int F, int FB, int64_t K,
                                                          int EB, int axis) {
    constexpr int NP = np3(N);
    extern __shared__ double smem_d[];
    T *sx = reinterpret_cast<T *>(smem_d);
    int *sijk = reinterpret_cast<int *>(sx + EB * FB * NP);
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int f0 = blockIdx.y * FB;
    const int ne = (int)min((int64_t)EB, K - e0);
    const int nf = min(FB, F - f0);
    for (int t = threadIdx.x; t < NP; t += blockDim.x) {
        int i, j, k;
        decode3<N>(t, i, j, k);
        sijk[t] = i | (j << 8) | (k << 16);
    }
    const bool fin = lx.f < lx.n;
    const int nit = ne * nf * NP;
    for (int t = threadIdx.x; t < nit; t += blockDim.x) {
        int e, f, n;
        split3(t, fin, nf, NP, e, f, n);
        sx[(e * FB + f) * NP + n] = x[(e0 + e) * lx.e + (int64_t)n * lx.n + (int64_t)(f0 + f) * lx.f];
    }
    __syncthreads();
    const bool fout = lo.f < lo.n;
    for (int t = threadIdx.x; t < nit; t += blockDim.x) {
        int e, f, n;
        split3(t, fout, nf, NP, e, f, n);
        const int p = sijk[n];
        T dr, ds, dt;
        grad_ref<N, T>(sx + (e * FB + f) * NP, n, p & 255, (p >> 8) & 255, p >> 16, dr, ds, dt);
        const int64_t oa = (e0 + e) * lo.e + (int64_t)n * lo.n + (int64_t)(f0 + f) * lo.f;
        if constexpr (GRAD) {
            const double *g = geo + (e0 + e) * 9;
            o0[oa] = T(g[0]) * dr + T(g[1]) * ds + T(g[2]) * dt;
            o1[oa] = T(g[3]) * dr + T(g[4]) * ds + T(g[5]) * dt;
            o2[oa] = T(g[6]) * dr + T(g[7]) * ds + T(g[8]) * dt;
        } else {
            o0[oa] = axis == 0 ? dr : axis == 1 ? ds : dt;
        }
    }
}

// Compact layouts ([K][F][Np] field-major, or [K][Np][F] field-fastest, same for input and
// outputs): a block's EB whole elements are one contiguous chunk of at most kChunkBytes, so
// the chunk is copied into LDS with a fixed number of 16-byte loads per thread, all issued
// before the first LDS store, and every output lands at the same chunk position as its
// input (coalesced stores).  The geometric factors of the block's elements are staged once.
constexpr int kChunkBytes = 16384;
constexpr int kChunkMaxElems = 64;

template <typename T, int VEC> struct VecT;
template <> struct VecT<double, 1> { using type = double; };
template <> struct VecT<double, 2> { using type = double2; };
template <> struct VecT<float, 1> { using type = float; };
template <> struct VecT<float, 4> { using type = float4; };

template <int N, typename T, int VEC, bool AOS, bool GRAD>
__global__ __launch_bounds__(kDerivThreads) void k_deriv_chunk(const T *__restrict__ x, T *__restrict__ o0,
                                                                T *__restrict__ o1, T *__restrict__ o2,
                                                                const double *__restrict__ geo, int F, int64_t K,
                                                                int EB, uint32_t magicF, int axis) {
    constexpr int NP = np3(N);
    constexpr int VALS = kChunkBytes / sizeof(T);
    constexpr int PER = VALS / kDerivThreads;  // values per thread
    constexpr int LPT = PER / VEC;             // vector loads per thread
    using V = typename VecT<T, VEC>::type;
    __shared__ __align__(16) T sx[VALS];
    __shared__ T sg[GRAD ? kChunkMaxElems * 9 : 1];
    __shared__ int sijk[NP];
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int ne = (int)min((int64_t)EB, K - e0);
    const int C = ne * F * NP;
    const int64_t base = e0 * F * NP;
    V r[LPT];
#pragma unroll
    for (int u = 0; u < LPT; ++u) {  // clamped, unconditional loads keep r[] in registers
        const int p = (threadIdx.x + u * kDerivThreads) * VEC;
        r[u] = *reinterpret_cast<const V *>(x + base + (p < C ? p : 0));
    }
    constexpr int GPT = GRAD ? (kChunkMaxElems * 9 + kDerivThreads - 1) / kDerivThreads : 0;
    double g[GPT > 0 ? GPT : 1];
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
        const int t = threadIdx.x + u * kDerivThreads;
        if (t < ne * 9) g[u] = geo[e0 * 9 + t];
    }
    for (int t = threadIdx.x; t < NP; t += kDerivThreads) {
        int i, j, k;
        decode3<N>(t, i, j, k);
        sijk[t] = i | (j << 8) | (k << 16);
    }
#pragma unroll
    for (int u = 0; u < LPT; ++u) {
        const int p = (threadIdx.x + u * kDerivThreads) * VEC;
        if (p < C) *reinterpret_cast<V *>(sx + p) = r[u];
    }
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
        const int t = threadIdx.x + u * kDerivThreads;
        if (t < ne * 9) sg[t] = T(g[u]);
    }
    __syncthreads();
#pragma unroll 2
    for (int u = 0; u < PER; ++u) {
        const int t = threadIdx.x + u * kDerivThreads;
        if (t >= C) break;
        int e, f, n;
        if (AOS) {
            const int q = (int)(((uint32_t)t * magicF) >> 20);  // t / F, exact for t < 2^20 / F
            f = t - q * F;
            e = q / NP;
            n = q - e * NP;
        } else {
            const int c = t / NP;
            n = t - c * NP;
            e = (int)(((uint32_t)c * magicF) >> 20);
            f = c - e * F;
        }
        const int pk = sijk[n];
        T dr, ds, dt;
        if (AOS) {
            const T *xb = sx + e * F * NP + f;
            grad_ref_ld<N, T>([xb, F](int m) { return xb[m * F]; }, n, pk & 255, (pk >> 8) & 255, pk >> 16, dr, ds, dt);
        } else {
            const T *xb = sx + (e * F + f) * NP;
            grad_ref_ld<N, T>([xb](int m) { return xb[m]; }, n, pk & 255, (pk >> 8) & 255, pk >> 16, dr, ds, dt);
        }
        const int64_t oa = base + t;
        if constexpr (GRAD) {
            const T *gg = sg + e * 9;
            o0[oa] = gg[0] * dr + gg[1] * ds + gg[2] * dt;
            o1[oa] = gg[3] * dr + gg[4] * ds + gg[5] * dt;
            o2[oa] = gg[6] * dr + gg[7] * ds + gg[8] * dt;
        } else {
            o0[oa] = axis == 0 ? dr : axis == 1 ? ds : dt;
        }
    }
}

// Field-major compact layout [K][F][Np]: lanes are nodes.  A wave owns one group of 64
// consecutive nodes and walks the block's columns (element, field) at a fixed stride, so
// the node's stencil offsets and weights are computed once per thread and every LDS read
// of a column is the thread's offset plus a compile-time multiple of Np.  The geometric
// factors of a column are wave-uniform (broadcast LDS reads).  The weights carry the 1/2
// of the reference stencil (an exact power-of-two scaling of every partial sum).
template <int N> struct NodeLanes {
    static constexpr int NP = np3(N);
    static constexpr int G = (NP + 63) / 64;                           // node groups
    static constexpr int WPG = G >= 4 ? 1 : (G == 3 ? 2 : 4 / G);      // waves per group
    static constexpr int NT = 64 * G * WPG;
    static constexpr bool fits = NT <= 1024;
};

template <int N, typename T, int VEC, int AX>  // AX = 0, 1, 2: one axis; 3: gradient
__global__ __launch_bounds__(NodeLanes<N>::fits ? NodeLanes<N>::NT : 64) void k_deriv_nodes(
    const T *__restrict__ x, T *__restrict__ o0, T *__restrict__ o1, T *__restrict__ o2,
    const double *__restrict__ geo, int F, int64_t K, int EB) {
    using NL = NodeLanes<N>;
    constexpr int NP = NL::NP, G = NL::G, WPG = NL::WPG, NT = NL::NT;
    constexpr int VALS = kChunkBytes / sizeof(T);
    constexpr int LPT = (VALS / VEC + NT - 1) / NT;
    constexpr int GS = 16 / sizeof(T) * ((9 * sizeof(T) + 15) / 16);  // padded factors per element
    using V = typename VecT<T, VEC>::type;
    __shared__ __align__(16) T sx[VALS];
    __shared__ __align__(16) T sg[AX == 3 ? kChunkMaxElems * GS : 1];
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int ne = (int)min((int64_t)EB, K - e0);
    const int C = ne * F * NP;
    const int64_t base = e0 * F * NP;
    V r[LPT];
#pragma unroll
    for (int u = 0; u < LPT; ++u) {
        const int p = (threadIdx.x + u * NT) * VEC;
        r[u] = *reinterpret_cast<const V *>(x + base + (p < C ? p : 0));
    }
    constexpr int GPT = AX == 3 ? (kChunkMaxElems * 9 + NT - 1) / NT : 0;
    double g[GPT > 0 ? GPT : 1];
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
        const int t = threadIdx.x + u * NT;
        g[u] = geo[e0 * 9 + (t < ne * 9 ? t : 0)];
    }
#pragma unroll
    for (int u = 0; u < LPT; ++u) {
        const int p = (threadIdx.x + u * NT) * VEC;
        if (p < C) *reinterpret_cast<V *>(sx + p) = r[u];
    }
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
        const int t = threadIdx.x + u * NT;
        if (t < ne * 9) sg[(t / 9) * GS + t % 9] = T(g[u]);
    }
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int grp = w % G, cw = w / G;
    const int n = grp * 64 + (threadIdx.x & 63);
    if (n >= NP) return;
    int i, j, k;
    decode3<N>(n, i, j, k);
    const int l = N - i - j - k;
    const int rl = N - k - j + 1, sk = (N - k + 1) * (N - k + 2) / 2, skm = (N - k + 2) * (N - k + 3) / 2;
    const int km0 = n - skm + j, kp0 = n + sk - j;
    const int o_im = i > 0 ? n - 1 : n, o_ip = l > 0 ? n + 1 : n;
    const int o_jm = j > 0 ? n - rl - 1 : n, o_jmip = j > 0 ? n - rl : n;
    const int o_jp = l > 0 ? n + rl : n, o_jpim = i > 0 ? n + rl - 1 : n;
    const int o_km = k > 0 ? km0 : n, o_kmip = k > 0 ? km0 + 1 : n, o_kmjp = k > 0 ? km0 + rl + 1 : n;
    const int o_kp = l > 0 ? kp0 : n, o_kpim = i > 0 ? kp0 - 1 : n, o_kpjm = j > 0 ? kp0 - rl : n;
    const T hi = T(0.5 * i), hj = T(0.5 * j), hk = T(0.5 * k), hl = T(0.5 * l);
    const int ncol = ne * F;
    constexpr int UF = sizeof(T) == 8 ? 1 : 4;  // measured: FP64 runs best rolled, FP32 unrolled by 4
#pragma unroll UF
    for (int c = cw; c < ncol; c += WPG) {
        const T *xb = sx + c * NP;
        const T xc = xb[n], xim = xb[o_im];
        T dr = T(0), ds = T(0), dt = T(0);
        if (AX == 0 || AX == 3) {
            const T xip = xb[o_ip], xjm = xb[o_jm], xjmip = xb[o_jmip], xkm = xb[o_km], xkmip = xb[o_kmip];
            dr += hi * (xc - xim);
            dr += hj * (xjmip - xjm);
            dr += hk * (xkmip - xkm);
            dr += hl * (xip - xc);
        }
        if (AX == 1 || AX == 3) {
            const T xjm = xb[o_jm], xjp = xb[o_jp], xjpim = xb[o_jpim], xkm = xb[o_km], xkmjp = xb[o_kmjp];
            ds += hi * (xjpim - xim);
            ds += hj * (xc - xjm);
            ds += hk * (xkmjp - xkm);
            ds += hl * (xjp - xc);
        }
        if (AX == 2 || AX == 3) {
            const T xjm = xb[o_jm], xkm = xb[o_km], xkp = xb[o_kp], xkpim = xb[o_kpim], xkpjm = xb[o_kpjm];
            dt += hi * (xkpim - xim);
            dt += hj * (xkpjm - xjm);
            dt += hk * (xc - xkm);
            dt += hl * (xkp - xc);
        }
        const int64_t oa = base + c * NP + n;
        if constexpr (AX == 3) {
            const T *gg = sg + (c / F) * GS;
            o0[oa] = gg[0] * dr + gg[1] * ds + gg[2] * dt;
            o1[oa] = gg[3] * dr + gg[4] * ds + gg[5] * dt;
            o2[oa] = gg[6] * dr + gg[7] * ds + gg[8] * dt;
        } else {
            o0[oa] = AX == 0 ? dr : AX == 1 ? ds : dt;
        }
    }
}

// 0 = not compact, 1 = field-major [K][F][Np], 2 = field-fastest [K][Np][F]
inline int compact_kind(bb_layout l, int F, int NP) {
    if (l.elem != (int64_t)F * NP) return 0;
    if (l.node == 1 && (l.field == NP || F == 1)) return 1;
    if (l.field == 1 && l.node == F) return 2;
    return 0;
}

template <int N, typename T, bool GRAD>
bool try_deriv_chunk(bb_ctx *ctx, int axis, int F, int64_t K, const void *x, bb_layout lx, const double *geo,
                     void *o0, void *o1, void *o2, bb_layout lo) {
    constexpr int NP = np3(N);
    constexpr int VALS = kChunkBytes / sizeof(T);
    constexpr int VW = 16 / sizeof(T);
    const int kind = compact_kind(lx, F, NP);
    if (!kind || kind != compact_kind(lo, F, NP) || F * NP > VALS) return false;
    if (lx.elem != lo.elem || lx.node != lo.node || lx.field != lo.field) return false;
    int EB = VALS / (F * NP);
    if (EB > kChunkMaxElems) EB = kChunkMaxElems;
    const uint32_t magicF = (uint32_t)(((1u << 20) + F - 1) / F);
    const bool vec = (F * NP) % VW == 0 && ((uintptr_t)x % 16) == 0;
    const dim3 grid(grid1d(K, EB));
    if constexpr (NodeLanes<N>::fits) {
        if (kind == 1) {
            auto gon = [&](auto vc, auto axc) {
                constexpr int VEC = decltype(vc)::value;
                constexpr int AX = decltype(axc)::value;
                k_deriv_nodes<N, T, VEC, AX><<<grid, NodeLanes<N>::NT, 0, ctx->stream>>>(
                    (const T *)x, (T *)o0, (T *)o1, (T *)o2, geo, F, K, EB);
            };
            auto ax = [&](auto vc) {
                if (GRAD) gon(vc, std::integral_constant<int, 3>{});
                else if (axis == 0) gon(vc, std::integral_constant<int, 0>{});
                else if (axis == 1) gon(vc, std::integral_constant<int, 1>{});
                else gon(vc, std::integral_constant<int, 2>{});
            };
            if (vec) ax(std::integral_constant<int, VW>{});
            else ax(std::integral_constant<int, 1>{});
            return true;
        }
    }
    auto go = [&](auto vc, auto ac) {
        constexpr int VEC = decltype(vc)::value;
        constexpr bool AOS = decltype(ac)::value;
        k_deriv_chunk<N, T, VEC, AOS, GRAD><<<grid, kDerivThreads, 0, ctx->stream>>>(
            (const T *)x, (T *)o0, (T *)o1, (T *)o2, geo, F, K, EB, magicF, axis);
    };
    using I1 = std::integral_constant<int, 1>;
    using IV = std::integral_constant<int, VW>;
    using BT = std::true_type;
    using BF = std::false_type;
    if (kind == 2 && F > 1) {
        if (vec) go(IV{}, BT{}); else go(I1{}, BT{});
    } else {
        if (vec) go(IV{}, BF{}); else go(I1{}, BF{});
    }
    return true;
}

template <typename T, bool GRAD>
int launch_deriv(bb_ctx *ctx, int axis, int N, int F, int64_t K, const void *x, bb_layout lx, const double *geo,
                 void *o0, void *o1, void *o2, bb_layout lo) {
    if (K == 0) return BB_OK;
    return dispatch_n<1, 20>(N, [&](auto nc) -> int {
        constexpr int NN = decltype(nc)::value;
        constexpr int NP = np3(NN);
        if (try_deriv_chunk<NN, T, GRAD>(ctx, axis, F, K, x, lx, geo, o0, o1, o2, lo)) {
            BB_LAUNCH_CHECK(ctx);
            return BB_OK;
        }
        int FB = F < kDerivTarget / NP ? F : (kDerivTarget / NP > 0 ? kDerivTarget / NP : 1);
        int EB = kDerivTarget / (FB * NP);
        if (EB < 1) EB = 1;
        dim3 grid(grid1d(K, EB), (unsigned)((F + FB - 1) / FB));
        size_t lds = (size_t)EB * FB * NP * sizeof(T) + NP * sizeof(int);
        k_deriv<NN, T, GRAD><<<grid, kDerivThreads, lds, ctx->stream>>>(
            (const T *)x, lay(lx), (T *)o0, (T *)o1, (T *)o2, lay(lo), geo, F, FB, K, EB, axis);
        BB_LAUNCH_CHECK(ctx);
        return BB_OK;
    });
}

// ElevationMatrix{N} * x  (degree N-1 -> N), elevation_multiply! order of operations
template <typename T>
__global__ void k_elevate(int N, const T *__restrict__ x, Lay lx, T *__restrict__ out, Lay lo, int F, int64_t K) {
    const int NF = nfp2(N);
    const int64_t total = K * F * NF;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(t % NF);
        const int64_t c = t / NF;
        const int64_t e = c / F;
        const int f = (int)(c % F);
        int a, b;
        decode2(N, m, a, b);
        const int cc = N - a - b;
        const T *xc = x + e * lx.e + (int64_t)f * lx.f;
        T v = T(0);
        if (a > 0) v = T(a) * xc[(int64_t)lin2(N - 1, a - 1, b) * lx.n];
        if (b > 0) v = fma(T(b), xc[(int64_t)lin2(N - 1, a, b - 1) * lx.n], v);
        if (cc > 0) v = fma(T(cc), xc[(int64_t)lin2(N - 1, a, b) * lx.n], v);
        out[e * lo.e + (int64_t)m * lo.n + (int64_t)f * lo.f] = v / T(N);
    }
}

// uM .= q[Fmask, :]
template <typename T>
__global__ void k_gather(int N, const T *__restrict__ q, Lay lq, T *__restrict__ uM, Lay lm, int F, int64_t K) {
    const int NF = nfp2(N);
    const int64_t total = K * F * 4 * NF;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(t % (4 * NF));
        const int64_t c = t / (4 * NF);
        const int64_t e = c / F;
        const int f = (int)(c % F);
        int a, b;
        decode2(N, m % NF, a, b);
        const int v = face_vidx(N, m / NF, a, b);
        uM[e * lm.e + (int64_t)m * lm.n + (int64_t)f * lm.f] = q[e * lq.e + (int64_t)v * lq.n + (int64_t)f * lq.f];
    }
}

struct Sign6 {
    double s[6];
};

template <int PDE>
__global__ void k_interface_flux(int kind, int N, const double *__restrict__ uM, Lay lm, const int32_t *__restrict__ mapP,
                                 const int8_t *__restrict__ bc, Sign6 sg, const double *__restrict__ nx,
                                 const double *__restrict__ ny, const double *__restrict__ nz, double *__restrict__ flux,
                                 Lay lf, int64_t K) {
    constexpr int F = Pde<PDE>::F;
    const int M4 = 4 * nfp2(N);
    const int64_t total = K * M4;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / M4;
        const int m = (int)(idx % M4);
        double um[F], d[F], o[F];
        const double *pm = uM + e * lm.e + (int64_t)m * lm.n;
#pragma unroll
        for (int f = 0; f < F; ++f) um[f] = pm[(int64_t)f * lm.f];
        if (bc && bc[idx]) {
#pragma unroll
            for (int f = 0; f < F; ++f) d[f] = sg.s[f] * um[f] - um[f];
        } else {
            int64_t p = mapP[idx];
            if (p < 0 || p >= total) p = idx;  // an invalid index reads the own trace (zero jump)
            const double *pp = uM + (p / M4) * lm.e + (int64_t)(p % M4) * lm.n;
#pragma unroll
            for (int f = 0; f < F; ++f) d[f] = pp[(int64_t)f * lm.f] - um[f];
        }
        const double n0 = nx[idx], n1 = ny[idx], n2 = nz[idx];
        const double Jf = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        Pde<PDE>::flux(kind, d, n0, n1, n2, Jf, 1.0 / Jf, o);
        double *po = flux + e * lf.e + (int64_t)m * lf.n;
#pragma unroll
        for (int f = 0; f < F; ++f) po[(int64_t)f * lf.f] = o[f];
    }
}

// ---- plain-path helpers (solver layout [K][F][Np], FP64)

// fx(q), fy(q), fz(q) at every node in one pass (the three SVector maps of
// examples/3d_wave.jl:59-61).  The direction is a compile-time constant in each call of
// Pde::fdir: a run-time three-way select over the directions was miscompiled for gfx950
// (the negated entries of the third direction were dropped), so no such select is used.
template <int PDE>
__global__ void k_pde_fdir3(int NP, const double *__restrict__ q, double *__restrict__ fx, double *__restrict__ fy,
                            double *__restrict__ fz, int64_t K) {
    constexpr int F = Pde<PDE>::F;
    const int64_t total = K * NP;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t / NP;
        const int n = (int)(t % NP);
        double qv[F], o[F];
#pragma unroll
        for (int f = 0; f < F; ++f) qv[f] = q[(e * F + f) * NP + n];
        Pde<PDE>::fdir(0, qv, o);
#pragma unroll
        for (int f = 0; f < F; ++f) fx[(e * F + f) * NP + n] = o[f];
        Pde<PDE>::fdir(1, qv, o);
#pragma unroll
        for (int f = 0; f < F; ++f) fy[(e * F + f) * NP + n] = o[f];
        Pde<PDE>::fdir(2, qv, o);
#pragma unroll
        for (int f = 0; f < F; ++f) fz[(e * F + f) * NP + n] = o[f];
    }
}

struct Ptr9 {
    const double *d[9];
};

// du -= (or +=) rxJ*dfxdr + sxJ*dfxds + ... + tzJ*dfzdt; du /= J  (examples/3d_wave.jl:75-83)
__global__ void k_plain_combine(bool plus, int FNP, double *__restrict__ du, Ptr9 d, const double *__restrict__ geo,
                                const double *__restrict__ J, int64_t K) {
    const int64_t total = K * FNP;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t / FNP;
        const double *g = geo + e * 9;
        double s = g[0] * d.d[0][t] + g[1] * d.d[1][t] + g[2] * d.d[2][t] + g[3] * d.d[3][t] + g[4] * d.d[4][t] +
                   g[5] * d.d[5][t] + g[6] * d.d[6][t] + g[7] * d.d[7][t] + g[8] * d.d[8][t];
        double v = plus ? du[t] + s : du[t] - s;
        du[t] = v / J[e];
    }
}

// res = a res + dt du (res not read in the first stage), q = q + b res.  Pairs of values
// per thread as 16-byte accesses (the arrays are allocations of the library, 16-byte
// aligned), the odd last value by thread 0; du and res are read once and res written
// once, as non-temporal accesses.
__global__ void k_lsrk_update(int64_t n, double *__restrict__ q, double *__restrict__ res, const double *__restrict__ du,
                              double a, double b, double dt, int first) {
    typedef double V2 __attribute__((ext_vector_type(2)));
    const int64_t n2 = n >> 1;
    V2 *q2 = reinterpret_cast<V2 *>(q), *r2 = reinterpret_cast<V2 *>(res);
    const V2 *d2 = reinterpret_cast<const V2 *>(du);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (int64_t)gridDim.x * blockDim.x) {
        const V2 d = __builtin_nontemporal_load(&d2[t]);
        V2 r = dt * d;
        if (!first) r = a * __builtin_nontemporal_load(&r2[t]) + r;
        __builtin_nontemporal_store(r, &r2[t]);
        q2[t] = q2[t] + b * r;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const int64_t t = n - 1;
        const double r = first ? dt * du[t] : a * res[t] + dt * du[t];
        res[t] = r;
        q[t] = q[t] + b * r;
    }
}

inline unsigned grid_cap(int64_t work) {
    int64_t g = (work + 255) / 256;
    return (unsigned)(g > 262144 ? 262144 : (g < 1 ? 1 : g));
}

}  // namespace bb

using namespace bb;

static int check_layout(bb_ctx *ctx, bb_layout l, const char *what) {
    if (l.elem < 0 || l.node < 0 || l.field < 0)
        return bb_fail(ctx, BB_ERR_ARG, "negative stride in layout of %s", what);
    return BB_OK;
}

extern "C" {

int bb_deriv_apply(bb_ctx *ctx, int axis, int N, int F, int64_t K, int dtype, const void *x, bb_layout lx, void *out,
                   bb_layout lo) {
    int rc = bb_check_common(ctx, N, F, K, 1, 20);
    if (rc) return rc;
    if (axis < 0 || axis > 2) return bb_fail(ctx, BB_ERR_ARG, "axis %d not in {0,1,2}", axis);
    if (K && (!x || !out)) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    if ((rc = check_layout(ctx, lx, "x")) || (rc = check_layout(ctx, lo, "out"))) return rc;
    if (dtype == BB_F64) return launch_deriv<double, false>(ctx, axis, N, F, K, x, lx, nullptr, out, out, out, lo);
    if (dtype == BB_F32) return launch_deriv<float, false>(ctx, axis, N, F, K, x, lx, nullptr, out, out, out, lo);
    return bb_fail(ctx, BB_ERR_ARG, "dtype %d", dtype);
}

int bb_grad_apply(bb_ctx *ctx, int N, int F, int64_t K, int dtype, const void *x, bb_layout lx, const double *geo,
                  void *dx, void *dy, void *dz, bb_layout lo) {
    int rc = bb_check_common(ctx, N, F, K, 1, 20);
    if (rc) return rc;
    if (K && (!x || !geo || !dx || !dy || !dz)) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    if ((rc = check_layout(ctx, lx, "x")) || (rc = check_layout(ctx, lo, "out"))) return rc;
    if (dtype == BB_F64) return launch_deriv<double, true>(ctx, 0, N, F, K, x, lx, geo, dx, dy, dz, lo);
    if (dtype == BB_F32) return launch_deriv<float, true>(ctx, 0, N, F, K, x, lx, geo, dx, dy, dz, lo);
    return bb_fail(ctx, BB_ERR_ARG, "dtype %d", dtype);
}

int bb_elevate_apply(bb_ctx *ctx, int N, int F, int64_t K, int dtype, const void *x, bb_layout lx, void *out,
                     bb_layout lo) {
    int rc = bb_check_common(ctx, N, F, K, 1, 20);
    if (rc) return rc;
    if (K && (!x || !out)) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    if ((rc = check_layout(ctx, lx, "x")) || (rc = check_layout(ctx, lo, "out"))) return rc;
    if (K == 0) return BB_OK;
    unsigned g = grid_cap(K * F * nfp2(N));
    if (dtype == BB_F64)
        k_elevate<double><<<g, 256, 0, ctx->stream>>>(N, (const double *)x, lay(lx), (double *)out, lay(lo), F, K);
    else if (dtype == BB_F32)
        k_elevate<float><<<g, 256, 0, ctx->stream>>>(N, (const float *)x, lay(lx), (float *)out, lay(lo), F, K);
    else
        return bb_fail(ctx, BB_ERR_ARG, "dtype %d", dtype);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

int bb_face_gather(bb_ctx *ctx, int N, int F, int64_t K, int dtype, const void *q, bb_layout lq, void *uM,
                   bb_layout lm) {
    int rc = bb_check_common(ctx, N, F, K, 1, 20);
    if (rc) return rc;
    if (K && (!q || !uM)) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    if ((rc = check_layout(ctx, lq, "q")) || (rc = check_layout(ctx, lm, "uM"))) return rc;
    if (K == 0) return BB_OK;
    unsigned g = grid_cap(K * F * 4 * nfp2(N));
    if (dtype == BB_F64)
        k_gather<double><<<g, 256, 0, ctx->stream>>>(N, (const double *)q, lay(lq), (double *)uM, lay(lm), F, K);
    else if (dtype == BB_F32)
        k_gather<float><<<g, 256, 0, ctx->stream>>>(N, (const float *)q, lay(lq), (float *)uM, lay(lm), F, K);
    else
        return bb_fail(ctx, BB_ERR_ARG, "dtype %d", dtype);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

int bb_interface_flux(bb_ctx *ctx, int pde, int flux_kind, int N, int64_t K, const double *uM, bb_layout lm,
                      const int32_t *mapP, const int8_t *bc, const double *bc_sign_host, const double *nxJ,
                      const double *nyJ, const double *nzJ, double *flux, bb_layout lf) {
    if (pde != BB_PDE_WAVE && pde != BB_PDE_MAXWELL) return bb_fail(ctx, BB_ERR_ARG, "pde %d", pde);
    int F = pde == BB_PDE_WAVE ? 4 : 6;
    int rc = bb_check_common(ctx, N, F, K, 1, 20);
    if (rc) return rc;
    if (flux_kind != BB_FLUX_REFERENCE && flux_kind != BB_FLUX_UPWIND) return bb_fail(ctx, BB_ERR_ARG, "flux %d", flux_kind);
    if (K && (!uM || !mapP || !nxJ || !nyJ || !nzJ || !flux)) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    if (bc && !bc_sign_host) return bb_fail(ctx, BB_ERR_NULL, "bc given without bc_sign");
    if (K == 0) return BB_OK;
    Sign6 sg{};
    for (int f = 0; f < F; ++f) sg.s[f] = bc ? bc_sign_host[f] : 1.0;
    unsigned g = grid_cap(K * 4 * nfp2(N));
    if (pde == BB_PDE_WAVE)
        k_interface_flux<BB_PDE_WAVE><<<g, 256, 0, ctx->stream>>>(flux_kind, N, uM, lay(lm), mapP, bc, sg, nxJ, nyJ, nzJ,
                                                                  flux, lay(lf), K);
    else
        k_interface_flux<BB_PDE_MAXWELL><<<g, 256, 0, ctx->stream>>>(flux_kind, N, uM, lay(lm), mapP, bc, sg, nxJ, nyJ,
                                                                     nzJ, flux, lay(lf), K);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

}  // extern "C"

int bb_pde_fdir3(bb_ctx *ctx, int pde, int N, int64_t K, const double *q, double *fx, double *fy, double *fz) {
    int NP = np3(N);
    unsigned g = grid_cap(K * NP);
    if (pde == BB_PDE_WAVE) k_pde_fdir3<BB_PDE_WAVE><<<g, 256, 0, ctx->stream>>>(NP, q, fx, fy, fz, K);
    else k_pde_fdir3<BB_PDE_MAXWELL><<<g, 256, 0, ctx->stream>>>(NP, q, fx, fy, fz, K);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

int bb_plain_combine(bb_ctx *ctx, int pde, int N, int64_t K, double *du, double *const d[9], const double *geo,
                     const double *J) {
    int F = pde == BB_PDE_WAVE ? 4 : 6;
    int FNP = F * np3(N);
    Ptr9 p;
    for (int m = 0; m < 9; ++m) p.d[m] = d[m];
    k_plain_combine<<<grid_cap(K * FNP), 256, 0, ctx->stream>>>(pde == BB_PDE_MAXWELL, FNP, du, p, geo, J, K);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

int bb_lsrk_update(bb_ctx *ctx, int64_t n, double *q, double *res, const double *du, double a, double b, double dt,
                   int first) {
    if ((((uintptr_t)q | (uintptr_t)res | (uintptr_t)du) & 15u) != 0)
        return bb_fail(ctx, BB_ERR_ARG, "the RK update needs 16-byte aligned arrays");
    k_lsrk_update<<<grid_cap((n + 1) / 2), 256, 0, ctx->stream>>>(n, q, res, du, a, b, dt, first);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}
