// bb_stage.hpp -- one fused kernel per low-storage RK stage:
//
//   trace gather + interface flux -> 4-face lift (L0 + reduction chain) -> volume
//   derivatives + geometric combine -> / J -> res = a res + dt rhs -> q_out = q_in + b res
//
// i.e. examples/3d_wave.jl:36-84 (or 3d_maxwell.jl:180-237) followed by one stage of a
// five-stage 2N-storage RK45, for a block of EB whole elements per workgroup.
//
// Data in HBM (per element, all contiguous, element-major):
//   q_in, q_out [K][F][Np] FP64 (ping-pong: neighbours read the old stage while the new
//   one is written), res [K][F][Np] (FP64, or FP32 in mixed precision), geo [K][9],
//   J [K], nJ [K][4][3], conn [K][4] = {neighbour, face | perm << 2 | bc << 5}.
//
// LDS (T = FP64, or FP32 in mixed precision):
//   s1: the element's coefficients, later chain buffer A
//   s2: interface flux (4 faces x Nfp per field), later chain buffer B
// Each thread owns one volume node: its F volume terms and F lift accumulators stay in
// registers (accumulators always FP64 -- the signed binomially weighted sum of the four
// faces' chains, where cancellation happens), so no LDS scatter and no atomics.
#pragma once
#include <type_traits>

#include "bb_pde.hpp"
#include "bb_stencil.hpp"

namespace bb {

struct StageArgs {
    const double *q_in;
    double *q_out;
    void *res;
    const double *geo;
    const double *J;
    const double *nJ;
    const int2 *conn;
    int64_t K;
    double ca, cb, dt;
    int first;     // stage 1: res is not read (a_1 = 0)
    int rhs_only;  // write rhs into res, leave q alone
    int flux_kind;
    double sign[6];
};

template <int N, int PDE, bool MIXED, int EB> struct StageCfg {
    using T = std::conditional_t<MIXED, float, double>;
    static constexpr int F = Pde<PDE>::F;
    static constexpr int NP = np3(N), NFP = nfp2(N);
    static constexpr int NT = ((EB * NP + 63) / 64) * 64;
    static constexpr int R1 = (EB * F * NP > EB * F * 4 * NFP) ? EB * F * NP : EB * F * 4 * NFP;
    static constexpr int R2 = EB * F * 4 * NFP;
    static constexpr size_t lds = (size_t)(R1 + R2) * sizeof(T) + (size_t)NP * 2;
    static constexpr bool valid = NT <= 1024 && lds <= 64 * 1024;
};

template <int N> struct FaceSteps {
    int step[4];
    int idx[4];
};

template <int N, int PDE, bool MIXED, int EB>
__global__ __launch_bounds__((StageCfg<N, PDE, MIXED, EB>::NT)) void k_stage(StageArgs A) {
    using Cfg = StageCfg<N, PDE, MIXED, EB>;
    using T = typename Cfg::T;
    using P = Pde<PDE>;
    constexpr int F = Cfg::F, NP = Cfg::NP, NFP = Cfg::NFP, NT = Cfg::NT;
    __shared__ T s1[Cfg::R1];
    __shared__ T s2[Cfg::R2];
    __shared__ uint16_t dec[NP];

    const int tid = threadIdx.x;
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int ne = (int)min((int64_t)EB, A.K - e0);
    const bool own = tid < ne * NP;
    const int el = own ? tid / NP : 0;
    const int n = own ? tid % NP : 0;
    const int64_t eg = e0 + el;

    for (int t = tid; t < NP; t += NT) {
        int d = 0, u = t;
        while (u >= nfp2(d)) { u -= nfp2(d); ++d; }
        int a, b;
        decode2(d, u, a, b);
        dec[t] = (uint16_t)(a | (b << 8));
    }
    // ---- load the element's state (coalesced: consecutive threads, consecutive nodes)
    double qv[F];
    if (own) {
#pragma unroll
        for (int f = 0; f < F; ++f) {
            qv[f] = A.q_in[(eg * F + f) * NP + n];
            s1[(el * F + f) * NP + n] = T(qv[f]);
        }
    }
    __syncthreads();

    // ---- volume term: 3 reference derivatives per field, geometric combine, PDE flux Jacobian
    T vol[F];
#pragma unroll
    for (int f = 0; f < F; ++f) vol[f] = T(0);
    int i = 0, j = 0, k = 0;
    decode3<N>(n, i, j, k);
    if (own) {
        const double *g = A.geo + eg * 9;
        const T g0 = T(g[0]), g1 = T(g[1]), g2 = T(g[2]), g3 = T(g[3]), g4 = T(g[4]), g5 = T(g[5]), g6 = T(g[6]),
                g7 = T(g[7]), g8 = T(g[8]);
#pragma unroll
        for (int f = 0; f < F; ++f) {
            T dr, ds, dt;
            grad_ref<N, T>(s1 + (el * F + f) * NP, n, i, j, k, dr, ds, dt);
            const T gx = g0 * dr + g1 * ds + g2 * dt;
            const T gy = g3 * dr + g4 * ds + g5 * dt;
            const T gz = g6 * dr + g7 * ds + g8 * dt;
            P::template vol_add<T>(f, gx, gy, gz, vol);
        }
    }

    // ---- interface flux at every face node of the block's elements -> s2
    for (int t = tid; t < ne * 4 * NFP; t += NT) {
        const int e2 = t / (4 * NFP), phi = (t / NFP) & 3, m = t % NFP;
        const int ab = dec[np3(N - 1) + m];
        const int a = ab & 255, b = ab >> 8, c = N - a - b;
        const int vi = face_vidx(N, phi, a, b);
        T um[F], d[F], o[F];
#pragma unroll
        for (int f = 0; f < F; ++f) um[f] = s1[(e2 * F + f) * NP + vi];
        const int64_t ef = e0 + e2;
        const int2 cn = A.conn[ef * 4 + phi];
        if (((cn.y >> 5) & 7) == 0) {
            const int nface = cn.y & 3, p = (cn.y >> 2) & 7;
            const int abc[3] = {a, b, c};
            const int nvi = face_vidx(N, nface, abc[perm_entry(p, 0)], abc[perm_entry(p, 1)]);
            const double *qn = A.q_in + (int64_t)cn.x * F * NP + nvi;
#pragma unroll
            for (int f = 0; f < F; ++f) d[f] = T(qn[f * NP]) - um[f];
        } else {
#pragma unroll
            for (int f = 0; f < F; ++f) d[f] = T(A.sign[f]) * um[f] - um[f];
        }
        const double *nj = A.nJ + (ef * 4 + phi) * 3;
        P::template flux<T>(A.flux_kind, d, T(nj[0]), T(nj[1]), T(nj[2]), o);
#pragma unroll
        for (int f = 0; f < F; ++f) s2[((e2 * F + f) * 4 + phi) * NFP + m] = o[f];
    }
    __syncthreads();

    // ---- L0 on all chains (element, field, face): s2 -> s1
    for (int t = tid; t < ne * F * 4 * NFP; t += NT) {
        const int q = t / NFP, beta = t % NFP;
        const int ab = dec[np3(N - 1) + beta];
        s1[q * NFP + beta] = l0_apply<N, T>(s2 + q * NFP, ab & 255, ab >> 8);
    }
    // where each face's chain lands on this thread's node
    const int l = N - i - j - k;
    const int st[4] = {j, l, i, k};
    const int ix[4] = {lin2(N - j, i, k), lin2(N - l, j, k), lin2(N - i, j, k), lin2(N - k, i, j)};
    double acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0;
    __syncthreads();

    auto accumulate = [&](int s, const T *buf, double w) {
        if (!own) return;
        const int order[4] = {3, 0, 1, 2};  // face 4 first, then G1, G2, G3 (BernsteinLift.jl:202-217)
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const int ff = order[o];
            if (st[ff] == s) {
#pragma unroll
                for (int f = 0; f < F; ++f) acc[f] += w * (double)buf[((el * F + f) * 4 + ff) * NFP + ix[ff]];
            }
        }
    };
    accumulate(0, s1, 1.0);
    T *cur = s1, *nxt = s2;
#pragma unroll
    for (int s = 1; s <= N; ++s) {
        const int dd = N - s;
        const T inv = T(1) / T(dd + 1);
        for (int t = tid; t < ne * F * 4 * nfp2(dd); t += NT) {
            const int q = t / nfp2(dd), beta = t % nfp2(dd);
            const int ab = dec[np3(dd - 1) + beta];
            nxt[q * NFP + beta] = reduce_apply<T>(cur + q * NFP, dd, ab & 255, ab >> 8, inv);
        }
        __syncthreads();
        accumulate(s, nxt, lj(N, s));
        T *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }

    // ---- epilogue: rhs = (lift + volume) / J, then the RK stage
    if (own) {
        const double Je = A.J[eg];
        using TR = T;
        TR *res = reinterpret_cast<TR *>(A.res);
#pragma unroll
        for (int f = 0; f < F; ++f) {
            const double rhs = (acc[f] + (double)vol[f]) / Je;
            const int64_t g = (eg * F + f) * NP + n;
            if (A.rhs_only) {
                res[g] = TR(rhs);
            } else {
                const double r = A.first ? A.dt * rhs : A.ca * (double)res[g] + A.dt * rhs;
                const TR rs = TR(r);
                res[g] = rs;
                A.q_out[g] = qv[f] + A.cb * (double)rs;
            }
        }
    }
}

// launch one stage; returns BB_ERR_UNSUPPORTED for an invalid (N, EB) combination
using StageLaunchFn = int (*)(int N, int eb, const StageArgs &a, hipStream_t st);
int stage_launch_wave_f64(int N, int eb, const StageArgs &a, hipStream_t st);
int stage_launch_wave_mixed(int N, int eb, const StageArgs &a, hipStream_t st);
int stage_launch_maxwell_f64(int N, int eb, const StageArgs &a, hipStream_t st);
int stage_launch_maxwell_mixed(int N, int eb, const StageArgs &a, hipStream_t st);

constexpr int kStageMaxN = 10;
constexpr int kStageEB[] = {1, 2, 3, 4, 6, 8};

}  // namespace bb

#define BB_STAGE_DEFINE(NAME, PDE, MIXED)                                                              \
    namespace bb {                                                                                     \
    int NAME(int N, int eb, const StageArgs &a, hipStream_t st) {                                     \
        const int64_t K = a.K;                                                                         \
        return dispatch_n<1, kStageMaxN>(N, [&](auto nc) -> int {                                      \
            constexpr int NN = decltype(nc)::value;                                                    \
            int rc = BB_ERR_UNSUPPORTED;                                                               \
            auto one = [&](auto ebc) {                                                                 \
                constexpr int E = decltype(ebc)::value;                                                \
                if constexpr (StageCfg<NN, PDE, MIXED, E>::valid) {                                    \
                    if (eb == E) {                                                                     \
                        k_stage<NN, PDE, MIXED, E><<<grid1d(K, E), StageCfg<NN, PDE, MIXED, E>::NT, 0, st>>>(a); \
                        rc = hipGetLastError() == hipSuccess ? BB_OK : BB_ERR_HIP;                     \
                    }                                                                                  \
                }                                                                                      \
            };                                                                                         \
            one(std::integral_constant<int, 1>{});                                                     \
            one(std::integral_constant<int, 2>{});                                                     \
            one(std::integral_constant<int, 3>{});                                                     \
            one(std::integral_constant<int, 4>{});                                                     \
            one(std::integral_constant<int, 6>{});                                                     \
            one(std::integral_constant<int, 8>{});                                                     \
            return rc;                                                                                 \
        });                                                                                            \
    }                                                                                                  \
    }
