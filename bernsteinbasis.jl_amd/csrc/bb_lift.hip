// bb_lift.hip -- the O(N^3) Bernstein lift (Chan & Warburton factorisation).
//
// For each face of each column:
//   E_0 = L0 x            (L0 = (N+1)^2/2 E^T E, a 7-point stencil on the face)
//   E_s = (E_{N-s}^{N-s+1})^T E_{s-1}     s = 1..N   (degree reduction chain)
//   slab s of the face lift = l_s E_s      (l_s signed binomial weights)
// src/BernsteinLift.jl:139-164 (one face) and :191-220 (four faces + permuted sum).
//
// GPU organisation: a block owns C columns.  The chain runs in LDS with one barrier per
// degree; the chains of all 4 faces x C columns advance together, so a step has
// C*4*Nfp(d) independent items.  Only the current and next chain vectors are kept
// (two LDS buffers).  Each thread owns fixed volume nodes and, after every step, adds
// the slab values that land on its nodes (face f4 at slab k, f1 at slab j, f2 at slab l,
// f3 at slab i) into a private accumulator -- a deterministic gather, no atomics.
#include "bb_dispatch.hpp"
#include "bb_stencil.hpp"

namespace bb {

constexpr int kLiftThreads = 256;

template <int N> struct LiftCfg {
    static constexpr int NP = np3(N), NFP = nfp2(N);
    static constexpr int R = ((NP + kLiftThreads - 1) / kLiftThreads) > 4 ? ((NP + kLiftThreads - 1) / kLiftThreads) : 4;
    static constexpr int C = (kLiftThreads * R) / NP;  // columns per block (>= 1)
};

// dec[np3(d-1) + t] = (a | b << 8) for entry t of a degree-d triangle, d = 0..N
template <int N> __device__ __forceinline__ void build_dec(uint16_t *dec) {
    for (int t = threadIdx.x; t < np3(N); t += blockDim.x) {
        int d = 0, u = t;
        while (u >= nfp2(d)) { u -= nfp2(d); ++d; }
        int a, b;
        decode2(d, u, a, b);
        dec[t] = (uint16_t)(a | (b << 8));
    }
}

// per owned volume node: the slab (step) and face-node index where each face's chain lands
struct NodeFaces {
    int step[4];  // face order f1 f2 f3 f4
    int idx[4];
};
template <int N> __device__ __forceinline__ NodeFaces node_faces(int n) {
    int i, j, k;
    decode3<N>(n, i, j, k);
    const int l = N - i - j - k;
    NodeFaces r;
    r.step[0] = j; r.idx[0] = lin2(N - j, i, k);  // G1[(i,k,j)]
    r.step[1] = l; r.idx[1] = lin2(N - l, j, k);  // G2[(j,k,l)]
    r.step[2] = i; r.idx[2] = lin2(N - i, j, k);  // G3[(j,k,i)]
    r.step[3] = k; r.idx[3] = lin2(N - k, i, j);  // face 4, written first
    return r;
}

template <int N, typename TI, typename TC, typename TA>
__global__ __launch_bounds__(kLiftThreads) void k_lift4(const TI *__restrict__ x, Lay lx, TI *__restrict__ out, Lay lo,
                                                        int F, int64_t ncols) {
    using Cfg = LiftCfg<N>;
    constexpr int NP = Cfg::NP, NFP = Cfg::NFP, C = Cfg::C, R = Cfg::R;
    __shared__ TC bufA[C * 4 * NFP];
    __shared__ TC bufB[C * 4 * NFP];
    __shared__ uint16_t dec[NP];
    const int64_t c0 = (int64_t)blockIdx.x * C;
    const int nc = (int)min((int64_t)C, ncols - c0);
    build_dec<N>(dec);
    // load face data [c][4 Nfp] -> bufB
    for (int t = threadIdx.x; t < nc * 4 * NFP; t += kLiftThreads) {
        const int c = t / (4 * NFP), m = t % (4 * NFP);
        const int64_t col = c0 + c, e = col / F;
        const int f = (int)(col % F);
        bufB[t] = TC(x[e * lx.e + (int64_t)m * lx.n + (int64_t)f * lx.f]);
    }
    NodeFaces nf[R];
    TA acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int u = threadIdx.x + r * kLiftThreads;
        nf[r] = node_faces<N>(u < C * NP ? u % NP : 0);
        acc[r] = TA(0);
    }
    __syncthreads();
    // L0 on every face: bufB -> bufA
    for (int t = threadIdx.x; t < nc * 4 * NFP; t += kLiftThreads) {
        const int q = t / NFP, b = dec[np3(N - 1) + t % NFP];
        bufA[t] = l0_apply<N, TC>(bufB + q * NFP, b & 255, b >> 8);
    }
    __syncthreads();
    TC *cur = bufA, *nxt = bufB;
    auto accumulate = [&](int s, const TC *buf, double w) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int u = threadIdx.x + r * kLiftThreads;
            if (u < nc * NP) {
                const int c = u / NP;
                const TC *bc = buf + c * 4 * NFP;
                const int order[4] = {3, 0, 1, 2};
#pragma unroll
                for (int o = 0; o < 4; ++o) {
                    const int ff = order[o];
                    if (nf[r].step[ff] == s) acc[r] += TA(w) * TA(bc[ff * NFP + nf[r].idx[ff]]);
                }
            }
        }
    };
    accumulate(0, cur, 1.0);
#pragma unroll
    for (int s = 1; s <= N; ++s) {
        const int d = N - s;  // output degree
        const TC inv = TC(1) / TC(d + 1);
        for (int t = threadIdx.x; t < nc * 4 * nfp2(d); t += kLiftThreads) {
            const int q = t / nfp2(d), beta = t % nfp2(d);
            const int b = dec[np3(d - 1) + beta];
            nxt[q * NFP + beta] = reduce_apply<TC>(cur + q * NFP, d, b & 255, b >> 8, inv);
        }
        __syncthreads();
        accumulate(s, nxt, lj(N, s));
        TC *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int u = threadIdx.x + r * kLiftThreads;
        if (u < nc * NP) {
            const int64_t col = c0 + u / NP, e = col / F;
            const int f = (int)(col % F), n = u % NP;
            out[e * lo.e + (int64_t)n * lo.n + (int64_t)f * lo.f] = TI(acc[r]);
        }
    }
}

// Compact field-major layouts (x [K][F][4 Nfp], out [K][F][Np]): one block per element,
// kLiftW threads.  The element's face data are copied into LDS with 16-byte
// loads; L0 and every reduction step write their vectors into a per-(field, face) slab
// array [F][4][Np] (slab s at offset tet(N, s), i.e. the reference's G1..G3 arrays before
// the l_s weights); finally each lane gathers, for its volume nodes, the four face values
// in the order G4 + G1 + G2 + G3 with the l_s weights in the accumulation type TA.  Work
// items are (face, face node); the index and weight arithmetic of an item is shared by the
// F fields, which are processed four at a time (reads of four fields, then their stores).
constexpr int kLiftW = 128;  // measured: two waves per element beat one (64) and four (256)
constexpr int kLiftFB = 4;

template <int N, typename TI, typename TC, typename TA>
__global__ __launch_bounds__(kLiftW) void k_lift4w(const TI *__restrict__ x, TI *__restrict__ out, int F, int64_t K) {
    constexpr int NP = np3(N), NFP = nfp2(N);
    extern __shared__ __align__(16) unsigned char lift_smem[];
    TC *S = reinterpret_cast<TC *>(lift_smem);                                // [F][4][NP]
    TI *X = reinterpret_cast<TI *>(lift_smem + (size_t)F * 4 * NP * sizeof(TC));  // [F][4][NFP]
    __shared__ uint16_t dec[NP];
    __shared__ double lw[N + 1];
    const int tid = threadIdx.x;
    const int64_t e = blockIdx.x;
    // face data: F * 4 Nfp values, a multiple of 16 bytes for either type
    constexpr int VW = 16 / sizeof(TI);
    using V = std::conditional_t<sizeof(TI) == 8, double2, float4>;
    const int nv = F * 4 * NFP / VW;
    const V *xs = reinterpret_cast<const V *>(x + e * F * 4 * NFP);
    V *xd = reinterpret_cast<V *>(X);
    for (int t0 = 0; t0 < nv; t0 += 8 * kLiftW) {
        V r[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = t0 + tid + u * kLiftW;
            r[u] = xs[t < nv ? t : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = t0 + tid + u * kLiftW;
            if (t < nv) xd[t] = r[u];
        }
    }
    build_dec<N>(dec);
    if (tid <= N) {
#pragma unroll
        for (int q = 0; q <= N; ++q)
            if (tid == q) lw[q] = lj(N, q);
    }
    __syncthreads();
    // L0 on every face of every field -> slab 0
    for (int m = tid; m < 4 * NFP; m += kLiftW) {
        const int phi = m / NFP, beta = m - phi * NFP;
        const int ab = dec[np3(N - 1) + beta];
        const L0Pt<TC> pt = l0_prep<N, TC>(ab & 255, ab >> 8);
        for (int f0 = 0; f0 < F; f0 += kLiftFB) {
            TC v[kLiftFB];
#pragma unroll
            for (int u = 0; u < kLiftFB; ++u) {
                const int f = min(f0 + u, F - 1);
                const TI *xf = X + (f * 4 + phi) * NFP;
                TC s = TC(pt.w0) * TC(xf[pt.i0]);
                s += TC(pt.w1) * TC(xf[pt.i1]);
                s += TC(pt.w2) * TC(xf[pt.i2]);
                s += TC(pt.w3) * TC(xf[pt.i3]);
                s += TC(pt.w4) * TC(xf[pt.i4]);
                s += TC(pt.w5) * TC(xf[pt.i5]);
                s += TC(pt.w6) * TC(xf[pt.i6]);
                v[u] = TC(0.5) * s;
            }
#pragma unroll
            for (int u = 0; u < kLiftFB; ++u)
                if (f0 + u < F) S[((f0 + u) * 4 + phi) * NP + beta] = v[u];
        }
    }
    __syncthreads();
    // the reduction chain: slab s (degree N-s) from slab s-1 (degree N-s+1)
#pragma unroll
    for (int sl = 1; sl <= N; ++sl) {
        const int d = N - sl, nf = nfp2(d);
        const int so = tet(N, sl - 1), dof = tet(N, sl);
        const TC inv = TC(1) / TC(d + 1);
        for (int m = tid; m < 4 * nf; m += kLiftW) {
            const int phi = m / nf, beta = m - phi * nf;
            const int ab = dec[np3(d - 1) + beta];
            const int a = ab & 255, b = ab >> 8;
            const int idx = so + lin2(d + 1, a, b);
            const int o1 = idx + 1, o2 = idx + (d + 2 - b);
            const TC wa = TC(a + 1), wb = TC(b + 1), wc = TC(d + 1 - a - b);
            for (int f0 = 0; f0 < F; f0 += kLiftFB) {
                TC v[kLiftFB];
#pragma unroll
                for (int u = 0; u < kLiftFB; ++u) {
                    const int f = min(f0 + u, F - 1);
                    const TC *sf = S + (f * 4 + phi) * NP;
                    v[u] = fma(wa, sf[o1], fma(wb, sf[o2], wc * sf[idx])) * inv;
                }
#pragma unroll
                for (int u = 0; u < kLiftFB; ++u)
                    if (f0 + u < F) S[((f0 + u) * 4 + phi) * NP + dof + beta] = v[u];
            }
        }
        __syncthreads();
    }
    // gather: out[n] = G4 + G1 + G2 + G3 with the l_s weights
    TI *oe = out + e * F * NP;
    for (int n = tid; n < NP; n += kLiftW) {
        int i, j, k;
        decode3<N>(n, i, j, k);
        const int l = N - i - j - k;
        // faces in the order f4, f1, f2, f3: slab and in-slab index of this node
        const int s4 = k, s1 = j, s2 = l, s3 = i;
        const int p4 = 3 * NP + tet(N, s4) + lin2(N - k, i, j);
        const int p1 = 0 * NP + tet(N, s1) + lin2(N - j, i, k);
        const int p2 = 1 * NP + tet(N, s2) + lin2(N - l, j, k);
        const int p3 = 2 * NP + tet(N, s3) + lin2(N - i, j, k);
        const TA w4 = TA(lw[s4]), w1 = TA(lw[s1]), w2 = TA(lw[s2]), w3 = TA(lw[s3]);
        for (int f0 = 0; f0 < F; f0 += kLiftFB) {
            TA v[kLiftFB];
#pragma unroll
            for (int u = 0; u < kLiftFB; ++u) {
                const int f = min(f0 + u, F - 1);
                const TC *sf = S + f * 4 * NP;
                TA a = w4 * TA(sf[p4]);
                a += w1 * TA(sf[p1]);
                a += w2 * TA(sf[p2]);
                a += w3 * TA(sf[p3]);
                v[u] = a;
            }
#pragma unroll
            for (int u = 0; u < kLiftFB; ++u)
                if (f0 + u < F) oe[(f0 + u) * NP + n] = TI(v[u]);
        }
    }
}

// Compile-time index tables of degree N for the record kernel below:
//   ab[np3(d-1) + t] = a | b << 8 for entry t of a degree-d triangle (d = 0..N)
//   go[4 n + f]      = where face f's chain lands on volume node n: tet(N, slab) + index
//                      in the slab (f1 slab j at (i,k), f2 slab l at (j,k), f3 slab i at
//                      (j,k), f4 slab k at (i,j); src/BernsteinLift.jl:202-217)
//   sl[n]            = the four slabs j | l << 4 | i << 8 | k << 12
//   wt[4 n + f]      = lw_chain(N, slab of face f at n), faces f1 f2 f3 f4 (k_lift4x's gather)
template <int N> struct LiftTab {
    alignas(16) double wt[4 * np3(N)];
    uint16_t ab[np3(N)];
    uint16_t go[4 * np3(N)];
    uint16_t sl[np3(N)];
    constexpr LiftTab() : wt{}, ab{}, go{}, sl{} {
        int r = 0;
        for (int d = 0; d <= N; ++d)
            for (int b = 0; b <= d; ++b)
                for (int a = 0; a <= d - b; ++a) ab[r++] = (uint16_t)(a | (b << 8));
        int n = 0;
        for (int k = 0; k <= N; ++k)
            for (int j = 0; j <= N - k; ++j)
                for (int i = 0; i <= N - j - k; ++i, ++n) {
                    const int l = N - i - j - k;
                    go[4 * n + 0] = (uint16_t)(tet(N, j) + lin2(N - j, i, k));
                    go[4 * n + 1] = (uint16_t)(tet(N, l) + lin2(N - l, j, k));
                    go[4 * n + 2] = (uint16_t)(tet(N, i) + lin2(N - i, j, k));
                    go[4 * n + 3] = (uint16_t)(tet(N, k) + lin2(N - k, i, j));
                    sl[n] = (uint16_t)(j | (l << 4) | (i << 8) | (k << 12));
                    wt[4 * n + 0] = lw_chain(N, j);
                    wt[4 * n + 1] = lw_chain(N, l);
                    wt[4 * n + 2] = lw_chain(N, i);
                    wt[4 * n + 3] = lw_chain(N, k);
                }
    }
};
template <int N> __device__ constexpr LiftTab<N> kLiftTab = LiftTab<N>();

// Compact field-major layouts with F (1..4, 6) as a template parameter and 7 <= N <= 10: one
// block of four waves per element, wave f owning face f's chain for all F fields.  The
// face data become LDS records of F values (field fastest; FP64 with an even F in planes of
// two fields), so that a stencil point is one address and one to three 16-byte accesses for
// all fields.  The chain is unnormalised (slab
// 0 = 2 L0 x, no 1/(d+1) per step; lw_chain carries both factors), each lane keeps its node
// (a,b) of the degree N-1 triangle for the whole chain and steps its offsets by
// increments, and the steps of a chain need only wave-level ordering.  Every slab of every
// chain stays in LDS; each thread then gathers its volume nodes' four face values in the
// order G4 + G1 + G2 + G3 with the weights in TA.
constexpr int kLiftR = 256;
template <int N, int F, typename TI, typename TC> struct LiftRCfg {
    static constexpr int NP = np3(N), NFP = nfp2(N);
    // FP64 with an even field count: planes of two fields (16 bytes per node and plane);
    // otherwise whole records
    static constexpr int W = (sizeof(TC) == 8 && F % 2 == 0) ? 2 : F;
    static constexpr size_t lds = (size_t)4 * (NP + NFP) * F * sizeof(TC) + (N + 1) * 8;
    // below N = 7 the four waves per element leave most lanes idle and the two-wave kernel
    // (k_lift4w) is faster (measured at N = 3, 5: 1.5x)
    static constexpr bool valid = N >= 7 && N <= 10 && lds <= 64 * 1024;
};

template <int N, int F, typename TI, typename TC, typename TA>
__global__ __launch_bounds__(kLiftR) void k_lift4r(const TI *__restrict__ x, TI *__restrict__ out, int64_t K) {
    constexpr int NP = np3(N), NFP = nfp2(N), MI = (NFP + 63) / 64, W = LiftRCfg<N, F, TI, TC>::W;
    constexpr int PSS = 4 * NP * W, PSX = 4 * NFP * W;  // plane strides
    const LiftTab<N> &tb = kLiftTab<N>;
    __shared__ __align__(16) TC S[4 * NP * F];   // [plane][face][slab node][W fields]
    __shared__ __align__(16) TC X[4 * NFP * F];  // [plane][face][face node][W fields]
    __shared__ double lw[N + 1];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);  // this wave's face
    const int64_t e = blockIdx.x;
    // per-lane table entries, read first (L2-resident)
    int l0ab[MI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) {
        const int beta = mi * 64 + lane;
        l0ab[mi] = beta < NFP ? tb.ab[np3(N - 1) + beta] : 0;
    }
    const bool lact = lane < nfp2(N - 1);
    const int chab = lact ? tb.ab[np3(N - 2) + lane] : 0;
    // face w's data of every field -> records
    const TI *xe = x + e * F * 4 * NFP + w * NFP;
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) {
        const int m = mi * 64 + lane;
        if (m < NFP) {
            TC v[F];
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = TC(xe[f * 4 * NFP + m]);
            plane_store<F, W>(X + __mul24(w * NFP + m, W), PSX, v);
        }
    }
    if (tid <= N) {
#pragma unroll
        for (int s = 0; s <= N; ++s)
            if (tid == s) lw[s] = lw_chain(N, s);
    }
    wave_lds_sync();
    // L0 (without its 1/2) -> slab 0 of chain w
    const TC *xw = X + w * NFP * W;
    TC *sw = S + w * NP * W;
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) {
        const int beta = mi * 64 + lane;
        if (beta < NFP) {
            const L0Rec<TC> pt = l0_prep_rec<N, W, TC>(l0ab[mi] & 255, l0ab[mi] >> 8);
            TC v[F], y[F];
            plane_load<F, W>(xw + pt.o0, PSX, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = pt.w0 * y[f];
            plane_load<F, W>(xw + pt.o1, PSX, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w1, y[f], v[f]);
            plane_load<F, W>(xw + pt.o2, PSX, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w2, y[f], v[f]);
            plane_load<F, W>(xw + pt.o3, PSX, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w3, y[f], v[f]);
            plane_load<F, W>(xw + pt.o4, PSX, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w4, y[f], v[f]);
            plane_load<F, W>(xw + pt.o5, PSX, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w5, y[f], v[f]);
            plane_load<F, W>(xw + pt.o6, PSX, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w6, y[f], v[f]);
            plane_store<F, W>(sw + __mul24(beta, W), PSS, v);
        }
    }
    wave_lds_sync();
    // the chain: (a+1) x_{a+1,b} + (b+1) x_{a,b+1} + (d+1-a-b) x_{ab} (reduction_multiply!
    // without /(d+1)), lane -> (a,b) fixed, offsets stepped by increments
    {
        const int a = chab & 255, b = chab >> 8;
        const int bF = __mul24(b, W);
        int idxF = __mul24(lin2(N, a, b), W), dstF = __mul24(lin2(N - 1, a, b), W);
        int o2F = idxF + __mul24(N + 1 - b, W);
        TC wc = TC(N - a - b);
        const TC wa = TC(a + 1), wb = TC(b + 1);
#pragma unroll
        for (int s = 1; s <= N; ++s) {
            if (lact && a + b <= N - s) {
                const TC *src = sw + tet(N, s - 1) * W;
                TC x1[F], x2[F], x0[F], v[F];
                plane_load<F, W>(src + idxF + W, PSS, x1);
                plane_load<F, W>(src + o2F, PSS, x2);
                plane_load<F, W>(src + idxF, PSS, x0);
#pragma unroll
                for (int f = 0; f < F; ++f) v[f] = fma(wa, x1[f], fma(wb, x2[f], wc * x0[f]));
                plane_store<F, W>(sw + tet(N, s) * W + dstF, PSS, v);
            }
            wave_lds_sync();
            idxF -= bF;
            dstF -= bF;
            o2F -= bF + W;
            wc -= TC(1);
        }
    }
    __syncthreads();
    // gather: out[n] = G4 + G1 + G2 + G3 with the slab weights
    TI *oe = out + e * F * NP;
    for (int n = tid; n < NP; n += kLiftR) {
        const uint2 gp = reinterpret_cast<const uint2 *>(tb.go)[n];
        const int go[4] = {(int)(gp.x & 0xffff), (int)(gp.x >> 16), (int)(gp.y & 0xffff), (int)(gp.y >> 16)};
        const int sl = tb.sl[n];
        const int slab[4] = {sl & 15, (sl >> 4) & 15, (sl >> 8) & 15, sl >> 12};
        TA acc[F];
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = TA(0);
        constexpr int order[4] = {3, 0, 1, 2};
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const int ff = order[o];
            const TA wt = TA(lw[slab[ff]]);
            TC v[F];
            plane_load<F, W>(S + ff * NP * W + __mul24(go[ff], W), PSS, v);
#pragma unroll
            for (int f = 0; f < F; ++f) acc[f] = fma(wt, TA(v[f]), acc[f]);
        }
#pragma unroll
        for (int f = 0; f < F; ++f) oe[f * NP + n] = TI(acc[f]);
    }
}

// Compact field-major layouts, F as a template parameter, 5 <= N <= 9 (nfp(N) <= 64): face
// data and every slab as LDS records and a final gather, as in k_lift4r, organised to issue
// few instructions per byte, which is what bounds this lift (counters: the LDS costs about
// the same time per instruction whatever its width, and once its reads are cut the VALU
// issue binds).
//  - Lane L keeps face node (a,b) = L of a triangle for the whole chain and its node's
//    current value in registers: of the three points of a chain step only (a,b+1) is read
//    from LDS, (a,b) is the lane's own value and (a+1,b) that of lane L+1, taken by a DPP
//    wavefront shift.  L0 likewise takes (a-1,b) and (a+1,b) from lanes L-1 and L+1 and
//    reads the other four stencil points.  A step is one record store and one record read
//    (three reads and a store in k_lift4r).
//  - The chains are packed as they shrink: one chain per wave while a step's output has
//    more than 32 nodes, then two chains per wave, then all four in one wave (16-lane
//    groups), each phase re-reading its first step's points from the slab records.
//  - Two waves per element (faces w and w + 2 until the chains are packed), so that twice
//    as many elements as with four waves are in flight where the LDS allows it.
//  - The gather's slab weights come per node from the constant table.
// Lanes outside a triangle, and lanes whose node has left the shrinking chain, compute
// finite values that no lane in the chain reads with a non-zero weight and that are never
// stored; their LDS reads go to a valid record.
template <int N, int F, typename TI, typename TC> struct LiftXCfg {
    static constexpr int NP = np3(N), NFP = nfp2(N);
    static constexpr int W = LiftRCfg<N, F, TI, TC>::W;
    static constexpr size_t lds = (size_t)4 * NP * F * sizeof(TC);  // the face records live in the slabs
    static constexpr bool valid = N >= 5 && NFP <= 64 && lds <= 64 * 1024;
};

// 2 L0 = (N+1)^2 (E_N^{N+1})^T E_N^{N+1}, the dense Nfp x Nfp matrix of the 7-point stencil
// of l0_prep_rec (integer entries, exact in FP32), as A operands of v_mfma_f32_16x16x4f32:
// row tile mt, lane L, k step kk holds L0[16 mt + L % 16][4 kk + L / 16] at a[(mt 64 + L) KT + kk]
// (zero outside the matrix); klo/khi bound each row tile's k steps that hold non-zero entries.
template <int N> struct L0MTab {
    static constexpr int NFP = nfp2(N), MT = (NFP + 15) / 16, KT = (NFP + 3) / 4;
    float a[MT * 64 * KT];
    int klo[MT], khi[MT];
    constexpr L0MTab() : a{}, klo{}, khi{} {
        for (int mt = 0; mt < MT; ++mt) {
            klo[mt] = KT;
            khi[mt] = -1;
        }
        for (int b = 0; b <= N; ++b)
            for (int aa = 0; aa <= N - b; ++aa) {
                const int c = N - aa - b, r = lin2(N, aa, b);
                const int pts[7][3] = {{aa, b, (aa + 1) * (aa + 1) + (b + 1) * (b + 1) + (c + 1) * (c + 1)},
                                       {aa + 1, b - 1, (aa + 1) * b},     {aa + 1, b, (aa + 1) * c},
                                       {aa - 1, b + 1, (b + 1) * aa},     {aa, b + 1, (b + 1) * c},
                                       {aa - 1, b, (c + 1) * aa},         {aa, b - 1, (c + 1) * b}};
                for (int q = 0; q < 7; ++q) {
                    if (pts[q][2] == 0) continue;
                    const int col = lin2(N, pts[q][0], pts[q][1]);
                    const int mt = r / 16, kk = col / 4, L = (r % 16) + 16 * (col % 4);
                    a[(mt * 64 + L) * KT + kk] += (float)pts[q][2];
                    klo[mt] = kk < klo[mt] ? kk : klo[mt];
                    khi[mt] = kk > khi[mt] ? kk : khi[mt];
                }
            }
    }
};
template <int N> __device__ constexpr L0MTab<N> kL0MTab = L0MTab<N>();

// chain steps of k_lift4x, packed as the chain shrinks: one chain per wave while a step's
// output has more than 32 nodes, then two chains per wave (32-lane halves, waves 0 and 1),
// then all four chains in wave 0 (16-lane groups) once it has at most 16
template <int N> struct LiftXPhases {
    static constexpr int last_step_over(int g) {
        int s = 0;
        while (s + 1 <= N && nfp2(N - s - 1) > g) ++s;
        return s;
    }
    static constexpr int A = last_step_over(32);  // steps 1..A: one chain per wave
    static constexpr int B = last_step_over(16);  // steps A+1..B: two chains per wave
};

// Steps s0..s1 of one chain in a group of lanes whose lane g holds node (a,b) = g of the
// degree-(N - s0) triangle (the output of step s0): step s0 reads its three points from
// the records of slab s0 - 1, later steps keep (a,b) in registers and take (a+1,b) from
// lane g+1 (a DPP wavefront shift: an active lane's right neighbour is in its own group).
template <int N, int F, int W, int S0, int S1, typename TC>
__device__ __forceinline__ void lift_chain_group(TC *sw, int PSS, int g, const LiftTab<N> &tb) {
    constexpr int d0 = N - S0;
    const bool valid = g < nfp2(d0);
    const int ab = valid ? tb.ab[np3(d0 - 1) + g] : 0;
    const int a = ab & 255, b = ab >> 8;
    const TC wa = TC(a + 1), wb = TC(b + 1);
    TC wc = TC(d0 + 1 - a - b);
    int idx = lin2(d0 + 1, a, b);  // (a,b) in the input triangle of the current step
    TC xs[F];
#pragma unroll
    for (int s = S0; s <= S1; ++s) {
        const int d = N - s;
        const bool act = valid && a + b <= d;
        const int o2 = act ? idx + d + 2 - b : 0;  // (a,b+1); inactive lanes read slab record 0
        const TC *src = sw + tet(N, s - 1) * W;
        TC x2[F];
        if (s == S0) {
            TC x0[F], x1[F];
            plane_load<F, W>(src + __mul24(act ? idx : 0, W), PSS, x0);
            plane_load<F, W>(src + __mul24(act ? idx + 1 : 0, W), PSS, x1);
            plane_load<F, W>(src + __mul24(o2, W), PSS, x2);
#pragma unroll
            for (int f = 0; f < F; ++f) xs[f] = fma(wa, x1[f], fma(wb, x2[f], wc * x0[f]));
        } else {
            wave_lds_sync();
            plane_load<F, W>(src + __mul24(o2, W), PSS, x2);
#pragma unroll
            for (int f = 0; f < F; ++f) xs[f] = fma(wa, xl_next(xs[f]), fma(wb, x2[f], wc * xs[f]));
        }
        if (act) plane_store<F, W>(sw + __mul24(tet(N, s) + idx - b, W), PSS, xs);
        idx -= b;
        wc -= TC(1);
    }
}

template <int N, int F, typename TI, typename TC, typename TA, int NW, bool L0M = false>
__global__ __launch_bounds__(64 * NW) void k_lift4x(const TI *__restrict__ x, TI *__restrict__ out, int64_t K) {
    constexpr int NP = np3(N), NFP = nfp2(N), W = LiftXCfg<N, F, TI, TC>::W;
    constexpr int PSS = 4 * NP * W;  // plane stride
    constexpr int PA = LiftXPhases<N>::A, PB = LiftXPhases<N>::B;
    constexpr int FQ = 4 / NW;  // faces per wave until the chains are packed
    static_assert(NFP <= 64 && (NW == 2 || NW == 4), "one lane per face node, 2 or 4 waves");
    static_assert(!L0M || (std::is_same_v<TC, float> && F == 4), "matrix-core L0: FP32, 4 fields (16 columns)");
    const LiftTab<N> &tb = kLiftTab<N>;
    __shared__ __align__(16) TC S[4 * NP * F];  // [plane][face][slab node][W fields]
    // face f's data, record m, sit at S + (f NP + Nfp + m) W: in the face's own slab 1..
    // records, which only its chain writes, and only after the face's L0 has read them
    // (2 Nfp <= Np for N >= 3)
    TC *const X = S + NFP * W;
    static_assert(2 * NFP <= NP, "face records inside slabs 1..");
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid / 64);  // faces w, w + NW, ...
    const int64_t e = blockIdx.x;
    const bool node = lane < NFP;
    const int ab = node ? tb.ab[np3(N - 1) + lane] : 0;
    const int a = ab & 255, b = ab >> 8;
    // the wave's faces' data of every field: registers (one face node per lane) and LDS records
    TC y[FQ][F];
#pragma unroll
    for (int q = 0; q < FQ; ++q) {
        const TI *xe = x + e * F * 4 * NFP + (w + q * NW) * NFP + (node ? lane : 0);
#pragma unroll
        for (int f = 0; f < F; ++f) {
            const TC v = TC(xe[f * 4 * NFP]);
            y[q][f] = node ? v : TC(0);
        }
    }
#pragma unroll
    for (int q = 0; q < FQ; ++q)
        if (node) plane_store<F, W>(X + __mul24((w + q * NW) * NP + lane, W), PSS, y[q]);
    wave_lds_sync();
    // L0 (without its 1/2): the 7-point stencil of l0_prep_rec, same weights and order.
    // (Values of lanes outside a triangle are never read with a non-zero weight: a lane in
    // the chain reads only nodes that were in it one step earlier.  They stay finite -- a
    // few steps of bounded weights -- and are never stored.)
    if constexpr (L0M) {
        // slab 0 of all four chains as one product 2 L0 X on the matrix cores: the element's
        // 16 columns (face, field) of X are the B operand, read from the face records; wave
        // w computes the row tiles mt = w, w + NW, .. over the k steps that hold non-zero
        // entries of L0 (the stencil couples face nodes at most N+2 apart)
        typedef float v4f __attribute__((ext_vector_type(4)));
        constexpr int MT = L0MTab<N>::MT, KT = L0MTab<N>::KT;
        const L0MTab<N> &lt = kL0MTab<N>;
        const int col = lane & 15, fc = col / F, fld = col % F;
        __syncthreads();  // the face records of every wave
        for (int mt = w; mt < MT; mt += NW) {
            v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
            const float *ap = lt.a + (mt * 64 + lane) * KT;
            for (int kk = lt.klo[mt]; kk <= lt.khi[mt]; ++kk) {
                const int k = 4 * kk + (lane >> 4);
                const float bv = k < NFP ? X[__mul24(fc * NP + k, W) + fld] : 0.f;
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[kk], bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * mt + 4 * (lane >> 4) + r;
                if (row < NFP) S[__mul24(fc * NP + row, W) + fld] = acc[r];
            }
        }
        __syncthreads();  // slab 0 of every chain
    }
    const L0Rec<TC> pt = l0_prep_rec<N, W, TC>(a, b);
#pragma unroll
    for (int q = 0; q < FQ; ++q) {
        const int fc = w + q * NW;
        const TC *xw = X + fc * NP * W;
        TC *sw = S + fc * NP * W;
        TC xs[F];
        if constexpr (L0M) {
            // the chain continues from this lane's node of slab 0 (needed in registers only
            // when the first steps run one chain per wave)
            if constexpr (PA >= 1) plane_load<F, W>(sw + __mul24(lane, W), PSS, xs);
        } else {
            TC y1[F], y3[F], y4[F], y6[F];
            plane_load<F, W>(xw + pt.o1, PSS, y1);
            plane_load<F, W>(xw + pt.o3, PSS, y3);
            plane_load<F, W>(xw + pt.o4, PSS, y4);
            plane_load<F, W>(xw + pt.o6, PSS, y6);
#pragma unroll
            for (int f = 0; f < F; ++f) {
                TC v = pt.w0 * y[q][f];
                v = fma(pt.w1, y1[f], v);
                v = fma(pt.w2, xl_next(y[q][f]), v);
                v = fma(pt.w3, y3[f], v);
                v = fma(pt.w4, y4[f], v);
                v = fma(pt.w5, xl_prev(y[q][f]), v);
                xs[f] = fma(pt.w6, y6[f], v);
            }
            if (node) plane_store<F, W>(sw + __mul24(lane, W), PSS, xs);
        }
        // the chain: (a+1) x_{a+1,b} + (b+1) x_{a,b+1} + (d+1-a-b) x_{ab}
        // (reduction_multiply! without /(d+1)).  Steps 1..PA: one chain per wave, lanes
        // still on the degree-N layout.
        if constexpr (PA >= 1) {
            const int bF = __mul24(b, W);
            int dstF = __mul24(lin2(N - 1, a, b), W);
            int o2F = __mul24(lin2(N, a, b) + N + 1 - b, W);
            TC wc = TC(N - a - b);
            const TC wa = TC(a + 1), wb = TC(b + 1);
#pragma unroll
            for (int s = 1; s <= PA; ++s) {
                const bool act = node && a + b <= N - s;
                wave_lds_sync();
                TC x2[F];
                plane_load<F, W>(sw + tet(N, s - 1) * W + (act ? o2F : 0), PSS, x2);
#pragma unroll
                for (int f = 0; f < F; ++f) xs[f] = fma(wa, xl_next(xs[f]), fma(wb, x2[f], wc * xs[f]));
                if (act) plane_store<F, W>(sw + tet(N, s) * W + dstF, PSS, xs);
                dstF -= bF;
                o2F -= bF + W;
                wc -= TC(1);
            }
        }
    }
    // steps PA+1..PB: two chains per wave (waves 0, 1; chain 2 w + half)
    if constexpr (PB > PA) {
        __syncthreads();
        if (w < 2)
            lift_chain_group<N, F, W, PA + 1, PB, TC>(S + (2 * w + (lane >> 5)) * NP * W, PSS, lane & 31, tb);
    }
    // steps PB+1..N: the four chains in the 16-lane groups of wave 0
    if constexpr (N > PB) {
        __syncthreads();
        if (w == 0) lift_chain_group<N, F, W, PB + 1, N, TC>(S + (lane >> 4) * NP * W, PSS, lane & 15, tb);
    }
    __syncthreads();
    // gather: out[n] = G4 + G1 + G2 + G3 with the slab weights (a per-node table)
    TI *oe = out + e * F * NP;
    for (int n = tid; n < NP; n += 64 * NW) {
        const uint2 gp = reinterpret_cast<const uint2 *>(tb.go)[n];
        const int go[4] = {(int)(gp.x & 0xffff), (int)(gp.x >> 16), (int)(gp.y & 0xffff), (int)(gp.y >> 16)};
        const double2 w01 = reinterpret_cast<const double2 *>(tb.wt)[2 * n];
        const double2 w23 = reinterpret_cast<const double2 *>(tb.wt)[2 * n + 1];
        const double wf[4] = {w01.x, w01.y, w23.x, w23.y};
        TA acc[F];
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = TA(0);
        constexpr int order[4] = {3, 0, 1, 2};
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const int ff = order[o];
            const TA wt = TA(wf[ff]);
            TC v[F];
            plane_load<F, W>(S + ff * NP * W + __mul24(go[ff], W), PSS, v);
#pragma unroll
            for (int f = 0; f < F; ++f) acc[f] = fma(wt, TA(v[f]), acc[f]);
        }
#pragma unroll
        for (int f = 0; f < F; ++f) oe[f * NP + n] = TI(acc[f]);
    }
}

template <int N, typename TI, typename TC, typename TA, bool L0M = false>
bool try_lift4x(bb_ctx *ctx, int F, int64_t K, const void *x, bb_layout lx, void *out, bb_layout lo) {
    constexpr int NP = np3(N), NFP = nfp2(N);
    if (lx.elem != (int64_t)F * 4 * NFP || lx.node != 1 || (F > 1 && lx.field != 4 * NFP)) return false;
    if (lo.elem != (int64_t)F * NP || lo.node != 1 || (F > 1 && lo.field != NP)) return false;
    bool done = false;
    auto one = [&](auto fc) {
        constexpr int FF = decltype(fc)::value;
        if constexpr (LiftXCfg<N, FF, TI, TC>::valid && (!L0M || FF == 4)) {
            if (F == FF) {
                // two waves per element: twice the elements of four-wave blocks in flight where
                // the LDS allows it (FP32 N=7 F=4: 518 -> 352 us), no slower elsewhere
                k_lift4x<N, FF, TI, TC, TA, 2, L0M><<<(unsigned)K, 128, 0, ctx->stream>>>((const TI *)x, (TI *)out, K);
                done = true;
            }
        }
    };
    one(std::integral_constant<int, 1>{});
    one(std::integral_constant<int, 2>{});
    one(std::integral_constant<int, 3>{});
    one(std::integral_constant<int, 4>{});
    one(std::integral_constant<int, 6>{});
    return done;
}

template <int N, typename TI, typename TC, typename TA>
bool try_lift4r(bb_ctx *ctx, int F, int64_t K, const void *x, bb_layout lx, void *out, bb_layout lo) {
    constexpr int NP = np3(N), NFP = nfp2(N);
    if (lx.elem != (int64_t)F * 4 * NFP || lx.node != 1 || (F > 1 && lx.field != 4 * NFP)) return false;
    if (lo.elem != (int64_t)F * NP || lo.node != 1 || (F > 1 && lo.field != NP)) return false;
    bool done = false;
    auto one = [&](auto fc) {
        constexpr int FF = decltype(fc)::value;
        if constexpr (LiftRCfg<N, FF, TI, TC>::valid) {
            if (F == FF) {
                k_lift4r<N, FF, TI, TC, TA><<<(unsigned)K, kLiftR, 0, ctx->stream>>>((const TI *)x, (TI *)out, K);
                done = true;
            }
        }
    };
    one(std::integral_constant<int, 1>{});
    one(std::integral_constant<int, 2>{});
    one(std::integral_constant<int, 3>{});
    one(std::integral_constant<int, 4>{});
    one(std::integral_constant<int, 6>{});
    return done;
}

template <int N, typename TI, typename TC, typename TA>
bool try_lift4w(bb_ctx *ctx, int F, int64_t K, const void *x, bb_layout lx, void *out, bb_layout lo) {
    constexpr int NP = np3(N), NFP = nfp2(N);
    if (lx.elem != (int64_t)F * 4 * NFP || lx.node != 1 || (F > 1 && lx.field != 4 * NFP)) return false;
    if (lo.elem != (int64_t)F * NP || lo.node != 1 || (F > 1 && lo.field != NP)) return false;
    if ((uintptr_t)x % 16) return false;
    const size_t lds = (size_t)F * 4 * (NP * sizeof(TC) + NFP * sizeof(TI));
    if (lds + NP * 2 + (N + 1) * 8 > 64 * 1024) return false;
    k_lift4w<N, TI, TC, TA><<<(unsigned)K, kLiftW, lds, ctx->stream>>>((const TI *)x, (TI *)out, F, K);
    return true;
}

// ---- float-float (compensated) arithmetic for the chain of the conditioning study
struct FF {
    float hi, lo;
};
__device__ __forceinline__ FF ff_two_sum(float a, float b) {
    float s = a + b;
    float bb_ = s - a;
    float e = (a - (s - bb_)) + (b - bb_);
    return FF{s, e};
}
__device__ __forceinline__ FF ff_quick(float a, float b) {
    float s = a + b;
    return FF{s, b - (s - a)};
}
__device__ __forceinline__ FF ff_add(FF x, FF y) {
    FF s = ff_two_sum(x.hi, y.hi);
    float e = s.lo + x.lo + y.lo;
    return ff_quick(s.hi, e);
}
__device__ __forceinline__ FF ff_mul_int(FF x, float w) {  // w a small integer, exact in float
    float p = w * x.hi;
    float e = fmaf(w, x.hi, -p);
    return ff_quick(p, fmaf(w, x.lo, e));
}
__device__ __forceinline__ FF ff_div_int(FF x, float D) {
    float q = x.hi / D;
    float r = fmaf(-q, D, x.hi);  // exact remainder
    float ql = (r + x.lo) / D;
    return ff_quick(q, ql);
}
__device__ __forceinline__ FF ff_mul_d(FF x, double w) {
    float wh = (float)w, wl = (float)(w - (double)wh);
    float p = wh * x.hi;
    float e = fmaf(wh, x.hi, -p);
    e = fmaf(wh, x.lo, fmaf(wl, x.hi, e));
    return ff_quick(p, e);
}

template <typename TC> struct Chain {
    __device__ static TC from_f(float v) { return TC(v); }
    __device__ static TC from_d(double v) { return TC(v); }
    __device__ static TC reduce(const TC *x, int d, int a, int b) {
        return reduce_apply<TC>(x, d, a, b, TC(1) / TC(d + 1));
    }
    __device__ static double scaled(TC v, double w) { return (double)(TC(w) * v); }
};
template <> struct Chain<FF> {
    __device__ static FF from_f(float v) { return FF{v, 0.f}; }
    __device__ static FF from_d(double v) {
        float h = (float)v;
        return FF{h, (float)(v - (double)h)};
    }
    __device__ static FF reduce(const FF *x, int d, int a, int b) {
        const int idx = lin2(d + 1, a, b);
        FF s = ff_mul_int(x[idx + 1], float(a + 1));
        s = ff_add(s, ff_mul_int(x[idx + (d + 2 - b)], float(b + 1)));
        s = ff_add(s, ff_mul_int(x[idx], float(d + 1 - a - b)));
        return ff_div_int(s, float(d + 1));
    }
    __device__ static double scaled(FF v, double w) {
        FF p = ff_mul_d(v, w);
        return (double)p.hi + (double)p.lo;
    }
};
template <> struct Chain<float> {
    __device__ static float from_f(float v) { return v; }
    __device__ static float from_d(double v) { return (float)v; }
    __device__ static float reduce(const float *x, int d, int a, int b) {
        return reduce_apply<float>(x, d, a, b, 1.0f / float(d + 1));
    }
    __device__ static double scaled(float v, double w) { return (double)((float)w * v); }
};

// single face: TL = L0 arithmetic type, TC = chain type, TO = output type
template <int N, typename TI, typename TL, typename TC, typename TO>
__global__ __launch_bounds__(kLiftThreads) void k_face_lift(const TI *__restrict__ x, Lay lx, TO *__restrict__ out,
                                                            Lay lo, int F, int64_t ncols, int C) {
    constexpr int NP = np3(N), NFP = nfp2(N);
    extern __shared__ double smem_d[];
    TC *bufA = reinterpret_cast<TC *>(smem_d);
    TC *bufB = bufA + C * NFP;
    TL *xin = reinterpret_cast<TL *>(bufB + C * NFP);
    uint16_t *dec = reinterpret_cast<uint16_t *>(xin + C * NFP);
    const int64_t c0 = (int64_t)blockIdx.x * C;
    const int nc = (int)min((int64_t)C, ncols - c0);
    build_dec<N>(dec);
    for (int t = threadIdx.x; t < nc * NFP; t += kLiftThreads) {
        const int c = t / NFP, m = t % NFP;
        const int64_t col = c0 + c, e = col / F;
        const int f = (int)(col % F);
        xin[t] = TL(x[e * lx.e + (int64_t)m * lx.n + (int64_t)f * lx.f]);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nc * NFP; t += kLiftThreads) {
        const int c = t / NFP, beta = t % NFP;
        const int b = dec[np3(N - 1) + beta];
        const TL v = l0_apply<N, TL>(xin + c * NFP, b & 255, b >> 8);
        bufA[t] = Chain<TC>::from_d((double)v);
        const int64_t col = c0 + c, e = col / F;
        const int f = (int)(col % F);
        out[e * lo.e + (int64_t)beta * lo.n + (int64_t)f * lo.f] = TO(v);
    }
    __syncthreads();
    TC *cur = bufA, *nxt = bufB;
#pragma unroll
    for (int s = 1; s <= N; ++s) {
        const int d = N - s;
        const double w = lj(N, s);
        for (int t = threadIdx.x; t < nc * nfp2(d); t += kLiftThreads) {
            const int c = t / nfp2(d), beta = t % nfp2(d);
            const int b = dec[np3(d - 1) + beta];
            const TC v = Chain<TC>::reduce(cur + c * NFP, d, b & 255, b >> 8);
            nxt[c * NFP + beta] = v;
            const int64_t col = c0 + c, e = col / F;
            const int f = (int)(col % F);
            out[e * lo.e + (int64_t)(tet(N, s) + beta) * lo.n + (int64_t)f * lo.f] = TO(Chain<TC>::scaled(v, w));
        }
        __syncthreads();
        TC *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
}

template <int N, typename TI, typename TC, typename TA, bool L0M = false>
int launch_lift4(bb_ctx *ctx, int F, int64_t K, const void *x, bb_layout lx, void *out, bb_layout lo) {
    if (K <= 0x7fffffff && (try_lift4x<N, TI, TC, TA, L0M>(ctx, F, K, x, lx, out, lo) ||
                            try_lift4r<N, TI, TC, TA>(ctx, F, K, x, lx, out, lo) ||
                            try_lift4w<N, TI, TC, TA>(ctx, F, K, x, lx, out, lo))) {
        BB_LAUNCH_CHECK(ctx);
        return BB_OK;
    }
    const int64_t ncols = K * F;
    k_lift4<N, TI, TC, TA><<<grid1d(ncols, LiftCfg<N>::C), kLiftThreads, 0, ctx->stream>>>(
        (const TI *)x, lay(lx), (TI *)out, lay(lo), F, ncols);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

template <int N, typename TI, typename TL, typename TC, typename TO>
int launch_face_lift(bb_ctx *ctx, int F, int64_t K, const void *x, bb_layout lx, void *out, bb_layout lo) {
    constexpr int NFP = nfp2(N);
    const int64_t ncols = K * F;
    int C = 2048 / NFP;
    if (C < 1) C = 1;
    size_t lds = (size_t)C * NFP * (2 * sizeof(TC) + sizeof(TL)) + np3(N) * sizeof(uint16_t) + 16;
    k_face_lift<N, TI, TL, TC, TO><<<grid1d(ncols, C), kLiftThreads, lds, ctx->stream>>>(
        (const TI *)x, lay(lx), (TO *)out, lay(lo), F, ncols, C);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

}  // namespace bb

using namespace bb;

extern "C" {

int bb_lift_apply(bb_ctx *ctx, int N, int F, int64_t K, int dtype, int mode, const void *x, bb_layout lx, void *out,
                  bb_layout lo) {
    int rc = bb_check_common(ctx, N, F, K, 1, 20);
    if (rc) return rc;
    if (K && (!x || !out)) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    if (lx.elem < 0 || lx.node < 0 || lx.field < 0 || lo.elem < 0 || lo.node < 0 || lo.field < 0)
        return bb_fail(ctx, BB_ERR_ARG, "negative stride");
    if (K == 0) return BB_OK;
    return dispatch_n<1, 20>(N, [&](auto nc) -> int {
        constexpr int NN = decltype(nc)::value;
        if (dtype == BB_F64) return launch_lift4<NN, double, double, double>(ctx, F, K, x, lx, out, lo);
        if (dtype != BB_F32) return bb_fail(ctx, BB_ERR_ARG, "dtype %d", dtype);
        if (mode == BB_LIFT_PLAIN) return launch_lift4<NN, float, float, float>(ctx, F, K, x, lx, out, lo);
        if (mode == BB_LIFT_ACC64) return launch_lift4<NN, float, float, double>(ctx, F, K, x, lx, out, lo);
        if (mode == BB_LIFT_ACC64_MFMA) return launch_lift4<NN, float, float, double, true>(ctx, F, K, x, lx, out, lo);
        if (mode == BB_LIFT_TWOSUM)
            return bb_fail(ctx, BB_ERR_UNSUPPORTED, "two-sum mode is implemented for the single-face lift only");
        return bb_fail(ctx, BB_ERR_ARG, "lift mode %d", mode);
    });
}

int bb_face_lift_apply(bb_ctx *ctx, int N, int F, int64_t K, int x_dtype, int out_dtype, int mode, const void *x,
                       bb_layout lx, void *out, bb_layout lo) {
    int rc = bb_check_common(ctx, N, F, K, 1, 20);
    if (rc) return rc;
    if (K && (!x || !out)) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    if (lx.elem < 0 || lx.node < 0 || lx.field < 0 || lo.elem < 0 || lo.node < 0 || lo.field < 0)
        return bb_fail(ctx, BB_ERR_ARG, "negative stride");
    if ((x_dtype != BB_F64 && x_dtype != BB_F32) || (out_dtype != BB_F64 && out_dtype != BB_F32))
        return bb_fail(ctx, BB_ERR_ARG, "dtype %d/%d", x_dtype, out_dtype);
    if (mode < BB_LIFT_PLAIN || mode > BB_LIFT_ACC64_MFMA) return bb_fail(ctx, BB_ERR_ARG, "lift mode %d", mode);
    if (K == 0) return BB_OK;
    return dispatch_n<1, 20>(N, [&](auto nc) -> int {
        constexpr int NN = decltype(nc)::value;
        auto go = [&](auto to) -> int {
            using TO = decltype(to);
            if (x_dtype == BB_F64) return launch_face_lift<NN, double, double, double, TO>(ctx, F, K, x, lx, out, lo);
            if (mode == BB_LIFT_PLAIN) return launch_face_lift<NN, float, float, float, TO>(ctx, F, K, x, lx, out, lo);
            if (mode == BB_LIFT_ACC64 || mode == BB_LIFT_ACC64_MFMA)
                return launch_face_lift<NN, float, float, double, TO>(ctx, F, K, x, lx, out, lo);
            return launch_face_lift<NN, float, float, FF, TO>(ctx, F, K, x, lx, out, lo);
        };
        return out_dtype == BB_F64 ? go(double(0)) : go(float(0));
    });
}

}  // extern "C"
