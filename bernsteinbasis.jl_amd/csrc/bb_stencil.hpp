// bb_stencil.hpp -- the barycentric derivative stencils and face-lift building blocks.
#pragma once
#include <type_traits>

#include "bb_common.hpp"

namespace bb {

// LDS written by one lane and read by another lane of the same wave: a wave executes its
// LDS instructions in order; the fences keep the compiler from moving accesses across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// linear index -> (i,j,k) for degree N (k outer, j, i inner)
template <int N> __device__ __forceinline__ void decode3(int n, int &i, int &j, int &k) {
    int kk = 0;
    while (n >= nfp2(N - kk)) { n -= nfp2(N - kk); ++kk; }
    int jj = 0;
    while (n >= N - kk - jj + 1) { n -= N - kk - jj + 1; ++jj; }
    i = n; j = jj; k = kk;
}
// linear index -> (a,b) for a degree-n triangle (b outer, a inner)
__host__ __device__ __forceinline__ void decode2(int n, int m, int &a, int &b) {
    int bb_ = 0;
    while (m >= n - bb_ + 1) { m -= n - bb_ + 1; ++bb_; }
    a = m; b = bb_;
}

// All three reference-coordinate derivatives of one coefficient vector x (degree N,
// linear order) at node idx = (i,j,k,l):
//   Dr = 1/2 [ i(x_a - x_{a-e0+e3}) + j(x_{a-e1+e0} - x_{a-e1+e3}) + k(...) + l(x_{a-e3+e0} - x_a) ]
// exactly the four weighted differences of fast_Dr_multiply! (src/BernsteinDerivativeMatrix_3D.jl:96-132),
// fast_Ds_multiply! (:138-174) and fast_Dt_multiply! (:180-216), accumulated in the same
// order.  A term whose weight is zero reads the centre value instead of an index that
// does not exist (branch-free), and contributes 0 * finite = 0.  ld(m) returns the
// coefficient of linear node m (a strided LDS read in the compact-layout kernels).
template <int N, typename T, typename LD>
__device__ __forceinline__ void grad_ref_ld(LD ld, int idx, int i, int j, int k, T &dr, T &ds, T &dt) {
    const int l = N - i - j - k;
    const int rl = N - k - j + 1;                    // length of row j in slab k
    const int sk = (N - k + 1) * (N - k + 2) / 2;    // size of slab k
    const int skm = (N - k + 2) * (N - k + 3) / 2;   // size of slab k-1
    const T xc = ld(idx);
    const T xim = ld(i > 0 ? idx - 1 : idx);          // (i-1,j,k)
    const T xip = ld(l > 0 ? idx + 1 : idx);          // (i+1,j,k)
    const T xjm = ld(j > 0 ? idx - rl - 1 : idx);     // (i,j-1,k)
    const T xjmip = ld(j > 0 ? idx - rl : idx);       // (i+1,j-1,k)
    const T xjp = ld(l > 0 ? idx + rl : idx);         // (i,j+1,k)
    const T xjpim = ld(i > 0 ? idx + rl - 1 : idx);   // (i-1,j+1,k)
    const int km0 = idx - skm + j;                    // (i,j,k-1)
    const T xkm = ld(k > 0 ? km0 : idx);
    const T xkmip = ld(k > 0 ? km0 + 1 : idx);        // (i+1,j,k-1)
    const T xkmjp = ld(k > 0 ? km0 + rl + 1 : idx);   // (i,j+1,k-1)
    const int kp0 = idx + sk - j;                     // (i,j,k+1)
    const T xkp = ld(l > 0 ? kp0 : idx);
    const T xkpim = ld(i > 0 ? kp0 - 1 : idx);        // (i-1,j,k+1)
    const T xkpjm = ld(j > 0 ? kp0 - rl : idx);       // (i,j-1,k+1)
    const T ti = T(i), tj = T(j), tk = T(k), tl = T(l), h = T(0.5);
    T v = T(0);
    v += ti * (xc - xim);
    v += tj * (xjmip - xjm);
    v += tk * (xkmip - xkm);
    v += tl * (xip - xc);
    dr = h * v;
    v = T(0);
    v += ti * (xjpim - xim);
    v += tj * (xc - xjm);
    v += tk * (xkmjp - xkm);
    v += tl * (xjp - xc);
    ds = h * v;
    v = T(0);
    v += ti * (xkpim - xim);
    v += tj * (xkpjm - xjm);
    v += tk * (xc - xkm);
    v += tl * (xkp - xc);
    dt = h * v;
}
template <int N, typename T>
__device__ __forceinline__ void grad_ref(const T *x, int idx, int i, int j, int k, T &dr, T &ds, T &dt) {
    grad_ref_ld<N, T>([x](int m) { return x[m]; }, idx, i, j, k, dr, ds, dt);
}

// L0 x at face node (a,b) of degree N:  L0 = (N+1)^2/2 (E_N^{N+1})^T E_N^{N+1}
// (src/BernsteinLift.jl:141-145), which is the 7-point stencil
//   1/2 [ ((a+1)^2+(b+1)^2+(c+1)^2) x_{ab} + (a+1) b x_{a+1,b-1} + (a+1) c x_{a+1,b}
//         + (b+1) a x_{a-1,b+1} + (b+1) c x_{a,b+1} + (c+1) a x_{a-1,b} + (c+1) b x_{a,b-1} ]
// (degree-N face layout; a weight is zero exactly when its neighbour does not exist).
template <int N, typename T> __device__ __forceinline__ T l0_apply(const T *x, int a, int b) {
    const int c = N - a - b;
    const int idx = lin2(N, a, b);
    const int rl = N - b + 1;  // row b length
    const T x0 = x[idx];
    const T xap_bm = x[b > 0 ? idx - rl : idx];       // (a+1,b-1)
    const T xap = x[c > 0 ? idx + 1 : idx];           // (a+1,b)
    const T xam_bp = x[a > 0 ? idx + rl - 1 : idx];   // (a-1,b+1)
    const T xbp = x[c > 0 ? idx + rl : idx];          // (a,b+1)
    const T xam = x[a > 0 ? idx - 1 : idx];           // (a-1,b)
    const T xbm = x[b > 0 ? idx - rl - 1 : idx];      // (a,b-1)
    T s = T((a + 1) * (a + 1) + (b + 1) * (b + 1) + (c + 1) * (c + 1)) * x0;
    s += T((a + 1) * b) * xap_bm;
    s += T((a + 1) * c) * xap;
    s += T((b + 1) * a) * xam_bp;
    s += T((b + 1) * c) * xbp;
    s += T((c + 1) * a) * xam;
    s += T((c + 1) * b) * xbm;
    return T(0.5) * s;
}

// The same L0 stencil split into an index/weight part (computed once per face node) and
// an evaluation part (repeated for every field that shares the face node).
template <typename T> struct L0Pt {
    int i0, i1, i2, i3, i4, i5, i6;
    T w0, w1, w2, w3, w4, w5, w6;
};
template <int N, typename T> __device__ __forceinline__ L0Pt<T> l0_prep(int a, int b) {
    const int c = N - a - b;
    const int idx = lin2(N, a, b);
    const int rl = N - b + 1;
    L0Pt<T> p;
    p.i0 = idx;
    p.i1 = b > 0 ? idx - rl : idx;
    p.i2 = c > 0 ? idx + 1 : idx;
    p.i3 = a > 0 ? idx + rl - 1 : idx;
    p.i4 = c > 0 ? idx + rl : idx;
    p.i5 = a > 0 ? idx - 1 : idx;
    p.i6 = b > 0 ? idx - rl - 1 : idx;
    p.w0 = T((a + 1) * (a + 1) + (b + 1) * (b + 1) + (c + 1) * (c + 1));
    p.w1 = T((a + 1) * b);
    p.w2 = T((a + 1) * c);
    p.w3 = T((b + 1) * a);
    p.w4 = T((b + 1) * c);
    p.w5 = T((c + 1) * a);
    p.w6 = T((c + 1) * b);
    return p;
}
template <typename T> __device__ __forceinline__ T l0_eval(const T *x, const L0Pt<T> &p) {
    T s = p.w0 * x[p.i0];
    s += p.w1 * x[p.i1];
    s += p.w2 * x[p.i2];
    s += p.w3 * x[p.i3];
    s += p.w4 * x[p.i4];
    s += p.w5 * x[p.i5];
    s += p.w6 * x[p.i6];
    return T(0.5) * s;
}

// The F fields of one node stored next to each other (LDS records of F values, RS values
// apart): one address per node, read and written as 16-byte words where the record stride
// and the access are multiples of 16 bytes, else as 8-byte words where they are multiples
// of 8, else value by value (p is then aligned to that word size)
template <int F, int RS, typename T> constexpr int recv_word() {
    constexpr int B = RS * (int)sizeof(T), A = F * (int)sizeof(T);
    return (B % 16 == 0 && A % 16 == 0) ? 16 / (int)sizeof(T) : (B % 8 == 0 && A % 8 == 0) ? 8 / (int)sizeof(T) : 1;
}
// Record stride for F values of type T.  Records of 32 bytes (four FP64 fields) meet
// two-way LDS bank conflicts under 16-byte accesses; padding them to 40 bytes removes the
// conflicts but halves the access width, and measured slower (the 4-face lift at N=7:
// 1087 -> 1350 us), so records are not padded.
template <typename T> constexpr int rec_stride(int F) { return F; }
template <int F, int RS = F, typename T> __device__ __forceinline__ void recv_load(const T *p, T (&v)[F]) {
    constexpr int W = recv_word<F, RS, T>();
    if constexpr (W == 1 || F % W != 0) {
#pragma unroll
        for (int f = 0; f < F; ++f) v[f] = p[f];
    } else {
        typedef T V __attribute__((ext_vector_type(W)));
        const V *q = reinterpret_cast<const V *>(p);
#pragma unroll
        for (int c = 0; c < F / W; ++c) {
            const V x = q[c];
#pragma unroll
            for (int w = 0; w < W; ++w) v[c * W + w] = x[w];
        }
    }
}
template <int F, int RS = F, typename T> __device__ __forceinline__ void recv_store(T *p, const T (&v)[F]) {
    constexpr int W = recv_word<F, RS, T>();
    if constexpr (W == 1 || F % W != 0) {
#pragma unroll
        for (int f = 0; f < F; ++f) p[f] = v[f];
    } else {
        typedef T V __attribute__((ext_vector_type(W)));
        V *q = reinterpret_cast<V *>(p);
#pragma unroll
        for (int c = 0; c < F / W; ++c) {
            V x;
#pragma unroll
            for (int w = 0; w < W; ++w) x[w] = v[c * W + w];
            q[c] = x;
        }
    }
}

// The F values of a node split over F / W planes of W values each, plane c at p + c PS:
// consecutive nodes are W values apart in every plane (with 16-byte planes, consecutive
// lanes then meet distinct LDS banks where a whole 32-byte record would meet two-way
// conflicts).  W = F is the plain record.
template <int F, int W, typename T> __device__ __forceinline__ void plane_load(const T *p, int PS, T (&v)[F]) {
    static_assert(F % W == 0, "planes of whole W");
#pragma unroll
    for (int c = 0; c < F / W; ++c) {
        T t[W];
        recv_load<W, W>(p + c * PS, t);
#pragma unroll
        for (int w = 0; w < W; ++w) v[c * W + w] = t[w];
    }
}
template <int F, int W, typename T> __device__ __forceinline__ void plane_store(T *p, int PS, const T (&v)[F]) {
    static_assert(F % W == 0, "planes of whole W");
#pragma unroll
    for (int c = 0; c < F / W; ++c) {
        T t[W];
#pragma unroll
        for (int w = 0; w < W; ++w) t[w] = v[c * W + w];
        recv_store<W, W>(p + c * PS, t);
    }
}

// The 13 stencil points of node idx = (i,j,k) of a degree-N coefficient vector (record
// indices, in the order grad2_pts reads them): the centre, the i-1, j-1 and k-1 neighbours,
// then for d/dr the points (i+1,j-1), (i+1,k-1), (i+1); for d/ds (j+1,i-1), (j+1,k-1),
// (j+1); for d/dt (k+1,i-1), (k+1,j-1), (k+1) -- each replaced by the centre where its
// weight (i, j, k or l = N-i-j-k) is zero.  Shared by the kernels' run-time offsets and
// the per-degree tables built on the host.
__host__ __device__ constexpr void grad_points(int N, int idx, int i, int j, int k, int (&p)[13]) {
    const int l = N - i - j - k;
    const int rl = N - k - j + 1;
    const int sk = (N - k + 1) * (N - k + 2) / 2, skm = (N - k + 2) * (N - k + 3) / 2;
    const int km = idx + j - skm, kp = idx + sk - j;
    p[0] = idx;
    p[1] = i > 0 ? idx - 1 : idx;
    p[2] = j > 0 ? idx - rl - 1 : idx;
    p[3] = k > 0 ? km : idx;
    p[4] = j > 0 ? idx - rl : idx;
    p[5] = k > 0 ? km + 1 : idx;
    p[6] = l > 0 ? idx + 1 : idx;
    p[7] = i > 0 ? idx + rl - 1 : idx;
    p[8] = k > 0 ? km + rl + 1 : idx;
    p[9] = l > 0 ? idx + rl : idx;
    p[10] = i > 0 ? kp - 1 : idx;
    p[11] = j > 0 ? kp - rl : idx;
    p[12] = l > 0 ? kp : idx;
}

// grad2_rec's arithmetic on given record offsets o[13] (in values, i.e. index times F) and
// weights ti, tj, tk, tl: the four weighted differences per derivative in a fixed order
template <int F, int G, typename T>
__device__ __forceinline__ void grad2_pts(const T *x, const int (&o)[13], T ti, T tj, T tk, T tl, T (&dr)[G],
                                          T (&ds)[G], T (&dt)[G]) {
    T xc[G], xim[G], xjm[G], xkm[G], y[G], z[G];
    recv_load<G, F>(x + o[0], xc);
    recv_load<G, F>(x + o[1], xim);
    recv_load<G, F>(x + o[2], xjm);
    recv_load<G, F>(x + o[3], xkm);
    // d/dr: i (x_c - x_{i-1}) + j (x_{i+1,j-1} - x_{j-1}) + k (x_{i+1,k-1} - x_{k-1}) + l (x_{i+1} - x_c)
    recv_load<G, F>(x + o[4], y);
    recv_load<G, F>(x + o[5], z);
#pragma unroll
    for (int f = 0; f < G; ++f) dr[f] = fma(tk, z[f] - xkm[f], fma(tj, y[f] - xjm[f], ti * (xc[f] - xim[f])));
    recv_load<G, F>(x + o[6], y);
#pragma unroll
    for (int f = 0; f < G; ++f) dr[f] = fma(tl, y[f] - xc[f], dr[f]);
    // d/ds: i (x_{i-1,j+1} - x_{i-1}) + j (x_c - x_{j-1}) + k (x_{j+1,k-1} - x_{k-1}) + l (x_{j+1} - x_c)
    recv_load<G, F>(x + o[7], y);
    recv_load<G, F>(x + o[8], z);
#pragma unroll
    for (int f = 0; f < G; ++f) ds[f] = fma(tk, z[f] - xkm[f], fma(tj, xc[f] - xjm[f], ti * (y[f] - xim[f])));
    recv_load<G, F>(x + o[9], y);
#pragma unroll
    for (int f = 0; f < G; ++f) ds[f] = fma(tl, y[f] - xc[f], ds[f]);
    // d/dt: i (x_{i-1,k+1} - x_{i-1}) + j (x_{j-1,k+1} - x_{j-1}) + k (x_c - x_{k-1}) + l (x_{k+1} - x_c)
    recv_load<G, F>(x + o[10], y);
    recv_load<G, F>(x + o[11], z);
#pragma unroll
    for (int f = 0; f < G; ++f) dt[f] = fma(tk, xc[f] - xkm[f], fma(tj, z[f] - xjm[f], ti * (y[f] - xim[f])));
    recv_load<G, F>(x + o[12], y);
#pragma unroll
    for (int f = 0; f < G; ++f) dt[f] = fma(tl, y[f] - xc[f], dt[f]);
}

// Twice the three reference-coordinate derivatives of G fields of node records F values
// apart (x + F m holds node m; x points at the group's first field): the 13 stencil
// points of grad_ref_ld, each read once for the group, the four weighted differences per
// derivative in the same order (as explicit fma, so that every kernel rounds alike),
// without the factor 1/2 (the caller folds it into the geometric factors)
template <int N, int F, int G, typename T>
__device__ __forceinline__ void grad2_rec(const T *x, int idx, int i, int j, int k, T (&dr)[G], T (&ds)[G],
                                          T (&dt)[G]) {
    const int l = N - i - j - k;
    int p[13];
    grad_points(N, idx, i, j, k, p);
    // record offsets (index times F), 24-bit multiplies only
    int o[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) o[q] = __mul24(p[q], F);
    grad2_pts<F, G, T>(x, o, T(i), T(j), T(k), T(l), dr, ds, dt);
}

// The L0 stencil of l0_prep for node records of F values: record offsets (index times F)
// of the 7 points and their integer weights (without the factor 1/2 of L0)
template <typename T> struct L0Rec {
    int o0, o1, o2, o3, o4, o5, o6;
    T w0, w1, w2, w3, w4, w5, w6;
};
// The 7 points of the L0 stencil at face node (a,b) of degree N (face-record indices:
// the node, then (a+1,b-1), (a+1,b), (a-1,b+1), (a,b+1), (a-1,b), (a,b-1) in linear order,
// each replaced by the node where its weight is zero) and their integer weights, without
// the factor 1/2 of L0.  Shared by the kernels' run-time offsets and the per-degree tables.
__host__ __device__ constexpr void l0_points(int N, int a, int b, int (&p)[7], int (&w)[7]) {
    const int c = N - a - b;
    const int ci = lin2(N, a, b), rl = N - b + 1;
    p[0] = ci;
    p[1] = b > 0 ? ci - rl : ci;
    p[2] = c > 0 ? ci + 1 : ci;
    p[3] = a > 0 ? ci + rl - 1 : ci;
    p[4] = c > 0 ? ci + rl : ci;
    p[5] = a > 0 ? ci - 1 : ci;
    p[6] = b > 0 ? ci - rl - 1 : ci;
    w[0] = (a + 1) * (a + 1) + (b + 1) * (b + 1) + (c + 1) * (c + 1);
    w[1] = (a + 1) * b;
    w[2] = (a + 1) * c;
    w[3] = (b + 1) * a;
    w[4] = (b + 1) * c;
    w[5] = (c + 1) * a;
    w[6] = (c + 1) * b;
}

template <int N, int F, typename T> __device__ __forceinline__ L0Rec<T> l0_prep_rec(int a, int b) {
    int p[7], w[7];
    l0_points(N, a, b, p, w);
    L0Rec<T> r;
    r.o0 = __mul24(p[0], F);
    r.o1 = __mul24(p[1], F);
    r.o2 = __mul24(p[2], F);
    r.o3 = __mul24(p[3], F);
    r.o4 = __mul24(p[4], F);
    r.o5 = __mul24(p[5], F);
    r.o6 = __mul24(p[6], F);
    r.w0 = T(w[0]);
    r.w1 = T(w[1]);
    r.w2 = T(w[2]);
    r.w3 = T(w[3]);
    r.w4 = T(w[4]);
    r.w5 = T(w[5]);
    r.w6 = T(w[6]);
    return r;
}

// Weight of slab s of the unnormalised lift chain (stage and lift kernels): slab 0 is 2 L0 x
// and step t (degree N-t+1 -> N-t) omits the 1/(N-t+1) of reduction_multiply!, so slab s
// is 2 N!/(N-s)! times the slab of face_lift_multiply! (src/BernsteinLift.jl:139-179);
// its l_s (src/constants.jl:10-12) divided by that integer, rounded once
__host__ __device__ constexpr double lw_chain(int N, int s) {
    double d = 2.0;
    for (int t = 0; t < s; ++t) d *= (N - t);
    return lj(N, s) / d;
}

// Cross-lane moves within a wave (no LDS storage): the value of lane + 1 (DPP wavefront
// shift left; 0 in lane 63), of lane - 1 (shift right; 0 in lane 0), and of the lane whose
// index is addr / 4 (ds_bpermute; the lane index is taken modulo 64).  Every lane of the
// wave must execute these with the same control flow (full exec), since a disabled source
// lane does not supply its value.
// (mov_dpp with bound_ctrl: a lane whose source is outside the wave gets 0, with no old value
// to materialise)
__device__ __forceinline__ int xl_next_i(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ int xl_prev_i(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int xl_from_i(int addr, int v) { return __builtin_amdgcn_ds_bpermute(addr, v); }
__device__ __forceinline__ float xl_next(float v) { return __int_as_float(xl_next_i(__float_as_int(v))); }
__device__ __forceinline__ float xl_prev(float v) { return __int_as_float(xl_prev_i(__float_as_int(v))); }
__device__ __forceinline__ float xl_from(int addr, float v) { return __int_as_float(xl_from_i(addr, __float_as_int(v))); }
__device__ __forceinline__ double xl_join(int lo, int hi) {
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double xl_next(double v) {
    const long long u = __double_as_longlong(v);
    return xl_join(xl_next_i((int)u), xl_next_i((int)(u >> 32)));
}
__device__ __forceinline__ double xl_prev(double v) {
    const long long u = __double_as_longlong(v);
    return xl_join(xl_prev_i((int)u), xl_prev_i((int)(u >> 32)));
}
__device__ __forceinline__ double xl_from(int addr, double v) {
    const long long u = __double_as_longlong(v);
    return xl_join(xl_from_i(addr, (int)u), xl_from_i(addr, (int)(u >> 32)));
}

// (E_{d}^{d+1})^T x: degree d+1 -> degree d at (a,b), reduction_multiply!
// (src/BernsteinLift.jl:66-90): ((a+1) x_{a+1,b} + (b+1) x_{a,b+1} + (d+1-a-b) x_{ab}) / (d+1)
template <typename T> __device__ __forceinline__ T reduce_apply(const T *x, int d, int a, int b, T inv) {
    const int idx = lin2(d + 1, a, b);
    const T v = fma(T(a + 1), x[idx + 1], fma(T(b + 1), x[idx + (d + 2 - b)], T(d + 1 - a - b) * x[idx]));
    return v * inv;
}

}  // namespace bb
