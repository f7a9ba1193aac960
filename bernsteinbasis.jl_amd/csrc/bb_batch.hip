// bb_batch.hip -- many independent meshes whose elements carry their own degree, stepped
// together: the elements are stored grouped by degree (group N: all elements of degree N,
// state [K_N][F][Np(N)] FP64, groups concatenated).  Every LSRK45 stage runs the transfer
// pass (k_xtrace: for every face whose neighbour has another degree, the neighbour's face
// trace -- written to a send slot by the neighbour's previous stage -- brought to the face's
// degree by a host-built transfer matrix on the matrix cores: degree elevation from below, L2
// projection from above) and then each degree group's fused stage kernel, which reads the
// transferred traces like same-degree neighbour traces and writes its own outgoing traces
// with its new state.  The group kernels of a stage run on four streams and all read the
// previous stage's state (ping-pong buffers).  Steps are captured into hipGraphs; the tiling
// of each degree group is autotuned separately.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "bb_dispatch.hpp"
#include "bb_host.hpp"
#include "bb_stage.hpp"

using namespace bb;

struct BatchGroup {
    int N = 0;
    int64_t start = 0, count = 0, qoff = 0;  // first element, elements, state offset (doubles)
    uint16_t *tab = nullptr;
    float *l0a = nullptr;
    int tiling[2] = {0, 0};
    float cost[2] = {0.f, 0.f};  // the autotuned stage kernel's time (ms), 0 before tuning
};

// Each stage runs the transfer pass over all cross faces on the context's stream, then the
// groups' stage kernels on kLanes streams ("lanes"), forked from the context's stream and
// joined back into it: the groups of one stage are independent (each writes its own state,
// residual and send slots), so they overlap, and the lanes are balanced by cost (longest
// first).  (One pass per group, each before its group's kernel on the group's lane, was ~1%
// slower on the 64-mesh workload: 14.07 against 13.92 ms per mixed step.)
constexpr int kLanes = 4;

struct bb_batch {
    bb_ctx *ctx = nullptr;
    int pde = 0, F = 0, flux_kind = 1;
    int64_t K = 0, words = 0;
    double sign[6] = {1, 1, 1, 1, 1, 1};
    std::vector<BatchGroup> groups;
    double *geo = nullptr, *invJ = nullptr, *nJ = nullptr, *dte = nullptr;
    float *geo32 = nullptr;  // geo / 2 in FP32 (mixed precision); geo holds geo / 2 in FP64
    int2 *conn = nullptr;
    int64_t *nboff = nullptr;
    // the traces between elements of different degree (see BatchHost): the transfer pass's
    // chunks and pair matrices (A-operand layout), the trace buffer the stage kernels read,
    // the send buffers (traces of q[0] / q[1], written by the stage kernels that produce that
    // state), and the per-face arrays of the pass that fills a send buffer from a state
    int64_t nchunk = 0, nx = 0;
    XChunk *xchunk = nullptr;
    int64_t *xoff = nullptr, *soff = nullptr, *xnbq = nullptr, *xsrc = nullptr;
    int32_t *xcode = nullptr;
    double *xa64 = nullptr;
    float *xa32 = nullptr;
    uint64_t *xmask = nullptr;                // [121][2] the pairs' nonzero A tiles
    double *xbuf = nullptr;                   // xwords values: FP64, or FP32 in the first half
    double *send[2] = {nullptr, nullptr};     // swords values each, likewise
    int send_prec = -1;                       // the precision send[cur] holds (-1: none)
    hipEvent_t fork = nullptr;
    hipStream_t lane[kLanes] = {};
    hipEvent_t lane_done[kLanes] = {};
    double *q[2] = {nullptr, nullptr};
    int cur = 0;
    double *res64 = nullptr;
    float *res32 = nullptr;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int g_flip = 0, g_prec = 0;
    bool g_ok = false;
};

static StageLaunchFn batch_fn(int prec) {
    return prec == BB_PREC_MIXED ? stage_launch_wave_md_mixed : stage_launch_wave_md_f64;
}

static bool batch_valid(int prec, int N, int tiling) {
    StageArgs a{};
    a.K = 0;
    return batch_fn(prec)(N, tiling, a, nullptr, true) == BB_OK;
}

// the stage arguments of elements [e0, e0 + ne) of group g (ne < 0: the whole group)
static StageArgs batch_args(bb_batch *b, const BatchGroup &g, int prec, int cur, int64_t e0 = 0, int64_t ne = -1) {
    StageArgs a{};
    const int64_t s = g.start + e0, qo = g.qoff + e0 * b->F * np3(g.N);
    a.q_in = b->q[cur] + qo;
    a.q_out = b->q[cur ^ 1] + qo;
    a.res = prec == BB_PREC_MIXED ? (void *)(b->res32 + qo) : (void *)(b->res64 + qo);
    a.geoh = prec == BB_PREC_MIXED ? (const void *)(b->geo32 + s * 9) : (const void *)(b->geo + s * 9);
    a.invJ = b->invJ + s;
    a.nJ = b->nJ + s * 20;
    a.conn = b->conn + s * 4;
    a.tab = g.tab;
    a.l0a = g.l0a;
    a.K = ne < 0 ? g.count : ne;
    a.flux_kind = b->flux_kind;
    for (int f = 0; f < 6; ++f) a.sign[f] = b->sign[f];
    a.q_nb = b->q[cur];
    a.nboff = b->nboff + s * 4;
    a.dte = b->dte + s;
    a.xbuf = b->xbuf;
    a.xoff = b->xoff + s * 4;
    a.xsend = b->send[cur ^ 1];
    a.soff = b->soff + s * 4;
    return a;
}

__device__ constexpr PermFwd kPermFwd = PermFwd();

// The 16x16x4 matrix-core products: lane L supplies A[L % 16][L / 16] and B[L / 16][L % 16]
// and holds four entries of column L % 16 of D: rows 4 (L / 16) + e in FP32, rows 4 e + L / 16
// in FP64 (the FP64 product interleaves its output rows)
template <typename T> struct XMfma;
template <> struct XMfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int L, int e) { return 4 * (L >> 4) + e; }
};
template <> struct XMfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int L, int e) { return 4 * e + (L >> 4); }
};

// The transfer pass (one launch per stage, before the degree groups): for every cross face,
// the neighbour's face trace (degree Nn, its face and node order: send buffer, written by
// the neighbour's previous stage) brought to the face's degree N and written in the face's
// own node order, xbuf[out + m F + f], in the kernels' type T.  A workgroup takes a chunk of
// at most kXChunk faces of one degree pair, whose traces are one contiguous range:
//   1. the pair's A tiles (W, Nfp(N) x Nfp(Nn)) straight into LDS, shared by the 4 waves;
//   2. wave w takes faces 4w .. 4w + 3, the 16 columns n (face 4w + n / 4, field n % 4) of
//      the product: each lane loads its B operands of every k step from global memory at
//      once (row j = 4 ks + L / 16 of column L % 16: 64 contiguous bytes per face and k
//      step; zero for the k-step padding j >= Nfp(Nn) and for faces past the chunk), then
//      multiplies on the matrix cores (16x16x4, FP32 or FP64 accumulation), RT row tiles
//      over the k steps;
//   3. the products through LDS (over the A tiles) to the chunk's output range, one 16-byte
//      record per face node: node m of face k is row kPermFwd[N][perm_k][m] of its column.
// Bounds: the host builds every offset from the arrays it uploads (N, Nn in 1..10); a chunk
// reads nc Nfp(Nn) F values from src0 and writes nc Nfp(N) F values from out0.
constexpr int xa_tiles_max() {
    int m = 0;
    for (int N = 1; N <= 10; ++N)
        for (int Nn = 1; Nn <= 10; ++Nn) {
            const int t = (nfp2(N) + 15) / 16 * ((nfp2(Nn) + 3) / 4);
            if (Nn != N && t > m) m = t;
        }
    return m;
}
constexpr int kXATiles = xa_tiles_max();  // 70: N = 10 <- 9
constexpr int kXKS = (kTraceMax + 3) / 4;  // k steps at most (17)
typedef __attribute__((address_space(3))) void lds_void;
// Register bound: 8 waves per SIMD for FP32 (64 VGPRs, no spills; the A tiles' LDS admits 8
// workgroups per CU; at 7 waves the mixed batch steps were ~1% slower), 4 for FP64, whose
// A tiles admit 4 workgroups.  Unbounded, the compiler hoists every k step's LDS reads of A
// and the FP32 kernel takes 120 VGPRs, 3 waves per SIMD.
template <bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MIXED ? 8 : 4))) void k_xtrace(
    const XChunk *__restrict__ chunk, int64_t nchunk, const void *__restrict__ send, const void *__restrict__ xa,
    const uint64_t *__restrict__ xmask, void *__restrict__ xbuf) {
    using T = std::conditional_t<MIXED, float, double>;
    using M = XMfma<T>;
    constexpr int F = 4;
    constexpr int PW = 6 * kPermRow / 4;  // the degree's PermFwd rows in 32-bit words
    __shared__ __align__(16) T as[kXATiles * 64];
    __shared__ uint32_t pfw[PW];
    const int tid = threadIdx.x, w = tid >> 6, L = tid & 63;
    // Chunks in workgroup order, i.e. dealt round-robin to the 8 XCDs: the chunks are sorted
    // by degree pair and a chunk's cost grows with the degrees, so giving each XCD a contiguous
    // range (as the stage kernels do) would leave the last XCD 1.4x the mean bytes and 1.7x the
    // mean matrix work.  Only the A tiles are reused (every pair's fit in each XCD's L2).
    for (int64_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const XChunk cd = chunk[c];
        const int nc = cd.code & 255, N = (cd.code >> 8) & 15, Nn = (cd.code >> 12) & 15;
        const int nf = nfp2(N), nfn = nfp2(Nn);
        const int KS = (nfn + 3) / 4, RT = (nf + 15) / 16, NA = RT * KS * 64;
        // the A tiles of the pair that are not all zero: bit ks RT + rt (degree elevations by
        // one or two have mostly zero tiles)
        const uint64_t mk0 = xmask[2 * (11 * N + Nn)], mk1 = xmask[2 * (11 * N + Nn) + 1];
        __syncthreads();  // the previous chunk's reads of as and pfw are done
        {
            // 16-byte pieces straight into LDS, lane L of a wave to its destination + 16 L
            const char *ag = reinterpret_cast<const char *>(xa) + (size_t)cd.aoff * sizeof(T);
            const int P = NA * (int)sizeof(T) / 16;
            for (int p0 = 64 * w; p0 < P; p0 += 256)
                if (p0 + L < P)
                    __builtin_amdgcn_global_load_lds(ag + (size_t)(p0 + L) * 16, (lds_void *)(as + p0 * 16 / sizeof(T)),
                                                     16, 0, 0);
        }
        // this lane's B operands: face k, field L & 3, rows j = 4 ks + L / 16
        const int k = 4 * w + ((L & 15) >> 2);
        const T *src = reinterpret_cast<const T *>(send) + cd.src0 + (int64_t)k * nfn * F + 4 * (L >> 4) + (L & 3);
        T b[kXKS];
#pragma unroll
        for (int ks = 0; ks < kXKS; ++ks) b[ks] = (ks < KS && k < nc && 4 * ks + (L >> 4) < nfn) ? src[16 * ks] : T(0);
        // the degree's node permutations, for the output
        uint32_t pv = 0;
        if (tid < PW) pv = reinterpret_cast<const uint32_t *>(kPermFwd.v[N])[tid];
        if (tid < PW) pfw[tid] = pv;
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this thread's pieces are in LDS
        __syncthreads();
        typename M::v4 acc[5];
#pragma unroll
        for (int rt = 0; rt < 5; ++rt) acc[rt] = typename M::v4{T(0), T(0), T(0), T(0)};
        if (4 * w < nc) {
            const T *ap = as + L;
#pragma unroll
            for (int ks = 0; ks < kXKS; ++ks)
                if (ks < KS) {
#pragma unroll
                    for (int rt = 0; rt < 5; ++rt) {
                        const int bit = ks * RT + rt;  // < 85
                        if (rt < RT && (((bit < 64 ? mk0 >> bit : mk1 >> (bit - 64)) & 1) != 0))
                            acc[rt] = M::op(ap[(ks * RT + rt) * 64], b[ks], acc[rt]);
                    }
                }
        }
        // the products -> LDS (over the A tiles, once every wave is done with them) as
        // os[face][row][f], rows in the neighbour's orientation; then the chunk's output, one
        // contiguous range, in 16-byte records: record (face k, node m) <- row
        // kPermFwd[N][perm_k][m]
        __syncthreads();
        T *os = as;
        if (k < nc) {
#pragma unroll
            for (int rt = 0; rt < 5; ++rt)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = 16 * rt + M::row(L, e);
                    if (rt < RT && r < nf) os[(k * nf + r) * F + (L & 3)] = acc[rt][e];
                }
        }
        __syncthreads();
        typedef T v4t __attribute__((ext_vector_type(4)));
        v4t *out = reinterpret_cast<v4t *>(reinterpret_cast<T *>(xbuf) + cd.out0);
        const uint8_t *pb = reinterpret_cast<const uint8_t *>(pfw);
        // record t = kk nf + m, kk and m stepped along with t
        int kk = tid / nf, m = tid - kk * nf;
        const int dk = 256 / nf, dm = 256 - dk * nf;
        for (int t = tid; t < nc * nf; t += 256) {
            const int r = pb[((cd.perms >> (3 * kk)) & 7) * kPermRow + m];
            out[t] = *reinterpret_cast<const v4t *>(os + (kk * nf + r) * F);
            kk += dk;
            m += dm;
            if (m >= nf) {
                m -= nf;
                ++kk;
            }
        }
    }
}

// The send buffer from a state (after the state was set, or in another precision): one
// thread per (cross face i, node j of the neighbour's face): node j of face fn of the
// neighbour (state at q + xnbq[i], degree Nn, code fn | Nn << 2) -> send[xsrc[i] + j F + f].
template <bool MIXED>
__global__ __launch_bounds__(256) void k_xsend_gather(const double *__restrict__ q, int64_t nx,
                                                      const int64_t *__restrict__ xnbq,
                                                      const int64_t *__restrict__ xsrc,
                                                      const int32_t *__restrict__ xcode, void *__restrict__ send) {
    using T = std::conditional_t<MIXED, float, double>;
    constexpr int F = 4;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nx * kTraceMax; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / kTraceMax;
        const int j = (int)(t - i * kTraceMax);
        const int code = xcode[i], fn = code & 3, Nn = (code >> 2) & 15;
        if (j >= nfp2(Nn)) continue;
        int a, b;
        decode2(Nn, j, a, b);
        const double *qj = q + xnbq[i] + face_vidx(Nn, fn, a, b);
        T *d = reinterpret_cast<T *>(send) + xsrc[i] + j * F;
#pragma unroll
        for (int f = 0; f < F; ++f) d[f] = T(qj[f * np3(Nn)]);
    }
}

// the traces of q[cur] in the send buffer, in precision prec (a no-op where they are)
static int batch_send_ready(bb_batch *b, int prec) {
    if (b->send_prec == prec || !b->nx) return BB_OK;
    const int64_t items = b->nx * kTraceMax;
    const int64_t blocks = (items + 255) / 256;
    const unsigned grid = (unsigned)(blocks < (1 << 20) ? blocks : (1 << 20));
    if (prec == BB_PREC_MIXED)
        k_xsend_gather<true><<<grid, 256, 0, b->ctx->stream>>>(b->q[b->cur], b->nx, b->xnbq, b->xsrc, b->xcode,
                                                               b->send[b->cur]);
    else
        k_xsend_gather<false><<<grid, 256, 0, b->ctx->stream>>>(b->q[b->cur], b->nx, b->xnbq, b->xsrc, b->xcode,
                                                                b->send[b->cur]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bb_fail(b->ctx, BB_ERR_HIP, "trace gather launch failed: %s", hipGetErrorString(e));
    b->send_prec = prec;
    return BB_OK;
}

// the transfer pass on the traces of q[cur], chunks [c0, c0 + nc), on stream st
static int batch_xtrace(bb_batch *b, int prec, int cur, int64_t c0, int64_t nc, hipStream_t st) {
    if (nc <= 0) return BB_OK;
    const unsigned grid = (unsigned)(nc < (1 << 22) ? nc : (1 << 22));
    if (prec == BB_PREC_MIXED)
        k_xtrace<true><<<grid, 256, 0, st>>>(b->xchunk + c0, nc, b->send[cur], b->xa32, b->xmask, b->xbuf);
    else
        k_xtrace<false><<<grid, 256, 0, st>>>(b->xchunk + c0, nc, b->send[cur], b->xa64, b->xmask, b->xbuf);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bb_fail(b->ctx, BB_ERR_HIP, "transfer pass launch failed: %s", hipGetErrorString(e));
    return BB_OK;
}

static int batch_launch(bb_batch *b, const BatchGroup &g, int prec, int tiling, const StageArgs &a,
                        hipStream_t st = nullptr) {
    int rc = batch_fn(prec)(g.N, tiling, a, st ? st : b->ctx->stream, false);
    if (rc == BB_ERR_UNSUPPORTED) return bb_fail(b->ctx, rc, "no fused kernel for N=%d with tiling %d", g.N, tiling);
    if (rc) return bb_fail(b->ctx, rc, "fused stage launch failed: %s", hipGetErrorString(t_launch_err));
    return BB_OK;
}

static int batch_default_tiling(int prec, int N) {
    const int NP = np3(N);
    for (int lm : {LIFT_GATHER, LIFT_PINGPONG, LIFT_WAVE, LIFT_WAVE_GT, LIFT_WAVE1}) {
        int best = 0;
        for (int e : kStageEB) {
            if (e * NP > 1024) break;
            best = e;
            if (e * NP >= 192) break;
        }
        for (int e = best; e >= 1; e /= 2)
            if (batch_valid(prec, N, tiling_code(e, lm))) return tiling_code(e, lm);
    }
    return 0;
}

static int batch_prepare(bb_batch *b, int prec) {
    if (prec != BB_PREC_F64 && prec != BB_PREC_MIXED) return bb_fail(b->ctx, BB_ERR_ARG, "prec %d", prec);
    bb_ctx *ctx = b->ctx;
    if (prec == BB_PREC_F64 && !b->res64) {
        BB_HIP(ctx, hipMalloc(&b->res64, b->words * sizeof(double)));
        int rc = bb_fill_zero(ctx, b->res64, b->words * sizeof(double));
        if (rc) return rc;
    }
    if (prec == BB_PREC_MIXED && !b->res32) {
        BB_HIP(ctx, hipMalloc(&b->res32, b->words * sizeof(float)));
        int rc = bb_fill_zero(ctx, b->res32, b->words * sizeof(float));
        if (rc) return rc;
    }
    for (auto &g : b->groups) {
        if (!g.tiling[prec]) g.tiling[prec] = batch_default_tiling(prec, g.N);
        if (!g.tiling[prec]) return bb_fail(ctx, BB_ERR_UNSUPPORTED, "no fused kernel fits N=%d", g.N);
    }
    return BB_OK;
}

// the lanes' groups in issue order (longest processing time first, each to the least loaded
// lane); cost: the autotuned time, else the group's state size.  (Cutting the largest group
// into element ranges on several lanes was slower on the 64-mesh workload, 15.0-15.1 against
// 14.6 ms per step: concurrent kernels share the chip's bandwidth, and the cut group's parts
// then hold two lanes for the whole stage.)
static void batch_schedule(const bb_batch *b, int prec, std::vector<int> (&lanes)[kLanes]) {
    std::vector<std::pair<double, int>> c;
    for (size_t i = 0; i < b->groups.size(); ++i) {
        const BatchGroup &g = b->groups[i];
        c.push_back({g.cost[prec] > 0.f ? (double)g.cost[prec] : 1e-9 * (double)g.count * np3(g.N), (int)i});
    }
    std::stable_sort(c.begin(), c.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
    double load[kLanes] = {};
    for (auto &l : lanes) l.clear();
    for (const auto &ci : c) {
        int m = 0;
        for (int l = 1; l < kLanes; ++l)
            if (load[l] < load[m]) m = l;
        load[m] += ci.first;
        lanes[m].push_back(ci.second);
    }
}

// enqueue nsteps steps (capturable): per stage, the transfer pass, then each degree group's
// stage kernel on its lane; the send buffer must hold the traces of q[*cur] (batch_send_ready)
static int batch_enqueue(bb_batch *b, int prec, double dt_scale, int nsteps, int *cur) {
    int rc;
    bb_ctx *ctx = b->ctx;
    std::vector<int> lanes[kLanes];
    batch_schedule(b, prec, lanes);
    for (int st = 0; st < nsteps; ++st)
        for (int stage = 0; stage < 5; ++stage) {
            if ((rc = batch_xtrace(b, prec, *cur, 0, b->nchunk, ctx->stream))) return rc;
            BB_HIP(ctx, hipEventRecord(b->fork, ctx->stream));
            for (int l = 0; l < kLanes; ++l) {
                if (lanes[l].empty()) continue;
                BB_HIP(ctx, hipStreamWaitEvent(b->lane[l], b->fork, 0));
                for (int gi : lanes[l]) {
                    const BatchGroup &g = b->groups[gi];
                    StageArgs a = batch_args(b, g, prec, *cur);
                    a.ca = kLsrkA[stage];
                    a.cb = kLsrkB[stage];
                    a.dt = dt_scale;
                    a.first = stage == 0;
                    if ((rc = batch_launch(b, g, prec, g.tiling[prec], a, b->lane[l]))) return rc;
                }
                BB_HIP(ctx, hipEventRecord(b->lane_done[l], b->lane[l]));
            }
            for (int l = 0; l < kLanes; ++l)
                if (!lanes[l].empty()) BB_HIP(ctx, hipStreamWaitEvent(ctx->stream, b->lane_done[l], 0));
            *cur ^= 1;
        }
    return BB_OK;
}

static void batch_drop_graphs(bb_batch *b) {
    for (auto &g : b->gexec)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
    b->g_ok = false;
}

extern "C" {

static int batch_create(bb_ctx *ctx, int pde, int flux_kind, int64_t K, const int32_t *deg_host,
                        const double *geo_host, const double *J_host, const double *nJ_host,
                        const int32_t *nbr_host, const int32_t *nbr_face_host, const int32_t *perm_host,
                        const int32_t *bc_host, const double *bc_sign_host, const double *dt_host,
                        const double *transfer_host, const int64_t *transfer_off_host, int64_t transfer_len,
                        bb_batch **out) {
    BatchHost h;
    int rc = batch_host_setup(ctx, pde, flux_kind, K, deg_host, geo_host, J_host, nJ_host, nbr_host, nbr_face_host,
                              perm_host, bc_host, dt_host, transfer_host, transfer_off_host, transfer_len, h);
    if (rc) return rc;
    // the per-degree index tables, before anything is allocated on the device
    std::vector<std::vector<uint16_t>> tabs;
    std::vector<std::vector<float>> l0as;
    for (const auto &hg : h.groups) {
        tabs.push_back(stage_tables(hg.N));
        l0as.push_back(l0_mfma_table(hg.N));
    }
    std::vector<BatchGroup> groups(h.groups.size());
    for (size_t i = 0; i < groups.size(); ++i) {
        groups[i].N = h.groups[i].N;
        groups[i].start = h.groups[i].start;
        groups[i].count = h.groups[i].count;
        groups[i].qoff = h.groups[i].qoff;
    }
    bb_batch *b = new (std::nothrow) bb_batch();
    if (!b) return bb_fail(ctx, BB_ERR_NOMEM, "host memory allocation failed");
    b->ctx = ctx;
    b->pde = pde;
    b->F = h.F;
    b->flux_kind = flux_kind;
    b->K = K;
    b->words = h.words;
    b->groups.swap(groups);
    if (bc_sign_host)
        for (int f = 0; f < h.F; ++f) b->sign[f] = bc_sign_host[f];
    auto fail = [&](hipError_t err) {
        int r = bb_fail(ctx, BB_ERR_HIP, "batch allocation: %s", hipGetErrorString(err));
        bb_batch_destroy(b);
        return r;
    };
    hipError_t err;
    const int64_t words = h.words;
    const size_t nx = h.xnbq.size(), nch = h.xchunk.size();
    if ((err = hipMalloc(&b->geo, K * 9 * sizeof(double))) || (err = hipMalloc(&b->invJ, K * sizeof(double))) ||
        (err = hipMalloc(&b->geo32, K * 9 * sizeof(float))) ||
        (err = hipMalloc(&b->nJ, K * 20 * sizeof(double))) || (err = hipMalloc(&b->dte, K * sizeof(double))) ||
        (err = hipMalloc(&b->conn, K * 4 * sizeof(int2))) || (err = hipMalloc(&b->nboff, K * 4 * sizeof(int64_t))) ||
        (err = hipMalloc(&b->q[0], words * sizeof(double))) ||
        (err = hipMalloc(&b->xoff, K * 4 * sizeof(int64_t))) ||
        (err = hipMalloc(&b->soff, K * 4 * sizeof(int64_t))) ||
        (err = hipMalloc(&b->xchunk, (nch + 1) * sizeof(XChunk))) ||
        (err = hipMalloc(&b->xnbq, (nx + 1) * sizeof(int64_t))) ||
        (err = hipMalloc(&b->xsrc, (nx + 1) * sizeof(int64_t))) ||
        (err = hipMalloc(&b->xcode, (nx + 1) * sizeof(int32_t))) ||
        (err = hipMalloc(&b->send[0], (h.swords + 1) * sizeof(double))) ||
        (err = hipMalloc(&b->send[1], (h.swords + 1) * sizeof(double))) ||
        (err = hipMalloc(&b->xa64, (h.xa64.size() + 1) * sizeof(double))) ||
        (err = hipMalloc(&b->xa32, (h.xa32.size() + 1) * sizeof(float))) ||
        (err = hipMalloc(&b->xmask, h.xmask.size() * sizeof(uint64_t))) ||
        (err = hipMalloc(&b->xbuf, (h.xwords + 1) * sizeof(double))) ||
        (err = hipMalloc(&b->q[1], words * sizeof(double))))
        return fail(err);
    if ((err = hipMemcpy(b->geo, h.geoh.data(), K * 9 * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->geo32, h.geoh32.data(), K * 9 * sizeof(float), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->invJ, h.invJ.data(), K * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->nJ, h.nj5.data(), K * 20 * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->dte, dt_host, K * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->conn, h.conn.data(), K * 4 * sizeof(int2), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->nboff, h.nboff.data(), K * 4 * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->xoff, h.xoff.data(), K * 4 * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->soff, h.soff.data(), K * 4 * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->xchunk, h.xchunk.data(), nch * sizeof(XChunk), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->xnbq, h.xnbq.data(), nx * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->xsrc, h.xsrc.data(), nx * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->xcode, h.xcode.data(), nx * sizeof(int32_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->xa64, h.xa64.data(), h.xa64.size() * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->xa32, h.xa32.data(), h.xa32.size() * sizeof(float), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->xmask, h.xmask.data(), h.xmask.size() * sizeof(uint64_t), hipMemcpyHostToDevice)))
        return fail(err);
    b->nchunk = (int64_t)nch;
    b->nx = (int64_t)nx;
    if ((err = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming))) return fail(err);
    for (int l = 0; l < kLanes; ++l)
        if ((err = hipStreamCreateWithFlags(&b->lane[l], hipStreamNonBlocking)) ||
            (err = hipEventCreateWithFlags(&b->lane_done[l], hipEventDisableTiming)))
            return fail(err);
    for (size_t i = 0; i < b->groups.size(); ++i) {
        BatchGroup &g = b->groups[i];
        const std::vector<uint16_t> &tab = tabs[i];
        const std::vector<float> &l0a = l0as[i];
        if ((err = hipMalloc(&g.tab, tab.size() * sizeof(uint16_t))) ||
            (err = hipMemcpy(g.tab, tab.data(), tab.size() * sizeof(uint16_t), hipMemcpyHostToDevice)))
            return fail(err);
        if (!l0a.empty() && ((err = hipMalloc(&g.l0a, l0a.size() * sizeof(float))) ||
                             (err = hipMemcpy(g.l0a, l0a.data(), l0a.size() * sizeof(float), hipMemcpyHostToDevice))))
            return fail(err);
    }
    if ((rc = bb_fill_zero(ctx, b->q[0], words * sizeof(double))) || (rc = bb_fill_zero(ctx, b->q[1], words * sizeof(double)))) {
        bb_batch_destroy(b);
        return rc;
    }
    *out = b;
    return BB_OK;
}

int bb_batch_destroy(bb_batch *b) {
    if (!b) return BB_ERR_NULL;
    hipStreamSynchronize(b->ctx->stream);
    for (auto &l : b->lane)
        if (l) hipStreamSynchronize(l);
    batch_drop_graphs(b);
    void *ptrs[] = {b->geo,   b->geo32, b->invJ, b->nJ,   b->dte,  b->conn, b->nboff, b->q[0],   b->q[1],
                    b->res64, b->res32, b->xoff, b->soff, b->xchunk, b->xnbq, b->xsrc, b->xcode, b->xa64,
                    b->xa32,  b->xmask, b->xbuf,  b->send[0], b->send[1]};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto &g : b->groups) {
        if (g.tab) hipFree(g.tab);
        if (g.l0a) hipFree(g.l0a);
    }
    for (int l = 0; l < kLanes; ++l) {
        if (b->lane_done[l]) hipEventDestroy(b->lane_done[l]);
        if (b->lane[l]) hipStreamDestroy(b->lane[l]);
    }
    if (b->fork) hipEventDestroy(b->fork);
    delete b;
    return BB_OK;
}

int bb_batch_state_words(bb_batch *b, int64_t *words) {
    if (!b || !words) return BB_ERR_NULL;
    *words = b->words;
    return BB_OK;
}

int bb_batch_set_state(bb_batch *b, const double *q_host) {
    if (!b || !q_host) return BB_ERR_NULL;
    int rc = check_finite_host(b->ctx, "state", q_host, b->words);
    if (rc) return rc;
    BB_HIP(b->ctx, hipMemcpyAsync(b->q[b->cur], q_host, b->words * sizeof(double), hipMemcpyHostToDevice,
                                  b->ctx->stream));
    BB_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    b->send_prec = -1;  // the traces of the new state are gathered before the next use
    return BB_OK;
}

int bb_batch_get_state(bb_batch *b, double *q_host) {
    if (!b || !q_host) return BB_ERR_NULL;
    BB_HIP(b->ctx, hipMemcpyAsync(q_host, b->q[b->cur], b->words * sizeof(double), hipMemcpyDeviceToHost,
                                  b->ctx->stream));
    BB_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return BB_OK;
}

int bb_batch_check_state(bb_batch *b, int64_t *count) {
    if (!b || !count) return BB_ERR_NULL;
    const int rc = bb_nonfinite_device(b->ctx, BB_F64, b->q[b->cur], b->words, count);
    if (rc) return rc;
    if (*count) return bb_fail(b->ctx, BB_ERR_NONFINITE, "the state holds %lld NaN or Inf values", (long long)*count);
    return BB_OK;
}

static int batch_rhs(bb_batch *b, int prec, double *du_host) {
    if (!b || !du_host) return BB_ERR_NULL;
    int rc = batch_prepare(b, prec);
    if (rc) return rc;
    bb_ctx *ctx = b->ctx;
    if ((rc = batch_send_ready(b, prec)) || (rc = batch_xtrace(b, prec, b->cur, 0, b->nchunk, ctx->stream)))
        return rc;
    for (auto &g : b->groups) {
        StageArgs a = batch_args(b, g, prec, b->cur);
        a.q_out = nullptr;
        a.rhs_only = 1;
        a.first = 1;
        a.dt = 1.0;
        if ((rc = batch_launch(b, g, prec, g.tiling[prec], a))) return rc;
    }
    if (prec == BB_PREC_F64) {
        BB_HIP(ctx, hipMemcpyAsync(du_host, b->res64, b->words * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> tmp(b->words);
        BB_HIP(ctx, hipMemcpyAsync(tmp.data(), b->res32, b->words * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int64_t t = 0; t < b->words; ++t) du_host[t] = tmp[t];
    }
    return BB_OK;
}

int bb_batch_step(bb_batch *b, int prec, double dt_scale, int nsteps) {
    if (!b) return BB_ERR_NULL;
    if (nsteps < 0) return bb_fail(b->ctx, BB_ERR_ARG, "nsteps %d", nsteps);
    int rc = batch_prepare(b, prec);
    if (rc || (rc = batch_send_ready(b, prec))) return rc;
    return batch_enqueue(b, prec, dt_scale, nsteps, &b->cur);
}

int bb_batch_graph_capture(bb_batch *b, int prec, double dt_scale, int nsteps) {
    if (!b) return BB_ERR_NULL;
    if (nsteps < 1) return bb_fail(b->ctx, BB_ERR_ARG, "nsteps %d", nsteps);
    int rc = batch_prepare(b, prec);
    if (rc) return rc;
    bb_ctx *ctx = b->ctx;
    BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    batch_drop_graphs(b);
    const int flip = (5 * nsteps) & 1;
    for (int g = 0; g < (flip ? 2 : 1); ++g) {
        int start = b->cur ^ g, c = start;
        hipGraph_t graph = nullptr;
        BB_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ctx->capturing = true;
        rc = batch_enqueue(b, prec, dt_scale, nsteps, &c);
        ctx->capturing = false;
        hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (rc) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return bb_fail(ctx, BB_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&b->gexec[start], graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return bb_fail(ctx, BB_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    b->g_flip = flip;
    b->g_prec = prec;
    b->g_ok = true;
    return BB_OK;
}

int bb_batch_graph_launch(bb_batch *b) {
    if (!b) return BB_ERR_NULL;
    if (!b->g_ok || !b->gexec[b->cur]) return bb_fail(b->ctx, BB_ERR_STATE, "no graph captured for this buffer");
    const int rc = batch_send_ready(b, b->g_prec);  // outside the graph: only after a new state
    if (rc) return rc;
    BB_HIP(b->ctx, hipGraphLaunch(b->gexec[b->cur], b->ctx->stream));
    b->cur ^= b->g_flip;
    return BB_OK;
}

int bb_batch_set_tuning(bb_batch *b, int prec, int N, int tiling) {
    if (!b) return BB_ERR_NULL;
    if (prec != BB_PREC_F64 && prec != BB_PREC_MIXED) return bb_fail(b->ctx, BB_ERR_ARG, "prec %d", prec);
    if (!batch_valid(prec, N, tiling)) return bb_fail(b->ctx, BB_ERR_ARG, "tiling %d not compiled for N=%d", tiling, N);
    for (auto &g : b->groups)
        if (g.N == N) {
            g.tiling[prec] = tiling;
            g.cost[prec] = 0.f;
            batch_drop_graphs(b);
            return BB_OK;
        }
    return bb_fail(b->ctx, BB_ERR_ARG, "no element of degree %d", N);
}

int bb_batch_autotune(bb_batch *b, int prec, int *tilings) {
    if (!b) return BB_ERR_NULL;
    int rc = batch_prepare(b, prec);
    if (rc) return rc;
    bb_ctx *ctx = b->ctx;
    // the traces the group kernels read
    if ((rc = batch_send_ready(b, prec)) || (rc = batch_xtrace(b, prec, b->cur, 0, b->nchunk, ctx->stream)))
        return rc;
    for (auto &g : b->groups) {
        StageArgs a = batch_args(b, g, prec, b->cur);  // writes the other buffer only
        a.ca = kLsrkA[1];
        a.cb = kLsrkB[1];
        a.dt = 0.0;
        a.first = 0;
        float best = 1e30f;
        int best_t = 0;
        for (int lm : {LIFT_GATHER, LIFT_PINGPONG, LIFT_WAVE, LIFT_WAVE_GT, LIFT_WAVE1})
            for (int e : kStageEB) {
                const int t = tiling_code(e, lm);
                if (!batch_valid(prec, g.N, t)) continue;
                float ms = 0.f;
                if ((rc = batch_launch(b, g, prec, t, a))) return rc;
                BB_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
                for (int r = 0; r < 5; ++r)
                    if ((rc = batch_launch(b, g, prec, t, a))) return rc;
                BB_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
                BB_HIP(ctx, hipEventSynchronize(ctx->ev1));
                BB_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                if (ms < best) {
                    best = ms;
                    best_t = t;
                }
            }
        if (!best_t) return bb_fail(ctx, BB_ERR_UNSUPPORTED, "no fused kernel compiled for N=%d", g.N);
        g.tiling[prec] = best_t;
        g.cost[prec] = best;
    }
    batch_drop_graphs(b);
    if (tilings) {
        for (int n = 0; n <= kStageMaxN; ++n) tilings[n] = 0;
        for (auto &g : b->groups) tilings[g.N] = g.tiling[prec];
    }
    return BB_OK;
}

// Entry points that allocate on the host run under bb_guard: an exception becomes
// BB_ERR_NOMEM.
int bb_batch_create(bb_ctx *ctx, int pde, int flux_kind, int64_t K, const int32_t *deg_host, const double *geo_host,
                    const double *J_host, const double *nJ_host, const int32_t *nbr_host,
                    const int32_t *nbr_face_host, const int32_t *perm_host, const int32_t *bc_host,
                    const double *bc_sign_host, const double *dt_host, const double *transfer_host,
                    const int64_t *transfer_off_host, int64_t transfer_len, bb_batch **out) {
    if (!ctx || !out) return BB_ERR_NULL;
    *out = nullptr;
    return bb_guard(ctx, [&] {
        return batch_create(ctx, pde, flux_kind, K, deg_host, geo_host, J_host, nJ_host, nbr_host, nbr_face_host,
                            perm_host, bc_host, bc_sign_host, dt_host, transfer_host, transfer_off_host, transfer_len,
                            out);
    });
}

int bb_batch_rhs(bb_batch *b, int prec, double *du_host) {
    if (!b || !du_host) return BB_ERR_NULL;
    return bb_guard(b->ctx, [&] { return batch_rhs(b, prec, du_host); });
}

}  // extern "C"
