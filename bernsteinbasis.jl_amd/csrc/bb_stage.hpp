// bb_stage.hpp -- one fused kernel per low-storage RK stage:
//
//   trace gather + interface flux -> 4-face lift (L0 + reduction chain) -> volume
//   derivatives + geometric combine -> / J -> res = a res + dt rhs -> q_out = q_in + b res
//
// i.e. examples/3d_wave.jl:36-84 (or 3d_maxwell.jl:36-93) followed by one stage of a
// five-stage 2N-storage RK45, for a block of EB whole elements per workgroup.
//
// Data in HBM (per element, all contiguous, element-major):
//   q_in, q_out [K][F][Np] FP64 (ping-pong: neighbours read the old stage while the new
//   one is written), res [K][F][Np] (FP64, or FP32 in mixed precision), geo / 2 [K][9],
//   1/J [K], nJ [K][4][5], conn [K][4] = {neighbour, face | perm << 2 | bc << 5}.
//
// LDS (T = FP64, or FP32 in mixed precision), every array as node records of F values;
// the use of each buffer depends on the lift organisation (StageCfg below): the element's
// coefficients, the interface flux (4 faces x Nfp records), and the chain vectors (two
// buffers, or every slab of every chain; with wave-owned chains each chain's flux sits in
// its own chain region until the chain overwrites it)
// Each thread owns one volume node: its F volume terms and F lift accumulators stay in
// registers (accumulators always FP64 -- the signed binomially weighted sum of the four
// faces' chains, where cancellation happens), so no LDS scatter and no atomics.
#pragma once
#include <type_traits>
#include <vector>

#include "bb_pde.hpp"
#include "bb_stencil.hpp"

namespace bb {

struct StageArgs {
    const double *q_in;
    double *q_out;
    void *res;
    const void *geoh;      // [K][9] halved geometric factors rxJ..tzJ / 2 in the kernel's type
    const double *invJ;    // [K] 1/J
    const double *nJ;      // [K][4][5]: nxJ nyJ nzJ Jf 1/Jf per face
    const int2 *conn;
    const uint16_t *tab;   // index tables of degree N (stage_tables)
    const float *l0a;      // L0 in the A-operand layout of the FP32 MFMA (l0_mfma_table)
    int64_t K;
    double ca, cb, dt;
    int first;     // stage 1: res is not read (a_1 = 0)
    int rhs_only;  // write rhs into res, leave q alone
    int flux_kind;
    double sign[6];
    // elements of mixed degree (k_stage with MD = true): the neighbour across face f of
    // element e has its state at q_nb + nboff[4 e + f] and its degree Nn in bits 8..11 of
    // conn.y; element e steps by dt * dte[e].  Where Nn differs from the element's degree,
    // the neighbour's trace brought to this degree (by the transfer pass that runs before
    // the stage, k_xtrace) is at xbuf + xoff[4 e + f]: Nfp records of F values (the
    // kernel's type), in this face's node order.  A face f whose neighbour has another
    // degree also sends: with the new state, the stage writes the face's own trace (Nfp
    // records of F values in the kernel's type, its own face node order) to xsend +
    // soff[4 e + f] (soff < 0: no neighbour of another degree)
    const double *q_nb;
    const int64_t *nboff;
    const double *dte;
    const void *xbuf;
    const int64_t *xoff;
    void *xsend;
    const int64_t *soff;
};

// Five-stage fourth-order 2N-storage RK (Carpenter & Kennedy, NASA TM-109112, 1994)
constexpr double kLsrkA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                              -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
constexpr double kLsrkB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                              1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                              2277821191437.0 / 14882151754819.0};

constexpr int kTraceMax = 66;  // face nodes of the highest degree of a fused kernel (N = 10)

// m -> r: node m of a degree-d face triangle (d = 1..10) is node r of the same face in the
// neighbour's orientation, for face permutation p: (a, b, c) of m -> (a2, b2) = its entries
// perm_entry(p, 0) and perm_entry(p, 1), r = lin2(d, a2, b2)
constexpr int kPermRow = (kTraceMax + 3) / 4 * 4;  // 68: rows of whole 32-bit words
struct alignas(16) PermFwd {
    uint8_t v[11][6][kPermRow];
    constexpr PermFwd() : v{} {
        for (int d = 1; d <= 10; ++d)
            for (int p = 0; p < 6; ++p)
                for (int b = 0; b <= d; ++b)
                    for (int a = 0; a <= d - b; ++a) {
                        const int x[3] = {a, b, d - a - b};
                        v[d][p][lin2(d, a, b)] = (uint8_t)lin2(d, x[perm_entry(p, 0)], x[perm_entry(p, 1)]);
                    }
    }
};

// Lift organisation inside the stage kernel:
//  GATHER   : every chain vector of every face is kept in LDS (slab s of face f at the
//             volume offset of slab s, i.e. the reference's G1..G3 arrays without the l_s
//             weights); each thread then gathers its node's four face values once, in the
//             reference's order G4 + G1 + G2 + G3.  LDS: 4 F Np + 4 F Nfp values.
//  PINGPONG : only the current and next chain vectors are kept; after every step each
//             thread adds the values that land on its node.  LDS: ~8 F Nfp values, but the
//             per-step accumulation diverges across the lanes of a wave.
//  GATHER_MFMA (FP32 only): GATHER, with slab 0 = L0 x of all the block's chains computed
//             as one product on the matrix cores: L0 (Nfp x Nfp, zero-padded to 16-row
//             tiles and 4-deep k steps) times the chain columns (element, field, face) in
//             v_mfma_f32_16x16x4f32 tiles with FP32 accumulation.
//  WAVE     : GATHER, with each (element, face) chain owned by one wave from the flux to
//             its last reduction step, so that these phases need only wave-level ordering
//             (no workgroup barrier between them); the element state gets its own LDS
//             buffer, since fast waves overwrite chain slabs while slow ones still read it.
//  WAVE_GT  : WAVE with the index tables read from global memory (L2-resident; every
//             entry is read once per workgroup, and the chain steps by increments), which
//             leaves the LDS to the state and chains and admits more workgroups per CU
//             where the LDS limits them.
//  WAVE1    : WAVE_GT, with the chains' later steps packed: once a step's output has at
//             most 32 nodes (degree <= 6), a workgroup barrier hands the chains over and each
//             wave runs two chains in the two halves of its lanes (a wave-owned chain keeps
//             most lanes idle in those steps: 36 wave-steps per element at N = 9, 22 here).
enum { LIFT_GATHER = 0, LIFT_PINGPONG = 1, LIFT_GATHER_MFMA = 2, LIFT_WAVE = 3, LIFT_WAVE_GT = 4, LIFT_WAVE1 = 5 };


template <int N, int PDE, bool MIXED, int EB, int LM = LIFT_GATHER, bool MD = false> struct StageCfg {
    using T = std::conditional_t<MIXED, float, double>;
    static constexpr int F = Pde<PDE>::F;
    static constexpr int NP = np3(N), NFP = nfp2(N);
    static constexpr int NT = ((EB * NP + 63) / 64) * 64;
    static constexpr int RS = rec_stride<T>(F);  // LDS record stride (values)
    // PINGPONG: s1 = state, then chain A; s2 = flux, then chain B
    // GATHER:   s1 = state, then all chain vectors (EB 4 Np records); s2 = flux
    // WAVE*:    s0 = state, s1 = all chain vectors, s2 = flux -- or, with FXA, the flux of
    //           each chain inside that chain's own region of s1 (records NFP .. 2 NFP - 1,
    //           i.e. slab 1 onwards): the wave that owns the chain writes its flux, reads it
    //           in L0 (which writes slab 0) and only then starts the chain, so no other
    //           buffer is needed
    static constexpr bool WAVEF = LM == LIFT_WAVE || LM == LIFT_WAVE_GT || LM == LIFT_WAVE1;
    static constexpr bool GATHER = LM == LIFT_GATHER || LM == LIFT_GATHER_MFMA || WAVEF;
    static constexpr int R0 = WAVEF ? EB * RS * NP : 0;  // separate state buffer (WAVE)
    static constexpr int R1 = GATHER ? EB * RS * 4 * NP
                              : ((EB * RS * NP > EB * RS * 4 * NFP) ? EB * RS * NP : EB * RS * 4 * NFP);
    static constexpr bool FXA = WAVEF && 2 * NFP <= NP;
    static constexpr int R2 = FXA ? 0 : EB * RS * 4 * NFP;
    static constexpr int TABG = (2 * NP + 28 * NFP + 3) / 4 * 4;  // ijk, dec, own / neighbour face nodes
    static constexpr int TAB = TABG + 4 * NP;                      // then the gather offsets
    static constexpr int TAB16 = (TAB + 7) / 8;    // the same in 16-byte words

    // index tables in LDS, or read from global memory (WAVE_GT, WAVE1)
    static constexpr bool TABL = LM != LIFT_WAVE_GT && LM != LIFT_WAVE1;
    // prologue loads issued in the order they are consumed (FP64 single-degree kernels)
    static constexpr bool ORDERED = !MIXED && !MD;
    // uniform per-element data read in the prologue: geometric factors with one element
    // per workgroup, face scalings with wave-owned chains (mixed-degree kernels: with
    // wave-owned chains only, and without the pinning below)
    static constexpr bool PRE_GEO = EB == 1 && (!MD || WAVEF);
    static constexpr bool PRE_NJ = WAVEF;
    static constexpr size_t lds = (size_t)(R0 + R1 + R2) * sizeof(T) + (TABL ? (size_t)TAB16 * 16 : 0) + (N + 1) * 8;
    // MFMA tiling of L0: MT row tiles of 16, KT k steps of 4
    static constexpr int MT = (NFP + 15) / 16, KT = (NFP + 3) / 4;
    static constexpr bool valid = NT <= 1024 && lds <= 64 * 1024 && (LM != LIFT_GATHER_MFMA || MIXED);
    // waves per SIMD the register allocation must allow (1: no bound)
    static constexpr int WPE = (MD && N == 10) ? 7 : 1;
};

// L0 couples face nodes at most N+1 apart in linear order (the 7-point stencil reaches the
// neighbouring rows, whose offsets are N+1-b and N+2-b for row b >= 1): the k steps of row
// tile mt that can hold non-zero entries
__host__ __device__ constexpr int l0_klo(int N, int mt) { return (16 * mt - (N + 1)) > 0 ? (16 * mt - (N + 1)) / 4 : 0; }
__host__ __device__ constexpr int l0_khi(int N, int mt) {
    return ((16 * mt + 15 + (N + 1)) < nfp2(N) - 1 ? (16 * mt + 15 + (N + 1)) : nfp2(N) - 1) / 4;
}

// 2 L0 = (N+1)^2 (E_N^{N+1})^T E_N^{N+1} as a dense Nfp x Nfp matrix (the 7-point stencil
// of l0_apply without its factor 1/2: integer entries, exact in FP32), laid out as the A operand
// of v_mfma_f32_16x16x4f32: for row tile mt, lane L and k step kk the value
// L0[16 mt + L % 16][4 kk + L / 16] (zero outside the matrix), stored [mt][L][kk].
inline std::vector<float> l0_mfma_table(int N) {
    const int NFP = nfp2(N), MT = (NFP + 15) / 16, KT = (NFP + 3) / 4;
    std::vector<double> L0((size_t)NFP * NFP, 0.0);
    for (int b = 0; b <= N; ++b)
        for (int a = 0; a <= N - b; ++a) {
            const int c = N - a - b, r = lin2(N, a, b);
            const int nb[7][3] = {{a, b, (a + 1) * (a + 1) + (b + 1) * (b + 1) + (c + 1) * (c + 1)},
                                  {a + 1, b - 1, (a + 1) * b}, {a + 1, b, (a + 1) * c}, {a - 1, b + 1, (b + 1) * a},
                                  {a, b + 1, (b + 1) * c},     {a - 1, b, (c + 1) * a}, {a, b - 1, (c + 1) * b}};
            for (auto &q : nb)
                if (q[2] != 0) L0[(size_t)r * NFP + lin2(N, q[0], q[1])] += q[2];
        }
    std::vector<float> t((size_t)MT * 64 * KT, 0.f);
    bool outside = false;
    for (int mt = 0; mt < MT; ++mt)
        for (int L = 0; L < 64; ++L)
            for (int kk = 0; kk < KT; ++kk) {
                const int row = 16 * mt + L % 16, col = 4 * kk + L / 16;
                if (row < NFP && col < NFP) t[((size_t)mt * 64 + L) * KT + kk] = (float)L0[(size_t)row * NFP + col];
                if (t[((size_t)mt * 64 + L) * KT + kk] != 0.f && (kk < l0_klo(N, mt) || kk > l0_khi(N, mt)))
                    outside = true;
            }
    if (outside) t.clear();  // an entry the kernel would skip: refuse the table
    return t;
}

// Index tables of degree N for the stage kernel, built once on the host (uint16 each):
//   [0, Np)            node n -> i | j << 4 | k << 8
//   [Np, 2 Np)         entry t of a degree-d triangle (d = 0..N, at offset Np(d-1)) -> a | b << 8
//   [2 Np, +4 Nfp)     own face f, node m -> volume index
//   [.., +24 Nfp)      neighbour face g, permutation p, my node m -> neighbour's volume index
//   [G, G + 4 Np)      (G = the above rounded up to 4 entries) node n, face f -> where face f's
//                      chain lands on n: tet(N, slab) + index in the slab (f1 slab j at (i,k),
//                      f2 slab l at (j,k), f3 slab i at (j,k), f4 slab k at (i,j))
inline std::vector<uint16_t> stage_tables(int N) {
    const int NP = np3(N), NFP = nfp2(N);
    const int G = (2 * NP + 28 * NFP + 3) / 4 * 4;
    std::vector<uint16_t> t((G + 4 * NP + 7) / 8 * 8, 0);  // padded to 16-byte words
    int r = 0;
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N - k; ++j)
            for (int i = 0; i <= N - j - k; ++i) t[r++] = (uint16_t)(i | (j << 4) | (k << 8));
    for (int d = 0; d <= N; ++d)
        for (int b = 0; b <= d; ++b)
            for (int a = 0; a <= d - b; ++a) t[r++] = (uint16_t)(a | (b << 8));
    for (int f = 0; f < 4; ++f)
        for (int b = 0; b <= N; ++b)
            for (int a = 0; a <= N - b; ++a) t[r++] = (uint16_t)face_vidx(N, f, a, b);
    for (int g = 0; g < 4; ++g)
        for (int p = 0; p < 6; ++p)
            for (int b = 0; b <= N; ++b)
                for (int a = 0; a <= N - b; ++a) {
                    const int abc[3] = {a, b, N - a - b};
                    t[r++] = (uint16_t)face_vidx(N, g, abc[perm_entry(p, 0)], abc[perm_entry(p, 1)]);
                }
    r = G;
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N - k; ++j)
            for (int i = 0; i <= N - j - k; ++i) {
                const int l = N - i - j - k;
                t[r++] = (uint16_t)(tet(N, j) + lin2(N - j, i, k));
                t[r++] = (uint16_t)(tet(N, l) + lin2(N - l, j, k));
                t[r++] = (uint16_t)(tet(N, i) + lin2(N - i, j, k));
                t[r++] = (uint16_t)(tet(N, k) + lin2(N - k, i, j));
            }
    return t;
}

// the slab weights lw_chain(N, s), s = 0..N, as a table (one load per thread in the
// prologue instead of a select chain over the constants)
template <int N> struct LwTab {
    double v[N + 1];
    constexpr LwTab() : v{} {
        for (int s = 0; s <= N; ++s) v[s] = lw_chain(N, s);
    }
};
template <int N> __device__ constexpr LwTab<N> kLwTab = LwTab<N>();

constexpr int kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    const unsigned q = nb / kXcds, r = nb % kXcds;
    const unsigned x = b % kXcds, s = b / kXcds;
    return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}

template <int N, int PDE, bool MIXED, int EB, int LM, bool MD = false>
__global__ __launch_bounds__((StageCfg<N, PDE, MIXED, EB, LM, MD>::NT),
                             (StageCfg<N, PDE, MIXED, EB, LM, MD>::WPE)) void k_stage(StageArgs A) {
    using Cfg = StageCfg<N, PDE, MIXED, EB, LM, MD>;
    using T = typename Cfg::T;
    using P = Pde<PDE>;
    constexpr int F = Cfg::F, NP = Cfg::NP, NFP = Cfg::NFP, NT = Cfg::NT, RS = Cfg::RS;
    // every LDS array holds node records of F values (field fastest): see recv_load
    __shared__ __align__(16) T s1[Cfg::R1];
    __shared__ __align__(16) T s2[Cfg::R2 ? Cfg::R2 : 1];
    // flux record of chain c (= 4 element + face), face node m: fxb + (c FXS + m) RS
    T *const fxb = Cfg::FXA ? s1 + NFP * RS : s2;
    constexpr int FXS = Cfg::FXA ? NP : NFP;
    __shared__ __align__(16) T s0[Cfg::WAVEF ? Cfg::R0 : 1];
    T *sst = Cfg::WAVEF ? s0 : s1;  // the element state in LDS: [element][node][field]
    __shared__ uint4 tab16[Cfg::TABL ? Cfg::TAB16 : 1];

    const uint16_t *tab = Cfg::TABL ? reinterpret_cast<const uint16_t *>(tab16) : A.tab;
    __shared__ double lw[N + 1];  // slab weights of the unnormalised chain (lw_chain)
    const uint16_t *ijk = tab, *dec = tab + NP, *fvid = tab + 2 * NP, *nvid = tab + 2 * NP + 4 * NFP;

    const int tid = threadIdx.x;
    // Workgroups are dealt round-robin to the 8 XCDs (workgroup b -> XCD b mod 8), each
    // with its own L2.  Remap so that XCD x processes one contiguous range of element
    // blocks: face neighbours (mostly the same Kuhn cube or the next one) are then read
    // by the same XCD while their owner's lines are still in its L2.
    const int64_t e0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * EB;
    const int ne = (int)min((int64_t)EB, A.K - e0);
    const bool own = tid < ne * NP;
    const int el = own ? tid / NP : 0;
    const int n = own ? tid % NP : 0;
    // with one element per workgroup its index is uniform, so per-element data (1/J, the
    // geometric factors) can be read with scalar loads
    const int64_t eg = EB == 1 ? e0 : e0 + el;

    // ---- prologue: every global read the stage needs is issued here, before the first
    // LDS work, so that its latency overlaps the volume phase: index tables (into LDS, or
    // the per-lane entries of the later phases from global memory), the element state and
    // the RK residual of the owned node, and the neighbour traces of this thread's face
    // nodes.
    int pk = 0;   // this node's (i, j, k)
    uint2 gp{};   // where the four faces' chains land on this node
    if constexpr (Cfg::TABL) {
        for (int t = tid; t < Cfg::TAB16; t += NT) tab16[t] = reinterpret_cast<const uint4 *>(A.tab)[t];
    } else {
        pk = tab[n];
        if constexpr (Cfg::GATHER) gp = reinterpret_cast<const uint2 *>(tab + Cfg::TABG)[n];
    }
    if (tid <= N) lw[tid] = kLwTab<N>.v[tid];
    // matrix-core L0: this wave's first row tile of L0 as A operands, needed after the flux
    constexpr int MT = Cfg::MT, KT = Cfg::KT, NW = NT / 64;
    typedef float v4f __attribute__((ext_vector_type(4)));
    const int wave_u = __builtin_amdgcn_readfirstlane(tid / 64);
    float amf[LM == LIFT_GATHER_MFMA ? KT : 1];
    if constexpr (LM == LIFT_GATHER_MFMA) {
        if (wave_u < MT) {
            const float *ap = A.l0a + ((size_t)wave_u * 64 + (tid & 63)) * KT;
#pragma unroll
            for (int kk = 0; kk < KT; ++kk) amf[kk] = ap[kk];
        }
    }
    // face-node work items of this thread (at most RF of them).  WAVE: chain c = (element,
    // face) belongs to wave c mod NW, its nodes to the lanes; otherwise items are dealt to
    // the threads in order.
    constexpr int MI = (NFP + 63) / 64, CW = (EB * 4 + NW - 1) / NW;
    constexpr int RF = Cfg::WAVEF ? CW * MI : (EB * 4 * NFP + NT - 1) / NT;
    auto fitem = [&](int r, int &e2, int &phi, int &m) -> bool {
        if constexpr (Cfg::WAVEF) {
            const int c = wave_u + NW * (r / MI);
            m = (r % MI) * 64 + (tid & 63);
            e2 = c >> 2;
            phi = c & 3;
            return c < ne * 4 && m < NFP;
        } else {
            const int t = tid + r * NT;
            e2 = t / (4 * NFP);
            phi = (t / NFP) & 3;
            m = t % NFP;
            return t < ne * 4 * NFP;
        }
    };
    // WAVE_GT (tables in global memory): the per-lane table entries of the later phases are
    // read here with the other global reads -- own face node of each flux item, the face
    // node (a,b) of each L0 item and the chain lane's node of the degree N-1 triangle
    int fvp[RF], l0p[MI], chp = 0;
    using TR = T;
    TR *res = reinterpret_cast<TR *>(A.res);
    double qv[F];
    TR rv[F];
    double invJ = 0.0, dte = 0.0;  // read here: a load issued in the epilogue stalls the workgroup's end
    int64_t sof[4] = {-1, -1, -1, -1};
    int2 cnv[RF];
    T nb[RF][F];
    // per-element data with a uniform index, as scalar loads issued in the prologue rather
    // than at the start of the phase that uses them (each element's data is read once, so
    // the first read always goes to memory): the halved geometric factors with one element
    // per workgroup, the face scalings nJ of each item with wave-owned chains.  Scalar loads
    // return out of order, so any later wait for one waits for all; where they are issued
    // depends on the prologue order below.  (Mixed-degree kernels read them here only with
    // wave-owned chains, and do not pin them: there the pinning turned later uniform table
    // reads into vector loads, and the extra registers cost a wave per SIMD.)
    T gh[9];
    double njw[Cfg::PRE_NJ ? RF : 1][5];
    auto uniform_loads = [&]() {
        if constexpr (Cfg::PRE_GEO) {
            const T *g = reinterpret_cast<const T *>(A.geoh) + e0 * 9;
#pragma unroll
            for (int c = 0; c < 9; ++c) gh[c] = g[c];
        }
        if constexpr (Cfg::PRE_NJ) {
#pragma unroll
            for (int r = 0; r < RF; ++r) {
                // chain c = (element, face) of item r is uniform across the wave
                const int c = __builtin_amdgcn_readfirstlane(wave_u + NW * (r / MI));
                const int e2 = (c >> 2) < ne ? (c >> 2) : 0;
                const double *nj = A.nJ + ((e0 + e2) * 4 + (c & 3)) * 5;
#pragma unroll
                for (int c5 = 0; c5 < 5; ++c5) njw[r][c5] = nj[c5];
            }
        }
    };
    if constexpr (Cfg::ORDERED) {
        // Vector loads complete in issue order (one counter), so a wait for one load is a
        // wait for every load issued before it.  Here they are issued in the order they
        // are consumed -- tables (with the neighbour node index of each face item), state,
        // neighbour traces, residual -- without branches around them (a branch makes the
        // merged wait counts cover later loads too): the state barrier then waits for the
        // state alone and the traces are waited for in the flux phase.  This pays for the
        // FP64 single-degree kernels, whose occupancy LDS limits; elsewhere the longer-lived
        // registers cost more occupancy than the overlap gains.
        //
        // Items whose neighbour is an interior element (nbi) read its trace; the others read
        // their own element's node 0 and ignore it.
        int nvr[RF];
        bool nbi[RF];
        const double *nbase[RF];
        auto nb_index = [&](int r) {
            int e2, phi, m;
            const bool ok = fitem(r, e2, phi, m);
            e2 = e2 < ne ? e2 : 0;  // phi is always 0..3; with wave-owned chains e2 stays uniform
            cnv[r] = A.conn[(e0 + e2) * 4 + phi];
            nbi[r] = ok && ((cnv[r].y >> 5) & 7) == 0;
            nvr[r] = nvid[nbi[r] ? ((cnv[r].y & 3) * 6 + ((cnv[r].y >> 2) & 7)) * NFP + m : 0];
            nbase[r] = A.q_in + (nbi[r] ? (int64_t)cnv[r].x : eg) * F * NP;
        };
        if constexpr (!Cfg::TABL) {
#pragma unroll
            for (int r = 0; r < RF; ++r) nb_index(r);
#pragma unroll
            for (int r = 0; r < RF; ++r) {
                int e2, phi, m;
                fvp[r] = fvid[fitem(r, e2, phi, m) ? phi * NFP + m : 0];
            }
            // (entries of lanes without an item are read at index 0 and never used: no branch
            // around the loads, whose merged wait counts would otherwise cover later loads)
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) {
                const int beta = mi * 64 + (tid & 63);
                l0p[mi] = dec[np3(N - 1) + (beta < NFP ? beta : 0)];
            }
            chp = dec[np3(N - 2) + ((tid & 63) < nfp2(N - 1) ? (tid & 63) : 0)];
            __builtin_amdgcn_sched_barrier(0);
        }
        // every thread loads (threads past the block's nodes read node 0 of element e0): no
        // branch around the loads
#pragma unroll
        for (int f = 0; f < F; ++f) qv[f] = A.q_in[(eg * F + f) * NP + n];
        invJ = A.invJ[eg];
        dte = A.dt;
        if constexpr (Cfg::TABL) {
            __syncthreads();  // tables visible
#pragma unroll
            for (int r = 0; r < RF; ++r) nb_index(r);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < RF; ++r) {
            const double *qn = nbase[r] + (nbi[r] ? nvr[r] : 0);
#pragma unroll
            for (int f = 0; f < F; ++f) nb[r][f] = qn[f * NP];
        }
        __builtin_amdgcn_sched_barrier(0);
        // the residual is loaded also where it is not used (first stage, RHS only) and then
        // replaced, again so that no branch surrounds the loads
        {
            const bool zero = A.first || A.rhs_only;
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const TR x = __builtin_nontemporal_load(&res[(eg * F + f) * NP + n]);
                rv[f] = zero ? TR(0) : x;
            }
        }
        // last: an earlier wait for a scalar load (the connectivity above) would wait for
        // these too, and the state barrier waits for the state loads anyway
        uniform_loads();
    } else {
        if (own) {
#pragma unroll
            for (int f = 0; f < F; ++f) {
                qv[f] = A.q_in[(eg * F + f) * NP + n];
                rv[f] = (A.first || A.rhs_only) ? TR(0) : __builtin_nontemporal_load(&res[(eg * F + f) * NP + n]);
            }
            invJ = A.invJ[eg];
            dte = MD ? A.dt * A.dte[eg] : A.dt;
        }
        // the send slots of the element's faces (mixed degree, one element per workgroup:
        // uniform, read here rather than at the end where the stores need them)
        if constexpr (MD && EB == 1) {
#pragma unroll
            for (int fc = 0; fc < 4; ++fc) sof[fc] = A.soff[e0 * 4 + fc];
        }
        // the connectivity of each item (a scalar load with wave-owned chains) and the uniform
        // per-element data are issued with the state loads, so that their latencies overlap
        // rather than follow one another (mixed-degree kernels: with wave-owned chains only;
        // their other organisations hold per-thread connectivity in registers); the barrier
        // is needed only for index tables in LDS
        constexpr bool EARLY = !MD || Cfg::WAVEF;
        if constexpr (EARLY) {
#pragma unroll
            for (int r = 0; r < RF; ++r) {
                int e2, phi, m;
                fitem(r, e2, phi, m);
                e2 = e2 < ne ? e2 : 0;
                cnv[r] = A.conn[(e0 + e2) * 4 + phi];
            }
            uniform_loads();
        }
        if constexpr (Cfg::TABL || !EARLY) __syncthreads();  // tables visible
        if constexpr (!Cfg::TABL) {
#pragma unroll
            for (int r = 0; r < RF; ++r) {
                int e2, phi, m;
                fvp[r] = fitem(r, e2, phi, m) ? fvid[phi * NFP + m] : 0;
            }
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) {
                const int beta = mi * 64 + (tid & 63);
                l0p[mi] = beta < NFP ? dec[np3(N - 1) + beta] : 0;
            }
            if ((tid & 63) < nfp2(N - 1)) chp = dec[np3(N - 2) + (tid & 63)];
        }
#pragma unroll
        for (int r = 0; r < RF; ++r) {
            int e2, phi, m;
            if (fitem(r, e2, phi, m)) {
                if constexpr (!EARLY) cnv[r] = A.conn[(e0 + e2) * 4 + phi];
                if (((cnv[r].y >> 5) & 7) == 0 && (!MD || ((cnv[r].y >> 8) & 15) == N)) {
                    const int nvi = nvid[((cnv[r].y & 3) * 6 + ((cnv[r].y >> 2) & 7)) * NFP + m];
                    const double *qn = MD ? A.q_nb + A.nboff[(e0 + e2) * 4 + phi] + nvi
                                          : A.q_in + (int64_t)cnv[r].x * F * NP + nvi;
#pragma unroll
                    for (int f = 0; f < F; ++f) nb[r][f] = T(qn[f * NP]);
                } else if (MD && ((cnv[r].y >> 5) & 7) == 0) {
                    // a neighbour of another degree: its trace at this degree and node order,
                    // from the transfer pass's buffer
                    const T *xb = reinterpret_cast<const T *>(A.xbuf) + A.xoff[(e0 + e2) * 4 + phi] + m * F;
#pragma unroll
                    for (int f = 0; f < F; ++f) nb[r][f] = xb[f];
                }
            }
        }
    }
    // the loaded values are taken here (in scalar registers), so that the loads are not
    // moved into the later phases that use them
    auto pin_scalar = [](auto &x) { asm volatile("" ::"s"(x)); };
    if constexpr (MD && EB == 1) {
#pragma unroll
        for (int fc = 0; fc < 4; ++fc) pin_scalar(sof[fc]);
    }
    if constexpr (Cfg::PRE_GEO && !MD) {
#pragma unroll
        for (int c = 0; c < 9; ++c) pin_scalar(gh[c]);
    }
    if constexpr (Cfg::PRE_NJ && !MD) {
#pragma unroll
        for (int r = 0; r < RF; ++r)
#pragma unroll
            for (int c = 0; c < 5; ++c) pin_scalar(njw[r][c]);
    }
    if (own) {
        T v[F];
#pragma unroll
        for (int f = 0; f < F; ++f) v[f] = T(qv[f]);
        recv_store<F, RS>(sst + __mul24(el * NP + n, RS), v);
    }
    __syncthreads();  // state visible

    // ---- interface flux at every face node of the block's elements -> s2 (or the chain
    // buffer).  It runs before the volume term: the neighbour values it consumes then die
    // before the volume phase's registers are live (mixed Maxwell N=9: 87 -> 81 VGPRs)
#pragma unroll
    for (int r = 0; r < RF; ++r) {
        int e2, phi, m;
        if (fitem(r, e2, phi, m)) {
            const int vi = Cfg::TABL ? fvid[phi * NFP + m] : fvp[r];
            T um[F], d[F], o[F];
            recv_load<F, RS>(sst + __mul24(e2 * NP + vi, RS), um);
            if (((cnv[r].y >> 5) & 7) == 0) {
#pragma unroll
                for (int f = 0; f < F; ++f) d[f] = nb[r][f] - um[f];
            } else {
#pragma unroll
                for (int f = 0; f < F; ++f) d[f] = T(A.sign[f]) * um[f] - um[f];
            }
            double nj[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) nj[c] = Cfg::PRE_NJ ? njw[Cfg::PRE_NJ ? r : 0][c] : A.nJ[((e0 + e2) * 4 + phi) * 5 + c];
            P::template flux<T>(A.flux_kind, d, T(nj[0]), T(nj[1]), T(nj[2]), T(nj[3]), T(nj[4]), o);
            recv_store<F, RS>(fxb + __mul24((e2 * 4 + phi) * FXS + m, RS), o);
        }
    }
    // ---- volume term: 3 reference derivatives per field, geometric combine, PDE flux Jacobian
    T vol[F];
#pragma unroll
    for (int f = 0; f < F; ++f) vol[f] = T(0);
    if constexpr (Cfg::TABL) {
        pk = ijk[n];
        if constexpr (Cfg::GATHER) gp = reinterpret_cast<const uint2 *>(tab + Cfg::TABG)[n];
    }
    const int i = pk & 15, j = (pk >> 4) & 15, k = pk >> 8;
    if (own) {
        // halved geometric factors (grad2_rec returns twice the reference derivatives),
        // prepared on the host in T: with one element per workgroup they are scalar loads
        // (issued in the prologue by single-degree kernels) that the FMAs read from scalar
        // registers
        if constexpr (!Cfg::PRE_GEO) {
            const T *g = reinterpret_cast<const T *>(A.geoh) + eg * 9;
#pragma unroll
            for (int c = 0; c < 9; ++c) gh[c] = g[c];
        }
        // fields in groups of G = 2 (one 8- or 16-byte word per stencil point and group).
        // The 13 stencil points of a group are in registers at once, so the group size
        // sets this phase's register peak: FP32 groups of four fields cost the wave N=7
        // kernel two waves per SIMD against groups of two (84 -> 71 VGPRs, 727 -> 667 us),
        // and groups of one (70 VGPRs, 6 waves) did not speed up Maxwell N=9
        constexpr int G = (F % 2 == 0) ? 2 : 1;
        const T *xs = sst + __mul24(el, NP * RS);
#pragma unroll
        for (int g0 = 0; g0 < F; g0 += G) {
            T dr[G], ds[G], dt[G];
            grad2_rec<N, RS, G, T>(xs + g0, n, i, j, k, dr, ds, dt);
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const T gx = fma(gh[2], dt[q], fma(gh[1], ds[q], gh[0] * dr[q]));
                const T gy = fma(gh[5], dt[q], fma(gh[4], ds[q], gh[3] * dr[q]));
                const T gz = fma(gh[8], dt[q], fma(gh[7], ds[q], gh[6] * dr[q]));
                P::template vol_add<T>(g0 + q, gx, gy, gz, vol);
            }
            // one group at a time: hoisting the next group's reads above this group's
            // arithmetic would hold all 13 F values per node in registers at once
            __builtin_amdgcn_sched_barrier(0);
        }
    }

    if constexpr (Cfg::WAVEF) wave_lds_sync();
    else __syncthreads();

    // where each face's chain lands on this thread's node: face f1..f4 -> (slab, index)
    const int l = N - i - j - k;
    const int st[4] = {j, l, i, k};
    const int ix[4] = {lin2(N - j, i, k), lin2(N - l, j, k), lin2(N - i, j, k), lin2(N - k, i, j)};
    double acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0;
    const int order[4] = {3, 0, 1, 2};  // face 4 first, then G1, G2, G3 (BernsteinLift.jl:202-217)

    // chain storage, chain c = 4 element + face: GATHER keeps every slab of every chain
    // ([chain][Np][field], slab s at the volume offset tet(N,s)); PINGPONG keeps two
    // [chain][Nfp][field] buffers.  The chain is unnormalised: slab 0 is 2 L0 x and each
    // step omits its 1/(d+1); the slab weights lw carry both factors.
    constexpr int QS = Cfg::GATHER ? NP : NFP;
    // ---- L0 on every (element, face): flux s2 -> slab 0 (buffer s1)
    if constexpr (LM == LIFT_GATHER_MFMA) {
        // chain column col = F (4 element + face) + field reads s2 and writes s1 at stride F.
        // Wave w computes the row tiles mt = w, w + NW, .. for all column tiles at once:
        // the B operands of every column tile are read first, then the k steps advance
        // the column tiles' independent accumulators in turn.
        constexpr int NTLM = (EB * F * 4 + 15) / 16;
        const int ncols = ne * F * 4;
        const int lane = tid & 63;
        for (int mt = wave_u; mt < MT; mt += NW) {
            if (mt != wave_u) {
                const float *ap = A.l0a + ((size_t)mt * 64 + lane) * KT;
#pragma unroll
                for (int kk = 0; kk < KT; ++kk) amf[kk] = ap[kk];
            }
            const int klo = l0_klo(N, mt), khi = l0_khi(N, mt);
            float b[NTLM][KT];
#pragma unroll
            for (int nt = 0; nt < NTLM; ++nt) {
                const int col = min(nt * 16 + (lane & 15), ncols - 1);
                const T *xb = s2 + (col / F) * NFP * RS + col % F;
#pragma unroll
                for (int kk = 0; kk < KT; ++kk) {
                    if (kk < klo || kk > khi) continue;
                    const int k = 4 * kk + (lane >> 4);
                    // rows of the padded k range meet zero columns of L0, but must be finite
                    b[nt][kk] = (4 * kk + 3 < NFP || k < NFP) ? xb[(k < NFP ? k : NFP - 1) * RS] : 0.f;
                }
            }
            v4f acc[NTLM];
#pragma unroll
            for (int nt = 0; nt < NTLM; ++nt) acc[nt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < KT; ++kk) {
                if (kk < klo || kk > khi) continue;  // an all-zero tile of L0
#pragma unroll
                for (int nt = 0; nt < NTLM; ++nt)
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(amf[kk], b[nt][kk], acc[nt], 0, 0, 0);
            }
#pragma unroll
            for (int nt = 0; nt < NTLM; ++nt) {
                const int col = nt * 16 + (lane & 15);
                if (col < ncols) {
                    T *yb = s1 + (col / F) * QS * RS + col % F;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int beta = 16 * mt + 4 * (lane >> 4) + r;
                        if (beta < NFP) yb[beta * RS] = acc[nt][r];
                    }
                }
            }
        }
    } else {
        auto l0_item = [&](int c, int beta, int ab) {
            const L0Rec<T> pt = l0_prep_rec<N, RS, T>(ab & 255, ab >> 8);
            const T *x = fxb + __mul24(c, FXS * RS);
            T v[F], y[F];
            // the 7-point sum in the order of l0_eval, as explicit fma
            recv_load<F, RS>(x + pt.o0, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = pt.w0 * y[f];
            recv_load<F, RS>(x + pt.o1, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w1, y[f], v[f]);
            recv_load<F, RS>(x + pt.o2, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w2, y[f], v[f]);
            recv_load<F, RS>(x + pt.o3, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w3, y[f], v[f]);
            recv_load<F, RS>(x + pt.o4, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w4, y[f], v[f]);
            recv_load<F, RS>(x + pt.o5, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w5, y[f], v[f]);
            recv_load<F, RS>(x + pt.o6, y);
#pragma unroll
            for (int f = 0; f < F; ++f) v[f] = fma(pt.w6, y[f], v[f]);
            recv_store<F, RS>(s1 + __mul24(c * QS + beta, RS), v);
        };
        if constexpr (Cfg::WAVEF) {
            for (int q = 0; q < CW; ++q) {
                const int c = wave_u + NW * q;
                if (c >= ne * 4) break;
#pragma unroll
                for (int mi = 0; mi < MI; ++mi) {
                    const int beta = mi * 64 + (tid & 63);
                    if (beta < NFP) l0_item(c, beta, Cfg::TABL ? (int)dec[np3(N - 1) + beta] : l0p[mi]);
                }
            }
        } else {
            for (int t = tid; t < ne * 4 * NFP; t += NT) l0_item(t / NFP, t % NFP, dec[np3(N - 1) + t % NFP]);
        }
    }
    if constexpr (Cfg::WAVEF) wave_lds_sync();
    else __syncthreads();
    // one unnormalised reduction step, degree dd+1 -> dd, at node (a,b) of chain c:
    //   (a+1) x_{a+1,b} + (b+1) x_{a,b+1} + (dd+1-a-b) x_{ab}   (reduction_multiply! without /(dd+1))
    // src points at the chain's source slab, dst at the output record; record offsets
    // idxF = RS lin2(dd+1, a, b) and o2F = RS (lin2(dd+1, a, b) + dd + 2 - b)
    auto chain_item = [&](const T *src, T *dst, int idxF, int o2F, T wa, T wb, T wc) {
        T x1[F], x2[F], x0[F], v[F];
        recv_load<F, RS>(src + idxF + RS, x1);
        recv_load<F, RS>(src + o2F, x2);
        recv_load<F, RS>(src + idxF, x0);
#pragma unroll
        for (int f = 0; f < F; ++f) v[f] = fma(wa, x1[f], fma(wb, x2[f], wc * x0[f]));
        recv_store<F, RS>(dst, v);
    };

    if constexpr (Cfg::WAVEF) {
        // lane -> node (a,b) of the degree N-1 triangle for the whole chain: at step dd the
        // lane is active while a + b <= dd, and its source index lin2(dd+1,a,b), destination
        // lin2(dd,a,b) and weight dd+1-a-b each drop by a constant from one step to the next
        const int lane = tid & 63;
        const bool lact = lane < nfp2(N - 1);
        int a = 0, b = 0;
        if (lact) {
            const int ab = Cfg::TABL ? (int)dec[np3(N - 2) + lane] : chp;
            a = ab & 255;
            b = ab >> 8;
        }
        const int bF = __mul24(b, RS);
        int idxF = __mul24(lin2(N, a, b), RS), dstF = __mul24(lin2(N - 1, a, b), RS);
        int o2F = idxF + __mul24(N + 1 - b, RS);
        T wc = T(N - a - b);
        const T wa = T(a + 1), wb = T(b + 1);
        // WAVE1: steps down to output degree DS + 1 as above, the rest two chains per wave
        constexpr int DS = LM == LIFT_WAVE1 ? (N - 1 < 6 ? N - 1 : 6) : -1;
#pragma unroll
        for (int s = 1; s <= N - 1 - DS; ++s) {
            const int dd = N - s;
            if (lact && a + b <= dd) {
                for (int q = 0; q < CW; ++q) {
                    const int c = wave_u + NW * q;
                    if (c >= ne * 4) break;
                    chain_item(s1 + (c * NP + tet(N, s - 1)) * RS, s1 + (c * NP + tet(N, s)) * RS + dstF, idxF, o2F,
                               wa, wb, wc);
                }
            }
            wave_lds_sync();
            idxF -= bF;
            dstF -= bF;
            o2F -= bF + RS;
            wc -= T(1);
        }
        if constexpr (DS >= 0) {
            __syncthreads();  // hand-over: every chain at degree DS + 1
            // lane half h of wave w runs chains 2 (w + NW q) + h, q = 0, 1, ..; lane -> node
            // (a,b) of the degree-DS triangle, stepped by increments as above
            constexpr int CP = (EB * 2 + NW - 1) / NW;  // chain pairs per wave
            const int h = lane >> 5, qn = lane & 31;
            const bool act2 = qn < nfp2(DS);
            int a2 = 0, b2 = 0;
            if (act2) {
                const int ab = (int)dec[np3(DS - 1) + qn];
                a2 = ab & 255;
                b2 = ab >> 8;
            }
            const int bF2 = __mul24(b2, RS);
            int idx2 = __mul24(lin2(DS + 1, a2, b2), RS), dst2 = __mul24(lin2(DS, a2, b2), RS);
            int o22 = idx2 + __mul24(DS + 2 - b2, RS);
            T wc2 = T(DS + 1 - a2 - b2);
            const T wa2 = T(a2 + 1), wb2 = T(b2 + 1);
#pragma unroll
            for (int s = N - DS; s <= N; ++s) {
                const int dd = N - s;
                if (act2 && a2 + b2 <= dd) {
                    for (int q = 0; q < CP; ++q) {
                        const int c = 2 * (wave_u + NW * q) + h;
                        if (c >= ne * 4) break;
                        const int cbo = __mul24(c, NP * RS);
                        chain_item(s1 + (cbo + tet(N, s - 1) * RS), s1 + (cbo + tet(N, s) * RS + dst2), idx2, o22,
                                   wa2, wb2, wc2);
                    }
                }
                wave_lds_sync();
                idx2 -= bF2;
                dst2 -= bF2;
                o22 -= bF2 + RS;
                wc2 -= T(1);
            }
        }
        __syncthreads();  // every chain complete before the gather
    } else {
        auto chain_step = [&](const T *src, T *dst, int src_off, int dst_off, int dd) {
            const int nf = nfp2(dd);
            for (int t = tid; t < ne * 4 * nf; t += NT) {
                const int c = t / nf, beta = t % nf;
                const int ab = dec[np3(dd - 1) + beta];
                const int a = ab & 255, b = ab >> 8;
                const int idxF = __mul24(lin2(dd + 1, a, b), RS);
                chain_item(src + __mul24(c * QS + src_off, RS), dst + __mul24(c * QS + dst_off + beta, RS), idxF,
                           idxF + __mul24(dd + 2 - b, RS), T(a + 1), T(b + 1), T(dd + 1 - a - b));
            }
        };
        if constexpr (Cfg::GATHER) {
#pragma unroll
            for (int s = 1; s <= N; ++s) {
                chain_step(s1, s1, tet(N, s - 1), tet(N, s), N - s);
                __syncthreads();
            }
        }
    }
    if constexpr (Cfg::GATHER) {
        // ---- each node gathers l_slab * E from its four faces
        if (own) {
            const int go[4] = {(int)(gp.x & 0xffff), (int)(gp.x >> 16), (int)(gp.y & 0xffff), (int)(gp.y >> 16)};
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const int ff = order[o];
                const double w = lw[st[ff]];
                T v[F];
                recv_load<F, RS>(s1 + __mul24(el * 4 + ff, NP * RS) + __mul24(go[ff], RS), v);
#pragma unroll
                for (int f = 0; f < F; ++f) acc[f] = fma(w, (double)v[f], acc[f]);
            }
        }
    } else {
        auto accumulate = [&](int s, const T *buf) {
            if (!own) return;
            const double w = lw[s];
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const int ff = order[o];
                if (st[ff] == s) {
                    T v[F];
                    recv_load<F, RS>(buf + __mul24((el * 4 + ff) * NFP + ix[ff], RS), v);
#pragma unroll
                    for (int f = 0; f < F; ++f) acc[f] = fma(w, (double)v[f], acc[f]);
                }
            }
        };
        accumulate(0, s1);
        T *cur = s1, *nxt = s2;
#pragma unroll
        for (int s = 1; s <= N; ++s) {
            const int dd = N - s;
            const int nf = nfp2(dd);
            for (int t = tid; t < ne * 4 * nf; t += NT) {
                const int c = t / nf, beta = t % nf;
                const int ab = dec[np3(dd - 1) + beta];
                const int a = ab & 255, b = ab >> 8;
                const int idxF = __mul24(lin2(dd + 1, a, b), RS);
                chain_item(cur + __mul24(c, NFP * RS), nxt + __mul24(c * NFP + beta, RS), idxF,
                           idxF + __mul24(dd + 2 - b, RS), T(a + 1), T(b + 1), T(dd + 1 - a - b));
            }
            __syncthreads();
            accumulate(s, nxt);
            T *tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
    }

    // ---- epilogue: rhs = (lift + volume) / J, then the RK stage
    if (own) {
        T qs[F];  // the new state in T, for the face traces sent to neighbours of another degree
#pragma unroll
        for (int f = 0; f < F; ++f) {
            const double rhs = (acc[f] + (double)vol[f]) * invJ;
            const int64_t g = (eg * F + f) * NP + n;
            if (A.rhs_only) {
                res[g] = TR(rhs);
            } else {
                // in the first stage rv is 0 (not loaded), so this is exactly dte * rhs
                const double r = fma(A.ca, (double)rv[f], dte * rhs);
                const TR rs = TR(r);
                const double qn = qv[f] + A.cb * (double)rs;
                __builtin_nontemporal_store(rs, &res[g]);
                __builtin_nontemporal_store(qn, &A.q_out[g]);
                qs[f] = T(qn);
            }
        }
        if constexpr (MD) {
            if (!A.rhs_only) {
                // this node is node ix[fc] of face fc where its slab index st[fc] is 0
#pragma unroll
                for (int fc = 0; fc < 4; ++fc) {
                    const int64_t o = EB == 1 ? sof[fc] : A.soff[eg * 4 + fc];
                    if (st[fc] == 0 && o >= 0) {
                        T *d = reinterpret_cast<T *>(A.xsend) + o + ix[fc] * F;
#pragma unroll
                        for (int f = 0; f < F; ++f) d[f] = qs[f];
                    }
                }
            }
        }
    }
}

// launch one stage with the given tiling code (or, with probe = true, only report whether
// that tiling is compiled); BB_ERR_UNSUPPORTED for a tiling that is not
using StageLaunchFn = int (*)(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe);
int stage_launch_wave_f64(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe);
int stage_launch_wave_mixed(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe);
int stage_launch_maxwell_f64(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe);
int stage_launch_maxwell_mixed(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe);
// elements of mixed degree (wave system)
int stage_launch_wave_md_f64(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe);
int stage_launch_wave_md_mixed(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe);

constexpr int kStageMaxN = 10;
// candidate tilings (elements per block, lift organisation), in autotuning order
constexpr int kStageEB[] = {1, 2, 3, 4};  // 3: mixed-degree kernels (a full wave of degree-3 nodes)
constexpr int kStageLM[] = {LIFT_GATHER, LIFT_PINGPONG, LIFT_GATHER_MFMA, LIFT_WAVE, LIFT_WAVE_GT, LIFT_WAVE1};

// tiling code: elements per block + 16 * lift organisation
constexpr int tiling_code(int eb, int lm) { return eb + 16 * lm; }

}  // namespace bb

#define BB_STAGE_DEFINE(NAME, PDE, MIXED)                                                              \
    namespace bb {                                                                                     \
    int NAME(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe) {                     \
        const int64_t K = a.K;                                                                         \
        return dispatch_n<1, kStageMaxN>(N, [&](auto nc) -> int {                                      \
            constexpr int NN = decltype(nc)::value;                                                    \
            int rc = BB_ERR_UNSUPPORTED;                                                               \
            auto one = [&](auto ebc, auto lmc) {                                                       \
                constexpr int E = decltype(ebc)::value, LMV = decltype(lmc)::value;                    \
                using Cfg = StageCfg<NN, PDE, MIXED, E, LMV>;                                          \
                if constexpr (Cfg::valid) {                                                            \
                    if (tiling == tiling_code(E, LMV)) {                                               \
                        if (probe) {                                                                   \
                            rc = BB_OK;                                                                \
                        } else {                                                                       \
                            k_stage<NN, PDE, MIXED, E, LMV><<<grid1d(K, E), Cfg::NT, 0, st>>>(a);      \
                            rc = launch_status();                                                      \
                        }                                                                              \
                    }                                                                                  \
                }                                                                                      \
            };                                                                                         \
            using G0 = std::integral_constant<int, LIFT_GATHER>;                                       \
            using P1 = std::integral_constant<int, LIFT_PINGPONG>;                                     \
            using M2 = std::integral_constant<int, LIFT_GATHER_MFMA>;                                  \
            using W3 = std::integral_constant<int, LIFT_WAVE>;                                         \
            using W4 = std::integral_constant<int, LIFT_WAVE_GT>;                                      \
            using W5 = std::integral_constant<int, LIFT_WAVE1>;                                        \
            one(std::integral_constant<int, 1>{}, W3{});                                               \
            one(std::integral_constant<int, 2>{}, W3{});                                               \
            one(std::integral_constant<int, 1>{}, W4{});                                               \
            one(std::integral_constant<int, 2>{}, W4{});                                               \
            one(std::integral_constant<int, 1>{}, W5{});                                               \
            one(std::integral_constant<int, 1>{}, M2{});                                               \
            one(std::integral_constant<int, 2>{}, M2{});                                               \
            one(std::integral_constant<int, 4>{}, M2{});                                               \
            one(std::integral_constant<int, 1>{}, G0{});                                               \
            one(std::integral_constant<int, 2>{}, G0{});                                               \
            one(std::integral_constant<int, 4>{}, G0{});                                               \
            one(std::integral_constant<int, 1>{}, P1{});                                               \
            one(std::integral_constant<int, 2>{}, P1{});                                               \
            one(std::integral_constant<int, 4>{}, P1{});                                               \
            return rc;                                                                                 \
        });                                                                                            \
    }                                                                                                  \
    }

#define BB_STAGE_DEFINE_MD(NAME, PDE, MIXED)                                                           \
    namespace bb {                                                                                     \
    int NAME(int N, int tiling, const StageArgs &a, hipStream_t st, bool probe) {                     \
        const int64_t K = a.K;                                                                         \
        return dispatch_n<1, kStageMaxN>(N, [&](auto nc) -> int {                                      \
            constexpr int NN = decltype(nc)::value;                                                    \
            int rc = BB_ERR_UNSUPPORTED;                                                               \
            auto one = [&](auto ebc, auto lmc) {                                                       \
                constexpr int E = decltype(ebc)::value, LMV = decltype(lmc)::value;                    \
                using Cfg = StageCfg<NN, PDE, MIXED, E, LMV, true>;                                    \
                if constexpr (Cfg::valid) {                                                            \
                    if (tiling == tiling_code(E, LMV)) {                                               \
                        if (probe) {                                                                   \
                            rc = BB_OK;                                                                \
                        } else {                                                                       \
                            k_stage<NN, PDE, MIXED, E, LMV, true><<<grid1d(K, E), Cfg::NT, 0, st>>>(a); \
                            rc = launch_status();                                                      \
                        }                                                                              \
                    }                                                                                  \
                }                                                                                      \
            };                                                                                         \
            using G0 = std::integral_constant<int, LIFT_GATHER>;                                       \
            using P1 = std::integral_constant<int, LIFT_PINGPONG>;                                     \
            using W3 = std::integral_constant<int, LIFT_WAVE>;                                         \
            using W4 = std::integral_constant<int, LIFT_WAVE_GT>;                                      \
            using W5 = std::integral_constant<int, LIFT_WAVE1>;                                        \
            one(std::integral_constant<int, 1>{}, G0{});                                               \
            one(std::integral_constant<int, 2>{}, G0{});                                               \
            one(std::integral_constant<int, 3>{}, G0{});                                               \
            one(std::integral_constant<int, 4>{}, G0{});                                               \
            one(std::integral_constant<int, 1>{}, P1{});                                               \
            one(std::integral_constant<int, 2>{}, P1{});                                               \
            one(std::integral_constant<int, 3>{}, P1{});                                               \
            one(std::integral_constant<int, 4>{}, P1{});                                               \
            one(std::integral_constant<int, 1>{}, W3{});                                               \
            one(std::integral_constant<int, 1>{}, W4{});                                               \
            one(std::integral_constant<int, 1>{}, W5{});                                               \
            return rc;                                                                                 \
        });                                                                                            \
    }                                                                                                  \
    }
