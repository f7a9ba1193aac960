// bb_host.hpp -- host-side set-up of a solver (one mesh, one degree) and of a batch of
// meshes with per-element degrees: argument validation and every host array the device
// copies are made from.  No HIP calls happen here, so these functions are also built
// host-only with the address and undefined-behaviour sanitizers and exercised without a
// GPU (their allocations may throw std::bad_alloc; the entry points that call them run
// under bb_guard).
#pragma once
#include <math.h>

#include <algorithm>
#include <tuple>
#include <vector>

#include "bb_stage.hpp"

namespace bb {

// connectivity word of one face: neighbour element, and face | perm << 2 | bc << 5 |
// neighbour degree << 8
inline int conn_word(int nf, int perm, int bc, int Nn) { return nf | (perm << 2) | (bc << 5) | (Nn << 8); }

// per-face scalings nxJ nyJ nzJ Jf 1/Jf from the three components of nJ; false for a zero normal
inline bool face_scalings(const double *n, double *o) {
    const double Jf = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (!(Jf > 0.0)) return false;
    o[0] = n[0];
    o[1] = n[1];
    o[2] = n[2];
    o[3] = Jf;
    o[4] = 1.0 / Jf;
    return true;
}

// index of the first NaN or Inf among x[0 .. n), or -1
inline int64_t first_nonfinite(const double *x, int64_t n) {
    for (int64_t t = 0; t < n; ++t)
        if (!isfinite(x[t])) return t;
    return -1;
}

// refuses a host array holding a NaN or Inf (what = its name in the message)
inline int check_finite_host(bb_ctx *ctx, const char *what, const double *x, int64_t n) {
    const int64_t t = first_nonfinite(x, n);
    if (t >= 0) return bb_fail(ctx, BB_ERR_NONFINITE, "%s[%lld] is not finite (%g)", what, (long long)t, x[t]);
    return BB_OK;
}

struct SolverHost {
    int F = 0, NP = 0, NFP = 0;
    std::vector<int2> conn;                // [K][4]
    std::vector<double> nJ3;               // [K][4][3] as given (the plain path's face normals)
    std::vector<double> nj5;               // [K][4][5]
    std::vector<double> invJ, geoh64;      // [K], [K][9] geo / 2
    std::vector<float> geoh32;             // [K][9]
    std::vector<uint16_t> tab;             // per-degree index tables of the stage kernel
    std::vector<float> l0a;                // L0 as matrix-core A operands
};

// validates the arguments of bb_solver_create and builds every host array it uploads
inline int solver_host_setup(bb_ctx *ctx, int pde, int flux_kind, int N, int64_t K, const double *geo,
                             const double *J, const double *nJ, const int32_t *nbr, const int8_t *nbr_face,
                             const int8_t *perm, const int8_t *bc, SolverHost &h) {
    if (!ctx) return BB_ERR_NULL;
    if (pde != BB_PDE_WAVE && pde != BB_PDE_MAXWELL) return bb_fail(ctx, BB_ERR_ARG, "pde %d", pde);
    if (flux_kind != BB_FLUX_REFERENCE && flux_kind != BB_FLUX_UPWIND)
        return bb_fail(ctx, BB_ERR_ARG, "flux %d", flux_kind);
    h.F = pde == BB_PDE_WAVE ? 4 : 6;
    int rc = bb_check_common(ctx, N, h.F, K, 1, kStageMaxN);
    if (rc) return rc;
    if (K < 1) return bb_fail(ctx, BB_ERR_ARG, "K must be >= 1");
    if (!geo || !J || !nJ || !nbr || !nbr_face || !perm || !bc) return bb_fail(ctx, BB_ERR_NULL, "null mesh array");
    if ((rc = check_finite_host(ctx, "geo", geo, K * 9)) || (rc = check_finite_host(ctx, "J", J, K)) ||
        (rc = check_finite_host(ctx, "nJ", nJ, K * 12)))
        return rc;
    h.NP = np3(N);
    h.NFP = nfp2(N);
    h.conn.resize(K * 4);
    h.nj5.resize(K * 20);
    for (int64_t t = 0; t < K * 4; ++t) {
        const int nb = nbr[t], nf = nbr_face[t], p = perm[t], b = bc[t];
        if (nb < 0 || nb >= K || nf < 0 || nf > 3 || p < 0 || p > 5 || b < 0 || b > 1)
            return bb_fail(ctx, BB_ERR_ARG, "bad connectivity at element %lld face %d", (long long)(t / 4),
                           (int)(t % 4));
        h.conn[t] = make_int2(nb, conn_word(nf, p, b, 0));
        if (!face_scalings(nJ + t * 3, h.nj5.data() + t * 5))
            return bb_fail(ctx, BB_ERR_ARG, "zero face normal at element %lld", (long long)(t / 4));
    }
    for (int64_t e = 0; e < K; ++e)
        if (!(J[e] > 0.0)) return bb_fail(ctx, BB_ERR_ARG, "non-positive Jacobian at element %lld", (long long)e);
    h.nJ3.assign(nJ, nJ + K * 12);
    h.tab = stage_tables(N);
    h.l0a = l0_mfma_table(N);
    if (h.l0a.empty())
        return bb_fail(ctx, BB_ERR_STATE, "L0 has entries outside the band the matrix-core kernel assumes (N=%d)", N);
    h.invJ.resize(K);
    h.geoh64.resize(K * 9);
    h.geoh32.resize(K * 9);
    for (int64_t e = 0; e < K; ++e) h.invJ[e] = 1.0 / J[e];
    for (int64_t t = 0; t < K * 9; ++t) {
        h.geoh64[t] = 0.5 * geo[t];
        h.geoh32[t] = (float)h.geoh64[t];
    }
    return BB_OK;
}

constexpr int kXChunk = 16;  // cross faces of one degree pair per transfer-pass workgroup

// a chunk of the transfer pass: its faces' traces at send[src0 ..] (Nfp(Nn) F values per
// face), results to xbuf[out0 ..] (Nfp(N) F values per face), face k's perm in bits 3k..3k+2
// of perms, nc faces | N << 8 | Nn << 12, the pair's A tiles at xa[aoff]
struct alignas(16) XChunk {
    int64_t src0, out0, perms;
    int32_t code, aoff;
};

struct BatchHostGroup {
    int N = 0;
    int64_t start = 0, count = 0, qoff = 0;  // first element, elements, state offset (doubles)
};

struct BatchHost {
    int F = 4;
    int64_t words = 0;                      // state length (doubles)
    std::vector<BatchHostGroup> groups;     // elements grouped by degree
    std::vector<int2> conn;                 // [K][4]
    std::vector<int64_t> nboff;             // [K][4] the neighbour's state offset
    // The traces between elements of different degree.  A face whose neighbour has another
    // degree (a "cross face") receives the neighbour's trace brought to its own degree: W
    // (Nfp(N) x Nfp(Nn), the transfer matrix of the degree pair N <- Nn) times the trace.
    //  * The cross faces are ordered by (N, Nn, neighbour state offset) and cut into chunks
    //    of kXChunk consecutive faces of one pair (XChunk; the last chunk of a pair may hold
    //    fewer); one pass runs every chunk.
    //  * The sending side: the neighbour's face trace, Nfp(Nn) records of F values in its own
    //    face node order, is written by the neighbour's stage kernel (with its new state) to
    //    the send buffer at soff[4 n + fn] (values; -1: that face sends nothing), the slots
    //    of the cross faces one after another in the order above -- so the traces of a chunk
    //    form one contiguous range.
    //  * The transfer pass multiplies each chunk's traces by W on the matrix cores: W of each
    //    pair present in the A-operand layout of the 16x16x4 product, xa[off + (ks RT + rt) 64
    //    + L] = W[16 rt + L % 16][4 ks + L / 16] (zero outside W; RT = ceil(Nfp(N) / 16) row
    //    tiles, ks over ceil(Nfp(Nn) / 4) k steps), in FP64 (xa64) and FP32 (xa32); and writes
    //    the result in the receiving face's node order (its perm) to the trace buffer, the
    //    cross faces' Nfp(N) records one after another in the order above: xoff[4 e + f] for
    //    the receiving stage kernel (0 for the other faces).
    //  * Per cross face, for the pass that fills the send buffer from a state (xsend_gather):
    //    the neighbour's state offset xnbq, its send slot xsrc, fn | Nn << 2 (xcode).
    std::vector<XChunk> xchunk;
    std::vector<int64_t> xnbq, xsrc, xoff, soff;
    std::vector<int32_t> xcode;
    int64_t xwords = 0, swords = 0;         // trace and send buffer lengths (values)
    std::vector<double> xa64;
    std::vector<float> xa32;
    std::vector<uint64_t> xmask;            // [121][2]: bit ks RT + rt of pair 11 N + Nn: tile not all zero
    std::vector<double> nj5;                // [K][4][5]
    std::vector<double> invJ, geoh;         // [K], [K][9] geo / 2
    std::vector<float> geoh32;              // [K][9]
};

// is the transfer matrix from degree Nn to N present in a table of transfer_len doubles?
inline bool transfer_present(const int64_t *toff, int64_t transfer_len, int N, int Nn) {
    const int64_t o = toff[11 * N + Nn];
    return o >= 0 && o + (int64_t)nfp2(N) * nfp2(Nn) <= transfer_len;
}

// validates the arguments of bb_batch_create and builds every host array it uploads
inline int batch_host_setup(bb_ctx *ctx, int pde, int flux_kind, int64_t K, const int32_t *deg, const double *geo,
                            const double *J, const double *nJ, const int32_t *nbr, const int32_t *nbr_face,
                            const int32_t *perm, const int32_t *bc, const double *dt, const double *transfer,
                            const int64_t *toff, int64_t transfer_len, BatchHost &h) {
    if (!ctx) return BB_ERR_NULL;
    if (pde != BB_PDE_WAVE) return bb_fail(ctx, BB_ERR_UNSUPPORTED, "batches of mixed degree: wave system only");
    if (flux_kind != BB_FLUX_REFERENCE && flux_kind != BB_FLUX_UPWIND)
        return bb_fail(ctx, BB_ERR_ARG, "flux %d", flux_kind);
    if (K < 1) return bb_fail(ctx, BB_ERR_ARG, "K must be >= 1");
    if (K > kMaxElements) return bb_fail(ctx, BB_ERR_ARG, "element count K=%lld too large", (long long)K);
    if (transfer_len < 0) return bb_fail(ctx, BB_ERR_ARG, "transfer_len %lld < 0", (long long)transfer_len);
    if (!deg || !geo || !J || !nJ || !nbr || !nbr_face || !perm || !bc || !dt || !transfer || !toff)
        return bb_fail(ctx, BB_ERR_NULL, "null batch array");
    int rc;
    if ((rc = check_finite_host(ctx, "geo", geo, K * 9)) || (rc = check_finite_host(ctx, "J", J, K)) ||
        (rc = check_finite_host(ctx, "nJ", nJ, K * 12)) || (rc = check_finite_host(ctx, "dt", dt, K)) ||
        (rc = check_finite_host(ctx, "transfer", transfer, transfer_len)))
        return rc;
    const int F = h.F;
    std::vector<int64_t> qoff_e(K);
    int64_t words = 0;
    for (int64_t e = 0; e < K; ++e) {
        const int N = deg[e];
        if (N < 1 || N > kStageMaxN) return bb_fail(ctx, BB_ERR_ARG, "degree %d of element %lld", N, (long long)e);
        if (e && N < deg[e - 1]) return bb_fail(ctx, BB_ERR_ARG, "elements must be grouped by increasing degree");
        if (!(J[e] > 0.0)) return bb_fail(ctx, BB_ERR_ARG, "non-positive Jacobian at element %lld", (long long)e);
        qoff_e[e] = words;
        words += (int64_t)F * np3(N);
    }
    h.words = words;
    h.conn.resize(K * 4);
    h.nboff.assign(K * 4, 0);
    h.nj5.resize(K * 20);
    for (int64_t t = 0; t < K * 4; ++t) {
        const int64_t e = t / 4;
        const int nb = nbr[t], nf = nbr_face[t], p = perm[t], b = bc[t];
        if (nb < 0 || nb >= K || nf < 0 || nf > 3 || p < 0 || p > 5 || b < 0 || b > 1)
            return bb_fail(ctx, BB_ERR_ARG, "bad connectivity at element %lld face %d", (long long)e, (int)(t % 4));
        const int N = deg[e], Nn = deg[nb];
        if (!b && Nn != N && !transfer_present(toff, transfer_len, N, Nn))
            return bb_fail(ctx, BB_ERR_ARG, "no transfer table from degree %d to %d", Nn, N);
        h.conn[t] = make_int2(nb, conn_word(nf, p, b, Nn));
        h.nboff[t] = b ? 0 : qoff_e[nb];
        if (!face_scalings(nJ + t * 3, h.nj5.data() + t * 5))
            return bb_fail(ctx, BB_ERR_ARG, "zero face normal at element %lld", (long long)e);
    }
    for (int64_t e = 0; e < K;) {
        BatchHostGroup g;
        g.N = deg[e];
        g.start = e;
        g.qoff = qoff_e[e];
        while (e < K && deg[e] == g.N) ++e;
        g.count = e - g.start;
        h.groups.push_back(g);
    }
    // every degree pair present, in the A-operand layout
    std::vector<int64_t> xaoff(121, -1);
    h.xmask.assign(242, 0);
    for (int64_t t = 0; t < K * 4; ++t) {
        const int N = deg[t / 4], Nn = deg[nbr[t]];
        if (bc[t] || Nn == N || xaoff[11 * N + Nn] >= 0) continue;
        const int nf = nfp2(N), nfn = nfp2(Nn), RT = (nf + 15) / 16, KS = (nfn + 3) / 4;
        const int64_t o = toff[11 * N + Nn], a0 = (int64_t)h.xa64.size();
        xaoff[11 * N + Nn] = a0;
        h.xa64.resize(a0 + (int64_t)KS * RT * 64, 0.0);
        for (int ks = 0; ks < KS; ++ks)
            for (int rt = 0; rt < RT; ++rt)
                for (int L = 0; L < 64; ++L) {
                    const int r = 16 * rt + L % 16, c = 4 * ks + L / 16;
                    if (r < nf && c < nfn) h.xa64[a0 + ((int64_t)ks * RT + rt) * 64 + L] = transfer[o + (int64_t)r * nfn + c];
                    if (h.xa64[a0 + ((int64_t)ks * RT + rt) * 64 + L] != 0.0) {
                        const int bit = ks * RT + rt;
                        h.xmask[2 * (11 * N + Nn) + bit / 64] |= uint64_t(1) << (bit % 64);
                    }
                }
    }
    if ((int64_t)h.xa64.size() > INT32_MAX) return bb_fail(ctx, BB_ERR_ARG, "transfer tables too large");
    h.xa32.assign(h.xa64.begin(), h.xa64.end());
    // the cross faces, in chunks of one degree pair
    {
        std::vector<int64_t> cross;
        for (int64_t t = 0; t < K * 4; ++t)
            if (!bc[t] && deg[nbr[t]] != deg[t / 4]) cross.push_back(t);
        auto key = [&](int64_t t) { return std::make_tuple(deg[t / 4], deg[nbr[t]], h.nboff[t], t); };
        std::sort(cross.begin(), cross.end(), [&](int64_t x, int64_t y) { return key(x) < key(y); });
        h.xoff.assign(K * 4, 0);
        h.soff.assign(K * 4, -1);
        h.xnbq.resize(cross.size());
        h.xsrc.resize(cross.size());
        h.xcode.resize(cross.size());
        int64_t w = 0, sw = 0;
        for (size_t i = 0; i < cross.size(); ++i) {
            const int64_t t = cross[i];
            const int N = deg[t / 4], Nn = deg[nbr[t]];
            const int64_t snd = (int64_t)nbr[t] * 4 + nbr_face[t];  // the sending face
            // each sending face sends to one face only: the connectivity must be mutual
            if (nbr[snd] != t / 4 || nbr_face[snd] != t % 4 || bc[snd])
                return bb_fail(ctx, BB_ERR_ARG, "face %d of element %lld and its neighbour do not match",
                               (int)(t % 4), (long long)(t / 4));
            h.xnbq[i] = h.nboff[t];
            h.xsrc[i] = sw;
            h.soff[snd] = sw;
            h.xoff[t] = w;
            h.xcode[i] = nbr_face[t] | (Nn << 2);
            // a chunk holds consecutive faces of one pair (N, Nn)
            const int64_t p = i ? cross[i - 1] : 0;
            const bool cont =
                i > 0 && (h.xchunk.back().code & 255) < kXChunk && deg[nbr[p]] == Nn && deg[p / 4] == N;
            if (cont) {
                XChunk &c = h.xchunk.back();
                c.perms |= (int64_t)perm[t] << (3 * (c.code & 255));
                ++c.code;
            } else {
                XChunk c;
                c.src0 = sw;
                c.out0 = w;
                c.perms = perm[t];
                c.code = 1 | (N << 8) | (Nn << 12);
                c.aoff = (int32_t)xaoff[11 * N + Nn];
                h.xchunk.push_back(c);
            }
            w += (int64_t)h.F * nfp2(N);
            sw += (int64_t)h.F * nfp2(Nn);
        }
        h.xwords = w;
        h.swords = sw;
    }
    h.invJ.resize(K);
    h.geoh.resize(K * 9);
    h.geoh32.resize(K * 9);
    for (int64_t e = 0; e < K; ++e) h.invJ[e] = 1.0 / J[e];
    for (int64_t t = 0; t < K * 9; ++t) {
        h.geoh[t] = 0.5 * geo[t];
        h.geoh32[t] = (float)h.geoh[t];
    }
    return BB_OK;
}

}  // namespace bb
