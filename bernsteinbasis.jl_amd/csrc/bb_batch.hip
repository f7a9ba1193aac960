// bb_batch.hip -- many independent meshes whose elements carry their own degree, stepped
// together: the elements are stored grouped by degree (group N: all elements of degree N,
// state [K_N][F][Np(N)] FP64, groups concatenated), and every LSRK45 stage is one launch of
// the fused stage kernel per degree present, all reading the previous stage's state of
// every group (ping-pong buffers) and writing the next.  A face between elements of
// different degree reads the neighbour's whole face trace and brings it to the element's
// degree with a host-built transfer matrix (degree elevation from below, L2 projection
// from above) before the usual upwind flux.  Steps are captured into hipGraphs; the
// tiling of each degree group is autotuned separately.
#include <math.h>
#include <string.h>

#include <vector>

#include "bb_dispatch.hpp"
#include "bb_stage.hpp"

using namespace bb;

struct BatchGroup {
    int N = 0;
    int64_t start = 0, count = 0, qoff = 0;  // first element, elements, state offset (doubles)
    uint16_t *tab = nullptr;
    float *l0a = nullptr;
    int tiling[2] = {0, 0};
};

struct bb_batch {
    bb_ctx *ctx = nullptr;
    int pde = 0, F = 0, flux_kind = 1;
    int64_t K = 0, words = 0;
    double sign[6] = {1, 1, 1, 1, 1, 1};
    std::vector<BatchGroup> groups;
    double *geo = nullptr, *invJ = nullptr, *nJ = nullptr, *dte = nullptr, *tr = nullptr, *evd = nullptr;
    float *geo32 = nullptr;  // geo / 2 in FP32 (mixed precision); geo holds geo / 2 in FP64
    float *trf = nullptr, *evf = nullptr;
    int32_t *erow = nullptr;
    uint8_t *ecol = nullptr;
    int2 *conn = nullptr;
    int64_t *nboff = nullptr, *troff = nullptr, *eoff = nullptr;
    double *q[2] = {nullptr, nullptr};
    int cur = 0;
    double *res64 = nullptr;
    float *res32 = nullptr;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int g_flip = 0;
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

static StageArgs batch_args(bb_batch *b, const BatchGroup &g, int prec, int cur) {
    StageArgs a{};
    a.q_in = b->q[cur] + g.qoff;
    a.q_out = b->q[cur ^ 1] + g.qoff;
    a.res = prec == BB_PREC_MIXED ? (void *)(b->res32 + g.qoff) : (void *)(b->res64 + g.qoff);
    a.geoh = prec == BB_PREC_MIXED ? (const void *)(b->geo32 + g.start * 9) : (const void *)(b->geo + g.start * 9);
    a.invJ = b->invJ + g.start;
    a.nJ = b->nJ + g.start * 20;
    a.conn = b->conn + g.start * 4;
    a.tab = g.tab;
    a.l0a = g.l0a;
    a.K = g.count;
    a.flux_kind = b->flux_kind;
    for (int f = 0; f < 6; ++f) a.sign[f] = b->sign[f];
    a.q_nb = b->q[cur];
    a.nboff = b->nboff + g.start * 4;
    a.tr = b->tr;
    a.troff = b->troff;
    a.dte = b->dte + g.start;
    a.trf = b->trf;
    a.erow = b->erow;
    a.eoff = b->eoff;
    a.ecol = b->ecol;
    a.evd = b->evd;
    a.evf = b->evf;
    return a;
}

static int batch_launch(bb_batch *b, const BatchGroup &g, int prec, int tiling, const StageArgs &a) {
    int rc = batch_fn(prec)(g.N, tiling, a, b->ctx->stream, false);
    if (rc == BB_ERR_UNSUPPORTED) return bb_fail(b->ctx, rc, "no fused kernel for N=%d with tiling %d", g.N, tiling);
    if (rc) return bb_fail(b->ctx, rc, "fused stage launch failed: %s", hipGetErrorString(hipGetLastError()));
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

// enqueue nsteps steps (capturable): per stage, one launch per degree group
static int batch_enqueue(bb_batch *b, int prec, double dt_scale, int nsteps, int *cur) {
    int rc;
    for (int st = 0; st < nsteps; ++st)
        for (int stage = 0; stage < 5; ++stage) {
            for (auto &g : b->groups) {
                StageArgs a = batch_args(b, g, prec, *cur);
                a.ca = kLsrkA[stage];
                a.cb = kLsrkB[stage];
                a.dt = dt_scale;
                a.first = stage == 0;
                if ((rc = batch_launch(b, g, prec, g.tiling[prec], a))) return rc;
            }
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

int bb_batch_create(bb_ctx *ctx, int pde, int flux_kind, int64_t K, const int32_t *deg_host, const double *geo_host,
                    const double *J_host, const double *nJ_host, const int32_t *nbr_host,
                    const int32_t *nbr_face_host, const int32_t *perm_host, const int32_t *bc_host,
                    const double *bc_sign_host, const double *dt_host, const double *transfer_host,
                    const int64_t *transfer_off_host, int64_t transfer_len, bb_batch **out) {
    if (!ctx || !out) return BB_ERR_NULL;
    *out = nullptr;
    if (pde != BB_PDE_WAVE) return bb_fail(ctx, BB_ERR_UNSUPPORTED, "batches of mixed degree: wave system only");
    if (flux_kind != BB_FLUX_REFERENCE && flux_kind != BB_FLUX_UPWIND) return bb_fail(ctx, BB_ERR_ARG, "flux %d", flux_kind);
    if (K < 1) return bb_fail(ctx, BB_ERR_ARG, "K must be >= 1");
    if (!deg_host || !geo_host || !J_host || !nJ_host || !nbr_host || !nbr_face_host || !perm_host || !bc_host ||
        !dt_host || !transfer_host || !transfer_off_host)
        return bb_fail(ctx, BB_ERR_NULL, "null batch array");
    const int F = 4;
    std::vector<int64_t> qoff_e(K);
    int64_t words = 0;
    for (int64_t e = 0; e < K; ++e) {
        const int N = deg_host[e];
        if (N < 1 || N > kStageMaxN) return bb_fail(ctx, BB_ERR_ARG, "degree %d of element %lld", N, (long long)e);
        if (e && N < deg_host[e - 1]) return bb_fail(ctx, BB_ERR_ARG, "elements must be grouped by increasing degree");
        if (!(J_host[e] > 0.0)) return bb_fail(ctx, BB_ERR_ARG, "non-positive Jacobian at element %lld", (long long)e);
        qoff_e[e] = words;
        words += (int64_t)F * np3(N);
    }
    std::vector<int2> conn(K * 4);
    std::vector<int64_t> nboff(K * 4, 0);
    std::vector<double> nj5(K * 20);
    for (int64_t t = 0; t < K * 4; ++t) {
        const int64_t e = t / 4;
        const int nb = nbr_host[t], nf = nbr_face_host[t], p = perm_host[t], bc = bc_host[t];
        if (nb < 0 || nb >= K || nf < 0 || nf > 3 || p < 0 || p > 5 || bc < 0 || bc > 1)
            return bb_fail(ctx, BB_ERR_ARG, "bad connectivity at element %lld face %d", (long long)e, (int)(t % 4));
        const int N = deg_host[e], Nn = deg_host[nb];
        if (!bc && Nn != N) {
            const int64_t o = transfer_off_host[11 * N + Nn];
            if (o < 0 || o + (int64_t)nfp2(N) * nfp2(Nn) > transfer_len)
                return bb_fail(ctx, BB_ERR_ARG, "no transfer table from degree %d to %d", Nn, N);
        }
        conn[t] = make_int2(nb, nf | (p << 2) | (bc << 5) | (Nn << 8));
        nboff[t] = bc ? 0 : qoff_e[nb];
        const double *n = nJ_host + t * 3;
        const double Jf = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (!(Jf > 0.0)) return bb_fail(ctx, BB_ERR_ARG, "zero face normal at element %lld", (long long)e);
        double *o = nj5.data() + t * 5;
        o[0] = n[0];
        o[1] = n[1];
        o[2] = n[2];
        o[3] = Jf;
        o[4] = 1.0 / Jf;
    }
    bb_batch *b = new bb_batch();
    b->ctx = ctx;
    b->pde = pde;
    b->F = F;
    b->flux_kind = flux_kind;
    b->K = K;
    b->words = words;
    if (bc_sign_host)
        for (int f = 0; f < F; ++f) b->sign[f] = bc_sign_host[f];
    for (int64_t e = 0; e < K;) {
        BatchGroup g;
        g.N = deg_host[e];
        g.start = e;
        g.qoff = qoff_e[e];
        while (e < K && deg_host[e] == g.N) ++e;
        g.count = e - g.start;
        b->groups.push_back(g);
    }
    // elevations (Nn < N) as sparse rows; every table also in FP32
    std::vector<int32_t> erow;
    std::vector<int64_t> eoff(121, -1);
    std::vector<uint8_t> ecol;
    std::vector<double> evd;
    for (int N = 1; N <= kStageMaxN; ++N)
        for (int Nn = 1; Nn < N; ++Nn) {
            const int64_t o = transfer_off_host[11 * N + Nn];
            if (o < 0 || o + (int64_t)nfp2(N) * nfp2(Nn) > transfer_len) continue;
            eoff[11 * N + Nn] = (int64_t)erow.size();
            for (int r = 0; r < nfp2(N); ++r) {
                erow.push_back((int32_t)evd.size());
                for (int c = 0; c < nfp2(Nn); ++c) {
                    const double v = transfer_host[o + (int64_t)r * nfp2(Nn) + c];
                    if (v != 0.0) {
                        ecol.push_back((uint8_t)c);
                        evd.push_back(v);
                    }
                }
            }
            erow.push_back((int32_t)evd.size());
        }
    std::vector<float> evf(evd.begin(), evd.end()), trf(transfer_host, transfer_host + transfer_len);
    auto fail = [&](hipError_t err) {
        int r = bb_fail(ctx, BB_ERR_HIP, "batch allocation: %s", hipGetErrorString(err));
        bb_batch_destroy(b);
        return r;
    };
    hipError_t err;
    std::vector<double> invJ(K), geoh(K * 9);
    std::vector<float> geoh32(K * 9);
    for (int64_t e = 0; e < K; ++e) invJ[e] = 1.0 / J_host[e];
    for (int64_t t = 0; t < K * 9; ++t) {
        geoh[t] = 0.5 * geo_host[t];
        geoh32[t] = (float)geoh[t];
    }
    if ((err = hipMalloc(&b->geo, K * 9 * sizeof(double))) || (err = hipMalloc(&b->invJ, K * sizeof(double))) ||
        (err = hipMalloc(&b->geo32, K * 9 * sizeof(float))) ||
        (err = hipMalloc(&b->nJ, K * 20 * sizeof(double))) || (err = hipMalloc(&b->dte, K * sizeof(double))) ||
        (err = hipMalloc(&b->conn, K * 4 * sizeof(int2))) || (err = hipMalloc(&b->nboff, K * 4 * sizeof(int64_t))) ||
        (err = hipMalloc(&b->tr, (transfer_len > 0 ? transfer_len : 1) * sizeof(double))) ||
        (err = hipMalloc(&b->troff, 121 * sizeof(int64_t))) || (err = hipMalloc(&b->q[0], words * sizeof(double))) ||
        (err = hipMalloc(&b->trf, (transfer_len > 0 ? transfer_len : 1) * sizeof(float))) ||
        (err = hipMalloc(&b->erow, (erow.size() + 1) * sizeof(int32_t))) ||
        (err = hipMalloc(&b->eoff, 121 * sizeof(int64_t))) ||
        (err = hipMalloc(&b->ecol, ecol.size() + 1)) || (err = hipMalloc(&b->evd, (evd.size() + 1) * sizeof(double))) ||
        (err = hipMalloc(&b->evf, (evf.size() + 1) * sizeof(float))) ||
        (err = hipMalloc(&b->q[1], words * sizeof(double))))
        return fail(err);
    if ((err = hipMemcpy(b->geo, geoh.data(), K * 9 * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->geo32, geoh32.data(), K * 9 * sizeof(float), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->invJ, invJ.data(), K * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->nJ, nj5.data(), K * 20 * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->dte, dt_host, K * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->conn, conn.data(), K * 4 * sizeof(int2), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->nboff, nboff.data(), K * 4 * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->tr, transfer_host, transfer_len * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->troff, transfer_off_host, 121 * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->trf, trf.data(), trf.size() * sizeof(float), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->erow, erow.data(), erow.size() * sizeof(int32_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->eoff, eoff.data(), 121 * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->ecol, ecol.data(), ecol.size(), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->evd, evd.data(), evd.size() * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->evf, evf.data(), evf.size() * sizeof(float), hipMemcpyHostToDevice)))
        return fail(err);
    for (auto &g : b->groups) {
        const std::vector<uint16_t> tab = stage_tables(g.N);
        const std::vector<float> l0a = l0_mfma_table(g.N);
        if ((err = hipMalloc(&g.tab, tab.size() * sizeof(uint16_t))) ||
            (err = hipMemcpy(g.tab, tab.data(), tab.size() * sizeof(uint16_t), hipMemcpyHostToDevice)))
            return fail(err);
        if (!l0a.empty() && ((err = hipMalloc(&g.l0a, l0a.size() * sizeof(float))) ||
                             (err = hipMemcpy(g.l0a, l0a.data(), l0a.size() * sizeof(float), hipMemcpyHostToDevice))))
            return fail(err);
    }
    int rc;
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
    batch_drop_graphs(b);
    void *ptrs[] = {b->geo, b->geo32, b->invJ, b->nJ, b->dte, b->tr, b->conn, b->nboff, b->troff, b->q[0], b->q[1], b->res64, b->res32,
                    b->trf, b->erow, b->eoff, b->ecol, b->evd, b->evf};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto &g : b->groups) {
        if (g.tab) hipFree(g.tab);
        if (g.l0a) hipFree(g.l0a);
    }
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
    BB_HIP(b->ctx, hipMemcpyAsync(b->q[b->cur], q_host, b->words * sizeof(double), hipMemcpyHostToDevice,
                                  b->ctx->stream));
    BB_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return BB_OK;
}

int bb_batch_get_state(bb_batch *b, double *q_host) {
    if (!b || !q_host) return BB_ERR_NULL;
    BB_HIP(b->ctx, hipMemcpyAsync(q_host, b->q[b->cur], b->words * sizeof(double), hipMemcpyDeviceToHost,
                                  b->ctx->stream));
    BB_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return BB_OK;
}

int bb_batch_rhs(bb_batch *b, int prec, double *du_host) {
    if (!b || !du_host) return BB_ERR_NULL;
    int rc = batch_prepare(b, prec);
    if (rc) return rc;
    bb_ctx *ctx = b->ctx;
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
    if (rc) return rc;
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
    b->g_ok = true;
    return BB_OK;
}

int bb_batch_graph_launch(bb_batch *b) {
    if (!b) return BB_ERR_NULL;
    if (!b->g_ok || !b->gexec[b->cur]) return bb_fail(b->ctx, BB_ERR_STATE, "no graph captured for this buffer");
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
    }
    batch_drop_graphs(b);
    if (tilings) {
        for (int n = 0; n <= kStageMaxN; ++n) tilings[n] = 0;
        for (auto &g : b->groups) tilings[g.N] = g.tiling[prec];
    }
    return BB_OK;
}

}  // extern "C"
