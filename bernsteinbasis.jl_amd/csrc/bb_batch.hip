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
#include "bb_host.hpp"
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
    int64_t *nboff = nullptr;
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
    a.dte = b->dte + g.start;
    a.trf = b->trf;
    a.erow = b->erow;
    a.ecol = b->ecol;
    a.evd = b->evd;
    a.evf = b->evf;
    return a;
}

static int batch_launch(bb_batch *b, const BatchGroup &g, int prec, int tiling, const StageArgs &a) {
    int rc = batch_fn(prec)(g.N, tiling, a, b->ctx->stream, false);
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
    const int64_t tl = transfer_len + kTransferPad;
    if ((err = hipMalloc(&b->geo, K * 9 * sizeof(double))) || (err = hipMalloc(&b->invJ, K * sizeof(double))) ||
        (err = hipMalloc(&b->geo32, K * 9 * sizeof(float))) ||
        (err = hipMalloc(&b->nJ, K * 20 * sizeof(double))) || (err = hipMalloc(&b->dte, K * sizeof(double))) ||
        (err = hipMalloc(&b->conn, K * 4 * sizeof(int2))) || (err = hipMalloc(&b->nboff, K * 4 * sizeof(int64_t))) ||
        (err = hipMalloc(&b->tr, tl * sizeof(double))) ||
        (err = hipMalloc(&b->q[0], words * sizeof(double))) ||
        (err = hipMalloc(&b->trf, tl * sizeof(float))) ||
        (err = hipMalloc(&b->erow, (h.erow.size() + 1) * sizeof(int32_t))) ||
        (err = hipMalloc(&b->ecol, h.ecol.size() + 1)) ||
        (err = hipMalloc(&b->evd, (h.evd.size() + 1) * sizeof(double))) ||
        (err = hipMalloc(&b->evf, (h.evf.size() + 1) * sizeof(float))) ||
        (err = hipMalloc(&b->q[1], words * sizeof(double))))
        return fail(err);
    if ((err = hipMemcpy(b->geo, h.geoh.data(), K * 9 * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->geo32, h.geoh32.data(), K * 9 * sizeof(float), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->invJ, h.invJ.data(), K * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->nJ, h.nj5.data(), K * 20 * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->dte, dt_host, K * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->conn, h.conn.data(), K * 4 * sizeof(int2), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->nboff, h.nboff.data(), K * 4 * sizeof(int64_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->tr, transfer_host, transfer_len * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->trf, h.trf.data(), h.trf.size() * sizeof(float), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->erow, h.erow.data(), h.erow.size() * sizeof(int32_t), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->ecol, h.ecol.data(), h.ecol.size(), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->evd, h.evd.data(), h.evd.size() * sizeof(double), hipMemcpyHostToDevice)) ||
        (err = hipMemcpy(b->evf, h.evf.data(), h.evf.size() * sizeof(float), hipMemcpyHostToDevice)))
        return fail(err);
    {
        const int rc = bb_fill_zero(ctx, b->tr + transfer_len, kTransferPad * sizeof(double));
        if (rc) {
            bb_batch_destroy(b);
            return rc;
        }
    }
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
    batch_drop_graphs(b);
    void *ptrs[] = {b->geo, b->geo32, b->invJ, b->nJ, b->dte, b->tr, b->conn, b->nboff, b->q[0], b->q[1], b->res64, b->res32,
                    b->trf, b->erow, b->ecol, b->evd, b->evf};
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
    int rc = check_finite_host(b->ctx, "state", q_host, b->words);
    if (rc) return rc;
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
