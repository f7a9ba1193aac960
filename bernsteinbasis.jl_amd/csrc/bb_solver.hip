// bb_solver.hip -- device-resident mesh + state, the two RHS paths, LSRK45 stepping,
// hipGraph capture of whole time steps, and the elements-per-workgroup autotuner.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "bb_dispatch.hpp"
#include "bb_host.hpp"
#include "bb_stage.hpp"

using namespace bb;

struct bb_solver {
    bb_ctx *ctx = nullptr;
    int pde = 0, F = 0, N = 0, flux_kind = 1, NP = 0, NFP = 0;
    int64_t K = 0;
    double sign[6] = {1, 1, 1, 1, 1, 1};
    // mesh
    double *geo = nullptr, *J = nullptr, *invJ = nullptr, *nJ = nullptr;  // nJ: [K][4][5] nxJ nyJ nzJ Jf 1/Jf
    double *geoh64 = nullptr;  // geo / 2, the fused kernels' operands (FP64 and FP32)
    float *geoh32 = nullptr;
    int2 *conn = nullptr;
    uint16_t *tab = nullptr;
    float *l0a = nullptr;  // L0 as MFMA A operands (mixed precision, matrix-core L0 tilings)
    std::vector<int2> conn_host;
    std::vector<double> nJ_host;
    // plain-path mesh arrays (StartUpDG layout) and scratch, allocated on first use
    int32_t *mapP = nullptr;
    int8_t *bcn = nullptr;
    double *nx = nullptr, *ny = nullptr, *nz = nullptr;
    double *uM = nullptr, *flux = nullptr, *fu[3] = {nullptr, nullptr, nullptr}, *d9[9] = {}, *du = nullptr;
    bool plain_ready = false;  // every plain-path array allocated and filled
    // state
    double *q[2] = {nullptr, nullptr};
    int cur = 0;
    double *res64 = nullptr;
    float *res32 = nullptr;
    int tiling[2] = {0, 0};  // fused-kernel tiling code per precision (0 = not chosen)
    // graphs
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int g_path = -1, g_start = 0, g_flip = 0;
};

static StageLaunchFn stage_fn(int pde, int prec) {
    if (pde == BB_PDE_WAVE) return prec == BB_PREC_MIXED ? stage_launch_wave_mixed : stage_launch_wave_f64;
    return prec == BB_PREC_MIXED ? stage_launch_maxwell_mixed : stage_launch_maxwell_f64;
}

// is a tiling compiled for this (pde, prec, N)?  a launch over zero elements is a no-op
// grid that only resolves the template
static bool stage_valid(bb_solver *s, int prec, int tiling) {
    StageArgs a{};
    a.K = 0;
    return stage_fn(s->pde, prec)(s->N, tiling, a, nullptr, true) == BB_OK;
}

static size_t state_words(const bb_solver *s) { return (size_t)s->K * s->F * s->NP; }

static int ensure_res(bb_solver *s, int prec) {
    bb_ctx *ctx = s->ctx;
    size_t n = state_words(s);
    if (prec == BB_PREC_F64 && !s->res64) {
        BB_HIP(ctx, hipMalloc(&s->res64, n * sizeof(double)));
        int rc = bb_fill_zero(ctx, s->res64, n * sizeof(double));
        if (rc) return rc;
    }
    if (prec == BB_PREC_MIXED && !s->res32) {
        BB_HIP(ctx, hipMalloc(&s->res32, n * sizeof(float)));
        int rc = bb_fill_zero(ctx, s->res32, n * sizeof(float));
        if (rc) return rc;
    }
    return BB_OK;
}

// default tiling: gather lift organisation, the fewest elements that give >= 192 owned
// threads, falling back to the ping-pong organisation where the gather one does not fit
static int default_tiling(bb_solver *s, int prec) {
    for (int lm : kStageLM) {
        int best = 0;
        for (int e : kStageEB) {
            if (e * s->NP > 1024) break;
            best = e;
            if (e * s->NP >= 192) break;
        }
        for (int e = best; e >= 1; e /= 2)
            if (stage_valid(s, prec, tiling_code(e, lm))) return tiling_code(e, lm);
    }
    return 0;
}

static int launch_stage(bb_solver *s, int prec, int tiling, const StageArgs &a) {
    int rc = stage_fn(s->pde, prec)(s->N, tiling, a, s->ctx->stream, false);
    if (rc == BB_ERR_UNSUPPORTED)
        return bb_fail(s->ctx, rc, "no fused kernel for N=%d with tiling %d", s->N, tiling);
    if (rc) return bb_fail(s->ctx, rc, "fused stage launch failed: %s", hipGetErrorString(t_launch_err));
    return BB_OK;
}

static StageArgs stage_args(bb_solver *s, int prec) {
    StageArgs a{};
    a.geoh = prec == BB_PREC_MIXED ? (const void *)s->geoh32 : (const void *)s->geoh64;
    a.invJ = s->invJ;
    a.nJ = s->nJ;
    a.conn = s->conn;
    a.tab = s->tab;
    a.l0a = s->l0a;
    a.K = s->K;
    a.flux_kind = s->flux_kind;
    for (int f = 0; f < 6; ++f) a.sign[f] = s->sign[f];
    a.res = prec == BB_PREC_MIXED ? (void *)s->res32 : (void *)s->res64;
    return a;
}

// ---- plain path: one kernel per operator of the reference RHS
// frees the plain path's arrays (after a failed set-up, before it is tried again)
static void free_plain(bb_solver *s) {
    void **ptrs[] = {(void **)&s->mapP, (void **)&s->bcn, (void **)&s->nx, (void **)&s->ny, (void **)&s->nz,
                     (void **)&s->uM, (void **)&s->flux, (void **)&s->du};
    for (void **p : ptrs)
        if (*p) {
            hipFree(*p);
            *p = nullptr;
        }
    for (auto &p : s->fu)
        if (p) {
            hipFree(p);
            p = nullptr;
        }
    for (auto &p : s->d9)
        if (p) {
            hipFree(p);
            p = nullptr;
        }
}

static int ensure_plain(bb_solver *s) {
    if (s->plain_ready) return BB_OK;
    free_plain(s);
    bb_ctx *ctx = s->ctx;
    const int NFP = s->NFP, N = s->N;
    const int64_t M = s->K * 4 * NFP;
    std::vector<int32_t> mp(M);
    std::vector<int8_t> bc(M);
    std::vector<double> nx(M), ny(M), nz(M);
    std::vector<int> ab(NFP * 2);
    for (int m = 0; m < NFP; ++m) {
        int a, b;
        decode2(N, m, a, b);
        ab[2 * m] = a;
        ab[2 * m + 1] = b;
    }
    for (int64_t e = 0; e < s->K; ++e)
        for (int f = 0; f < 4; ++f) {
            const int2 cn = s->conn_host[e * 4 + f];
            const int code = (cn.y >> 5) & 7, nf = cn.y & 3, p = (cn.y >> 2) & 7;
            for (int m = 0; m < NFP; ++m) {
                const int64_t idx = (e * 4 + f) * NFP + m;
                const int abc[3] = {ab[2 * m], ab[2 * m + 1], N - ab[2 * m] - ab[2 * m + 1]};
                if (code) {
                    mp[idx] = (int32_t)idx;
                } else {
                    const int a2 = abc[perm_entry(p, 0)], b2 = abc[perm_entry(p, 1)];
                    mp[idx] = (int32_t)(((int64_t)cn.x * 4 + nf) * NFP + lin2(N, a2, b2));
                }
                bc[idx] = (int8_t)code;
                nx[idx] = s->nJ_host[(e * 4 + f) * 3 + 0];
                ny[idx] = s->nJ_host[(e * 4 + f) * 3 + 1];
                nz[idx] = s->nJ_host[(e * 4 + f) * 3 + 2];
            }
        }
    const size_t fw = (size_t)M * s->F, vw = state_words(s);
    BB_HIP(ctx, hipMalloc(&s->mapP, M * sizeof(int32_t)));
    BB_HIP(ctx, hipMalloc(&s->bcn, M));
    BB_HIP(ctx, hipMalloc(&s->nx, M * sizeof(double)));
    BB_HIP(ctx, hipMalloc(&s->ny, M * sizeof(double)));
    BB_HIP(ctx, hipMalloc(&s->nz, M * sizeof(double)));
    BB_HIP(ctx, hipMemcpy(s->mapP, mp.data(), M * sizeof(int32_t), hipMemcpyHostToDevice));
    BB_HIP(ctx, hipMemcpy(s->bcn, bc.data(), M, hipMemcpyHostToDevice));
    BB_HIP(ctx, hipMemcpy(s->nx, nx.data(), M * sizeof(double), hipMemcpyHostToDevice));
    BB_HIP(ctx, hipMemcpy(s->ny, ny.data(), M * sizeof(double), hipMemcpyHostToDevice));
    BB_HIP(ctx, hipMemcpy(s->nz, nz.data(), M * sizeof(double), hipMemcpyHostToDevice));
    BB_HIP(ctx, hipMalloc(&s->uM, fw * sizeof(double)));
    BB_HIP(ctx, hipMalloc(&s->flux, fw * sizeof(double)));
    for (auto &p : s->fu) BB_HIP(ctx, hipMalloc(&p, vw * sizeof(double)));
    for (auto &p : s->d9) BB_HIP(ctx, hipMalloc(&p, vw * sizeof(double)));
    BB_HIP(ctx, hipMalloc(&s->du, vw * sizeof(double)));
    s->plain_ready = true;
    return BB_OK;
}

static int plain_rhs(bb_solver *s, const double *q) {
    bb_ctx *ctx = s->ctx;
    const int F = s->F, NP = s->NP, NFP = s->NFP, N = s->N;
    const int64_t K = s->K;
    const bb_layout lv{(int64_t)F * NP, 1, NP};        // [K][F][Np]
    const bb_layout lf{(int64_t)F * 4 * NFP, 1, 4 * NFP};  // [K][F][4Nfp]
    int rc;
    if ((rc = bb_face_gather(ctx, N, F, K, BB_F64, q, lv, s->uM, lf))) return rc;
    if ((rc = bb_interface_flux(ctx, s->pde, s->flux_kind, N, K, s->uM, lf, s->mapP, s->bcn, s->sign, s->nx, s->ny,
                                s->nz, s->flux, lf)))
        return rc;
    if ((rc = bb_pde_fdir3(ctx, s->pde, N, K, q, s->fu[0], s->fu[1], s->fu[2]))) return rc;
    for (int c = 0; c < 3; ++c)
        for (int ax = 0; ax < 3; ++ax)
            if ((rc = bb_deriv_apply(ctx, ax, N, F, K, BB_F64, s->fu[c], lv, s->d9[3 * c + ax], lv))) return rc;
    if ((rc = bb_lift_apply(ctx, N, F, K, BB_F64, BB_LIFT_PLAIN, s->flux, lf, s->du, lv))) return rc;
    return bb_plain_combine(ctx, s->pde, N, K, s->du, s->d9, s->geo, s->J);
}

// enqueue nsteps steps (no synchronisation; capturable)
static int enqueue_steps(bb_solver *s, int path, int prec, double dt, int nsteps, int *cur) {
    int rc;
    const int64_t nw = (int64_t)state_words(s);
    for (int st = 0; st < nsteps; ++st)
        for (int stage = 0; stage < 5; ++stage) {
            if (path == BB_PATH_PLAIN) {
                if ((rc = plain_rhs(s, s->q[*cur]))) return rc;
                if ((rc = bb_lsrk_update(s->ctx, nw, s->q[*cur], s->res64, s->du, kLsrkA[stage], kLsrkB[stage], dt, stage == 0)))
                    return rc;
            } else {
                StageArgs a = stage_args(s, prec);
                a.q_in = s->q[*cur];
                a.q_out = s->q[*cur ^ 1];
                a.ca = kLsrkA[stage];
                a.cb = kLsrkB[stage];
                a.dt = dt;
                a.first = stage == 0;
                if ((rc = launch_stage(s, prec, s->tiling[prec], a))) return rc;
                *cur ^= 1;
            }
        }
    return BB_OK;
}

static int check_path(bb_solver *s, int path, int prec) {
    if (path != BB_PATH_PLAIN && path != BB_PATH_FUSED) return bb_fail(s->ctx, BB_ERR_ARG, "path %d", path);
    if (prec != BB_PREC_F64 && prec != BB_PREC_MIXED) return bb_fail(s->ctx, BB_ERR_ARG, "prec %d", prec);
    if (path == BB_PATH_PLAIN && prec != BB_PREC_F64)
        return bb_fail(s->ctx, BB_ERR_UNSUPPORTED, "the plain path is FP64 only");
    int rc;
    if ((rc = ensure_res(s, path == BB_PATH_PLAIN ? BB_PREC_F64 : prec))) return rc;
    if (path == BB_PATH_PLAIN) return ensure_plain(s);
    if (!s->tiling[prec]) s->tiling[prec] = default_tiling(s, prec);
    if (!s->tiling[prec]) return bb_fail(s->ctx, BB_ERR_UNSUPPORTED, "no fused kernel fits N=%d", s->N);
    return BB_OK;
}

static void drop_graphs(bb_solver *s) {
    for (auto &g : s->gexec)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
    s->g_path = -1;
}

extern "C" {

static int solver_create(bb_ctx *ctx, int pde, int flux_kind, int N, int64_t K, const double *geo_host,
                         const double *J_host, const double *nJ_host, const int32_t *nbr_host,
                         const int8_t *nbr_face_host, const int8_t *perm_host, const int8_t *bc_host,
                         const double *bc_sign_host, bb_solver **out) {
    SolverHost h;
    int rc = solver_host_setup(ctx, pde, flux_kind, N, K, geo_host, J_host, nJ_host, nbr_host, nbr_face_host,
                               perm_host, bc_host, h);
    if (rc) return rc;
    bb_solver *s = new (std::nothrow) bb_solver();
    if (!s) return bb_fail(ctx, BB_ERR_NOMEM, "host memory allocation failed");
    s->ctx = ctx;
    s->pde = pde;
    s->F = h.F;
    s->N = N;
    s->K = K;
    s->flux_kind = flux_kind;
    s->NP = h.NP;
    s->NFP = h.NFP;
    if (bc_sign_host)
        for (int f = 0; f < h.F; ++f) s->sign[f] = bc_sign_host[f];
    auto fail = [&](hipError_t e) {
        int r = bb_fail(ctx, BB_ERR_HIP, "solver allocation: %s", hipGetErrorString(e));
        bb_solver_destroy(s);
        return r;
    };
    hipError_t e;
    const size_t vw = state_words(s);
    if ((e = hipMalloc(&s->geo, K * 9 * sizeof(double))) || (e = hipMalloc(&s->J, K * sizeof(double))) ||
        (e = hipMalloc(&s->invJ, K * sizeof(double))) || (e = hipMalloc(&s->geoh64, K * 9 * sizeof(double))) ||
        (e = hipMalloc(&s->geoh32, K * 9 * sizeof(float))) ||
        (e = hipMalloc(&s->nJ, K * 20 * sizeof(double))) || (e = hipMalloc(&s->conn, K * 4 * sizeof(int2))) ||
        (e = hipMalloc(&s->tab, h.tab.size() * sizeof(uint16_t))) ||
        (e = hipMalloc(&s->l0a, h.l0a.size() * sizeof(float))) ||
        (e = hipMalloc(&s->q[0], vw * sizeof(double))) || (e = hipMalloc(&s->q[1], vw * sizeof(double))))
        return fail(e);
    if ((e = hipMemcpy(s->geo, geo_host, K * 9 * sizeof(double), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(s->J, J_host, K * sizeof(double), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(s->invJ, h.invJ.data(), K * sizeof(double), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(s->geoh64, h.geoh64.data(), K * 9 * sizeof(double), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(s->geoh32, h.geoh32.data(), K * 9 * sizeof(float), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(s->nJ, h.nj5.data(), K * 20 * sizeof(double), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(s->conn, h.conn.data(), K * 4 * sizeof(int2), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(s->tab, h.tab.data(), h.tab.size() * sizeof(uint16_t), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(s->l0a, h.l0a.data(), h.l0a.size() * sizeof(float), hipMemcpyHostToDevice)))
        return fail(e);
    if ((rc = bb_fill_zero(ctx, s->q[0], vw * sizeof(double))) || (rc = bb_fill_zero(ctx, s->q[1], vw * sizeof(double)))) {
        bb_solver_destroy(s);
        return rc;
    }
    // kept for the plain path's face-node arrays (built on its first use); swaps do not allocate
    s->conn_host.swap(h.conn);
    s->nJ_host.swap(h.nJ3);
    *out = s;
    return BB_OK;
}

int bb_solver_destroy(bb_solver *s) {
    if (!s) return BB_ERR_NULL;
    hipStreamSynchronize(s->ctx->stream);
    drop_graphs(s);
    void *ptrs[] = {s->geo, s->geoh64, s->geoh32, s->J, s->invJ, s->nJ, s->conn, s->tab, s->l0a, s->mapP, s->bcn, s->nx, s->ny, s->nz, s->uM, s->flux, s->du,
                    s->q[0], s->q[1], s->res64, s->res32};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto p : s->fu)
        if (p) hipFree(p);
    for (auto p : s->d9)
        if (p) hipFree(p);
    delete s;
    return BB_OK;
}

int bb_solver_set_state(bb_solver *s, const double *q_host) {
    if (!s || !q_host) return BB_ERR_NULL;
    int rc = check_finite_host(s->ctx, "state", q_host, (int64_t)state_words(s));
    if (rc) return rc;
    BB_HIP(s->ctx, hipMemcpyAsync(s->q[s->cur], q_host, state_words(s) * sizeof(double), hipMemcpyHostToDevice,
                                  s->ctx->stream));
    BB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return BB_OK;
}

int bb_solver_get_state(bb_solver *s, double *q_host) {
    if (!s || !q_host) return BB_ERR_NULL;
    BB_HIP(s->ctx, hipMemcpyAsync(q_host, s->q[s->cur], state_words(s) * sizeof(double), hipMemcpyDeviceToHost,
                                  s->ctx->stream));
    BB_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return BB_OK;
}

int bb_solver_state_ptr(bb_solver *s, double **q_dev) {
    if (!s || !q_dev) return BB_ERR_NULL;
    *q_dev = s->q[s->cur];
    return BB_OK;
}

int bb_solver_check_state(bb_solver *s, int64_t *count) {
    if (!s || !count) return BB_ERR_NULL;
    const int rc = bb_nonfinite_device(s->ctx, BB_F64, s->q[s->cur], (int64_t)state_words(s), count);
    if (rc) return rc;
    if (*count) return bb_fail(s->ctx, BB_ERR_NONFINITE, "the state holds %lld NaN or Inf values", (long long)*count);
    return BB_OK;
}

static int solver_rhs(bb_solver *s, int path, int prec, double *du_host) {
    if (!s || !du_host) return BB_ERR_NULL;
    int rc = check_path(s, path, prec);
    if (rc) return rc;
    bb_ctx *ctx = s->ctx;
    const size_t vw = state_words(s);
    if (path == BB_PATH_PLAIN) {
        if ((rc = plain_rhs(s, s->q[s->cur]))) return rc;
        BB_HIP(ctx, hipMemcpyAsync(du_host, s->du, vw * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return BB_OK;
    }
    StageArgs a = stage_args(s, prec);
    a.q_in = s->q[s->cur];
    a.q_out = nullptr;
    a.rhs_only = 1;
    a.first = 1;
    a.dt = 1.0;
    if ((rc = launch_stage(s, prec, s->tiling[prec], a))) return rc;
    if (prec == BB_PREC_F64) {
        BB_HIP(ctx, hipMemcpyAsync(du_host, s->res64, vw * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> tmp(vw);
        BB_HIP(ctx, hipMemcpyAsync(tmp.data(), s->res32, vw * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t t = 0; t < vw; ++t) du_host[t] = tmp[t];
    }
    return BB_OK;
}

static int solver_step(bb_solver *s, int path, int prec, double dt, int nsteps) {
    if (!s) return BB_ERR_NULL;
    if (nsteps < 0) return bb_fail(s->ctx, BB_ERR_ARG, "nsteps %d", nsteps);
    int rc = check_path(s, path, prec);
    if (rc) return rc;
    return enqueue_steps(s, path, prec, dt, nsteps, &s->cur);
}

static int solver_graph_capture(bb_solver *s, int path, int prec, double dt, int nsteps) {
    if (!s) return BB_ERR_NULL;
    if (nsteps < 1) return bb_fail(s->ctx, BB_ERR_ARG, "nsteps %d", nsteps);
    int rc = check_path(s, path, prec);
    if (rc) return rc;
    bb_ctx *ctx = s->ctx;
    BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drop_graphs(s);
    // a fused step flips the ping-pong buffers once per stage; with an odd number of
    // stages per graph, capture one graph per starting buffer
    const int flip = path == BB_PATH_FUSED ? (5 * nsteps) & 1 : 0;
    for (int g = 0; g < (flip ? 2 : 1); ++g) {
        int start = s->cur ^ g, c = start;
        hipGraph_t graph = nullptr;
        BB_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ctx->capturing = true;
        rc = enqueue_steps(s, path, prec, dt, nsteps, &c);
        ctx->capturing = false;
        hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (rc) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return bb_fail(ctx, BB_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&s->gexec[start], graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return bb_fail(ctx, BB_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    s->g_path = path;
    s->g_flip = flip;
    return BB_OK;
}

int bb_solver_graph_launch(bb_solver *s) {
    if (!s) return BB_ERR_NULL;
    if (s->g_path < 0 || !s->gexec[s->cur]) return bb_fail(s->ctx, BB_ERR_STATE, "no graph captured for this buffer");
    BB_HIP(s->ctx, hipGraphLaunch(s->gexec[s->cur], s->ctx->stream));
    s->cur ^= s->g_flip;
    return BB_OK;
}

int bb_solver_set_tuning(bb_solver *s, int prec, int tiling) {
    if (!s) return BB_ERR_NULL;
    if (prec != BB_PREC_F64 && prec != BB_PREC_MIXED) return bb_fail(s->ctx, BB_ERR_ARG, "prec %d", prec);
    if (!stage_valid(s, prec, tiling))
        return bb_fail(s->ctx, BB_ERR_ARG, "tiling %d not compiled for N=%d", tiling, s->N);
    s->tiling[prec] = tiling;
    drop_graphs(s);
    return BB_OK;
}

static int solver_autotune(bb_solver *s, int prec, int *tiling) {
    if (!s) return BB_ERR_NULL;
    int rc = check_path(s, BB_PATH_FUSED, prec);
    if (rc) return rc;
    bb_ctx *ctx = s->ctx;
    StageArgs a = stage_args(s, prec);
    a.q_in = s->q[s->cur];
    a.q_out = s->q[s->cur ^ 1];  // the current state is only read
    a.ca = kLsrkA[1];
    a.cb = kLsrkB[1];
    a.dt = 0.0;
    a.first = 0;
    float best = 1e30f;
    int best_t = 0;
    for (int lm : kStageLM)
        for (int e : kStageEB) {
            const int t = tiling_code(e, lm);
            if (!stage_valid(s, prec, t)) continue;
            float ms = 0.f;
            if ((rc = launch_stage(s, prec, t, a))) return rc;  // warm
            BB_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
            for (int r = 0; r < 5; ++r)
                if ((rc = launch_stage(s, prec, t, a))) return rc;
            BB_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
            BB_HIP(ctx, hipEventSynchronize(ctx->ev1));
            BB_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            if (ms < best) {
                best = ms;
                best_t = t;
            }
        }
    if (!best_t) return bb_fail(ctx, BB_ERR_UNSUPPORTED, "no fused kernel compiled for N=%d", s->N);
    s->tiling[prec] = best_t;
    drop_graphs(s);
    if (tiling) *tiling = best_t;
    return BB_OK;
}

int bb_solver_stage_bytes(bb_solver *s, int prec, double *bytes, double *dofs) {
    if (!s) return BB_ERR_NULL;
    const double w = (double)state_words(s);
    const double sres = prec == BB_PREC_MIXED ? 4.0 : 8.0;
    // q in + q out (FP64), residual in + out, per-element geometry / normals / connectivity
    if (bytes) *bytes = w * (8.0 + 8.0 + 2.0 * sres) + (double)s->K * (9 * 8 + 8 + 20 * 8 + 4 * 8);
    if (dofs) *dofs = w;
    return BB_OK;
}

// The fused steps of bb_solver_step on the caller's state q, with the second state buffer
// and the residual in the context's scratch array (the device address of this device's
// copy of the library's static scratch, taken at context creation).  Calls nothing that
// allocates: the launchers, bb_fail (fixed buffer) and one device-to-device copy.
int bb_solver_step_noalloc(bb_solver *s, int prec, double *q, double dt, int nsteps) {
    if (!s || !q) return BB_ERR_NULL;
    bb_ctx *ctx = s->ctx;
    if (prec != BB_PREC_F64 && prec != BB_PREC_MIXED) return bb_fail(ctx, BB_ERR_ARG, "prec %d", prec);
    if (nsteps < 0) return bb_fail(ctx, BB_ERR_ARG, "nsteps %d", nsteps);
    if (!ctx->scratch) return bb_fail(ctx, BB_ERR_STATE, "no scratch array on device %d", ctx->device);
    const size_t vw = state_words(s);
    const size_t qbytes = (vw * sizeof(double) + 255) / 256 * 256;
    const size_t need = qbytes + vw * (prec == BB_PREC_MIXED ? sizeof(float) : sizeof(double));
    if (need > ctx->scratch_bytes)
        return bb_fail(ctx, BB_ERR_ARG, "K*F*Np = %zu needs %zu bytes of scratch; the library holds %zu per GPU", vw,
                       need, ctx->scratch_bytes);
    if (!s->tiling[prec]) s->tiling[prec] = default_tiling(s, prec);
    if (!s->tiling[prec]) return bb_fail(ctx, BB_ERR_UNSUPPORTED, "no fused kernel fits N=%d", s->N);
    double *buf[2] = {q, (double *)ctx->scratch};
    void *res = (unsigned char *)ctx->scratch + qbytes;
    int cur = 0, rc;
    for (int st = 0; st < nsteps; ++st)
        for (int stage = 0; stage < 5; ++stage) {
            StageArgs a = stage_args(s, prec);
            a.res = res;
            a.q_in = buf[cur];
            a.q_out = buf[cur ^ 1];
            a.ca = kLsrkA[stage];
            a.cb = kLsrkB[stage];
            a.dt = dt;
            a.first = stage == 0;
            if ((rc = launch_stage(s, prec, s->tiling[prec], a))) return rc;
            cur ^= 1;
        }
    // an odd number of stages leaves the state in the scratch buffer
    if (cur) BB_HIP(ctx, hipMemcpyAsync(q, buf[1], vw * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return BB_OK;
}

// Entry points that allocate on the host (set-up tables, the plain path's face-node arrays,
// widened copies) run under bb_guard: an exception becomes BB_ERR_NOMEM.
int bb_solver_create(bb_ctx *ctx, int pde, int flux_kind, int N, int64_t K, const double *geo_host,
                     const double *J_host, const double *nJ_host, const int32_t *nbr_host,
                     const int8_t *nbr_face_host, const int8_t *perm_host, const int8_t *bc_host,
                     const double *bc_sign_host, bb_solver **out) {
    if (!ctx || !out) return BB_ERR_NULL;
    *out = nullptr;
    return bb_guard(ctx, [&] {
        return solver_create(ctx, pde, flux_kind, N, K, geo_host, J_host, nJ_host, nbr_host, nbr_face_host,
                             perm_host, bc_host, bc_sign_host, out);
    });
}

int bb_solver_rhs(bb_solver *s, int path, int prec, double *du_host) {
    if (!s || !du_host) return BB_ERR_NULL;
    return bb_guard(s->ctx, [&] { return solver_rhs(s, path, prec, du_host); });
}

int bb_solver_step(bb_solver *s, int path, int prec, double dt, int nsteps) {
    if (!s) return BB_ERR_NULL;
    return bb_guard(s->ctx, [&] { return solver_step(s, path, prec, dt, nsteps); });
}

int bb_solver_graph_capture(bb_solver *s, int path, int prec, double dt, int nsteps) {
    if (!s) return BB_ERR_NULL;
    return bb_guard(s->ctx, [&] { return solver_graph_capture(s, path, prec, dt, nsteps); });
}

int bb_solver_autotune(bb_solver *s, int prec, int *tiling) {
    if (!s) return BB_ERR_NULL;
    return bb_guard(s->ctx, [&] { return solver_autotune(s, prec, tiling); });
}

}  // extern "C"
