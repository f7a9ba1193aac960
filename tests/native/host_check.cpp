// Host-side set-up of libbbdg under the address and undefined-behaviour sanitizers, with
// no GPU: the stage kernel's index tables and matrix-core L0 tables of every degree, the
// solver and batch set-up on small valid inputs and on each input they must reject, the
// exception guard of the entry points, and the fixed-size error buffer.  Built host-only
// (no device code) by tests/test_sanitizers.py; prints "host checks passed" and exits 0.
#include <stdexcept>
#include <string>

#include "bb_host.hpp"

using namespace bb;

static int g_fail = 0, g_checks = 0;
#define CHECK(cond)                                                            \
    do {                                                                       \
        ++g_checks;                                                            \
        if (!(cond)) {                                                         \
            ++g_fail;                                                          \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
        }                                                                      \
    } while (0)

static void check_tables() {
    for (int N = 1; N <= kStageMaxN; ++N) {
        const std::vector<uint16_t> t = stage_tables(N);
        const int NP = np3(N), NFP = nfp2(N);
        CHECK(t.size() % 8 == 0 && (int)t.size() >= 2 * NP + 28 * NFP + 4 * NP);
        for (int n = 0; n < NP; ++n) {  // node -> (i, j, k) decodes back to n
            const int i = t[n] & 15, j = (t[n] >> 4) & 15, k = t[n] >> 8;
            CHECK(lin3(N, i, j, k) == n);
        }
        for (int f = 0; f < 4; ++f)  // own face nodes
            for (int m = 0; m < NFP; ++m) {
                const int ab = t[NP + np3(N - 1) + m];
                CHECK(t[2 * NP + f * NFP + m] == face_vidx(N, f, ab & 255, ab >> 8));
            }
        const std::vector<float> l0 = l0_mfma_table(N);
        CHECK(!l0.empty());
        double sum = 0;
        for (float v : l0) {
            CHECK(v == (float)(int)v && v >= 0);  // integer entries of 2 L0, exact in FP32
            sum += v;
        }
        // 2 L0 = (N+1)^2 E^T E with E the elevation N -> N+1, whose rows sum to 1, so the
        // entries of E^T E sum to the number of rows of E, nfp(N+1)
        CHECK(sum == (double)(N + 1) * (N + 1) * nfp2(N + 1));
    }
}

// the transfer pass's node permutations (node m of a face -> its node in the neighbour's
// orientation) against the single-degree kernels' neighbour-node table: the same volume
// index of the neighbour, and a bijection of the face's nodes
static void check_face_perm() {
    static const PermFwd t;
    for (int d = 1; d <= kStageMaxN; ++d) {
        const std::vector<uint16_t> tab = stage_tables(d);
        const uint16_t *nvid = tab.data() + 2 * np3(d) + 4 * nfp2(d);
        for (int p = 0; p < 6; ++p) {
            std::vector<int> seen(nfp2(d), 0);
            for (int m = 0; m < nfp2(d); ++m) {
                const int r = t.v[d][p][m];
                CHECK(r < nfp2(d));
                if (r >= nfp2(d)) continue;
                ++seen[r];
                if (p == 0) CHECK(r == m);
                int a, b;
                decode2(d, r, a, b);
                for (int g = 0; g < 4; ++g) CHECK(face_vidx(d, g, a, b) == nvid[(g * 6 + p) * nfp2(d) + m]);
            }
            for (int r = 0; r < nfp2(d); ++r) CHECK(seen[r] == 1);
        }
    }
}

static void check_solver_setup() {
    bb_ctx ctx;
    const int N = 3;
    const int64_t K = 1;
    double geo[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, J[1] = {0.5};
    double nJ[12] = {0, -1, 0, 1, 1, 1, -1, 0, 0, 0, 0, -1};
    int32_t nbr[4] = {0, 0, 0, 0};
    int8_t nf[4] = {0, 1, 2, 3}, perm[4] = {0, 0, 0, 0}, bc[4] = {1, 1, 1, 1};
    auto run = [&](int pde, int flux, int n, int64_t k, const double *g, const int32_t *nb, const int8_t *f,
                   const int8_t *p, const int8_t *b, const double *j, const double *nj) {
        SolverHost h;
        return solver_host_setup(&ctx, pde, flux, n, k, g, j, nj, nb, f, p, b, h);
    };
    {
        SolverHost h;
        CHECK(solver_host_setup(&ctx, BB_PDE_MAXWELL, BB_FLUX_UPWIND, N, K, geo, J, nJ, nbr, nf, perm, bc, h) == BB_OK);
        CHECK(h.F == 6 && h.NP == 20 && h.NFP == 10 && h.conn.size() == 4 && h.nj5.size() == 20);
        CHECK(h.conn[2].x == 0 && h.conn[2].y == (2 | (1 << 5)));
        CHECK(h.nj5[5 + 3] == sqrt(3.0) && h.nj5[4] == 1.0 && h.invJ[0] == 2.0 && h.geoh64[0] == 0.5);
        CHECK(h.tab == stage_tables(N) && !h.l0a.empty());
    }
    {
        SolverHost h;
        CHECK(solver_host_setup(nullptr, 0, 1, N, K, geo, J, nJ, nbr, nf, perm, bc, h) == BB_ERR_NULL);
    }
    CHECK(run(2, 1, N, K, geo, nbr, nf, perm, bc, J, nJ) == BB_ERR_ARG);
    CHECK(run(0, 2, N, K, geo, nbr, nf, perm, bc, J, nJ) == BB_ERR_ARG);
    CHECK(run(0, 1, 0, K, geo, nbr, nf, perm, bc, J, nJ) == BB_ERR_ARG);
    CHECK(run(0, 1, kStageMaxN + 1, K, geo, nbr, nf, perm, bc, J, nJ) == BB_ERR_ARG);
    CHECK(run(0, 1, N, 0, geo, nbr, nf, perm, bc, J, nJ) == BB_ERR_ARG);
    CHECK(run(0, 1, N, -1, geo, nbr, nf, perm, bc, J, nJ) == BB_ERR_ARG);
    CHECK(run(0, 1, N, kMaxElements + 1, geo, nbr, nf, perm, bc, J, nJ) == BB_ERR_ARG);
    CHECK(run(0, 1, N, K, nullptr, nbr, nf, perm, bc, J, nJ) == BB_ERR_NULL);
    CHECK(run(0, 1, N, K, geo, nbr, nf, perm, bc, nullptr, nJ) == BB_ERR_NULL);
    {
        int32_t bad_nbr[4] = {0, 1, 0, 0};  // neighbour 1 of a one-element mesh
        CHECK(run(0, 1, N, K, geo, bad_nbr, nf, perm, bc, J, nJ) == BB_ERR_ARG);
        int8_t bad_nf[4] = {0, 4, 2, 3}, bad_perm[4] = {0, 0, 6, 0}, bad_bc[4] = {1, 2, 1, 1};
        CHECK(run(0, 1, N, K, geo, nbr, bad_nf, perm, bc, J, nJ) == BB_ERR_ARG);
        CHECK(run(0, 1, N, K, geo, nbr, nf, bad_perm, bc, J, nJ) == BB_ERR_ARG);
        CHECK(run(0, 1, N, K, geo, nbr, nf, perm, bad_bc, J, nJ) == BB_ERR_ARG);
        double J0[1] = {0.0}, Jnan[1] = {NAN}, nJ0[12] = {0, -1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1};
        CHECK(run(0, 1, N, K, geo, nbr, nf, perm, bc, J0, nJ) == BB_ERR_ARG);
        CHECK(run(0, 1, N, K, geo, nbr, nf, perm, bc, Jnan, nJ) == BB_ERR_NONFINITE);
        double geo_inf[9] = {1, 0, 0, 0, INFINITY, 0, 0, 0, 1}, nJ_nan[12] = {0, -1, 0, 1, 1, NAN, -1, 0, 0, 0, 0, -1};
        CHECK(run(0, 1, N, K, geo_inf, nbr, nf, perm, bc, J, nJ) == BB_ERR_NONFINITE);
        CHECK(std::string(ctx.err).find("geo[4]") != std::string::npos);
        CHECK(run(0, 1, N, K, geo, nbr, nf, perm, bc, J, nJ_nan) == BB_ERR_NONFINITE);
        CHECK(std::string(ctx.err).find("nJ[5]") != std::string::npos);
        CHECK(run(0, 1, N, K, geo, nbr, nf, perm, bc, J, nJ0) == BB_ERR_ARG);
        CHECK(std::string(ctx.err).find("zero face normal") != std::string::npos);
    }
}

static void check_batch_setup() {
    bb_ctx ctx;
    // two elements of degrees 3 and 4 sharing face 0, all other faces walls
    const int64_t K = 2;
    int32_t deg[2] = {3, 4};
    double geo[18], J[2] = {1, 1}, dt[2] = {1e-3, 1e-3}, nJ[24];
    for (int t = 0; t < 18; ++t) geo[t] = (t % 4 == 0) ? 1.0 : 0.0;
    for (int t = 0; t < 8; ++t) {
        nJ[3 * t] = 0;
        nJ[3 * t + 1] = -1;
        nJ[3 * t + 2] = 0;
    }
    int32_t nbr[8] = {1, 0, 0, 0, 0, 1, 1, 1}, nf[8] = {0, 1, 2, 3, 0, 1, 2, 3}, perm[8] = {0};
    int32_t bc[8] = {0, 1, 1, 1, 0, 1, 1, 1};
    // transfer tables 3 <- 4 (projection, dense) and 4 <- 3 (elevation, sparse)
    const int n3 = nfp2(3), n4 = nfp2(4);
    std::vector<double> tr((size_t)2 * n3 * n4, 0.0);
    std::vector<int64_t> toff(121, -1);
    toff[11 * 4 + 3] = 0;        // elevation 3 -> 4 first
    toff[11 * 3 + 4] = n3 * n4;  // then the projection 4 -> 3
    for (int r = 0; r < n3; ++r)
        for (int c = 0; c < n4; ++c) tr[(size_t)n3 * n4 + (size_t)r * n4 + c] = 1.0 / n4;
    for (int r = 0; r < n4; ++r) tr[(size_t)r * n3 + r % n3] = 1.0;  // one entry per row
    const int64_t tl = (int64_t)tr.size();
    auto run = [&](int pde, int64_t k, const int32_t *d, const int64_t *to, int64_t len) {
        BatchHost h;
        return batch_host_setup(&ctx, pde, BB_FLUX_UPWIND, k, d, geo, J, nJ, nbr, nf, perm, bc, dt, tr.data(), to, len,
                                h);
    };
    {
        BatchHost h;
        CHECK(batch_host_setup(&ctx, BB_PDE_WAVE, BB_FLUX_UPWIND, K, deg, geo, J, nJ, nbr, nf, perm, bc, dt, tr.data(),
                               toff.data(), tl, h) == BB_OK);
        CHECK(h.groups.size() == 2 && h.groups[1].qoff == 4 * np3(3) && h.words == 4 * (np3(3) + np3(4)));
        CHECK(h.nboff[0] == 4 * np3(3) && h.nboff[4] == 0 && (h.conn[0].y >> 8) == 4 && (h.conn[4].y >> 8) == 3);
        // the pair matrices in the A-operand layout of the 16x16x4 product: 3 <- 4 first (the
        // first cross face in element order is (0, 0)), then 4 <- 3
        const int RT3 = (n3 + 15) / 16, KS4 = (n4 + 3) / 4, RT4 = (n4 + 15) / 16, KS3 = (n3 + 3) / 4;
        CHECK(h.xa64.size() == (size_t)64 * (RT3 * KS4 + RT4 * KS3) && h.xa32.size() == h.xa64.size());
        for (int ks = 0; ks < KS4; ++ks)
            for (int rt = 0; rt < RT3; ++rt)
                for (int L = 0; L < 64; ++L) {
                    const int r = 16 * rt + L % 16, c = 4 * ks + L / 16;
                    const double want = (r < n3 && c < n4) ? tr[(size_t)n3 * n4 + (size_t)r * n4 + c] : 0.0;
                    CHECK(h.xa64[(ks * RT3 + rt) * 64 + L] == want);
                }
        const size_t a43 = (size_t)64 * RT3 * KS4;
        CHECK(h.xa64[a43 + 0] == tr[0] && h.xa64[a43 + 0] == 1.0);         // W[0][0]
        CHECK(h.xa64[a43 + 5 + 16 * 1] == tr[(size_t)5 * n3 + 1]);            // W[5][1]: lane 21
        CHECK(h.xa64[a43 + 64 + 3 + 16 * 2] == tr[(size_t)3 * n3 + 6]);       // W[3][6]: ks 1, lane 35
        CHECK(h.xa64[a43 + 128 + 14 + 16 * 1] == tr[(size_t)14 * n3 + 9]);    // W[14][9]: ks 2
        CHECK(h.xa64[a43 + 128 + 14 + 16 * 2] == 0.0);                        // column 10: padding
        // tiles not all zero: every one of the dense projection, the 3 k steps of the elevation
        CHECK(h.xmask.size() == 242 && h.xmask[2 * (11 * 3 + 4)] == 0xF && h.xmask[2 * (11 * 3 + 4) + 1] == 0);
        CHECK(h.xmask[2 * (11 * 4 + 3)] == 0x7 && h.xmask[0] == 0);
        // conn.x is the neighbour element, as in the single-degree kernels
        CHECK(h.conn[0].x == 1 && h.conn[4].x == 0);
        // the cross faces, ordered by the receiving face's degree first: face (0, 0) (degree 3,
        // from element 1) before face (1, 0) (degree 4, from element 0); one chunk per pair.
        // Send slots in that order: element 1 face 0 sends first (n4 nodes)
        CHECK(h.xnbq.size() == 2 && h.xnbq[0] == 4 * np3(3) && h.xnbq[1] == 0);
        CHECK(h.xcode[0] == (0 | (4 << 2)) && h.xcode[1] == (0 | (3 << 2)));
        CHECK(h.xsrc[0] == 0 && h.xsrc[1] == 4 * n4 && h.swords == 4 * (n3 + n4));
        CHECK(h.soff[4] == 0 && h.soff[0] == 4 * n4 && h.soff[1] == -1 && h.soff[7] == -1);
        CHECK(h.xchunk.size() == 2);
        CHECK(h.xchunk[0].src0 == 0 && h.xchunk[0].out0 == 0 && h.xchunk[0].perms == 0);
        CHECK(h.xchunk[0].code == (1 | (3 << 8) | (4 << 12)) && h.xchunk[0].aoff == 0);
        CHECK(h.xchunk[1].src0 == 4 * n4 && h.xchunk[1].out0 == 4 * n3);
        CHECK(h.xchunk[1].code == (1 | (4 << 8) | (3 << 12)) && h.xchunk[1].aoff == (int)a43);
        CHECK(h.xwords == 4 * (n3 + n4));
        CHECK(h.xoff[0] == 0 && h.xoff[4] == 4 * n3 && h.xoff[1] == 0 && h.xoff.size() == 8);
    }
    CHECK(run(BB_PDE_MAXWELL, K, deg, toff.data(), tl) == BB_ERR_UNSUPPORTED);
    CHECK(run(BB_PDE_WAVE, 0, deg, toff.data(), tl) == BB_ERR_ARG);
    CHECK(run(BB_PDE_WAVE, kMaxElements + 1, deg, toff.data(), tl) == BB_ERR_ARG);
    CHECK(run(BB_PDE_WAVE, K, deg, toff.data(), -1) == BB_ERR_ARG);
    CHECK(std::string(ctx.err).find("transfer_len") != std::string::npos);
    CHECK(run(BB_PDE_WAVE, K, deg, toff.data(), tl - 1) == BB_ERR_ARG);  // the 3 <- 4 table is cut short
    CHECK(run(BB_PDE_WAVE, K, deg, nullptr, tl) == BB_ERR_NULL);
    int32_t down[2] = {4, 3}, high[2] = {3, kStageMaxN + 1};
    CHECK(run(BB_PDE_WAVE, K, down, toff.data(), tl) == BB_ERR_ARG);
    CHECK(run(BB_PDE_WAVE, K, high, toff.data(), tl) == BB_ERR_ARG);
    std::vector<int64_t> missing = toff;
    missing[11 * 3 + 4] = -1;
    CHECK(run(BB_PDE_WAVE, K, deg, missing.data(), tl) == BB_ERR_ARG);
    CHECK(std::string(ctx.err).find("no transfer table from degree 4 to 3") != std::string::npos);
    // face (0, 0) names face 1 of element 1, a wall: the faces do not match
    nf[0] = 1;
    CHECK(run(BB_PDE_WAVE, K, deg, toff.data(), tl) == BB_ERR_ARG);
    CHECK(std::string(ctx.err).find("do not match") != std::string::npos);
    nf[0] = 0;
    // non-finite steps, transfer entries and geometry
    dt[1] = NAN;
    CHECK(run(BB_PDE_WAVE, K, deg, toff.data(), tl) == BB_ERR_NONFINITE);
    CHECK(std::string(ctx.err).find("dt[1]") != std::string::npos);
    dt[1] = 1e-3;
    tr[7] = -INFINITY;
    CHECK(run(BB_PDE_WAVE, K, deg, toff.data(), tl) == BB_ERR_NONFINITE);
    CHECK(std::string(ctx.err).find("transfer[7]") != std::string::npos);
    tr[7] = 0.0;
    geo[17] = NAN;
    CHECK(run(BB_PDE_WAVE, K, deg, toff.data(), tl) == BB_ERR_NONFINITE);
    geo[17] = 0.0;
    CHECK(run(BB_PDE_WAVE, K, deg, toff.data(), tl) == BB_OK);
}

static void check_nonfinite_scan() {
    bb_ctx ctx;
    std::vector<double> x(1000, 1.0);
    CHECK(first_nonfinite(x.data(), (int64_t)x.size()) == -1);
    CHECK(first_nonfinite(x.data(), 0) == -1);
    x[999] = INFINITY;
    CHECK(first_nonfinite(x.data(), (int64_t)x.size()) == 999);
    CHECK(first_nonfinite(x.data(), 999) == -1);
    x[3] = -NAN;
    CHECK(first_nonfinite(x.data(), (int64_t)x.size()) == 3);
    x[3] = 1e308 * 10.0;  // overflows to Inf
    CHECK(check_finite_host(&ctx, "state", x.data(), (int64_t)x.size()) == BB_ERR_NONFINITE);
    CHECK(std::string(ctx.err).find("state[3]") != std::string::npos);
    x[3] = x[999] = 1.0;
    CHECK(check_finite_host(&ctx, "state", x.data(), (int64_t)x.size()) == BB_OK);
}

static void check_guard_and_errors() {
    bb_ctx ctx;
    CHECK(bb_guard(&ctx, [] { return 7; }) == 7);
    CHECK(bb_guard(&ctx, []() -> int { throw std::bad_alloc(); }) == BB_ERR_NOMEM);
    CHECK(std::string(ctx.err) == "host memory allocation failed");
    CHECK(bb_guard(&ctx, []() -> int { throw std::runtime_error("boom"); }) == BB_ERR_NOMEM);
    CHECK(std::string(ctx.err).find("boom") != std::string::npos);
    CHECK(bb_guard(&ctx, []() -> int { throw 3; }) == BB_ERR_NOMEM);
    CHECK(bb_guard(nullptr, []() -> int { throw std::bad_alloc(); }) == BB_ERR_NOMEM);
    // an exception thrown by the standard library inside a guarded body (a request beyond
    // max_size() throws std::length_error before anything is allocated)
    CHECK(bb_guard(&ctx, [] {
              std::vector<double> v;
              v.reserve(v.max_size() + 1);
              return 0;
          }) == BB_ERR_NOMEM);
    std::string big(4000, 'x');
    CHECK(bb_fail(&ctx, BB_ERR_ARG, "%s", big.c_str()) == BB_ERR_ARG);
    CHECK(std::string(ctx.err).size() == sizeof ctx.err - 1);
    CHECK(bb_check_common(nullptr, 3, 4, 1, 1, 20) == BB_ERR_NULL);
    CHECK(bb_check_common(&ctx, 3, 65, 1, 1, 20) == BB_ERR_ARG);
}

int main() {
    check_tables();
    check_face_perm();
    check_solver_setup();
    check_batch_setup();
    check_guard_and_errors();
    check_nonfinite_scan();
    if (g_fail) {
        fprintf(stderr, "%d of %d host checks failed\n", g_fail, g_checks);
        return 1;
    }
    printf("host checks passed: %d\n", g_checks);
    return 0;
}
