/*
 * bbref.c -- CPU oracle (test infrastructure only, never shipped or measured).
 *
 * A plain-C, Float64 restatement of the reference package's Bernstein-Bezier
 * operators and of the example right-hand sides that call them.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
 * and only as a checker or as the timed CPU baseline.
 *
 * Every function cites the reference file:line it restates.  Index arguments
 * and tables are 0-based here (the reference is 1-based); the translation is
 * stated at each function.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   tables, l_j            : enumeration invariants + closed forms, bit-exact
 *   deriv_entry / deriv_mul: exact derivative of the Bernstein expansion
 *                            (rational arithmetic) + dense-vs-stencil
 *   elev_entry / elevation : polynomial identity V_N E = V_{N-1} (exact)
 *   reduction              : equals E^T (exact)
 *   face_lift / lift       : naive stack [L0; l_j E^T L0] and the exact
 *                            M^{-1}[M^f1..M^f4] from rational mass matrices
 *   rhs_wave / rhs_maxwell : invariants (constant state, linear fields,
 *                            energy decay, convergence to analytic modes)
 *   lsrk                   : exact stability polynomial of the published
 *                            Carpenter-Kennedy coefficients
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 20

/* ---------------------------------------------------------------- sizes */
int orc_np(int N) { return (N + 1) * (N + 2) * (N + 3) / 6; }
int orc_nfp(int N) { return (N + 1) * (N + 2) / 2; }

/* src/ordering.jl:8-16  tri_offsets(N): 21 entries, entry m = sum_{i=1..m}(N+2-i) */
void orc_tri_offsets(int N, int64_t out[21]) {
    int64_t count = 0;
    out[0] = 0;
    for (int i = 1; i <= 20; ++i) {
        count += N + 2 - i;
        out[i] = count;
    }
}

/* src/ordering.jl:18-26  tet_offsets(N): entry m = sum_{i=1..m} div((N+2-i)(N+3-i),2)
 * (Julia div truncates toward zero; the product of two consecutive integers is even,
 *  so the truncation never matters except for negative operands, which C also truncates.) */
void orc_tet_offsets(int N, int64_t out[21]) {
    int64_t count = 0;
    out[0] = 0;
    for (int i = 1; i <= 20; ++i) {
        count += ((int64_t)(N + 2 - i) * (N + 3 - i)) / 2;
        out[i] = count;
    }
}

/* src/ordering.jl:6  ij_to_linear(i,j,tri) = i + tri[j+1] + 1  (1-based result) */
int64_t orc_ij_to_linear(int i, int j, const int64_t tri[21]) { return i + tri[j] + 1; }

/* The same tables continued to 24 entries for internal use: reduction_multiply! of degree
 * N+1 (BernsteinLift.jl:142) reads row N+1 of tri_offsets(N+1), which for N = 20 lies past
 * the reference's 21-entry tuple (the reference raises a BoundsError there; the oracle
 * computes the mathematically defined value). */
static void tri_ext(int N, int64_t out[24]) {
    int64_t count = 0;
    out[0] = 0;
    for (int i = 1; i < 24; ++i) { count += N + 2 - i; out[i] = count; }
}
static void tet_ext(int N, int64_t out[24]) {
    int64_t count = 0;
    out[0] = 0;
    for (int i = 1; i < 24; ++i) { count += ((int64_t)(N + 2 - i) * (N + 3 - i)) / 2; out[i] = count; }
}

/* src/ordering.jl:33-35  ijk_to_linear = i + tri[j+1] + 1 + tet[k+1] - j*k  (1-based result) */
int64_t orc_ijk_to_linear(int i, int j, int k, const int64_t tri[21], const int64_t tet[21]) {
    return i + tri[j] + 1 + tet[k] - (int64_t)j * k;
}

/* src/ordering.jl:37-39  linear_to_ijkl_lookup(N): for k, for j, for i -> (i,j,k,N-i-j-k) */
void orc_linear_to_ijkl(int N, int32_t *out /* 4*Np */) {
    int r = 0;
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N - k; ++j)
            for (int i = 0; i <= N - j - k; ++i) {
                out[4 * r + 0] = i;
                out[4 * r + 1] = j;
                out[4 * r + 2] = k;
                out[4 * r + 3] = N - i - j - k;
                ++r;
            }
}

/* src/ordering.jl:47-51  2-D lookup: for j, for i -> (i,j,N-i-j) */
void orc_linear_to_ijk_2d(int N, int32_t *out /* 3*Nfp */) {
    int r = 0;
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N - j; ++i) {
            out[3 * r + 0] = i;
            out[3 * r + 1] = j;
            out[3 * r + 2] = N - i - j;
            ++r;
        }
}

static double binom(int n, int k) {
    /* exact for n <= 20 (values < 2^53); built from integers as Julia's binomial(N,j) */
    int64_t b = 1;
    for (int m = 1; m <= k; ++m) b = b * (n - k + m) / m;
    return (double)b;
}

/* src/constants.jl:10-12  l_j(N) = ntuple(j -> j<=N ? (isodd(j)?-1.0:1.0)*binomial(N,j)/(1.0+j) : 0.0, 20)
 * out[j-1] = l_j */
void orc_l_j(int N, double out[20]) {
    for (int j = 1; j <= 20; ++j) {
        if (j <= N) {
            double s = (j % 2) ? -1.0 : 1.0;
            out[j - 1] = (s * binom(N, j)) / (1.0 + j);
        } else {
            out[j - 1] = 0.0;
        }
    }
}

/* ------------------------------------------------------------ derivative */

/* src/BernsteinDerivativeMatrix_3D.jl:61-67  get_coeff */
static int get_coeff(int i1, int j1, int k1, int l1, int i2, int j2, int k2, int l2) {
    if (i1 == i2 && j1 == j2 && k1 == k2 && l1 == l2) return i1;
    if (i1 + 1 == i2 && j1 - 1 == j2 && k1 == k2 && l1 == l2) return j1;
    if (i1 + 1 == i2 && j1 == j2 && k1 - 1 == k2 && l1 == l2) return k1;
    if (i1 + 1 == i2 && j1 == j2 && k1 == k2 && l1 - 1 == l2) return l1;
    return 0;
}

/* src/BernsteinDerivativeMatrix_3D.jl:69-94  getindex(D{r,s,t}, m, n), m,n 0-based here.
 * axis 0 = r (:69-76), 1 = s (:78-85), 2 = t (:87-94). */
double orc_deriv_entry(int axis, int N, int m, int n) {
    int32_t a[4], b[4];
    /* linear_to_ijkl_lookup(N)[m] without building the whole table */
    int r = 0, found = 0;
    for (int k = 0; k <= N && found < 2; ++k)
        for (int j = 0; j <= N - k && found < 2; ++j)
            for (int i = 0; i <= N - j - k; ++i) {
                if (r == m) { a[0] = i; a[1] = j; a[2] = k; a[3] = N - i - j - k; ++found; }
                if (r == n) { b[0] = i; b[1] = j; b[2] = k; b[3] = N - i - j - k; ++found; }
                ++r;
            }
    int i1 = a[0], j1 = a[1], k1 = a[2], l1 = a[3];
    int i2 = b[0], j2 = b[1], k2 = b[2], l2 = b[3];
    int first;
    if (axis == 0) first = get_coeff(i1, j1, k1, l1, i2, j2, k2, l2);
    else if (axis == 1) first = get_coeff(j1, i1, k1, l1, j2, i2, k2, l2);
    else first = get_coeff(k1, j1, i1, l1, k2, j2, i2, l2);
    int second = get_coeff(l1, j1, k1, i1, l2, j2, k2, i2);
    return 0.5 * (first - second);
}

/* ijk_to_linear, 0-based, on the tables of degree N */
#define LIN(i, j, k) ((i) + tri[(j)] + tet[(k)] - (int64_t)(j) * (k))

/* src/BernsteinDerivativeMatrix_3D.jl:96-132 fast_Dr_multiply!, :138-174 fast_Ds_multiply!,
 * :180-216 fast_Dt_multiply!  (one column; x, out of length Np, strided by `stride`
 * so that an SVector{F} column can be processed field by field). */
void orc_deriv_mul_strided(int axis, int N, const double *x, double *out, int stride) {
    int64_t tri[24], tet[24];
    tri_ext(N, tri);
    tet_ext(N, tet);
    int64_t row = 0;
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N - k; ++j)
            for (int i = 0; i <= N - j - k; ++i) {
                int l = N - i - j - k;
                double val = 0.0;
                double xr = x[row * stride];
                double x1, x4;
                if (axis == 0) {
                    x1 = xr;
                    x4 = (i > 0) ? x[LIN(i - 1, j, k) * stride] : 0.0;
                    val += i * (x1 - x4);
                    if (j > 0) {
                        x1 = x[LIN(i + 1, j - 1, k) * stride];
                        x4 = x[LIN(i, j - 1, k) * stride];
                        val += j * (x1 - x4);
                    }
                    if (k > 0) {
                        x1 = x[LIN(i + 1, j, k - 1) * stride];
                        x4 = x[LIN(i, j, k - 1) * stride];
                        val += k * (x1 - x4);
                    }
                    x1 = (l > 0) ? x[LIN(i + 1, j, k) * stride] : 0.0;
                    x4 = xr;
                    val += l * (x1 - x4);
                } else if (axis == 1) {
                    if (i > 0) {
                        x1 = x[LIN(i - 1, j + 1, k) * stride];
                        x4 = x[LIN(i - 1, j, k) * stride];
                        val += i * (x1 - x4);
                    }
                    x1 = xr;
                    x4 = (j > 0) ? x[LIN(i, j - 1, k) * stride] : 0.0;
                    val += j * (x1 - x4);
                    if (k > 0) {
                        x1 = x[LIN(i, j + 1, k - 1) * stride];
                        x4 = x[LIN(i, j, k - 1) * stride];
                        val += k * (x1 - x4);
                    }
                    x1 = (l > 0) ? x[LIN(i, j + 1, k) * stride] : 0.0;
                    x4 = xr;
                    val += l * (x1 - x4);
                } else {
                    if (i > 0) {
                        x1 = x[LIN(i - 1, j, k + 1) * stride];
                        x4 = x[LIN(i - 1, j, k) * stride];
                        val += i * (x1 - x4);
                    }
                    if (j > 0) {
                        x1 = x[LIN(i, j - 1, k + 1) * stride];
                        x4 = x[LIN(i, j - 1, k) * stride];
                        val += j * (x1 - x4);
                    }
                    x1 = xr;
                    x4 = (k > 0) ? x[LIN(i, j, k - 1) * stride] : 0.0;
                    val += k * (x1 - x4);
                    x1 = (l > 0) ? x[LIN(i, j, k + 1) * stride] : 0.0;
                    x4 = xr;
                    val += l * (x1 - x4);
                }
                out[row * stride] = 0.5 * val;
                ++row;
            }
}

void orc_deriv_mul(int axis, int N, const double *x, double *out) {
    orc_deriv_mul_strided(axis, N, x, out, 1);
}

/* src/BernsteinDerivativeMatrix_3D.jl:222-231  column-batched mul!: x, out are Np x ncols,
 * column-major (column c starts at c*Np). */
void orc_deriv_mul_cols(int axis, int N, int64_t ncols, const double *x, double *out) {
    int Np = orc_np(N);
    for (int64_t c = 0; c < ncols; ++c) orc_deriv_mul(axis, N, x + c * Np, out + c * Np);
}

/* ------------------------------------------------------------- elevation */

/* src/ElevationMatrix.jl:17-24  getindex(ElevationMatrix{N}, m, n), 0-based m<Nfp(N), n<Nfp(N-1) */
double orc_elev_entry(int N, int m, int n) {
    int r = 0, i1 = 0, j1 = 0, k1 = 0, i2 = 0, j2 = 0, k2 = 0;
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N - j; ++i) {
            if (r == m) { i1 = i; j1 = j; k1 = N - i - j; }
            ++r;
        }
    r = 0;
    for (int j = 0; j <= N - 1; ++j)
        for (int i = 0; i <= N - 1 - j; ++i) {
            if (r == n) { i2 = i; j2 = j; k2 = N - 1 - i - j; }
            ++r;
        }
    if (i1 == i2 + 1 && j1 == j2 && k1 == k2) return (double)(i2 + 1) / N;
    if (i1 == i2 && j1 == j2 + 1 && k1 == k2) return (double)(j2 + 1) / N;
    if (i1 == i2 && j1 == j2 && k1 == k2 + 1) return (double)(k2 + 1) / N;
    return 0.0;
}

#define IJ(i, j) ((i) + tri[(j)])

/* src/BernsteinLift.jl:100-125  elevation_multiply!(out, N, x, tri_offsets(N)):
 * degree N (Nfp(N)) -> degree N+1 (Nfp(N+1)) */
void orc_elevation_multiply(int N, const double *x, double *out) {
    int64_t tri[24];
    tri_ext(N, tri);
    int row = 0;
    for (int j = 0; j <= N + 1; ++j)
        for (int i = 0; i <= N + 1 - j; ++i) {
            int k = N + 1 - i - j;
            double val = 0.0;
            if (i > 0) val = i * x[IJ(i - 1, j)];
            if (j > 0) val = fma((double)j, x[IJ(i, j - 1)], val);
            if (k > 0) val = fma((double)k, x[IJ(i, j)], val);
            out[row++] = val / (N + 1);
        }
}

/* src/BernsteinLift.jl:66-90  reduction_multiply!(out, N, x, tri_offsets(N)):
 * degree N -> degree N-1.  Safe in place (out == x), as the reference uses it. */
void orc_reduction_multiply(int N, const double *x, double *out) {
    int64_t tri[24];
    tri_ext(N, tri);
    int row = 0;
    for (int j = 0; j <= N - 1; ++j)
        for (int i = 0; i <= N - 1 - j; ++i) {
            double val = fma((double)(i + 1), x[IJ(i + 1, j)],
                             fma((double)(j + 1), x[IJ(i, j + 1)], (double)(N - i - j) * x[IJ(i, j)]));
            out[row++] = val / N;
        }
}

/* src/BernsteinLift.jl:139-164  face_lift_multiply!(out, N, x, E, F):
 * x (Nfp) -> out (Np, 3-D order, slab k=j holds l_j (E^N_{N-j})^T L0 x).
 * Scratch E has Nfp(N) entries, F has Nfp(N+1). */
void orc_face_lift(int N, const double *x, double *out, double *E, double *Fs) {
    double lj[20];
    orc_l_j(N, lj);
    orc_elevation_multiply(N, x, Fs);
    orc_reduction_multiply(N + 1, Fs, E);
    int nfp = orc_nfp(N);
    double scale = (double)((N + 1) * (N + 1)) / 2.0; /* (N + 1)^2/2: Int^Int then / 2 */
    for (int m = 0; m < nfp; ++m) E[m] *= scale;
    int index1 = nfp;
    for (int m = 0; m < nfp; ++m) out[m] = E[m];
    orc_reduction_multiply(N, E, E);
    for (int j = 1; j <= N; ++j) {
        int diff = (N + 1 - j) * (N + 2 - j) / 2;
        for (int i = 0; i < diff; ++i) out[index1 + i] = lj[j - 1] * E[i];
        if (j < N) {
            index1 += diff;
            orc_reduction_multiply(N - j, E, E);
        }
    }
}

/* src/BernsteinLift.jl:191-220  mul!(out, L::BernsteinLift, x): x = [f1; f2; f3; f4],
 * each Nfp(N); out Np.  scratch: Nfp(N) + Nfp(N+1) + 3*Np doubles. */
void orc_lift(int N, const double *x, double *out, double *scratch) {
    int64_t tri[24], tet[24];
    tri_ext(N, tri);
    tet_ext(N, tet);
    int nfp = orc_nfp(N), np = orc_np(N);
    double *E = scratch, *Fs = E + nfp, *G1 = Fs + orc_nfp(N + 1), *G2 = G1 + np, *G3 = G2 + np;
    orc_face_lift(N, x + 3 * nfp, out, E, Fs);
    orc_face_lift(N, x + 0 * nfp, G1, E, Fs);
    orc_face_lift(N, x + 1 * nfp, G2, E, Fs);
    orc_face_lift(N, x + 2 * nfp, G3, E, Fs);
    int row = 0;
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N - k; ++j)
            for (int i = 0; i <= N - k - j; ++i) {
                int l = N - i - j - k;
                out[row] += G1[LIN(i, k, j)];
                out[row] += G2[LIN(j, k, l)];
                out[row] += G3[LIN(j, k, i)];
                ++row;
            }
}

int orc_lift_scratch_len(int N) { return orc_nfp(N) + orc_nfp(N + 1) + 3 * orc_np(N); }

/* src/BernsteinLift.jl:264-269  column-batched lift: x is (4Nfp) x ncols, out Np x ncols */
void orc_lift_cols(int N, int64_t ncols, const double *x, double *out) {
    int nfp = orc_nfp(N), np = orc_np(N);
    double *s = (double *)malloc(sizeof(double) * orc_lift_scratch_len(N));
    for (int64_t c = 0; c < ncols; ++c) orc_lift(N, x + c * 4 * nfp, out + c * np, s);
    free(s);
}

/* single-face lift over columns (face_lift_multiply! on matrices, BernsteinLift.jl:174-179) */
void orc_face_lift_cols(int N, int64_t ncols, const double *x, double *out) {
    int nfp = orc_nfp(N), np = orc_np(N);
    double *E = (double *)malloc(sizeof(double) * (orc_nfp(N) + orc_nfp(N + 1)));
    for (int64_t c = 0; c < ncols; ++c) orc_face_lift(N, x + c * nfp, out + c * np, E, E + nfp);
    free(E);
}

/* ---------------------------------------------------------- face traces */

/* Face f (0-based: 0=f1 s=-1/j=0, 1=f2 r+s+t=-1/l=0, 2=f3 r=-1/i=0, 3=f4 t=-1/k=0),
 * face node (a,b) in 2-D order (b outer, a inner), c = N-a-b.  Returns the 0-based volume
 * index of the coefficient whose multi-index lies on the face, in the 2-D face ordering
 * that BernsteinLift.jl:207-217 expects for that face:
 *   f1: (i,k) = (a,b), j=0   -> LIN(a, 0, b)
 *   f2: (j,k) = (a,b), l=0   -> LIN(c, a, b)
 *   f3: (j,k) = (a,b), i=0   -> LIN(0, a, b)
 *   f4: (i,j) = (a,b), k=0   -> LIN(a, b, 0)
 * This plays the role of rd.Fmask (examples/3d_wave.jl:41): the trace of a Bernstein
 * polynomial on a face is the sub-vector of coefficients with zero exponent there. */
int orc_face_node_volume_index(int N, int f, int a, int b) {
    int64_t tri[24], tet[24];
    tri_ext(N, tri);
    tet_ext(N, tet);
    int c = N - a - b;
    switch (f) {
    case 0: return (int)LIN(a, 0, b);
    case 1: return (int)LIN(c, a, b);
    case 2: return (int)LIN(0, a, b);
    default: return (int)LIN(a, b, 0);
    }
}

void orc_fmask(int N, int32_t *out /* 4*Nfp */) {
    int nfp = orc_nfp(N);
    for (int f = 0; f < 4; ++f) {
        int m = 0;
        for (int b = 0; b <= N; ++b)
            for (int a = 0; a <= N - b; ++a) out[f * nfp + m++] = orc_face_node_volume_index(N, f, a, b);
    }
}

/* ------------------------------------------------------------- RHS */

/* Layouts (Julia column-major with SVector{F} elements, i.e. field fastest):
 *   q, du          : [K][Np][F]
 *   mapP           : [K][4Nfp] 0-based index into the flattened face array [K*4Nfp]
 *   bc             : [K][4Nfp] 0 = use mapP; 1 = wall (uP = sign .* uM)  (extension)
 *   nxJ, nyJ, nzJ  : [K][4Nfp]
 *   geo            : [K][9]  rxJ sxJ txJ ryJ syJ tyJ rzJ szJ tzJ
 *   J              : [K]
 * examples/3d_wave.jl:83 divides by md.J[1,1] for every element; the configs use
 * jittered meshes, so the element's own J[e] is used (as examples/3d_advec_scalar.jl:42). */

typedef struct {
    int N, F, pde; /* pde: 0 wave (p,u,v,w), 1 maxwell (Hx,Hy,Hz,Ex,Ey,Ez) */
    int flux_kind; /* 0 reference-verbatim, 1 Jf-consistent upwind */
    int64_t K;
    const int32_t *mapP;
    const int8_t *bc;
    const double *nxJ, *nyJ, *nzJ, *geo, *J;
    const double *bc_sign; /* F entries */
} orc_mesh_t;

/* examples/3d_wave.jl:20-33 */
static void wave_fx(const double *q, double *o) { o[0] = q[1]; o[1] = q[0]; o[2] = 0; o[3] = 0; }
static void wave_fy(const double *q, double *o) { o[0] = q[2]; o[1] = 0; o[2] = q[0]; o[3] = 0; }
static void wave_fz(const double *q, double *o) { o[0] = q[3]; o[1] = 0; o[2] = 0; o[3] = q[0]; }
/* examples/3d_maxwell.jl:20-33 (q = Hx Hy Hz Ex Ey Ez) */
static void mx_fx(const double *q, double *o) { o[0] = 0; o[1] = q[5]; o[2] = -q[4]; o[3] = 0; o[4] = -q[2]; o[5] = q[1]; }
static void mx_fy(const double *q, double *o) { o[0] = -q[5]; o[1] = 0; o[2] = q[3]; o[3] = q[2]; o[4] = 0; o[5] = -q[0]; }
static void mx_fz(const double *q, double *o) { o[0] = q[4]; o[1] = -q[3]; o[2] = 0; o[3] = -q[1]; o[4] = q[0]; o[5] = 0; }

/* examples/3d_wave.jl:45-54 (wave) and examples/3d_maxwell.jl:45-63 (maxwell):
 * the interface flux at one face node, from the jump d = uP - uM.
 * kind 0 restates the reference verbatim: its penalty terms carry Jf^0 (d_p, d) and
 * Jf^2 ((n.d) nJ) where nJ = n Jf.  kind 1 is the same upwind flux with every term
 * carrying Jf^1 (penalties d -> Jf d and (nJ.d) nJ -> (nJ.d) nJ / Jf); the reference's
 * scaling is unstable under explicit LSRK45 with dt = 0.5 h / N^2 on h <= 1/8 meshes. */
static void node_flux(int pde, int kind, const double *d, double nx, double ny, double nz, double *o) {
    double Jf = kind ? sqrt(nx * nx + ny * ny + nz * nz) : 1.0;
    if (pde == 0) {
        double ndotdU = nx * d[1] + ny * d[2] + nz * d[3];
        if (kind == 0) {
            o[0] = 0.5 * (d[0] - ndotdU);
            o[1] = 0.5 * ((ndotdU - d[0]) * nx);
            o[2] = 0.5 * ((ndotdU - d[0]) * ny);
            o[3] = 0.5 * ((ndotdU - d[0]) * nz);
        } else {
            double un = ndotdU / Jf;
            o[0] = 0.5 * (Jf * d[0] - ndotdU);
            o[1] = 0.5 * ((un - d[0]) * nx);
            o[2] = 0.5 * ((un - d[0]) * ny);
            o[3] = 0.5 * ((un - d[0]) * nz);
        }
    } else {
        double ndotdH = nx * d[0] + ny * d[1] + nz * d[2];
        double ndotdE = nx * d[3] + ny * d[4] + nz * d[5];
        double fx[6], fy[6], fz[6];
        mx_fx(d, fx);
        mx_fy(d, fy);
        mx_fz(d, fz);
        if (kind) { ndotdH /= Jf; ndotdE /= Jf; }
        double pen[6] = {ndotdH * nx, ndotdH * ny, ndotdH * nz, ndotdE * nx, ndotdE * ny, ndotdE * nz};
        for (int f = 0; f < 6; ++f)
            o[f] = 0.5 * fx[f] * nx + 0.5 * fy[f] * ny + 0.5 * fz[f] * nz + 0.5 * (Jf * d[f]) - 0.5 * pen[f];
    }
}

/* examples/3d_wave.jl:41 uM .= view(q, rd.Fmask, :)  (q, uM in [K][*][F] layout) */
void orc_face_gather(int N, int F, int64_t K, const double *q, double *uM) {
    int np = orc_np(N), nfp = orc_nfp(N);
    int32_t *fm = (int32_t *)malloc(sizeof(int32_t) * 4 * nfp);
    orc_fmask(N, fm);
    for (int64_t e = 0; e < K; ++e)
        for (int m = 0; m < 4 * nfp; ++m)
            for (int f = 0; f < F; ++f) uM[(e * 4 * nfp + m) * F + f] = q[(e * np + fm[m]) * F + f];
    free(fm);
}

/* examples/3d_wave.jl:43-56 / 3d_maxwell.jl:43-65: interface flux over all face nodes */
void orc_interface_flux(const orc_mesh_t *M, const double *uM, double *flux) {
    int F = M->F, nfp = orc_nfp(M->N);
    int64_t nf = M->K * 4 * nfp;
    for (int64_t idx = 0; idx < nf; ++idx) {
        double d[6], o[6];
        const double *um = uM + idx * F;
        if (M->bc && M->bc[idx]) {
            for (int f = 0; f < F; ++f) d[f] = M->bc_sign[f] * um[f] - um[f];
        } else {
            const double *up = uM + (int64_t)M->mapP[idx] * F;
            for (int f = 0; f < F; ++f) d[f] = up[f] - um[f];
        }
        node_flux(M->pde, M->flux_kind, d, M->nxJ[idx], M->nyJ[idx], M->nzJ[idx], o);
        for (int f = 0; f < F; ++f) flux[idx * F + f] = o[f];
    }
}

/* per-element part of examples/3d_wave.jl:58-80 (wave) / 3d_maxwell.jl:67-89 (maxwell),
 * then du ./= J[e] (:83 / :92), on one element of degree N: qe, due [Np][F], fe [4Nfp][F],
 * g the element's 9 geometric factors.  scratch: 12*Np*F + lift scratch doubles. */
static void rhs_element_core(int pde, int N, int F, const double *qe, const double *fe, double *due,
                             const double *g, double Je, double *scratch) {
    int np = orc_np(N), nfp = orc_nfp(N);
    double *fxu = scratch, *fyu = fxu + np * F, *fzu = fyu + np * F;
    double *d[9];
    for (int m = 0; m < 9; ++m) d[m] = fzu + np * F + m * np * F;
    double *ls = d[8] + np * F;
    double *xcol = ls + orc_lift_scratch_len(N);      /* 4Nfp */
    double *ocol = xcol + 4 * nfp;                     /* Np */
    for (int n = 0; n < np; ++n) {
        if (pde == 0) {
            wave_fx(qe + n * F, fxu + n * F);
            wave_fy(qe + n * F, fyu + n * F);
            wave_fz(qe + n * F, fzu + n * F);
        } else {
            mx_fx(qe + n * F, fxu + n * F);
            mx_fy(qe + n * F, fyu + n * F);
            mx_fz(qe + n * F, fzu + n * F);
        }
    }
    /* nine mul!(..., D{r,s,t}, f{x,y,z}u) on SVector{F} columns = F strided scalar matvecs */
    double *fu[3] = {fxu, fyu, fzu};
    for (int c = 0; c < 3; ++c)
        for (int ax = 0; ax < 3; ++ax)
            for (int f = 0; f < F; ++f) orc_deriv_mul_strided(ax, N, fu[c] + f, d[3 * c + ax] + f, F);
    /* mul!(view(du,:,e), LIFT, view(interface_flux,:,e)) on SVector{F}: per field */
    for (int f = 0; f < F; ++f) {
        for (int m = 0; m < 4 * nfp; ++m) xcol[m] = fe[m * F + f];
        orc_lift(N, xcol, ocol, ls);
        for (int n = 0; n < np; ++n) due[n * F + f] = ocol[n];
    }
    double rxJ = g[0], sxJ = g[1], txJ = g[2], ryJ = g[3], syJ = g[4], tyJ = g[5], rzJ = g[6],
           szJ = g[7], tzJ = g[8];
    for (int n = 0; n < np; ++n)
        for (int f = 0; f < F; ++f) {
            int o = n * F + f;
            double s = rxJ * d[0][o] + sxJ * d[1][o] + txJ * d[2][o] + ryJ * d[3][o] + syJ * d[4][o] +
                       tyJ * d[5][o] + rzJ * d[6][o] + szJ * d[7][o] + tzJ * d[8][o];
            if (pde == 0) due[o] -= s;
            else due[o] += s;
        }
    for (int o = 0; o < np * F; ++o) due[o] = due[o] / Je;
}

static void rhs_element(const orc_mesh_t *M, int64_t e, const double *q, const double *flux,
                        double *du, double *scratch) {
    int np = orc_np(M->N), nfp = orc_nfp(M->N);
    rhs_element_core(M->pde, M->N, M->F, q + e * np * M->F, flux + e * 4 * nfp * M->F, du + e * np * M->F,
                     M->geo + e * 9, M->J[e], scratch);
}

static int rhs_scratch_len(int N, int F) {
    return 12 * orc_np(N) * F + orc_lift_scratch_len(N) + 4 * orc_nfp(N) + orc_np(N);
}

/* examples/3d_wave.jl:36-84 / examples/3d_maxwell.jl:36-93 : rhs_matvec!(du, q, params, t)
 * nthreads > 1 runs the per-element loop as examples/3d_advec_multithreaded.jl:31 does
 * (Threads.@threads over elements, per-thread scratch as MultithreadedBernsteinLift,
 * BernsteinLift.jl:40-56), with a static schedule. */
void orc_rhs(int pde, int flux_kind, int N, int64_t K, const double *q, const int32_t *mapP, const int8_t *bc,
             const double *bc_sign, const double *nxJ, const double *nyJ, const double *nzJ,
             const double *geo, const double *J, double *du, int nthreads) {
    int F = pde == 0 ? 4 : 6;
    int nfp = orc_nfp(N);
    orc_mesh_t M = {N, F, pde, flux_kind, K, mapP, bc, nxJ, nyJ, nzJ, geo, J, bc_sign};
    double *uM = (double *)malloc(sizeof(double) * K * 4 * nfp * F);
    double *flux = (double *)malloc(sizeof(double) * K * 4 * nfp * F);
    orc_face_gather(N, F, K, q, uM);
    orc_interface_flux(&M, uM, flux);
    int sl = rhs_scratch_len(N, F);
    if (nthreads <= 1) {
        double *s = (double *)malloc(sizeof(double) * sl);
        for (int64_t e = 0; e < K; ++e) rhs_element(&M, e, q, flux, du, s);
        free(s);
    } else {
#ifdef _OPENMP
        double *s = (double *)malloc(sizeof(double) * sl * (size_t)nthreads);
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int64_t e = 0; e < K; ++e) rhs_element(&M, e, q, flux, du, s + (size_t)omp_get_thread_num() * sl);
        free(s);
#else
        double *s = (double *)malloc(sizeof(double) * sl);
        for (int64_t e = 0; e < K; ++e) rhs_element(&M, e, q, flux, du, s);
        free(s);
#endif
    }
    free(uM);
    free(flux);
}

/* ------------------------------------------------ elements of mixed degree */

/* the 6 permutations of {0,1,2} in lexicographic order */
static const int PERM3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

/* Elements that carry their own degree (many meshes batched, each element of degree
 * deg[e] <= 20).  No counterpart in the reference, whose RHS has one degree; built from its
 * pieces: the face trace of the neighbour, of degree Nn, is brought to the element's
 * degree N by the matrix T_{N,Nn} (degree elevation if Nn < N, the L2 projection onto
 * degree N if Nn > N; T + toff[N * 21 + Nn], row-major nfp(N) x nfp(Nn); toff < 0 or
 * Nn == N: identity), in the neighbour's own face orientation, and then read through the
 * face permutation exactly as mapP does for equal degrees.  The interface flux, volume
 * term and lift are those of orc_rhs at degree N.
 *   q, du : element e's [Np(N_e)][F] block at qoff[e] (doubles)
 *   nbr, nbf, perm, bc : [K][4] neighbour, its face, permutation code, 0 interior / 1 wall
 *   nJ : [K][4][3], geo : [K][9], J : [K]  */
void orc_rhs_md(int pde, int flux_kind, int64_t K, const int32_t *deg, const int64_t *qoff, const double *q,
                const int32_t *nbr, const int32_t *nbf, const int32_t *perm, const int32_t *bc,
                const double *bc_sign, const double *nJ, const double *geo, const double *J, const double *T,
                const int64_t *toff, double *du, int nthreads) {
    int F = pde == 0 ? 4 : 6;
    int nmax = 1;
    for (int64_t e = 0; e < K; ++e)
        if (deg[e] > nmax) nmax = deg[e];
    int nf_max = orc_nfp(nmax);
    /* per thread: rhs scratch, element flux [4Nfp][F], neighbour trace, transferred trace,
     * two fmask tables */
    int64_t sl = rhs_scratch_len(nmax, F) + 4 * nf_max * F + 2 * nf_max * F;
    int64_t il = 8 * nf_max;
#ifdef _OPENMP
    int nt = nthreads > 1 ? nthreads : 1;
#else
    int nt = 1;
#endif
    double *scr = (double *)malloc(sizeof(double) * sl * nt);
    int32_t *iscr = (int32_t *)malloc(sizeof(int32_t) * il * nt);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t e = 0; e < K; ++e) {
#ifdef _OPENMP
        int th = omp_get_thread_num();
#else
        int th = 0;
#endif
        double *s = scr + sl * th, *fe = s + rhs_scratch_len(nmax, F), *tn = fe + 4 * nf_max * F,
               *tt = tn + nf_max * F;
        int32_t *fm = iscr + il * th, *fmn = fm + 4 * nf_max;
        int N = deg[e], nfp = orc_nfp(N);
        const double *qe = q + qoff[e];
        orc_fmask(N, fm);
        for (int f = 0; f < 4; ++f) {
            int64_t t = e * 4 + f;
            const double *nj = nJ + t * 3;
            const double *up = NULL;
            if (!bc[t]) {
                int nb = nbr[t], Nn = deg[nb], nfpn = orc_nfp(Nn);
                orc_fmask(Nn, fmn);
                const double *qn = q + qoff[nb];
                for (int j = 0; j < nfpn; ++j)
                    for (int c = 0; c < F; ++c) tn[j * F + c] = qn[(int64_t)fmn[nbf[t] * nfpn + j] * F + c];
                if (Nn == N || toff[N * 21 + Nn] < 0) {
                    up = tn;
                } else {
                    const double *Tm = T + toff[N * 21 + Nn];
                    for (int r = 0; r < nfp; ++r)
                        for (int c = 0; c < F; ++c) {
                            double v = 0.0;
                            for (int j = 0; j < nfpn; ++j) v += Tm[(int64_t)r * nfpn + j] * tn[j * F + c];
                            tt[r * F + c] = v;
                        }
                    up = tt;
                }
            }
            int m = 0;
            for (int b = 0; b <= N; ++b)
                for (int a = 0; a <= N - b; ++a, ++m) {
                    const double *um = qe + (int64_t)fm[f * nfp + m] * F;
                    double d[6], o[6];
                    if (bc[t]) {
                        for (int c = 0; c < F; ++c) d[c] = bc_sign[c] * um[c] - um[c];
                    } else {
                        int abc[3] = {a, b, N - a - b}, p = perm[t];
                        int a2 = abc[PERM3[p][0]], b2 = abc[PERM3[p][1]];
                        int mm = a2 + b2 * (2 * N + 3 - b2) / 2;
                        for (int c = 0; c < F; ++c) d[c] = up[mm * F + c] - um[c];
                    }
                    node_flux(pde, flux_kind, d, nj[0], nj[1], nj[2], o);
                    for (int c = 0; c < F; ++c) fe[(f * nfp + m) * F + c] = o[c];
                }
        }
        rhs_element_core(pde, N, F, qe, fe, du + qoff[e], geo + e * 9, J[e], s);
    }
    free(scr);
    free(iscr);
}

/* ----------------------------------------------------------- time stepping */

/* Five-stage, fourth-order, 2N-storage low-storage Runge-Kutta (Carpenter & Kennedy,
 * NASA TM-109112, 1994).  The reference examples drive their RHS through OrdinaryDiffEq
 * (examples/3d_wave.jl:133-134); the configurations here ask for this integrator. */
static const double LSRK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                 -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double LSRK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                 1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                 2277821191437.0 / 14882151754819.0};

void orc_lsrk_coeffs(double *a, double *b) {
    for (int s = 0; s < 5; ++s) { a[s] = LSRK_A[s]; b[s] = LSRK_B[s]; }
}

/* one time step: for s: res = a_s res + dt rhs(q); q += b_s res.  res zeroed on entry. */
void orc_lsrk_step(int pde, int flux_kind, int N, int64_t K, double *q, double *res, double dt, const int32_t *mapP,
                   const int8_t *bc, const double *bc_sign, const double *nxJ, const double *nyJ,
                   const double *nzJ, const double *geo, const double *J, int nthreads) {
    int F = pde == 0 ? 4 : 6;
    int64_t n = K * orc_np(N) * F;
    double *du = (double *)malloc(sizeof(double) * n);
    memset(res, 0, sizeof(double) * n);
    for (int s = 0; s < 5; ++s) {
        orc_rhs(pde, flux_kind, N, K, q, mapP, bc, bc_sign, nxJ, nyJ, nzJ, geo, J, du, nthreads);
        for (int64_t m = 0; m < n; ++m) {
            res[m] = LSRK_A[s] * res[m] + dt * du[m];
            q[m] = q[m] + LSRK_B[s] * res[m];
        }
    }
    free(du);
}

/* the interface-flux loop alone (examples/3d_wave.jl:43-56 / 3d_maxwell.jl:43-65),
 * uM and flux in [K][4Nfp][F] */
void orc_flux(int pde, int flux_kind, int N, int64_t K, const double *uM, const int32_t *mapP, const int8_t *bc,
              const double *bc_sign, const double *nxJ, const double *nyJ, const double *nzJ, double *flux) {
    int F = pde == 0 ? 4 : 6;
    orc_mesh_t M = {N, F, pde, flux_kind, K, mapP, bc, nxJ, nyJ, nzJ, NULL, NULL, bc_sign};
    orc_interface_flux(&M, uM, flux);
}

/* ---------------------------------------------------------- extended precision */
/* face_lift_multiply! (src/BernsteinLift.jl:139-164) carried out in long double (x87
 * 64-bit mantissa): the reference against which the FP64 lift's own rounding is measured
 * in the lift-conditioning study.  x, out are double; ncols columns. */
void orc_face_lift_ld_cols(int N, int64_t ncols, const double *x, double *out) {
    int64_t tri[24];
    int nfp = orc_nfp(N), np = orc_np(N), nf1 = orc_nfp(N + 1);
    long double *E = (long double *)malloc(sizeof(long double) * (nfp + nf1));
    long double *Fs = E + nfp;
    for (int64_t c = 0; c < ncols; ++c) {
        const double *xc = x + c * nfp;
        double *oc = out + c * np;
        /* elevation N -> N+1 */
        tri_ext(N, tri);
        int row = 0;
        for (int j = 0; j <= N + 1; ++j)
            for (int i = 0; i <= N + 1 - j; ++i) {
                int k = N + 1 - i - j;
                long double v = 0.0L;
                if (i > 0) v += (long double)i * xc[i - 1 + tri[j]];
                if (j > 0) v += (long double)j * xc[i + tri[j - 1]];
                if (k > 0) v += (long double)k * xc[i + tri[j]];
                Fs[row++] = v / (N + 1);
            }
        /* reductions: degree n -> n-1 in place on a long double vector */
        #define RED_LD(n, src, dst)                                                            \
            do {                                                                               \
                int64_t t_[24];                                                                \
                tri_ext((n), t_);                                                              \
                int r_ = 0;                                                                    \
                for (int j_ = 0; j_ <= (n) - 1; ++j_)                                          \
                    for (int i_ = 0; i_ <= (n) - 1 - j_; ++i_) {                               \
                        long double v_ = (long double)(i_ + 1) * (src)[i_ + 1 + t_[j_]] +      \
                                         (long double)(j_ + 1) * (src)[i_ + t_[j_ + 1]] +      \
                                         (long double)((n) - i_ - j_) * (src)[i_ + t_[j_]];     \
                        (dst)[r_++] = v_ / (n);                                                \
                    }                                                                          \
            } while (0)
        RED_LD(N + 1, Fs, E);
        long double scale = (long double)((N + 1) * (N + 1)) / 2.0L;
        for (int m = 0; m < nfp; ++m) { E[m] *= scale; oc[m] = (double)E[m]; }
        int index1 = nfp;
        RED_LD(N, E, E);
        for (int j = 1; j <= N; ++j) {
            int diff = (N + 1 - j) * (N + 2 - j) / 2;
            long double lj = ((j % 2) ? -1.0L : 1.0L) * (long double)binom(N, j) / (1.0L + j);
            for (int i = 0; i < diff; ++i) oc[index1 + i] = (double)(lj * E[i]);
            if (j < N) {
                index1 += diff;
                RED_LD(N - j, E, E);
            }
        }
        #undef RED_LD
    }
    free(E);
}

/* one LSRK45 step of the mixed-degree system; element e advances with its own step dt[e]
 * (many independent meshes, each with its own step).  res zeroed on entry. */
void orc_lsrk_step_md(int pde, int flux_kind, int64_t K, const int32_t *deg, const int64_t *qoff, double *q,
                      double *res, const double *dt, const int32_t *nbr, const int32_t *nbf, const int32_t *perm,
                      const int32_t *bc, const double *bc_sign, const double *nJ, const double *geo,
                      const double *J, const double *T, const int64_t *toff, int nthreads) {
    int F = pde == 0 ? 4 : 6;
    int64_t n = qoff[K - 1] + (int64_t)orc_np(deg[K - 1]) * F;
    double *du = (double *)malloc(sizeof(double) * n);
    memset(res, 0, sizeof(double) * n);
    for (int s = 0; s < 5; ++s) {
        orc_rhs_md(pde, flux_kind, K, deg, qoff, q, nbr, nbf, perm, bc, bc_sign, nJ, geo, J, T, toff, du, nthreads);
        for (int64_t e = 0; e < K; ++e) {
            int64_t w = (int64_t)orc_np(deg[e]) * F;
            for (int64_t m = qoff[e]; m < qoff[e] + w; ++m) {
                res[m] = LSRK_A[s] * res[m] + dt[e] * du[m];
                q[m] = q[m] + LSRK_B[s] * res[m];
            }
        }
    }
    free(du);
}
