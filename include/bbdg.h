/*
 * bbdg.h -- C ABI of libbbdg.so: Bernstein-Bezier DG operators on MI355X (gfx950).
 *
 * The entry points replace the Julia method table that BernsteinBasis.jl exports
 * (src/BernsteinBasis.jl:19-23) and the right-hand side that the examples assemble
 * from it (examples/3d_wave.jl:36-84, examples/3d_maxwell.jl:36-93).  Each
 * declaration names the reference interface it replaces.
 *
 * Conventions
 *  - All operator pointers are DEVICE pointers unless the name says _host.
 *  - Indices are 0-based.  The reference's 1-based tables are exported as they are.
 *  - A "column" is one element x one field.  A bb_layout gives the strides (in
 *    array elements, not bytes) of the element, node and field indices, so both the
 *    reference's SVector{F} column-major layout ([K][Np][F]: field 1, node F,
 *    element Np*F) and the solver's SoA layout ([K][F][Np]: node 1, field Np,
 *    element F*Np) are accepted without copies.
 *  - Every function returns BB_OK (0) or an error code; nothing aborts or throws
 *    across the ABI.  bb_last_error() gives the message of the last failure on a
 *    context.  The reference performs no checks on the hot path (@inbounds).
 *  - A context owns one HIP stream; all work is ordered on it.  Contexts are not
 *    re-entrant: use one per host thread.
 */
#ifndef BBDG_H
#define BBDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum bb_status {
    BB_OK = 0,
    BB_ERR_ARG = 1,         /* bad enum / degree / count / stride */
    BB_ERR_NULL = 2,        /* null pointer */
    BB_ERR_HIP = 3,         /* HIP runtime error (message in bb_last_error) */
    BB_ERR_UNSUPPORTED = 4, /* valid request this build does not implement */
    BB_ERR_STATE = 5,       /* call out of order (e.g. graph launch before capture) */
    BB_ERR_NOMEM = 6,       /* host memory allocation (or other host-side) failure */
    BB_ERR_NONFINITE = 7    /* NaN or Inf in an input array or in a solver state */
};

enum bb_dtype { BB_F64 = 0, BB_F32 = 1 };
enum bb_axis { BB_AXIS_R = 0, BB_AXIS_S = 1, BB_AXIS_T = 2 };
enum bb_pde { BB_PDE_WAVE = 0, BB_PDE_MAXWELL = 1 };
/* interface flux: the reference's formula verbatim, or the same upwind flux with all
 * terms scaled by the face Jacobian Jf = |nJ| (stable under explicit RK on fine meshes) */
enum bb_flux { BB_FLUX_REFERENCE = 0, BB_FLUX_UPWIND = 1 };
/* lift arithmetic for FP32 data: everything FP32; FP32 chain + FP64 weighted
 * accumulation; FP32 chain carried as float-float (compensated, two-sum); as ACC64 with
 * L0 on the matrix cores (FP32 MFMA, FP32 accumulation) where a kernel for it is built
 * (4-face lift, compact field-major layout, F = 4, 5 <= N <= 9), else as ACC64 */
enum bb_lift_mode { BB_LIFT_PLAIN = 0, BB_LIFT_ACC64 = 1, BB_LIFT_TWOSUM = 2, BB_LIFT_ACC64_MFMA = 3 };
/* solver arithmetic: FP64 everywhere; mixed = FP32 derivative/flux/L0/chain, FP64 lift
 * accumulation and FP64 solution accumulator (FP32 RK residual register) */
enum bb_prec { BB_PREC_F64 = 0, BB_PREC_MIXED = 1 };
/* solver right-hand-side path: one kernel per operator (the reference's structure,
 * FP64 only) or one fused RHS + LSRK stage kernel */
enum bb_path { BB_PATH_PLAIN = 0, BB_PATH_FUSED = 1 };

typedef struct bb_layout {
    int64_t elem;  /* stride between elements */
    int64_t node;  /* stride between nodes (volume) or face nodes (face data) */
    int64_t field; /* stride between fields */
} bb_layout;

typedef struct bb_ctx bb_ctx;
typedef struct bb_solver bb_solver;
typedef struct bb_batch bb_batch;

/* ------------------------------------------------------------------ context */
int bb_ctx_create(int device, bb_ctx **out);
int bb_ctx_destroy(bb_ctx *ctx);
const char *bb_last_error(const bb_ctx *ctx);
const char *bb_version(void);
int bb_device_count(int *count);
/* gcnArchName, name and CU count of the context's device */
int bb_device_info(bb_ctx *ctx, char *arch, int arch_len, char *name, int name_len, int *cu_count);
/* HIP stream (hipStream_t) of the context, for callers that order their own work on it */
int bb_ctx_stream(bb_ctx *ctx, void **stream);

/* ------------------------------------------------------------------ memory (device) */
int bb_malloc(bb_ctx *ctx, size_t bytes, void **dptr);
int bb_free(bb_ctx *ctx, void *dptr);
int bb_memcpy_h2d(bb_ctx *ctx, void *dst, const void *src_host, size_t bytes);
int bb_memcpy_d2h(bb_ctx *ctx, void *dst_host, const void *src, size_t bytes);
int bb_memcpy_d2d(bb_ctx *ctx, void *dst, const void *src, size_t bytes);
int bb_fill_zero(bb_ctx *ctx, void *dst, size_t bytes); /* kernel fill, capturable */
/* dst = src by a plain copy kernel (16-byte aligned): the measured-bandwidth reference of
 * the roofline and the calibration kernel of the HBM byte counters.  No reference
 * counterpart. */
int bb_stream_copy(bb_ctx *ctx, void *dst, const void *src, size_t bytes);
int bb_synchronize(bb_ctx *ctx);
/* Counts the NaN and Inf values among n values of dtype at the device pointer x (waits for
 * the context's stream).  Returns BB_ERR_NONFINITE when the count is non-zero (the count is
 * written either way).  The operator entry points do not scan their inputs (the reference
 * checks nothing on its hot path, @inbounds); callers that need it scan with this.  No
 * reference counterpart. */
int bb_count_nonfinite(bb_ctx *ctx, int dtype, const void *x, int64_t n, int64_t *count);

/* HIP-event timing on the context stream.  bb_timer_stop returns the milliseconds
 * between the matching start and stop. */
int bb_timer_start(bb_ctx *ctx);
int bb_timer_stop(bb_ctx *ctx, float *ms);

/* ------------------------------------------------------------------ host tables
 * (no device work; bit-exact with the reference) */

/* src/ordering.jl:8-16 tri_offsets(N) and :18-26 tet_offsets(N) (21 entries each) and
 * src/constants.jl:10-12 l_j(N) (20 entries).  1 <= N <= 20. */
int bb_tables(int N, int64_t tri_offsets[21], int64_t tet_offsets[21], double l_j[20]);
/* src/ordering.jl:37-39 linear_to_ijkl_lookup(N): out[4*n+0..3] = (i,j,k,l), 0 <= N <= 20 */
int bb_linear_to_ijkl(int N, int32_t *out_host);
/* src/ordering.jl:47-51 bernstein_2d_scalar_multiindex_lookup(N)[1]: out[3*m..] = (i,j,k) */
int bb_linear_to_ijk_2d(int N, int32_t *out_host);
/* face traces (the role of rd.Fmask at examples/3d_wave.jl:41): out[f*Nfp + m] is the
 * volume index of face f's node m, faces f1..f4 in the order the lift expects
 * (src/BernsteinLift.jl:207-217) */
int bb_fmask(int N, int32_t *out_host);
/* dense entries: src/BernsteinDerivativeMatrix_3D.jl:69-94 getindex(D{r,s,t}, m, n) and
 * src/ElevationMatrix.jl:17-24 getindex(ElevationMatrix{N}, m, n) */
int bb_deriv_entry(int axis, int N, int m, int n, double *value);
int bb_elevation_entry(int N, int m, int n, double *value);

/* ------------------------------------------------------------------ operators
 * All act on K elements x F fields ("columns"), device pointers, ordered on ctx. */

/* mul!(out, D{r,s,t}, x), src/BernsteinDerivativeMatrix_3D.jl:134-136, :176-178,
 * :218-220 (vector) and :222-231 (column-batched); x, out: Np values per column. */
int bb_deriv_apply(bb_ctx *ctx, int axis, int N, int F, int64_t K, int dtype, const void *x, bb_layout lx,
                   void *out, bb_layout lo);

/* The three derivative matvecs combined with per-element geometric factors
 * (examples/3d_wave.jl:63-71 + :75-78): dx = rxJ Dr x + sxJ Ds x + txJ Dt x, etc.
 * geo: [K][9] FP64 = rxJ sxJ txJ ryJ syJ tyJ rzJ szJ tzJ. */
int bb_grad_apply(bb_ctx *ctx, int N, int F, int64_t K, int dtype, const void *x, bb_layout lx,
                  const double *geo, void *dx, void *dy, void *dz, bb_layout lo);

/* mul!(out, L::BernsteinLift, x), src/BernsteinLift.jl:191-220 (vector), :264-269
 * (matrix), threaded_mul! :233-262: x holds 4*Nfp face values per column in the order
 * [f1; f2; f3; f4] (node stride lx.node), out Np values.  mode applies to BB_F32. */
int bb_lift_apply(bb_ctx *ctx, int N, int F, int64_t K, int dtype, int mode, const void *x, bb_layout lx,
                  void *out, bb_layout lo);

/* face_lift_multiply!(out, L, x), src/BernsteinLift.jl:139-179: one face, Nfp -> Np
 * (slab k = j holds l_j (E^N_{N-j})^T L0 x).  x_dtype / out_dtype may differ (FP32 in,
 * FP64 out measures the chain precision); mode applies to FP32 input. */
int bb_face_lift_apply(bb_ctx *ctx, int N, int F, int64_t K, int x_dtype, int out_dtype, int mode,
                       const void *x, bb_layout lx, void *out, bb_layout lo);

/* ElevationMatrix{N} * x, src/ElevationMatrix.jl:11-24 (degree N-1 -> N on a face;
 * the same product as elevation_multiply!(out, N-1, x), src/BernsteinLift.jl:100-125).
 * x: Nfp(N-1) values per column, out: Nfp(N).  1 <= N <= 20. */
int bb_elevate_apply(bb_ctx *ctx, int N, int F, int64_t K, int dtype, const void *x, bb_layout lx, void *out,
                     bb_layout lo);

/* uM .= view(q, rd.Fmask, :), examples/3d_wave.jl:41 (out: 4*Nfp face values per column) */
int bb_face_gather(bb_ctx *ctx, int N, int F, int64_t K, int dtype, const void *q, bb_layout lq, void *uM,
                   bb_layout lm);

/* interface flux loop, examples/3d_wave.jl:43-56 (wave) / 3d_maxwell.jl:43-65 (Maxwell).
 * uM, flux: 4*Nfp face values per element x F fields; mapP: [K][4Nfp] indices into the
 * flattened face-node array (StartUpDG md.mapP, 0-based); bc: [K][4Nfp] 0 = interior,
 * 1 = wall (uP = bc_sign .* uM, may be NULL); nxJ/nyJ/nzJ: [K][4Nfp] FP64. FP64 only. */
int bb_interface_flux(bb_ctx *ctx, int pde, int flux_kind, int N, int64_t K, const double *uM, bb_layout lm,
                      const int32_t *mapP, const int8_t *bc, const double *bc_sign_host, const double *nxJ,
                      const double *nyJ, const double *nzJ, double *flux, bb_layout lf);

/* ------------------------------------------------------------------ solver
 * Device-resident mesh + state for the wave (F=4: p,u,v,w) or Maxwell (F=6: Hx,Hy,Hz,
 * Ex,Ey,Ez) system on K affine tets of degree N (1 <= N <= 10).  Replaces the
 * rhs_matvec! + ODEProblem pair of examples/3d_wave.jl:36-134 / 3d_maxwell.jl:36-138
 * with a five-stage low-storage RK45 (Carpenter & Kennedy 1994).
 *
 * Host arrays:  geo [K][9], J [K], nJ [K][4][3] (nxJ nyJ nzJ per face, faces f1..f4),
 *               nbr [K][4] neighbour element, nbr_face [K][4] (0..3),
 *               perm [K][4] (0..5: neighbour face vertex m = my face vertex p[m], p in
 *               lexicographic order of the permutations of {0,1,2}), bc [K][4]
 *               (0 interior, 1 wall), bc_sign [F] (uP = bc_sign .* uM on walls).
 * State layout: [K][F][Np] FP64.  Non-finite mesh values (geo, J, nJ) are refused with
 * BB_ERR_NONFINITE, as is a state with a NaN or Inf in bb_solver_set_state. */
int bb_solver_create(bb_ctx *ctx, int pde, int flux_kind, int N, int64_t K, const double *geo_host,
                     const double *J_host, const double *nJ_host, const int32_t *nbr_host,
                     const int8_t *nbr_face_host, const int8_t *perm_host, const int8_t *bc_host,
                     const double *bc_sign_host, bb_solver **out);
int bb_solver_destroy(bb_solver *s);
int bb_solver_set_state(bb_solver *s, const double *q_host);
int bb_solver_get_state(bb_solver *s, double *q_host);
/* device pointer of the current state ([K][F][Np] FP64) */
int bb_solver_state_ptr(bb_solver *s, double **q_dev);
/* count of NaN / Inf values in the current device state (a diverged run); BB_ERR_NONFINITE
 * when it is non-zero.  No reference counterpart. */
int bb_solver_check_state(bb_solver *s, int64_t *count);
/* du = rhs(q) of the current state, copied to du_host ([K][F][Np] FP64).  For
 * BB_PREC_MIXED the values are the FP32 residual register widened to FP64. */
int bb_solver_rhs(bb_solver *s, int path, int prec, double *du_host);
/* advance nsteps LSRK45 steps of size dt */
int bb_solver_step(bb_solver *s, int path, int prec, double dt, int nsteps);
/* capture nsteps steps into a hipGraph, then replay it with bb_solver_graph_launch */
int bb_solver_graph_capture(bb_solver *s, int path, int prec, double dt, int nsteps);
int bb_solver_graph_launch(bb_solver *s);
/* Tiling of the fused stage kernel: code = elements per workgroup (1, 2, 4) + 16 x lift
 * organisation (0: all chain vectors kept in LDS and gathered once per node; 1: two
 * chain buffers, per-step accumulation -- smaller LDS; 2, mixed precision only: as 0 with
 * L0 on the matrix cores, FP32 MFMA with FP32 accumulation; 3: as 0 with each face chain
 * owned by one wave, no workgroup barrier between flux, L0 and the chain; 4: as 3 with the
 * index tables read from global memory instead of LDS; 5: as 4 until a step's output fits
 * half a wave, then, after one workgroup barrier, each wave runs two chains in the two
 * halves of its lanes).  bb_solver_autotune times every
 * compiled tiling on the current state (which it leaves unchanged) and keeps the fastest. */
int bb_solver_autotune(bb_solver *s, int prec, int *tiling);
int bb_solver_set_tuning(bb_solver *s, int prec, int tiling);
/* HBM bytes one fused stage moves by the algorithmic count (state in/out, residual
 * in/out, neighbour traces, geometry, connectivity) and the DOF count K*Np*F */
int bb_solver_stage_bytes(bb_solver *s, int prec, double *bytes, double *dofs);

/* ---- stepping that allocates nothing -------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * fused path of bb_solver_step, with the state in the caller's device array q ([K][F][Np]
 * FP64, advanced in place by nsteps steps of size dt) and the kernels' scratch -- the
 * second state buffer of the stage ping-pong and the RK residual (FP64, or FP32 for
 * BB_PREC_MIXED) -- in the library's fixed scratch array of BB_SCRATCH_BYTES per GPU.
 * K*F*Np*8 bytes (rounded up to 256) plus K*F*Np*(8 or 4) bytes must fit in it, else
 * BB_ERR_ARG.  No memory is allocated.  Every call on one GPU uses that GPU's single
 * scratch array, which keeps the last call's contents: calls that overlap on one GPU (from
 * other contexts or host threads) race.  Nothing in the scratch has to start at zero: the
 * first stage of each step writes the residual before any stage reads it. */
#define BB_SCRATCH_BYTES ((size_t)64 << 20)
int bb_solver_step_noalloc(bb_solver *s, int prec, double *q, double dt, int nsteps);

/* ---- dense nodal comparator (extras/3d_nodal_wave.jl:57-73) ------------------------
 * The reference's comparison RHS applies dense Np x Np derivative matrices and a dense
 * Np x 4Nfp lift to nodal values (mul!(dfxdr, Dr, fxu), ..., mul!(du, LIFT, flux, 1, -1)).
 * bb_gemm is that product over all columns of all elements: C = alpha A B + beta C,
 * column-major, no transposes, on the context's stream (vendor BLAS, FP64 or FP32).
 * bb_pde_flux_dirs computes fx.(q), fy.(q), fz.(q) (examples/3d_wave.jl:59-61) on
 * [K][F][Np]; bb_rhs_combine does du -= (wave) or += (Maxwell) the nine geometric
 * products of d[0..8] = (dfx/dr, dfx/ds, dfx/dt, dfy/dr, ..., dfz/dt), then du ./= J
 * (examples/3d_wave.jl:75-83), all [K][F][Np] FP64. */
int bb_gemm(bb_ctx *ctx, int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void *A, int64_t lda,
            const void *B, int64_t ldb, double beta, void *C, int64_t ldc);
int bb_pde_flux_dirs(bb_ctx *ctx, int pde, int N, int64_t K, const double *q, double *fx, double *fy, double *fz);
int bb_rhs_combine(bb_ctx *ctx, int pde, int N, int64_t K, double *du, const double *const d[9], const double *geo,
                   const double *J);

/* ---- batches of elements of mixed degree ------------------------------------------
 * No single counterpart in the reference, whose examples run one mesh at one degree
 * (examples/3d_wave.jl:36-134); this is that RHS + LSRK45 for many independent meshes
 * at once, each element with its own degree 1 <= N <= 10 (wave system, F = 4).
 * Elements are numbered grouped by degree (deg non-decreasing); the state is the
 * concatenation over degree groups of [K_N][F][Np(N)] FP64.  nbr holds neighbours in
 * that numbering; nbr_face, perm, bc, nJ, geo, J as for bb_solver_create (int32 arrays);
 * dt [K] is each element's step (one per mesh), scaled by the dt_scale of a step call.
 * A neighbour trace of degree Nn != N is brought to degree N by the nfp(N) x nfp(Nn)
 * row-major matrix at transfer + transfer_off[11 N + Nn] (degree elevation for Nn < N,
 * the L2 projection for Nn > N), then the upwind flux is that of bb_solver_create.
 * Faces must match mutually (the neighbour across a face names that face back).
 * Each stage launches a transfer pass for the faces whose neighbour has another degree,
 * then the fused stage kernel of each degree present, those on four streams joined into
 * the context's stream (also inside captured graphs). */
int bb_batch_create(bb_ctx *ctx, int pde, int flux_kind, int64_t K, const int32_t *deg_host, const double *geo_host,
                    const double *J_host, const double *nJ_host, const int32_t *nbr_host,
                    const int32_t *nbr_face_host, const int32_t *perm_host, const int32_t *bc_host,
                    const double *bc_sign_host, const double *dt_host, const double *transfer_host,
                    const int64_t *transfer_off_host /* [11][11] */, int64_t transfer_len, bb_batch **out);
int bb_batch_destroy(bb_batch *b);
int bb_batch_state_words(bb_batch *b, int64_t *words);
int bb_batch_set_state(bb_batch *b, const double *q_host);
int bb_batch_get_state(bb_batch *b, double *q_host);
/* as bb_solver_check_state, over every degree group (non-finite mesh values, steps, transfer
 * entries and states are refused as for a solver) */
int bb_batch_check_state(bb_batch *b, int64_t *count);
int bb_batch_rhs(bb_batch *b, int prec, double *du_host);
int bb_batch_step(bb_batch *b, int prec, double dt_scale, int nsteps);
int bb_batch_graph_capture(bb_batch *b, int prec, double dt_scale, int nsteps);
int bb_batch_graph_launch(bb_batch *b);
/* tilings[N] (11 entries) receives the chosen tiling code of each degree present */
int bb_batch_autotune(bb_batch *b, int prec, int *tilings);
int bb_batch_set_tuning(bb_batch *b, int prec, int N, int tiling);

#ifdef __cplusplus
}
#endif
#endif /* BBDG_H */
