// bb_dense.hip -- the dense nodal comparator of extras/3d_nodal_wave.jl:57-73: the same
// wave/Maxwell RHS with dense Np x Np derivative matrices and a dense Np x 4Nfp lift applied
// to nodal values.  Those products are plain GEMMs (matrix x all columns of all elements),
// so they go to the vendor BLAS on the context's stream; the pointwise pieces reuse the
// plain path's kernels (face gather, interface flux, flux directions, geometric combine).
#include <rocblas/rocblas.h>

#include "bb_dispatch.hpp"

static int blas_handle(bb_ctx *ctx, rocblas_handle *h) {
    if (!ctx->blas) {
        rocblas_handle nh = nullptr;
        if (rocblas_create_handle(&nh) != rocblas_status_success)
            return bb_fail(ctx, BB_ERR_HIP, "rocblas_create_handle failed");
        ctx->blas = nh;
    }
    *h = (rocblas_handle)ctx->blas;
    if (rocblas_set_stream(*h, ctx->stream) != rocblas_status_success)
        return bb_fail(ctx, BB_ERR_HIP, "rocblas_set_stream failed");
    return BB_OK;
}

void bb_blas_release(bb_ctx *ctx) {
    if (ctx && ctx->blas) {
        rocblas_destroy_handle((rocblas_handle)ctx->blas);
        ctx->blas = nullptr;
    }
}

extern "C" {

int bb_gemm(bb_ctx *ctx, int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void *A, int64_t lda,
            const void *B, int64_t ldb, double beta, void *C, int64_t ldc) {
    if (!ctx) return BB_ERR_NULL;
    if (M < 0 || N < 0 || K < 0) return bb_fail(ctx, BB_ERR_ARG, "negative GEMM size");
    if (M == 0 || N == 0) return BB_OK;
    if (!A || !B || !C) return bb_fail(ctx, BB_ERR_NULL, "null GEMM operand");
    if (lda < M || ldb < K || ldc < M) return bb_fail(ctx, BB_ERR_ARG, "leading dimension too small");
    if (M > INT32_MAX || N > INT32_MAX || K > INT32_MAX || lda > INT32_MAX || ldb > INT32_MAX || ldc > INT32_MAX)
        return bb_fail(ctx, BB_ERR_ARG, "GEMM size beyond 32-bit indices");
    rocblas_handle h;
    int rc = blas_handle(ctx, &h);
    if (rc) return rc;
    rocblas_status st;
    if (dtype == BB_F64) {
        st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)M, (rocblas_int)N,
                           (rocblas_int)K, &alpha, (const double *)A, (rocblas_int)lda, (const double *)B,
                           (rocblas_int)ldb, &beta, (double *)C, (rocblas_int)ldc);
    } else if (dtype == BB_F32) {
        const float af = (float)alpha, bf = (float)beta;
        st = rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)M, (rocblas_int)N,
                           (rocblas_int)K, &af, (const float *)A, (rocblas_int)lda, (const float *)B,
                           (rocblas_int)ldb, &bf, (float *)C, (rocblas_int)ldc);
    } else {
        return bb_fail(ctx, BB_ERR_ARG, "dtype %d", dtype);
    }
    if (st != rocblas_status_success) return bb_fail(ctx, BB_ERR_HIP, "GEMM failed (rocblas status %d)", (int)st);
    return BB_OK;
}

int bb_pde_flux_dirs(bb_ctx *ctx, int pde, int N, int64_t K, const double *q, double *fx, double *fy, double *fz) {
    if (pde != BB_PDE_WAVE && pde != BB_PDE_MAXWELL) return bb_fail(ctx, BB_ERR_ARG, "pde %d", pde);
    int rc = bb_check_common(ctx, N, pde == BB_PDE_WAVE ? 4 : 6, K, 1, 20);
    if (rc) return rc;
    if (K == 0) return BB_OK;
    if (!q || !fx || !fy || !fz) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    return bb_pde_fdir3(ctx, pde, N, K, q, fx, fy, fz);
}

int bb_rhs_combine(bb_ctx *ctx, int pde, int N, int64_t K, double *du, const double *const d[9], const double *geo,
                   const double *J) {
    if (pde != BB_PDE_WAVE && pde != BB_PDE_MAXWELL) return bb_fail(ctx, BB_ERR_ARG, "pde %d", pde);
    int rc = bb_check_common(ctx, N, pde == BB_PDE_WAVE ? 4 : 6, K, 1, 20);
    if (rc) return rc;
    if (K == 0) return BB_OK;
    if (!du || !d || !geo || !J) return bb_fail(ctx, BB_ERR_NULL, "null operand");
    double *dd[9];
    for (int m = 0; m < 9; ++m) {
        if (!d[m]) return bb_fail(ctx, BB_ERR_NULL, "null derivative array %d", m);
        dd[m] = const_cast<double *>(d[m]);
    }
    return bb_plain_combine(ctx, pde, N, K, du, dd, geo, J);
}

}  // extern "C"
