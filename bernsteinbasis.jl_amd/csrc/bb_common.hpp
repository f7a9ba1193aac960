// bb_common.hpp -- index algebra, context and error plumbing shared by the kernels.
//
// Bernstein multi-indices (i,j,k,l), i+j+k+l = N, are stored in the reference's linear
// order: k outer, j, i inner (src/ordering.jl:37-39, docs/src/notation.md:34).  The
// closed forms below are the 0-based restatement of src/ordering.jl:6-35:
//   tri(n, j)     = j (2n + 3 - j) / 2            offset of row j of a degree-n triangle
//   tet(N, k)     = Np(N) - Np(N - k)              offset of slab k
//   lin3(N,i,j,k) = i + tri(N - k, j) + tet(N, k)  (= ijk_to_linear - 1, since
//                                                   tri_N(j) - j k = tri_{N-k}(j))
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <exception>
#include <new>

#include "bbdg.h"

// largest element count any entry point accepts (24-bit record offsets and int32 face-node
// indices of the plain path stay in range below it)
constexpr int64_t kMaxElements = (int64_t(1) << 31) / 4;

namespace bb {

__host__ __device__ constexpr int np3(int N) { return (N + 1) * (N + 2) * (N + 3) / 6; }
__host__ __device__ constexpr int nfp2(int N) { return (N + 1) * (N + 2) / 2; }
__host__ __device__ constexpr int tri(int n, int j) { return j * (2 * n + 3 - j) / 2; }
__host__ __device__ constexpr int tet(int N, int k) { return np3(N) - np3(N - k); }
__host__ __device__ constexpr int lin3(int N, int i, int j, int k) { return i + tri(N - k, j) + tet(N, k); }
__host__ __device__ constexpr int lin2(int n, int a, int b) { return a + tri(n, b); }

// Volume index of node (a,b) (c = N-a-b) of face f, faces in the order of
// src/BernsteinLift.jl:191-217 ([f1; f2; f3; f4]):
//   f1 (j=0): (i,k)=(a,b)   f2 (l=0): (j,k)=(a,b), i=c   f3 (i=0): (j,k)=(a,b)   f4 (k=0): (i,j)=(a,b)
__host__ __device__ constexpr int face_vidx(int N, int f, int a, int b) {
    return f == 0 ? lin3(N, a, 0, b) : f == 1 ? lin3(N, N - a - b, a, b) : f == 2 ? lin3(N, 0, a, b) : lin3(N, a, b, 0);
}

// l_j of src/constants.jl:10-12 (l_0 = 1 for the L0 slab); a constant expression, so the
// division is the correctly rounded IEEE one the reference performs at run time.
__host__ __device__ constexpr double binom_d(int n, int k) {
    double b = 1.0;
    for (int m = 1; m <= k; ++m) b = b * (n - k + m) / m;  // exact: partial products are integers < 2^53
    return b;
}
__host__ __device__ constexpr double lj(int N, int j) {
    return j == 0 ? 1.0 : (((j & 1) ? -1.0 : 1.0) * binom_d(N, j)) / (1.0 + j);
}

// the 6 permutations of {0,1,2} in lexicographic order (face orientation codes)
__host__ __device__ constexpr int perm_entry(int p, int m) {
    // rows: 012 021 102 120 201 210
    return (p == 0) ? m : (p == 1) ? (m == 0 ? 0 : 3 - m) : (p == 2) ? (m == 0 ? 1 : m == 1 ? 0 : 2)
         : (p == 3) ? (m == 0 ? 1 : m == 1 ? 2 : 0) : (p == 4) ? (m == 0 ? 2 : m == 1 ? 0 : 1)
                    : (2 - m);
}

struct Lay {
    int64_t e, n, f;
};
inline Lay lay(bb_layout l) { return Lay{l.elem, l.node, l.field}; }

template <typename T> struct DT;
template <> struct DT<double> { static constexpr int code = BB_F64; };
template <> struct DT<float> { static constexpr int code = BB_F32; };

}  // namespace bb

struct bb_ctx {
    int device = 0;
    int cus = 0;  // compute units of the device (grid sizes of persistent kernels)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool capturing = false;
    void *blas = nullptr;  // library GEMM handle, created on first use (bb_gemm)
    // this device's copy of the library's static scratch array (bb_runtime.hip): its device
    // address, taken once per device at context creation, and its size in bytes
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    unsigned long long *counter = nullptr;  // device counter of bb_count_nonfinite, allocated on first use
    char err[1024] = {0};  // message of the last failure (fixed size: recording it never allocates)
};

// error helpers (host).  bb_fail formats the message into the context's fixed buffer
// (truncating), so it cannot fail or throw.
__attribute__((format(printf, 3, 4))) inline int bb_fail(bb_ctx *ctx, int code, const char *fmt, ...) noexcept {
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->err, sizeof ctx->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

// Runs the body of an entry point that allocates on the host.  A failed host allocation,
// or any other C++ exception, becomes a status code: nothing throws across the C ABI.
template <class Fn> inline int bb_guard(bb_ctx *ctx, Fn &&fn) noexcept {
    try {
        return fn();
    } catch (const std::bad_alloc &) {
        return bb_fail(ctx, BB_ERR_NOMEM, "host memory allocation failed");
    } catch (const std::exception &e) {
        return bb_fail(ctx, BB_ERR_NOMEM, "host-side failure: %s", e.what());
    } catch (...) {
        return bb_fail(ctx, BB_ERR_NOMEM, "host-side failure (unknown exception)");
    }
}

// the HIP error of the last kernel launch made by a launcher that returns a status code,
// per host thread (hipGetLastError clears the error, so the launcher keeps it for the
// message its caller formats)
namespace bb {
inline thread_local hipError_t t_launch_err = hipSuccess;
inline int launch_status() {
    t_launch_err = hipGetLastError();
    return t_launch_err == hipSuccess ? BB_OK : BB_ERR_HIP;
}
}  // namespace bb

// count of NaN / Inf values among n values of dtype at device pointer x, into *count
// (bb_runtime.hip; waits for the context's stream)
int bb_nonfinite_device(bb_ctx *ctx, int dtype, const void *x, int64_t n, int64_t *count);

// releases the context's library GEMM handle (bb_dense.hip)
void bb_blas_release(bb_ctx *ctx);
#define BB_HIP(ctx, call)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return bb_fail((ctx), BB_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                                     \
    } while (0)
#define BB_LAUNCH_CHECK(ctx) BB_HIP(ctx, hipGetLastError())

// shared argument validation: degree in [nmin, nmax], 1 <= F <= 64, 0 <= K <= 2^29
inline int bb_check_common(bb_ctx *ctx, int N, int F, int64_t K, int nmin, int nmax) noexcept {
    if (!ctx) return BB_ERR_NULL;
    if (N < nmin || N > nmax) return bb_fail(ctx, BB_ERR_ARG, "degree N=%d outside [%d,%d]", N, nmin, nmax);
    if (F < 1 || F > 64) return bb_fail(ctx, BB_ERR_ARG, "field count F=%d outside [1,64]", F);
    if (K < 0) return bb_fail(ctx, BB_ERR_ARG, "element count K=%lld < 0", (long long)K);
    if (K > kMaxElements) return bb_fail(ctx, BB_ERR_ARG, "element count K=%lld too large", (long long)K);
    return BB_OK;
}
