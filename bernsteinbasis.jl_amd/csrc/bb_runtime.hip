// bb_runtime.hip -- context, device memory, timing and the host-side tables of the C ABI.
#include <string.h>

#include <vector>

#include <mutex>

#include "bb_common.hpp"

// The scratch of the entry points that allocate nothing (bb_solver_step_noalloc): a fixed
// array in the library's code object, one copy per GPU, held until the process ends.  In
// host code `scratch` names a host-side stand-in, never the device memory: its device
// address comes from hipGetSymbolAddress, once per device (with that device current), and
// only that address is passed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[BB_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// device address of this device's scratch (the current device must be `device`)
hipError_t scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[device] = p;
    }
    *out = g_scratch_dev[device];
    return hipSuccess;
}
}  // namespace

extern "C" {

const char *bb_version(void) { return "bbdg 0.1 gfx950"; }

int bb_device_count(int *count) {
    if (!count) return BB_ERR_NULL;
    if (hipGetDeviceCount(count) != hipSuccess) {
        *count = 0;
        return BB_ERR_HIP;
    }
    return BB_OK;
}

int bb_ctx_create(int device, bb_ctx **out) {
    if (!out) return BB_ERR_NULL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return BB_ERR_HIP;
    if (device < 0 || device >= n || device >= kMaxDevices) return BB_ERR_ARG;
    bb_ctx *c = new (std::nothrow) bb_ctx();
    if (!c) return BB_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        scratch_address(device, &c->scratch) != hipSuccess) {
        if (c->ev1) hipEventDestroy(c->ev1);
        if (c->ev0) hipEventDestroy(c->ev0);
        if (c->stream) hipStreamDestroy(c->stream);
        delete c;
        return BB_ERR_HIP;
    }
    c->scratch_bytes = BB_SCRATCH_BYTES;
    *out = c;
    return BB_OK;
}

int bb_ctx_destroy(bb_ctx *ctx) {
    if (!ctx) return BB_ERR_NULL;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    bb_blas_release(ctx);
    if (ctx->counter) hipFree(ctx->counter);
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return BB_OK;
}

const char *bb_last_error(const bb_ctx *ctx) { return ctx ? ctx->err : "null context"; }

int bb_ctx_stream(bb_ctx *ctx, void **stream) {
    if (!ctx || !stream) return BB_ERR_NULL;
    *stream = (void *)ctx->stream;
    return BB_OK;
}

int bb_device_info(bb_ctx *ctx, char *arch, int arch_len, char *name, int name_len, int *cu_count) {
    if (!ctx) return BB_ERR_NULL;
    hipDeviceProp_t p;
    BB_HIP(ctx, hipGetDeviceProperties(&p, ctx->device));
    if (arch && arch_len > 0) snprintf(arch, arch_len, "%s", p.gcnArchName);
    if (name && name_len > 0) snprintf(name, name_len, "%s", p.name);
    if (cu_count) *cu_count = p.multiProcessorCount;
    return BB_OK;
}

int bb_malloc(bb_ctx *ctx, size_t bytes, void **dptr) {
    if (!ctx || !dptr) return BB_ERR_NULL;
    BB_HIP(ctx, hipSetDevice(ctx->device));
    BB_HIP(ctx, hipMalloc(dptr, bytes ? bytes : 1));
    return BB_OK;
}

int bb_free(bb_ctx *ctx, void *dptr) {
    if (!ctx) return BB_ERR_NULL;
    if (dptr) BB_HIP(ctx, hipFree(dptr));
    return BB_OK;
}

int bb_memcpy_h2d(bb_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return BB_ERR_NULL;
    BB_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BB_OK;
}

int bb_memcpy_d2h(bb_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return BB_ERR_NULL;
    BB_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BB_OK;
}

int bb_memcpy_d2d(bb_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return BB_ERR_NULL;
    BB_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return BB_OK;
}

}  // extern "C"

__global__ void bb_fill_zero_kernel(uint32_t *p, size_t nwords) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += stride) p[i] = 0u;
}
__global__ void bb_fill_zero_bytes(uint8_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// dst = src in 16-byte words: the bandwidth reference for the roofline (read + write of
// `bytes` each) and the calibration kernel for the HBM byte counters
__global__ void bb_stream_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// *count += number of NaN / Inf among x[0 .. n): a grid-stride scan, each wave's count summed
// by ballot and one atomic per wave
template <typename T> __global__ void k_count_nonfinite(const T *__restrict__ x, int64_t n, unsigned long long *count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned c = 0;
    for (; i < n; i += stride) c += !isfinite(x[i]);
    // wave sum by a butterfly over the 64 lanes
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, (unsigned long long)c);
}

int bb_nonfinite_device(bb_ctx *ctx, int dtype, const void *x, int64_t n, int64_t *count) {
    if (!ctx || !count || (n > 0 && !x)) return BB_ERR_NULL;
    if (dtype != BB_F64 && dtype != BB_F32) return bb_fail(ctx, BB_ERR_ARG, "dtype %d", dtype);
    if (n < 0) return bb_fail(ctx, BB_ERR_ARG, "n=%lld < 0", (long long)n);
    *count = 0;
    if (!n) return BB_OK;
    if (!ctx->counter) BB_HIP(ctx, hipMalloc(&ctx->counter, sizeof(unsigned long long)));
    int rc = bb_fill_zero(ctx, ctx->counter, sizeof(unsigned long long));
    if (rc) return rc;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8 * 1024) blocks = 8 * 1024;
    if (dtype == BB_F64)
        k_count_nonfinite<double><<<(unsigned)blocks, 256, 0, ctx->stream>>>((const double *)x, n, ctx->counter);
    else
        k_count_nonfinite<float><<<(unsigned)blocks, 256, 0, ctx->stream>>>((const float *)x, n, ctx->counter);
    BB_LAUNCH_CHECK(ctx);
    unsigned long long c = 0;
    BB_HIP(ctx, hipMemcpyAsync(&c, ctx->counter, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
    BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *count = (int64_t)c;
    return BB_OK;
}

extern "C" {

int bb_count_nonfinite(bb_ctx *ctx, int dtype, const void *x, int64_t n, int64_t *count) {
    const int rc = bb_nonfinite_device(ctx, dtype, x, n, count);
    if (rc) return rc;
    if (*count) return bb_fail(ctx, BB_ERR_NONFINITE, "%lld of %lld values are NaN or Inf", (long long)*count, (long long)n);
    return BB_OK;
}

int bb_stream_copy(bb_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return BB_ERR_NULL;
    if ((((uintptr_t)dst | (uintptr_t)src | bytes) & 15u) != 0)
        return bb_fail(ctx, BB_ERR_ARG, "stream copy needs 16-byte aligned pointers and sizes");
    const size_t n = bytes / 16;
    if (!n) return BB_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 1024 * 256) blocks = 1024 * 256;
    bb_stream_copy_kernel<<<(unsigned)blocks, 256, 0, ctx->stream>>>((const uint4 *)src, (uint4 *)dst, n);
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

int bb_fill_zero(bb_ctx *ctx, void *dst, size_t bytes) {
    if (!ctx || (bytes && !dst)) return BB_ERR_NULL;
    if (!bytes) return BB_OK;
    if (((uintptr_t)dst & 3u) == 0 && (bytes & 3u) == 0) {
        size_t nw = bytes / 4;
        size_t blocks = (nw + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        bb_fill_zero_kernel<<<(unsigned)blocks, 256, 0, ctx->stream>>>((uint32_t *)dst, nw);
    } else {
        bb_fill_zero_bytes<<<(unsigned)((bytes + 255) / 256), 256, 0, ctx->stream>>>((uint8_t *)dst, bytes);
    }
    BB_LAUNCH_CHECK(ctx);
    return BB_OK;
}

int bb_synchronize(bb_ctx *ctx) {
    if (!ctx) return BB_ERR_NULL;
    BB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BB_OK;
}

int bb_timer_start(bb_ctx *ctx) {
    if (!ctx) return BB_ERR_NULL;
    BB_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    return BB_OK;
}

int bb_timer_stop(bb_ctx *ctx, float *ms) {
    if (!ctx || !ms) return BB_ERR_NULL;
    BB_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    BB_HIP(ctx, hipEventSynchronize(ctx->ev1));
    BB_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return BB_OK;
}

// ------------------------------------------------------------------ host tables

// src/ordering.jl:8-26 loops restated (the tables continue past N+1 exactly as the
// reference's loops do, so all 21 entries are bit-exact); src/constants.jl:10-12 for l_j.
int bb_tables(int N, int64_t tri_o[21], int64_t tet_o[21], double l_j[20]) {
    if (N < 1 || N > 20) return BB_ERR_ARG;
    if (tri_o) {
        int64_t c = 0;
        tri_o[0] = 0;
        for (int i = 1; i <= 20; ++i) tri_o[i] = (c += N + 2 - i);
    }
    if (tet_o) {
        int64_t c = 0;
        tet_o[0] = 0;
        for (int i = 1; i <= 20; ++i) tet_o[i] = (c += (int64_t)(N + 2 - i) * (N + 3 - i) / 2);
    }
    if (l_j)
        for (int j = 1; j <= 20; ++j) l_j[j - 1] = j <= N ? bb::lj(N, j) : 0.0;
    return BB_OK;
}

int bb_linear_to_ijkl(int N, int32_t *out) {
    if (N < 0 || N > 20) return BB_ERR_ARG;
    if (!out) return BB_ERR_NULL;
    int r = 0;
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N - k; ++j)
            for (int i = 0; i <= N - j - k; ++i, ++r) {
                out[4 * r] = i;
                out[4 * r + 1] = j;
                out[4 * r + 2] = k;
                out[4 * r + 3] = N - i - j - k;
            }
    return BB_OK;
}

int bb_linear_to_ijk_2d(int N, int32_t *out) {
    if (N < 0 || N > 20) return BB_ERR_ARG;
    if (!out) return BB_ERR_NULL;
    int r = 0;
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N - j; ++i, ++r) {
            out[3 * r] = i;
            out[3 * r + 1] = j;
            out[3 * r + 2] = N - i - j;
        }
    return BB_OK;
}

int bb_fmask(int N, int32_t *out) {
    if (N < 1 || N > 20) return BB_ERR_ARG;
    if (!out) return BB_ERR_NULL;
    int nf = bb::nfp2(N);
    for (int f = 0; f < 4; ++f) {
        int m = 0;
        for (int b = 0; b <= N; ++b)
            for (int a = 0; a <= N - b; ++a) out[f * nf + m++] = bb::face_vidx(N, f, a, b);
    }
    return BB_OK;
}

// src/BernsteinDerivativeMatrix_3D.jl:61-94: entry (m,n) of D^r, D^s, D^t.
// D^0 row alpha has alpha_t at column alpha - e_t + e_0; D = (D^axis - D^3) / 2.
static int coeff_first(int i1, int j1, int k1, int l1, int i2, int j2, int k2, int l2) {
    if (i1 == i2 && j1 == j2 && k1 == k2 && l1 == l2) return i1;
    if (i2 == i1 + 1 && j2 == j1 - 1 && k2 == k1 && l2 == l1) return j1;
    if (i2 == i1 + 1 && j2 == j1 && k2 == k1 - 1 && l2 == l1) return k1;
    if (i2 == i1 + 1 && j2 == j1 && k2 == k1 && l2 == l1 - 1) return l1;
    return 0;
}

int bb_deriv_entry(int axis, int N, int m, int n, double *value) {
    if (!value) return BB_ERR_NULL;
    if (N < 1 || N > 20 || axis < 0 || axis > 2) return BB_ERR_ARG;
    int np = bb::np3(N);
    if (m < 0 || n < 0 || m >= np || n >= np) return BB_ERR_ARG;
    int a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    int r = 0;
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N - k; ++j)
            for (int i = 0; i <= N - j - k; ++i, ++r) {
                if (r == m) { a[0] = i; a[1] = j; a[2] = k; a[3] = N - i - j - k; }
                if (r == n) { b[0] = i; b[1] = j; b[2] = k; b[3] = N - i - j - k; }
            }
    // swap the differentiated barycentric into position 0 (axis 0: i, 1: j, 2: k)
    int p = axis;
    int a1[4] = {a[p], a[1], a[2], a[3]}, b1[4] = {b[p], b[1], b[2], b[3]};
    a1[p] = a[0];
    b1[p] = b[0];
    if (p == 0) { a1[0] = a[0]; b1[0] = b[0]; }
    int first = coeff_first(a1[0], a1[1], a1[2], a1[3], b1[0], b1[1], b1[2], b1[3]);
    int second = coeff_first(a[3], a[1], a[2], a[0], b[3], b[1], b[2], b[0]);
    *value = 0.5 * (first - second);
    return BB_OK;
}

// src/ElevationMatrix.jl:17-24: rows degree N, columns degree N-1
int bb_elevation_entry(int N, int m, int n, double *value) {
    if (!value) return BB_ERR_NULL;
    if (N < 1 || N > 20) return BB_ERR_ARG;
    if (m < 0 || n < 0 || m >= bb::nfp2(N) || n >= bb::nfp2(N - 1)) return BB_ERR_ARG;
    int r = 0, i1 = 0, j1 = 0, k1 = 0, i2 = 0, j2 = 0, k2 = 0;
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N - j; ++i, ++r)
            if (r == m) { i1 = i; j1 = j; k1 = N - i - j; }
    r = 0;
    for (int j = 0; j <= N - 1; ++j)
        for (int i = 0; i <= N - 1 - j; ++i, ++r)
            if (r == n) { i2 = i; j2 = j; k2 = N - 1 - i - j; }
    double v = 0.0;
    if (i1 == i2 + 1 && j1 == j2 && k1 == k2) v = (double)(i2 + 1) / N;
    else if (i1 == i2 && j1 == j2 + 1 && k1 == k2) v = (double)(j2 + 1) / N;
    else if (i1 == i2 && j1 == j2 && k1 == k2 + 1) v = (double)(k2 + 1) / N;
    *value = v;
    return BB_OK;
}

}  // extern "C"
