// ctx.hip -- context, device memory, fills, events and error reporting.
#include "common.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace sable {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int ctx_scratch(sable_ctx_t* ctx, size_t bytes, void** out) {
    if (bytes == 0) bytes = 256;
    if (ctx->scratch_bytes < bytes) {
        if (ctx->scratch) {
            SABLE_HIP(hipStreamSynchronize(ctx->stream));
            SABLE_HIP(hipFree(ctx->scratch));
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
        }
        size_t want = bytes + bytes / 4;
        if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
            (void)hipGetLastError();
            ctx->scratch = nullptr;
            return fail(SABLE_ENOMEM, "scratch allocation of %zu bytes failed", want);
        }
        ctx->scratch_bytes = want;
    }
    *out = ctx->scratch;
    return SABLE_OK;
}

int ctx_pool_alloc(sable_ctx_t* ctx, size_t bytes, void** out) {
    if (bytes == 0) bytes = 256;
    *out = nullptr;
    if (!ctx->pool_tried) {
        ctx->pool_tried = true;
        hipMemPoolProps props;
        memset(&props, 0, sizeof(props));
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = ctx->device;
        if (hipMemPoolCreate(&ctx->pool, &props) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(ctx->pool, hipMemPoolAttrReleaseThreshold, &keep);
        } else {
            (void)hipGetLastError();
            ctx->pool = nullptr;
        }
    }
    hipError_t e = ctx->pool ? hipMallocFromPoolAsync(out, bytes, ctx->pool, ctx->stream) : hipMalloc(out, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(SABLE_ENOMEM, "temporary allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    return SABLE_OK;
}

void ctx_pool_free(sable_ctx_t* ctx, void* p) {
    if (!p) return;
    if (ctx->pool) (void)hipFreeAsync(p, ctx->stream);
    else (void)hipFree(p);
}

constexpr size_t STAGE_BYTES = (size_t)4 << 20;

static int ensure_stage(sable_ctx_t* ctx) {
    for (int k = 0; k < 2; k++) {
        if (!ctx->stage[k] && hipHostMalloc(&ctx->stage[k], STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            ctx->stage[k] = nullptr;
            return fail(SABLE_ENOMEM, "pinned staging buffer allocation failed");
        }
        if (!ctx->stage_ev[k]) SABLE_HIP(hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming));
    }
    return SABLE_OK;
}

int read_device_i64(sable_ctx_t* ctx, const void* dsrc, size_t count, int64_t* host_out) {
    SABLE_CHECK(count <= 64, SABLE_EINVAL, "read_device_i64: count %zu > 64", count);
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, dsrc, count * 8, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(host_out, ctx->pinned, count * 8);
    return SABLE_OK;
}

__global__ void k_fill_bytes(unsigned char* __restrict__ p, unsigned char v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
__global__ void k_fill_value(T* __restrict__ p, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_max_i32_init(int* out) {
    if (threadIdx.x == 0) *out = INT32_MIN;
}

__global__ void k_max_i32_reduce(int64_t n, const int32_t* __restrict__ a, int* __restrict__ out) {
    int m = INT32_MIN;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, a[i]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicMax(out, m);
}

template <typename T>
__global__ void k_vec_add(T* __restrict__ y, const T* __restrict__ w, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) y[i] = add_rn(y[i], w[i]);
}

}  // namespace sable

using namespace sable;

extern "C" {

SABLE_API const char* sable_last_error(void) { return g_err; }
SABLE_API const char* sable_version(void) { return "sablesparse-hip 0.1 gfx950"; }

SABLE_API int sable_ctx_create(int device, void* hip_stream, sable_ctx_t** out) {
    SABLE_CHECK(out != nullptr, SABLE_EINVAL, "sable_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    SABLE_HIP(hipGetDeviceCount(&ndev));
    SABLE_CHECK(device >= 0 && device < ndev, SABLE_EINVAL, "device %d out of range (%d devices)", device, ndev);
    SABLE_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    SABLE_HIP(hipGetDeviceProperties(&prop, device));
    auto* ctx = new sable_ctx_s();
    ctx->device = device;
    ctx->num_cu = prop.multiProcessorCount;
    if (hip_stream) {
        ctx->stream = (hipStream_t)hip_stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ctx;
            return fail(SABLE_EHIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
        }
        ctx->own_stream = true;
    }
    if (hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(SABLE_ENOMEM, "pinned host buffer allocation failed");
    }
    const int rc = spmv_noalloc_setup(device, &ctx->noalloc);
    if (rc) {
        (void)hipHostFree(ctx->pinned);
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return SABLE_OK;
}

SABLE_API void sable_ctx_destroy(sable_ctx_t* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->pool) (void)hipMemPoolDestroy(ctx->pool);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (int k = 0; k < 2; k++) {
        if (ctx->stage[k]) (void)hipHostFree(ctx->stage[k]);
        if (ctx->stage_ev[k]) (void)hipEventDestroy(ctx->stage_ev[k]);
    }
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

SABLE_API int sable_ctx_sync(sable_ctx_t* ctx) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "ctx is NULL");
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    return SABLE_OK;
}

SABLE_API void* sable_ctx_stream(sable_ctx_t* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

SABLE_API int sable_device_alloc(sable_ctx_t* ctx, size_t bytes, void** dptr) {
    SABLE_CHECK(ctx && dptr, SABLE_EINVAL, "sable_device_alloc: NULL argument");
    *dptr = nullptr;
    if (bytes == 0) bytes = 256;
    SABLE_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *dptr = nullptr;
        return fail(SABLE_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return SABLE_OK;
}

SABLE_API int sable_device_free(sable_ctx_t* ctx, void* dptr) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "ctx is NULL");
    if (!dptr) return SABLE_OK;
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    SABLE_HIP(hipFree(dptr));
    return SABLE_OK;
}

SABLE_API int sable_host_alloc(size_t bytes, void** out) {
    SABLE_CHECK(out, SABLE_EINVAL, "NULL argument");
    *out = nullptr;
    if (hipHostMalloc(out, bytes ? bytes : 64, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(SABLE_ENOMEM, "pinned host allocation of %zu bytes failed", bytes);
    }
    return SABLE_OK;
}

SABLE_API int sable_host_free(void* p) {
    if (p) SABLE_HIP(hipHostFree(p));
    return SABLE_OK;
}

SABLE_API int sable_memcpy_h2d(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "ctx is NULL");
    if (!bytes) return SABLE_OK;
    SABLE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    return SABLE_OK;
}

SABLE_API int sable_memcpy_d2h(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "ctx is NULL");
    if (!bytes) return SABLE_OK;
    SABLE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    return SABLE_OK;
}

SABLE_API int sable_memcpy_h2d_staged(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes) {
    SABLE_CHECK(ctx && (bytes == 0 || (dst && src)), SABLE_EINVAL, "NULL argument");
    if (!bytes) return SABLE_OK;
    int rc = ensure_stage(ctx);
    if (rc) return rc;
    int k = ctx->stage_next;
    for (size_t off = 0; off < bytes; off += STAGE_BYTES, k ^= 1) {
        const size_t n = bytes - off < STAGE_BYTES ? bytes - off : STAGE_BYTES;
        if (ctx->stage_used[k]) SABLE_HIP(hipEventSynchronize(ctx->stage_ev[k]));
        memcpy(ctx->stage[k], (const char*)src + off, n);
        SABLE_HIP(hipMemcpyAsync((char*)dst + off, ctx->stage[k], n, hipMemcpyHostToDevice, ctx->stream));
        SABLE_HIP(hipEventRecord(ctx->stage_ev[k], ctx->stream));
        ctx->stage_used[k] = true;
    }
    ctx->stage_next = k;
    return SABLE_OK;
}

SABLE_API int sable_memcpy_d2h_staged(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes) {
    SABLE_CHECK(ctx && (bytes == 0 || (dst && src)), SABLE_EINVAL, "NULL argument");
    if (!bytes) return SABLE_OK;
    int rc = ensure_stage(ctx);
    if (rc) return rc;
    const size_t nch = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
    auto len = [&](size_t c) { return c + 1 < nch ? STAGE_BYTES : bytes - c * STAGE_BYTES; };
    for (size_t c = 0; c <= nch; c++) {
        if (c < nch) {
            // chunk c into half c % 2 (its previous chunk was copied out in the last step)
            const int k = (int)(c & 1);
            if (ctx->stage_used[k]) SABLE_HIP(hipEventSynchronize(ctx->stage_ev[k]));
            SABLE_HIP(hipMemcpyAsync(ctx->stage[k], (const char*)src + c * STAGE_BYTES, len(c), hipMemcpyDeviceToHost,
                                     ctx->stream));
            SABLE_HIP(hipEventRecord(ctx->stage_ev[k], ctx->stream));
            ctx->stage_used[k] = true;
        }
        if (c >= 1) {
            // chunk c - 1 is copied out while chunk c is in flight
            const int k = (int)((c - 1) & 1);
            SABLE_HIP(hipEventSynchronize(ctx->stage_ev[k]));
            memcpy((char*)dst + (c - 1) * STAGE_BYTES, ctx->stage[k], len(c - 1));
            ctx->stage_used[k] = false;
        }
    }
    ctx->stage_next = 0;
    return SABLE_OK;
}

SABLE_API int sable_memcpy_d2d(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "ctx is NULL");
    if (!bytes) return SABLE_OK;
    SABLE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SABLE_OK;
}

SABLE_API int sable_fill_bytes(sable_ctx_t* ctx, void* dst, int byte, size_t bytes) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "ctx is NULL");
    if (!bytes) return SABLE_OK;
    unsigned char b = (unsigned char)byte;
    if (((uintptr_t)dst & 3) == 0 && (bytes & 3) == 0) {
        uint32_t w = 0x01010101u * b;
        size_t n = bytes / 4;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((int64_t)n, 256) > 8192 ? 8192 : grid_for((int64_t)n, 256)),
                           dim3(256), 0, ctx->stream, (uint32_t*)dst, w, n);
    } else {
        hipLaunchKernelGGL(k_fill_bytes, dim3(grid_for((int64_t)bytes, 256) > 8192 ? 8192 : grid_for((int64_t)bytes, 256)),
                           dim3(256), 0, ctx->stream, (unsigned char*)dst, b, bytes);
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_fill_value(sable_ctx_t* ctx, sable_dtype_t dtype, void* dst, int64_t n, double value) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "ctx is NULL");
    SABLE_CHECK(n >= 0, SABLE_EINVAL, "negative length");
    if (!n) return SABLE_OK;
    unsigned g = grid_for(n, 256);
    if (g > 8192) g = 8192;
    if (dtype == SABLE_F32)
        hipLaunchKernelGGL(k_fill_value<float>, dim3(g), dim3(256), 0, ctx->stream, (float*)dst, (float)value, n);
    else if (dtype == SABLE_F64)
        hipLaunchKernelGGL(k_fill_value<double>, dim3(g), dim3(256), 0, ctx->stream, (double*)dst, value, n);
    else
        return fail(SABLE_EINVAL, "bad dtype %d", (int)dtype);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_vec_add(sable_ctx_t* ctx, sable_dtype_t dtype, void* y, const void* w, int64_t n) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "ctx is NULL");
    SABLE_CHECK(n >= 0, SABLE_EINVAL, "negative length");
    if (!n) return SABLE_OK;
    unsigned g = grid_for(n, 256);
    if (g > 8192) g = 8192;
    if (dtype == SABLE_F32)
        hipLaunchKernelGGL(k_vec_add<float>, dim3(g), dim3(256), 0, ctx->stream, (float*)y, (const float*)w, n);
    else if (dtype == SABLE_F64)
        hipLaunchKernelGGL(k_vec_add<double>, dim3(g), dim3(256), 0, ctx->stream, (double*)y, (const double*)w, n);
    else
        return fail(SABLE_EINVAL, "bad dtype %d", (int)dtype);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_max_i32(sable_ctx_t* ctx, const int32_t* a, int64_t n, int32_t* out) {
    SABLE_CHECK(ctx && out && (a || n == 0), SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(n >= 0, SABLE_EINVAL, "negative length");
    *out = -1;
    if (n == 0) return SABLE_OK;
    void* s;
    int rc = ctx_scratch(ctx, 256, &s);
    if (rc) return rc;
    int* dm = (int*)s;
    // at most 1024 workgroups: one atomic per wave on one address
    unsigned g = grid_for(n, 256);
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(k_max_i32_init, dim3(1), dim3(64), 0, ctx->stream, dm);
    hipLaunchKernelGGL(k_max_i32_reduce, dim3(g), dim3(256), 0, ctx->stream, n, a, dm);
    SABLE_LAUNCH_CHECK();
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, dm, 4, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(out, ctx->pinned, 4);
    return SABLE_OK;
}

SABLE_API int sable_event_create(void** ev) {
    SABLE_CHECK(ev, SABLE_EINVAL, "NULL event");
    hipEvent_t e;
    SABLE_HIP(hipEventCreate(&e));
    *ev = (void*)e;
    return SABLE_OK;
}

SABLE_API int sable_event_destroy(void* ev) {
    if (ev) SABLE_HIP(hipEventDestroy((hipEvent_t)ev));
    return SABLE_OK;
}

SABLE_API int sable_event_record(sable_ctx_t* ctx, void* ev) {
    SABLE_CHECK(ctx && ev, SABLE_EINVAL, "NULL argument");
    SABLE_HIP(hipEventRecord((hipEvent_t)ev, ctx->stream));
    return SABLE_OK;
}

SABLE_API int sable_event_elapsed_ms(void* a, void* b, float* ms) {
    SABLE_CHECK(a && b && ms, SABLE_EINVAL, "NULL argument");
    SABLE_HIP(hipEventSynchronize((hipEvent_t)b));
    SABLE_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return SABLE_OK;
}

}  // extern "C"
