// ew.hip -- element-wise operations on stored values and eliminate_zeros,
// for matrices resident in HBM.
//
// Reference: seq/src/sparse32.js:671-1207 (out-of-place expm1 ... nearest,
// eliminate_zeros) and :200-660 (the in-place self_* methods), whose bodies
// are the sparse_opt_32.wat / sparse_opt_64.wat exports
//   <op>_coo / self_<op>_coo   (COO and CSR values; :597-1700)
//   <op>_dia / self_<op>_dia   (in-range slots of each diagonal; :2137-4976)
//   <op>_ell / self_<op>_ell   (all ncols*N slots; :4977-6347)
//   count_nonzeros_* / copy_nonzeros_*  (:500-596, :1948-2136, :5821-5950)
//
// Per value: expm1, log1p, sin, tan and pow are JavaScript Math functions
// that the WebAssembly module imports (seq/src/sparse32.js:8): the f32 value
// is widened, the function runs in double and the result is rounded back to
// f32.  The kernels do the same with the device's double-precision math
// library.  Everything else is the WebAssembly instruction: f32.sqrt,
// f32.ceil/floor/trunc/nearest, abs/neg (sign bit), and
//   deg2rad: v * (pi / 180), rad2deg: v / (pi / 180), with pi / 180 rounded
//            in the value type from the caller's pi (Math.PI);
//   sign:  v > 0 ? 1 : -1 in the COO/CSR exports, and
//   sign0: v == 0 ? 0 : (v > 0 ? 1 : -1) in the DIA/ELL exports.
// A value counts as nonzero when v > 0 || v < 0 (zeros and NaN do not).
#include "common.h"
#include "prims.h"

#include <math.h>
#include <string.h>

#include <algorithm>

using namespace sable;

namespace {

constexpr int EB = 256;

template <typename T>
__device__ __forceinline__ bool nonzero(T v) {
    return v > (T)0 || v < (T)0;
}

// Math.pow(x, y) where it differs from C's pow: a NaN exponent gives NaN
// and |x| == 1 with an infinite exponent gives NaN (ECMAScript Number::
// exponentiate); C returns 1 in both cases.
__device__ __forceinline__ double js_pow(double x, double y) {
    if (y != y) return __longlong_as_double(0x7ff8000000000000ll);
    if (fabs(x) == 1.0 && isinf(y)) return __longlong_as_double(0x7ff8000000000000ll);
    return pow(x, y);
}

template <typename T>
struct Op {
    int op;
    T param;  // pow exponent, pi / 180, or the scalar
    __device__ __forceinline__ T operator()(T v) const {
        switch (op) {
            case SABLE_EW_EXPM1: return (T)expm1((double)v);
            case SABLE_EW_LOG1P: return (T)log1p((double)v);
            case SABLE_EW_SIN: return (T)sin((double)v);
            case SABLE_EW_TAN: return (T)tan((double)v);
            case SABLE_EW_POW: return (T)js_pow((double)v, (double)param);
            case SABLE_EW_DEG2RAD: return mul_rn(v, param);
            case SABLE_EW_RAD2DEG: return div_rn(v, param);
            case SABLE_EW_SIGN: return v > (T)0 ? (T)1 : (T)-1;
            case SABLE_EW_SIGN0: return v == (T)0 ? (T)0 : (v > (T)0 ? (T)1 : (T)-1);
            case SABLE_EW_ABS: return fabs(v);
            case SABLE_EW_NEG: return -v;
            case SABLE_EW_SQRT: return sqrt_rn(v);
            case SABLE_EW_CEIL: return ceil(v);
            case SABLE_EW_FLOOR: return floor(v);
            case SABLE_EW_TRUNC: return trunc(v);
            case SABLE_EW_NEAREST: return rint(v);
            case SABLE_EW_MULTIPLY: return mul_rn(v, param);
            default: return v;
        }
    }
    // the sqrt builtins are IEEE-rounded (correctly rounded fp32 divide and
    // sqrt is clang's HIP default); __fsqrt_rn without
    // OCML_BASIC_ROUNDED_OPERATIONS is the approximate native instruction
    __device__ __forceinline__ static float sqrt_rn(float v) { return __builtin_sqrtf(v); }
    __device__ __forceinline__ static double sqrt_rn(double v) { return __builtin_sqrt(v); }
};

template <typename T>
__global__ void k_ew(Op<T> f, int64_t n, const T* __restrict__ in, T* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = f(in[i]);
}

// in-place: one pointer, so no restrict aliasing promise is made
template <typename T>
__global__ void k_ew_self(Op<T> f, int64_t n, T* v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] = f(v[i]);
}

// Bulk of a 16-byte aligned array: 16-byte vectors, EW_U of them loaded per
// thread before any is computed, with non-temporal loads and stores (the
// values are touched once), on a grid of up to 16384 workgroups; the
// steadiest of the shapes measured on MI355X, near the device-to-device
// copy rate.  in and out may be the same array (each vector is read, then
// written, by one thread).
constexpr int EW_U = 4;
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int n = 2;
};

template <typename T>
__global__ __launch_bounds__(256) void k_ew_vec(Op<T> f, int64_t nvec, const typename Vec16<T>::type* in,
                                                typename Vec16<T>::type* out) {
    typedef typename Vec16<T>::type V;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride * EW_U) {
        V v[EW_U];
#pragma unroll
        for (int u = 0; u < EW_U; ++u) {
            const int64_t k = i + u * stride;
            if (k < nvec) v[u] = __builtin_nontemporal_load(in + k);
        }
#pragma unroll
        for (int u = 0; u < EW_U; ++u) {
            const int64_t k = i + u * stride;
            if (k < nvec) {
#pragma unroll
                for (int c = 0; c < Vec16<T>::n; ++c) v[u][c] = f(v[u][c]);
                __builtin_nontemporal_store(v[u], out + k);
            }
        }
    }
}

// DIA slot range of diagonal d: rows [lo, hi), data index base + row.  The
// range is clamped to the diagonal's own stride slots, so an offset that
// does not fit the stride (|k| < N - stride) cannot reach another diagonal.
__device__ __forceinline__ void dia_range(int32_t k, int32_t N, int32_t stride, int32_t d, int64_t* base, int32_t* lo,
                                          int32_t* hi) {
    const int64_t b = (int64_t)d * stride - (k < 0 ? (int64_t)(N - stride) : 0);
    int64_t l = k < 0 ? -(int64_t)k : 0, h = k > 0 ? (int64_t)N - k : N;
    const int64_t s0 = (int64_t)d * stride - b, s1 = (int64_t)(d + 1) * stride - b;
    if (l < s0) l = s0;
    if (h > s1) h = s1;
    if (h < l) h = l;
    *base = b;
    *lo = (int32_t)l;
    *hi = (int32_t)h;
}

// Rows [r0, r1) of every diagonal (the parallel exports' worker range;
// 0..N for the serial ones), intersected with the diagonal's own rows.
template <typename T>
__global__ void k_ew_dia(Op<T> f, int32_t N, int32_t nd, int32_t stride, const int32_t* __restrict__ offset,
                         int32_t r0, int32_t r1, const T* in, T* out) {
    const int32_t d = blockIdx.y;
    if (d >= nd) return;
    int64_t base;
    int32_t lo, hi;
    dia_range(offset[d], N, stride, d, &base, &lo, &hi);
    if (lo < r0) lo = r0;
    if (hi > r1) hi = r1;
    for (int64_t r = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < hi; r += (int64_t)gridDim.x * blockDim.x)
        out[base + r] = f(in[base + r]);
}

// ELL slots k*N + i for rows i in [r0, r1), every slot column k.
template <typename T>
__global__ void k_ew_ell(Op<T> f, int32_t N, int32_t nc, int32_t r0, int32_t r1, const T* in, T* out) {
    const int32_t k = blockIdx.y;
    if (k >= nc) return;
    const int64_t base = (int64_t)k * N;
    for (int64_t i = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r1; i += (int64_t)gridDim.x * blockDim.x)
        out[base + i] = f(in[base + i]);
}

template <typename T>
__global__ void k_nz_flags(int64_t n, const T* __restrict__ v, int32_t* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = nonzero(v[i]) ? 1 : 0;
}

template <typename T>
__global__ void k_compact_coo(int64_t n, const int32_t* __restrict__ pos, const int32_t* __restrict__ row,
                              const int32_t* __restrict__ col, const T* __restrict__ val, int32_t* __restrict__ orow,
                              int32_t* __restrict__ ocol, T* __restrict__ oval) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T v = val[i];
    if (!nonzero(v)) return;
    const int32_t p = pos[i];
    if (orow) orow[p] = row[i];
    ocol[p] = col[i];
    oval[p] = v;
}

// out_row_ptr[i] = kept entries before row i's first entry
__global__ void k_csr_heads(int32_t N, const int32_t* __restrict__ rp, const int32_t* __restrict__ pos,
                            int32_t* __restrict__ orp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    orp[i] = pos[rp[i] - rp[0]];
}

// One workgroup per diagonal: nonzero count over the in-range slots.
template <typename T>
__global__ __launch_bounds__(EB) void k_dia_count(int32_t N, int32_t stride, const int32_t* __restrict__ offset,
                                                  const T* __restrict__ data, int32_t* __restrict__ counts) {
    const int32_t d = blockIdx.x;
    int64_t base;
    int32_t lo, hi;
    dia_range(offset[d], N, stride, d, &base, &lo, &hi);
    int c = 0;
    for (int32_t r = lo + threadIdx.x; r < hi; r += EB) c += nonzero(data[base + r]) ? 1 : 0;
    c = wave_sum(c);
    __shared__ int ws[EB / WAVE];
    if ((threadIdx.x & (WAVE - 1)) == 0) ws[threadIdx.x / WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < EB / WAVE; ++w) t += ws[w];
        counts[d] = t;
    }
}

// One workgroup per kept diagonal: its nonzeros, in row order, packed from
// slot dst_base + lo (copy_nonzeros_dia).
template <typename T>
__global__ __launch_bounds__(EB) void k_dia_pack(int32_t N, int32_t stride, const int32_t* __restrict__ offset,
                                                 const int32_t* __restrict__ kept, const T* __restrict__ data,
                                                 T* __restrict__ out) {
    const int32_t j = blockIdx.x;  // kept diagonal number
    const int32_t d = kept[j];
    int64_t base;
    int32_t lo, hi;
    dia_range(offset[d], N, stride, d, &base, &lo, &hi);
    int64_t obase, olo;
    {
        int32_t l2, h2;
        dia_range(offset[d], N, stride, j, &obase, &l2, &h2);
        olo = l2;
    }
    __shared__ int ws[EB / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    int carry = 0;
    for (int32_t r0 = lo; r0 < hi; r0 += EB) {
        const int32_t r = r0 + threadIdx.x;
        const T v = r < hi ? data[base + r] : (T)0;
        const int f = (r < hi && nonzero(v)) ? 1 : 0;
        int inc = f;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const int o = __shfl_up(inc, off, WAVE);
            if (lane >= off) inc += o;
        }
        if (lane == WAVE - 1) ws[w] = inc;
        __syncthreads();
        int before = carry, tot = 0;
        for (int k = 0; k < EB / WAVE; ++k) {
            if (k < w) before += ws[k];
            tot += ws[k];
        }
        if (f) out[obase + olo + before + inc - 1] = v;
        carry += tot;
        __syncthreads();
    }
}

// ELL: per-row nonzero counts; count_cols = max, nnz = total.
template <typename T>
__global__ void k_ell_count(int32_t N, int32_t nc, const T* __restrict__ data, unsigned long long* __restrict__ acc) {
    int m = 0;
    long long s = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        int c = 0;
        for (int32_t k = 0; k < nc; ++k) c += nonzero(data[(int64_t)k * N + i]) ? 1 : 0;
        m = max(m, c);
        s += c;
    }
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, WAVE));
    s = wave_sum(s);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        if (m) atomicMax(&acc[0], (unsigned long long)m);
        if (s) atomicAdd(&acc[1], (unsigned long long)s);
    }
}

// ELL compaction: row i's nonzeros in slot order into slots 0.. of the
// nc_out-wide output (copy_nonzeros_ell); other slots zero.
template <typename T>
__global__ void k_ell_pack(int32_t N, int32_t nc, int32_t nc_out, const T* __restrict__ data,
                           const int32_t* __restrict__ idx, T* __restrict__ odata, int32_t* __restrict__ oidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int32_t o = 0;
    for (int32_t k = 0; k < nc; ++k) {
        const T v = data[(int64_t)k * N + i];
        if (nonzero(v) && o < nc_out) {
            odata[(int64_t)o * N + i] = v;
            oidx[(int64_t)o * N + i] = idx[(int64_t)k * N + i];
            ++o;
        }
    }
    for (; o < nc_out; ++o) {
        odata[(int64_t)o * N + i] = (T)0;
        oidx[(int64_t)o * N + i] = 0;
    }
}

template <typename T>
__global__ void k_fill(int64_t n, T* __restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (T)0;
}

unsigned gs(int64_t n) { return grid_for(n, EB); }
unsigned gs_cap(int64_t n) {
    const int64_t g = (n + EB - 1) / EB;
    return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

bool op_valid(int op) { return op >= SABLE_EW_EXPM1 && op <= SABLE_EW_MULTIPLY; }

template <typename T>
Op<T> make_op(int op, double param) {
    Op<T> f{op, (T)param};
    if (op == SABLE_EW_DEG2RAD || op == SABLE_EW_RAD2DEG) f.param = (T)param / (T)180;  // pi / 180 in T
    return f;
}

template <typename T>
int ew_values_t(sable_ctx_t* ctx, int op, double param, int64_t n, const void* in, void* out) {
    const Op<T> f = make_op<T>(op, param);
    typedef typename Vec16<T>::type V;
    const int64_t per = Vec16<T>::n;
    if ((((uintptr_t)in | (uintptr_t)out) & 15) == 0 && n >= per) {
        const int64_t nvec = n / per, rest = n - nvec * per;
        const int64_t blocks = std::min<int64_t>((nvec + 256 * EW_U - 1) / (256 * EW_U), 16384);
        hipLaunchKernelGGL(k_ew_vec<T>, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, ctx->stream, f,
                           nvec, (const V*)in, (V*)out);
        SABLE_LAUNCH_CHECK();
        if (rest) {
            const T* ti = (const T*)in + nvec * per;
            T* to = (T*)out + nvec * per;
            if (in == out)
                hipLaunchKernelGGL(k_ew_self<T>, dim3(1), dim3(EB), 0, ctx->stream, f, rest, to);
            else
                hipLaunchKernelGGL(k_ew<T>, dim3(1), dim3(EB), 0, ctx->stream, f, rest, ti, to);
            SABLE_LAUNCH_CHECK();
        }
        return SABLE_OK;
    }
    if (in == out)
        hipLaunchKernelGGL(k_ew_self<T>, dim3(gs_cap(n)), dim3(EB), 0, ctx->stream, f, n, (T*)out);
    else
        hipLaunchKernelGGL(k_ew<T>, dim3(gs_cap(n)), dim3(EB), 0, ctx->stream, f, n, (const T*)in, (T*)out);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

template <typename T>
int ew_dia_t(sable_ctx_t* ctx, int op, double param, int32_t N, int32_t nd, int32_t stride, const int32_t* offset,
             int32_t r0, int32_t r1, const void* in, void* out) {
    const Op<T> f = make_op<T>(op, param);
    const unsigned gx = (unsigned)std::min<int64_t>(((int64_t)(r1 - r0) + EB - 1) / EB, 1024);
    hipLaunchKernelGGL(k_ew_dia<T>, dim3(gx ? gx : 1, (unsigned)nd), dim3(EB), 0, ctx->stream, f, N, nd, stride, offset,
                       r0, r1, (const T*)in, (T*)out);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

template <typename T>
int ew_ell_t(sable_ctx_t* ctx, int op, double param, int32_t N, int32_t nc, int32_t r0, int32_t r1, const void* in,
             void* out) {
    const Op<T> f = make_op<T>(op, param);
    const unsigned gx = (unsigned)std::min<int64_t>(((int64_t)(r1 - r0) + EB - 1) / EB, 1024);
    hipLaunchKernelGGL(k_ew_ell<T>, dim3(gx ? gx : 1, (unsigned)nc), dim3(EB), 0, ctx->stream, f, N, nc, r0, r1,
                       (const T*)in, (T*)out);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

template <typename T>
int nz_positions(sable_ctx_t* ctx, int64_t n, const T* val, int32_t** pos, int64_t* count) {
    void* s;
    const size_t need = Carver::round(4 * (size_t)(n + 1)) + 256 + scan_tmp_bytes(n);
    int rc = ctx_scratch(ctx, need, &s);
    if (rc) return rc;
    Carver cv(s);
    int32_t* p = cv.take<int32_t>(n + 1);
    void* tmp = cv.p + cv.used;
    hipLaunchKernelGGL(k_nz_flags<T>, dim3(gs(n)), dim3(EB), 0, ctx->stream, n, val, p);
    SABLE_LAUNCH_CHECK();
    rc = scan_exclusive_i32(ctx->stream, p, p, n, p + n, tmp);
    if (rc) return rc;
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, p + n, 4, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    int32_t c = 0;
    memcpy(&c, ctx->pinned, 4);
    *count = c;
    *pos = p;
    return SABLE_OK;
}

}  // namespace

#define EW_DT(dtype, F, ...)                                                   \
    ((dtype) == SABLE_F32 ? F<float>(__VA_ARGS__)                              \
     : (dtype) == SABLE_F64 ? F<double>(__VA_ARGS__)                           \
                            : fail(SABLE_EINVAL, "bad dtype"))

SABLE_API int sable_ew_values(sable_ctx_t* ctx, sable_dtype_t dtype, int op, double param, int64_t n, const void* in,
                              void* out) {
    SABLE_CHECK(ctx && in && out, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(op_valid(op), SABLE_EINVAL, "unknown element-wise op %d", op);
    SABLE_CHECK(n >= 0, SABLE_EINVAL, "negative size");
    if (n == 0) return SABLE_OK;
    return EW_DT(dtype, ew_values_t, ctx, op, param, n, in, out);
}

SABLE_API int sable_ew_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int op, double param, int32_t N, int32_t nd,
                           int32_t stride, const int32_t* offset, int32_t r0, int32_t r1, const void* in, void* out) {
    SABLE_CHECK(ctx && offset && in && out, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(op_valid(op), SABLE_EINVAL, "unknown element-wise op %d", op);
    SABLE_CHECK(N >= 0 && nd >= 0 && stride >= 0 && stride <= N, SABLE_EINVAL, "bad DIA sizes");
    if (r0 < 0) r0 = 0;
    if (r1 > N) r1 = N;
    if (N == 0 || nd == 0 || r0 >= r1) return SABLE_OK;
    SABLE_CHECK(nd <= 65535, SABLE_EINVAL, "more than 65535 diagonals");
    return EW_DT(dtype, ew_dia_t, ctx, op, param, N, nd, stride, offset, r0, r1, in, out);
}

SABLE_API int sable_ew_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int op, double param, int32_t N, int32_t nc,
                           int32_t r0, int32_t r1, const void* in, void* out) {
    SABLE_CHECK(ctx && in && out, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(op_valid(op), SABLE_EINVAL, "unknown element-wise op %d", op);
    SABLE_CHECK(N >= 0 && nc >= 0, SABLE_EINVAL, "bad ELL sizes");
    if (r0 < 0) r0 = 0;
    if (r1 > N) r1 = N;
    if (N == 0 || nc == 0 || r0 >= r1) return SABLE_OK;
    SABLE_CHECK(nc <= 65535, SABLE_EINVAL, "more than 65535 slot columns");
    return EW_DT(dtype, ew_ell_t, ctx, op, param, N, nc, r0, r1, in, out);
}

template <typename T>
static int count_nz_t(sable_ctx_t* ctx, int64_t n, const void* val, int64_t* count) {
    int32_t* pos;
    return nz_positions<T>(ctx, n, (const T*)val, &pos, count);
}

SABLE_API int sable_count_nonzeros(sable_ctx_t* ctx, sable_dtype_t dtype, int64_t n, const void* val, int64_t* count) {
    SABLE_CHECK(ctx && count, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(n >= 0 && n <= 0x7fffffffLL, SABLE_EINVAL, "bad size");
    *count = 0;
    if (n == 0) return SABLE_OK;
    return EW_DT(dtype, count_nz_t, ctx, n, val, count);
}

template <typename T>
static int elim_coo_t(sable_ctx_t* ctx, int64_t nnz, const int32_t* row, const int32_t* col, const void* val,
                      int32_t* orow, int32_t* ocol, void* oval, int64_t* count) {
    int32_t* pos;
    int rc = nz_positions<T>(ctx, nnz, (const T*)val, &pos, count);
    if (rc) return rc;
    hipLaunchKernelGGL(k_compact_coo<T>, dim3(gs(nnz)), dim3(EB), 0, ctx->stream, nnz, pos, row, col, (const T*)val,
                       orow, ocol, (T*)oval);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_eliminate_zeros_coo(sable_ctx_t* ctx, sable_dtype_t dtype, int64_t nnz, const int32_t* row,
                                        const int32_t* col, const void* val, int32_t* out_row, int32_t* out_col,
                                        void* out_val, int64_t* count) {
    SABLE_CHECK(ctx && count, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(nnz >= 0 && nnz <= 0x7fffffffLL, SABLE_EINVAL, "bad size");
    *count = 0;
    if (nnz == 0) return SABLE_OK;
    SABLE_CHECK(row && col && val && out_row && out_col && out_val, SABLE_EINVAL, "NULL array");
    return EW_DT(dtype, elim_coo_t, ctx, nnz, row, col, val, out_row, out_col, out_val, count);
}

template <typename T>
static int elim_csr_t(sable_ctx_t* ctx, int32_t N, const int32_t* rp, const int32_t* col, const void* val,
                      int32_t* orp, int32_t* ocol, void* oval, int64_t* count) {
    int64_t ends[2];
    int32_t e[2];
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, rp, 4, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipMemcpyAsync((char*)ctx->pinned + 4, rp + N, 4, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(e, ctx->pinned, 8);
    ends[0] = e[0];
    ends[1] = e[1];
    const int64_t nnz = ends[1] - ends[0];
    SABLE_CHECK(nnz >= 0, SABLE_EMALFORMED, "row_ptr[N] < row_ptr[0]");
    int32_t* pos;
    if (nnz == 0) {
        hipLaunchKernelGGL(k_fill<int32_t>, dim3(gs(N + 1)), dim3(EB), 0, ctx->stream, (int64_t)N + 1, orp);
        SABLE_LAUNCH_CHECK();
        *count = 0;
        return SABLE_OK;
    }
    int rc = nz_positions<T>(ctx, nnz, (const T*)val + ends[0], &pos, count);
    if (rc) return rc;
    hipLaunchKernelGGL(k_compact_coo<T>, dim3(gs(nnz)), dim3(EB), 0, ctx->stream, nnz, pos, (const int32_t*)nullptr,
                       col + ends[0], (const T*)val + ends[0], (int32_t*)nullptr, ocol, (T*)oval);
    SABLE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_csr_heads, dim3(gs((int64_t)N + 1)), dim3(EB), 0, ctx->stream, N, rp, pos, orp);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_eliminate_zeros_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                        const int32_t* col, const void* val, int32_t* out_row_ptr, int32_t* out_col,
                                        void* out_val, int64_t* count) {
    SABLE_CHECK(ctx && row_ptr && out_row_ptr && count, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0, SABLE_EINVAL, "negative size");
    *count = 0;
    return EW_DT(dtype, elim_csr_t, ctx, N, row_ptr, col, val, out_row_ptr, out_col, out_val, count);
}

template <typename T>
static int dia_counts(sable_ctx_t* ctx, int32_t N, int32_t nd, int32_t stride, const int32_t* offset, const T* data,
                      std::vector<int32_t>* counts) {
    void* s;
    int rc = ctx_scratch(ctx, 4 * (size_t)nd + 256, &s);
    if (rc) return rc;
    int32_t* c = (int32_t*)s;
    hipLaunchKernelGGL(k_dia_count<T>, dim3((unsigned)nd), dim3(EB), 0, ctx->stream, N, stride, offset, data, c);
    SABLE_LAUNCH_CHECK();
    counts->resize(nd);
    SABLE_HIP(hipMemcpyAsync(counts->data(), c, 4 * (size_t)nd, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    return SABLE_OK;
}

template <typename T>
static int count_dia_t(sable_ctx_t* ctx, int32_t N, int32_t nd, int32_t stride, const int32_t* offset, const void* data,
                       int32_t* count_diags, int64_t* count) {
    std::vector<int32_t> c;
    int rc = dia_counts<T>(ctx, N, nd, stride, offset, (const T*)data, &c);
    if (rc) return rc;
    int32_t kd = 0;
    int64_t t = 0;
    for (int32_t v : c) {
        if (v) ++kd;
        t += v;
    }
    *count_diags = kd;
    *count = t;
    return SABLE_OK;
}

static int dia_sizes_ok(int32_t N, int32_t nd, int32_t stride) {
    SABLE_CHECK(N >= 0 && nd >= 0 && stride >= 0 && stride <= N, SABLE_EINVAL, "bad DIA sizes");
    SABLE_CHECK(nd <= 0x7fffffff / 1, SABLE_EINVAL, "bad ndiags");
    return SABLE_OK;
}

SABLE_API int sable_count_nonzeros_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                                       const int32_t* offset, const void* data, int32_t* count_diags, int64_t* count) {
    SABLE_CHECK(ctx && count_diags && count, SABLE_EINVAL, "NULL argument");
    int rc = dia_sizes_ok(N, nd, stride);
    if (rc) return rc;
    *count_diags = 0;
    *count = 0;
    if (N == 0 || nd == 0) return SABLE_OK;
    return EW_DT(dtype, count_dia_t, ctx, N, nd, stride, offset, data, count_diags, count);
}

template <typename T>
static int elim_dia_t(sable_ctx_t* ctx, int32_t N, int32_t nd, int32_t stride, const int32_t* offset, const void* data,
                      int32_t* out_offset, void* out_data, int32_t* count_diags) {
    std::vector<int32_t> c;
    int rc = dia_counts<T>(ctx, N, nd, stride, offset, (const T*)data, &c);
    if (rc) return rc;
    std::vector<int32_t> kept, koff;
    std::vector<int32_t> hoff(nd);
    SABLE_HIP(hipMemcpy(hoff.data(), offset, 4 * (size_t)nd, hipMemcpyDeviceToHost));
    for (int32_t d = 0; d < nd; ++d)
        if (c[d]) {
            kept.push_back(d);
            koff.push_back(hoff[d]);
        }
    const int32_t kd = (int32_t)kept.size();
    *count_diags = kd;
    if (kd == 0) return SABLE_OK;
    void* s;
    rc = ctx_scratch(ctx, 4 * (size_t)kd + 256, &s);
    if (rc) return rc;
    int32_t* dk = (int32_t*)s;
    SABLE_HIP(hipMemcpyAsync(dk, kept.data(), 4 * (size_t)kd, hipMemcpyHostToDevice, ctx->stream));
    SABLE_HIP(hipMemcpyAsync(out_offset, koff.data(), 4 * (size_t)kd, hipMemcpyHostToDevice, ctx->stream));
    const int64_t slots = (int64_t)kd * stride;
    hipLaunchKernelGGL(k_fill<T>, dim3(gs(slots)), dim3(EB), 0, ctx->stream, slots, (T*)out_data);
    SABLE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dia_pack<T>, dim3((unsigned)kd), dim3(EB), 0, ctx->stream, N, stride, offset, dk,
                       (const T*)data, (T*)out_data);
    SABLE_LAUNCH_CHECK();
    SABLE_HIP(hipStreamSynchronize(ctx->stream));  // the host vectors above go out of scope
    return SABLE_OK;
}

SABLE_API int sable_eliminate_zeros_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                                        const int32_t* offset, const void* data, int32_t* out_offset, void* out_data,
                                        int32_t* count_diags) {
    SABLE_CHECK(ctx && offset && data && out_offset && out_data && count_diags, SABLE_EINVAL, "NULL argument");
    int rc = dia_sizes_ok(N, nd, stride);
    if (rc) return rc;
    *count_diags = 0;
    if (N == 0 || nd == 0) return SABLE_OK;
    return EW_DT(dtype, elim_dia_t, ctx, N, nd, stride, offset, data, out_offset, out_data, count_diags);
}

template <typename T>
static int count_ell_t(sable_ctx_t* ctx, int32_t N, int32_t nc, const void* data, int32_t* count_cols,
                       int64_t* count) {
    void* s;
    int rc = ctx_scratch(ctx, 256, &s);
    if (rc) return rc;
    unsigned long long* acc = (unsigned long long*)s;
    SABLE_HIP(hipMemsetAsync(acc, 0, 16, ctx->stream));
    // grid-stride on a capped grid: one atomic per wave on each counter
    hipLaunchKernelGGL(k_ell_count<T>, dim3(std::min(gs(N), 1024u)), dim3(EB), 0, ctx->stream, N, nc, (const T*)data,
                       acc);
    SABLE_LAUNCH_CHECK();
    int64_t h[2];
    rc = read_device_i64(ctx, acc, 2, h);
    if (rc) return rc;
    *count_cols = (int32_t)h[0];
    *count = h[1];
    return SABLE_OK;
}

SABLE_API int sable_count_nonzeros_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc, const void* data,
                                       int32_t* count_cols, int64_t* count) {
    SABLE_CHECK(ctx && count_cols && count, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0 && nc >= 0, SABLE_EINVAL, "negative size");
    *count_cols = 0;
    *count = 0;
    if (N == 0 || nc == 0) return SABLE_OK;
    return EW_DT(dtype, count_ell_t, ctx, N, nc, data, count_cols, count);
}

template <typename T>
static int elim_ell_t(sable_ctx_t* ctx, int32_t N, int32_t nc, int32_t nc_out, const void* data,
                      const int32_t* indices, void* out_data, int32_t* out_indices) {
    hipLaunchKernelGGL(k_ell_pack<T>, dim3(gs(N)), dim3(EB), 0, ctx->stream, N, nc, nc_out, (const T*)data, indices,
                       (T*)out_data, out_indices);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_eliminate_zeros_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc, int32_t nc_out,
                                        const void* data, const int32_t* indices, void* out_data,
                                        int32_t* out_indices) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0 && nc >= 0 && nc_out >= 0, SABLE_EINVAL, "negative size");
    if (N == 0 || nc_out == 0) return SABLE_OK;
    SABLE_CHECK(data && indices && out_data && out_indices, SABLE_EINVAL, "NULL array");
    return EW_DT(dtype, elim_ell_t, ctx, N, nc, nc_out, data, indices, out_data, out_indices);
}
