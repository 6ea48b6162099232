// refine.h -- one row of the triangular solve with one refinement step
// fused into the same sweep (shared by the single-matrix and batched level
// schedules).
#pragma once
#include "common.h"

namespace sable {

// s = fma(v_q, x_q, s) for the values of lanes q = 0 .. n-1 of an aligned
// group of G lanes, in lane order (the sequential chain's bits).
template <int G>
__device__ __forceinline__ double group_ordered_fma(double s, double v, double x, int n) {
    constexpr int W = G < 8 ? G : 8;
#pragma unroll
    for (int q0 = 0; q0 < G; q0 += W) {
        double vv[W], xx[W];
#pragma unroll
        for (int q = 0; q < W; q++) {
            vv[q] = __shfl(v, q0 + q, G);
            xx[q] = __shfl(x, q0 + q, G);
        }
#pragma unroll
        for (int q = 0; q < W; q++)
            if (q0 + q < n) s = fma(vv[q], xx[q], s);
    }
    return s;
}

// Row i (level-ordered buffers) of the fused refined solve, G lanes per row:
// XD[j] = (x0, d) of every row j of an earlier level (one 8-byte gather per
// entry brings both).
template <int G>
__device__ __forceinline__ void ref_row(int32_t i, bool valid, const int32_t* __restrict__ rp,
                                        const int32_t* __restrict__ col, const double* __restrict__ v64,
                                        const float* __restrict__ v32, const double* __restrict__ b, float2* XD,
                                        double* __restrict__ xo) {
    const int l = threadIdx.x & (G - 1);
    float s0 = 0.f, s2 = 0.f;
    double s = 0.0;
    int32_t a = 0, e = 0;
    if (valid) {
        a = rp[i];
        e = rp[i + 1] - 1;
    }
    // a and e are the same on the lanes of a group, so is every trip count
    for (int32_t j0 = a; j0 < e; j0 += G) {
        const int32_t j = j0 + l;
        float p0 = 0.f, p2 = 0.f;
        double v = 0.0, xd = 0.0;
        if (j < e) {
            const float w = v32[j];
            const float2 q = XD[col[j]];
            p0 = mul_rn(w, q.x);
            p2 = mul_rn(w, q.y);
            v = v64[j];
            xd = (double)q.x;
        }
        if (G == 1) {
            s0 = add_rn(s0, p0);
            s2 = add_rn(s2, p2);
            s = fma(v, xd, s);
        } else {
            const int n = min(G, e - j0);
            s0 = group_ordered_sum<false, G>(s0, p0, n);
            s2 = group_ordered_sum<false, G>(s2, p2, n);
            s = group_ordered_fma<G>(s, v, xd, n);
        }
    }
    if (valid) {
        const float w = v32[e];
        const double bi = b[i];
        const float x0 = div_rn(sub_rn((float)bi, s0), w);
        s = fma(v64[e], (double)x0, s);
        const float d = div_rn(sub_rn((float)(bi - s), s2), w);
        if (l == 0) {
            XD[i] = make_float2(x0, d);
            xo[i] = (double)x0 + (double)d;
        }
    }
}

}  // namespace sable
