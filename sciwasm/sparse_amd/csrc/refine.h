// refine.h -- one row of the triangular solve with one refinement step
// fused into the same sweep (shared by the single-matrix and batched level
// schedules).
#pragma once
#include "common.h"

namespace sable {

// One entry's inputs to the three sums of the fused refined solve.
struct RefEntry {
    double v;   // the FP64 value (its FP32 value is v rounded)
    float x0;   // the column's FP32 solution
    float d;    // the column's FP32 correction
};

// Row i (level-ordered buffers) of the fused refined solve, G lanes per row:
// XD[j] = (x0, d) of every row j of an earlier level (one 8-byte gather per
// entry brings both).  The FP32 values are the FP64 values rounded as they
// are read.  With G > 1 the lanes load G entries at once and stage them in
// `stage` (one RefEntry per thread of the workgroup); every lane of the
// group then reads them back in stored order (one 16-byte LDS read per
// entry, the group's lanes reading the same address) and forms the three
// sums itself, so all lanes hold the row's sums in sequential order.
template <int G>
__device__ __forceinline__ void ref_row(int32_t i, bool valid, const int32_t* __restrict__ rp,
                                        const int32_t* __restrict__ col, const double* __restrict__ v64,
                                        const double* __restrict__ b, float2* XD, double* __restrict__ xo,
                                        RefEntry* stage) {
    const int l = threadIdx.x & (G - 1);
    float s0 = 0.f, s2 = 0.f;
    double s = 0.0;
    int32_t a = 0, e = 0;
    if (valid) {
        a = rp[i];
        e = rp[i + 1] - 1;
    }
    RefEntry* grp = stage + (threadIdx.x - l);
    // a and e are the same on the lanes of a group, so is every trip count
    for (int32_t j0 = a; j0 < e; j0 += G) {
        const int32_t j = j0 + l;
        RefEntry t{0.0, 0.f, 0.f};
        if (j < e) {
            t.v = v64[j];
            const float2 q = XD[col[j]];
            t.x0 = q.x;
            t.d = q.y;
        }
        if (G == 1) {
            const float w = (float)t.v;
            s0 = add_rn(s0, mul_rn(w, t.x0));
            s2 = add_rn(s2, mul_rn(w, t.d));
            s = fma(t.v, (double)t.x0, s);
        } else {
            const int n = min(G, e - j0);
            stage[threadIdx.x] = t;
            __builtin_amdgcn_wave_barrier();
            for (int q = 0; q < n; q++) {
                const RefEntry u = grp[q];
                const float w = (float)u.v;
                s0 = add_rn(s0, mul_rn(w, u.x0));
                s2 = add_rn(s2, mul_rn(w, u.d));
                s = fma(u.v, (double)u.x0, s);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (valid) {
        const double vd = v64[e];
        const float w = (float)vd;
        const double bi = b[i];
        const float x0 = div_rn(sub_rn((float)bi, s0), w);
        s = fma(vd, (double)x0, s);
        const float d = div_rn(sub_rn((float)(bi - s), s2), w);
        if (l == 0) {
            XD[i] = make_float2(x0, d);
            xo[i] = (double)x0 + (double)d;
        }
    }
}

}  // namespace sable
