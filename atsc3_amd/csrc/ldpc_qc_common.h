// Device code shared by the QC min-sum engines (fused per-block kernel and
// per-iteration phase kernels): table layout, message codecs and the
// per-check / per-variable step bodies.  The semantics and the
// bit-exactness argument are stated in ldpc_qc.hip.
#pragma once
#include <hip/hip_fp16.h>

#include "fec_common.h"

#define QC_ROW 384

namespace atsc3 {

// Read-only code tables live in the constant address space: the compiler
// then reads them through the scalar cache when the index is uniform (every
// table index in the hot loops is), instead of spending vector-memory
// instructions on them.
#define QC_CONST __attribute__((address_space(4)))

struct QcArgs {
    const float *llr;
    int n, ncg, nvg, nbe;
    const QC_CONST int *cg_info;
    const QC_CONST int *be_off;   // pairs (s, t*384)
    const QC_CONST int *be_miss;
    const QC_CONST int *vg_info;
    const QC_CONST int *ve_off;   // pairs (b*384, u)
    const QC_CONST int *ve_miss;
    const QC_CONST int *c_steps, *v_steps;
    int nsc, nsv;
    int iters;
    float alpha;
    const float *msg_scale;
    float *tot_ws;
    void *msg_ws;
    uint8_t *bits, *conv;
    int *its, *bad;
};

// Message storage codecs.  FMT 0 is exact FP32.
template <int FMT>
struct Msg;
// ld = cvt(raw(p)): the engines issue every raw load of a step before the
// first conversion (a conversion right after its load would make the wave
// wait for that load before issuing the next one).
template <>
struct Msg<0> {
    using T = float;
    using R = float;
    __device__ static R raw(const T *p) { return *p; }
    __device__ static float cvt(R r, float) { return r; }
    __device__ static float ld(const T *p, float) { return *p; }
    __device__ static void st(T *p, float v, float) { *p = v; }
};
template <>
struct Msg<1> {
    using T = __half;
    using R = unsigned short;
    __device__ static R raw(const T *p) { return *reinterpret_cast<const unsigned short *>(p); }
    __device__ static float cvt(R r, float) { return __half2float(__ushort_as_half(r)); }
    __device__ static float ld(const T *p, float) { return __half2float(*p); }
    __device__ static void st(T *p, float v, float) {
        v = fminf(fmaxf(v, -65504.0f), 65504.0f);
        *p = __float2half_rn(v);
    }
};
template <>
struct Msg<2> {
    using T = int8_t;
    using R = int;
    __device__ static R raw(const T *p) { return (int)*p; }
    __device__ static float cvt(R r, float step) { return (float)r * step; }
    __device__ static float ld(const T *p, float step) { return (float)(*p) * step; }
    __device__ static void st(T *p, float v, float inv_step) {
        float q = rintf(v * inv_step);
        q = fminf(fmaxf(q, -127.0f), 127.0f);
        *p = (int8_t)q;
    }
};

// INT8 step when the caller gives none: 6 * median|llr| / 127 (the lower
// median, the ((n-1)/2)-th smallest |llr|).  A mean is pulled up by the heavy
// tail of faded high-order-constellation LLRs (too coarse a step), while the
// median tracks the bulk for every channel and constellation.  The median is
// found to 8 mantissa bits by two integer histograms in LDS -- the binary
// exponent, then the top 8 mantissa bits inside the median's exponent -- so
// the result does not depend on the order of the atomic updates.  Called by
// every thread of the workgroup; `hist` is 256 ints of LDS.
__device__ inline float int8_step_of_block(const float *lam, int n, int *hist) {
    __shared__ int sel[2];
    const int rank = (n - 1) / 2;
    for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    // 8 loads in flight per thread before the histogram updates
    const int step = blockDim.x;
    for (int v0 = threadIdx.x; v0 < n; v0 += 8 * step) {
        unsigned u[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            u[q] = v0 + q * step < n ? __float_as_uint(fabsf(lam[v0 + q * step])) : 0xffffffffu;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (u[q] != 0xffffffffu) atomicAdd(&hist[(u[q] >> 23) & 255], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0, e = 0;
        for (; e < 255 && c + hist[e] <= rank; ++e) c += hist[e];
        sel[0] = e;
        sel[1] = rank - c;
    }
    __syncthreads();
    const unsigned e = (unsigned)sel[0];
    const int r2 = sel[1];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int v0 = threadIdx.x; v0 < n; v0 += 8 * step) {
        unsigned u[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            u[q] = v0 + q * step < n ? __float_as_uint(fabsf(lam[v0 + q * step])) : 0xffffffffu;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (u[q] != 0xffffffffu && ((u[q] >> 23) & 255) == e)
                atomicAdd(&hist[(u[q] >> 15) & 255], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0, m = 0;
        for (; m < 255 && c + hist[m] <= r2; ++m) c += hist[m];
        sel[0] = m;
    }
    __syncthreads();
    // middle of the median's 8-bit mantissa bucket
    const float med = __uint_as_float((e << 23) | ((unsigned)sel[0] << 15) | (1u << 14));
    __syncthreads();
    return fmaxf(6.0f * med / 127.0f, 1e-30f);
}

__device__ __forceinline__ unsigned wrap360(unsigned p) { return min(p, p - 360u); }

// base[idx] with a wave-uniform base and a 32-bit per-lane element index:
// written as base + zext(32-bit byte offset) so the load/store can use the
// scalar-base + 32-bit vector-offset addressing form (one VGPR per access,
// no per-lane 64-bit address arithmetic).
template <typename T>
__device__ __forceinline__ const T &at(const T *base, unsigned idx) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                        (uint64_t)(idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ T &at(T *base, unsigned idx) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) +
                                  (uint64_t)(idx * (unsigned)sizeof(T)));
}

// The same with a byte offset (the caller has scaled the index).
template <typename T>
__device__ __forceinline__ const T &at(const T *base, unsigned, unsigned byte_off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                        (uint64_t)byte_off);
}

__device__ __forceinline__ float clamp_mag(float x) { return fminf(fabsf(x), INFINITY); }

// G checks (g_0..g_{G-1}, j), all of degree D: every load of the G checks
// is issued before the first conversion or store, then each check folds and
// emits.  FIRST (iteration 0: all messages zero) is a template parameter so
// the load sequence is straight-line code in both cases.
template <int FMT, int D, int G, bool FIRST>
__device__ __forceinline__ unsigned check_multi_impl(const QcArgs &a,
                                                     const float *__restrict__ tot,
                                                     typename Msg<FMT>::T *__restrict__ msg,
                                                     unsigned j, int start, bool update,
                                                     float step, float inv_step) {
    using M = Msg<FMT>;
    float tv[G][D];
    typename M::R mr[G][D];
    int b0[G];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        b0[gi] = __builtin_amdgcn_readfirstlane((a.cg_info[start + gi] >> 9) & 2047);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const int ex = a.be_off[2 * (b0[gi] + k)], ey = a.be_off[2 * (b0[gi] + k) + 1];
            const unsigned p = wrap360(j + (unsigned)ex);
            tv[gi][k] = at(tot + ey, p);
            if (!FIRST) mr[gi][k] = M::raw(&at(msg + (size_t)(b0[gi] + k) * QC_ROW, p));
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    float x[G][D];
    unsigned syn_total = 0u;
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        unsigned syn = 0u;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float mo = FIRST ? 0.0f : M::cvt(mr[gi][k], step);
            syn ^= (unsigned)(tv[gi][k] < 0.0f);
            x[gi][k] = tv[gi][k] - mo;
        }
        syn_total += syn;
    }
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        unsigned par = 0u;
        float m1 = INFINITY, m2 = INFINITY;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            par ^= (unsigned)(x[gi][k] < 0.0f);
            const float mag = clamp_mag(x[gi][k]);
            const float n2 = __builtin_amdgcn_fmed3f(m1, m2, mag);
            m1 = fminf(m1, mag);
            m2 = n2;
        }
        if (update) {
            const float a1 = m1 * a.alpha, a2 = m2 * a.alpha;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const unsigned p = wrap360(j + (unsigned)a.be_off[2 * (b0[gi] + k)]);
                float val = clamp_mag(x[gi][k]) == m1 ? a2 : a1;
                if (par ^ (unsigned)(x[gi][k] < 0.0f)) val = -val;
                M::st(&at(msg + (size_t)(b0[gi] + k) * QC_ROW, p), val, inv_step);
            }
        }
    }
    return syn_total;
}

template <int FMT, int D, int G>
__device__ __forceinline__ unsigned check_multi(const QcArgs &a, const float *__restrict__ tot,
                                                typename Msg<FMT>::T *__restrict__ msg,
                                                unsigned j, int start, bool first, bool update,
                                                float step, float inv_step) {
    return first ? check_multi_impl<FMT, D, G, true>(a, tot, msg, j, start, update, step,
                                                     inv_step)
                 : check_multi_impl<FMT, D, G, false>(a, tot, msg, j, start, update, step,
                                                      inv_step);
}

// Any degree, possibly with the missing corner entry: two passes, the
// second recomputes x from the same operands (identical FP32 value).
template <int FMT>
__device__ unsigned check_generic(const QcArgs &a, const float *__restrict__ tot,
                                  typename Msg<FMT>::T *__restrict__ msg, unsigned j, int b0,
                                  int dc, bool first, bool update, float step,
                                  float inv_step) {
    using M = Msg<FMT>;
    unsigned syn = 0u, par = 0u;
    float m1 = INFINITY, m2 = INFINITY;
    for (int k = 0; k < dc; ++k) {
        if (a.be_miss[b0 + k] == (int)j + 1) continue;
        const int ex = a.be_off[2 * (b0 + k)], ey = a.be_off[2 * (b0 + k) + 1];
        const unsigned p = wrap360(j + (unsigned)ex);
        const float tv = (tot + ey)[p];
        const float mo = first ? 0.0f : M::ld((msg + (size_t)(b0 + k) * QC_ROW) + p, step);
        const float xv = tv - mo;
        syn ^= (unsigned)(tv < 0.0f);
        par ^= (unsigned)(xv < 0.0f);
        const float mag = clamp_mag(xv);
        const float n2 = __builtin_amdgcn_fmed3f(m1, m2, mag);
        m1 = fminf(m1, mag);
        m2 = n2;
    }
    if (update) {
        const float a1 = m1 * a.alpha, a2 = m2 * a.alpha;
        for (int k = 0; k < dc; ++k) {
            if (a.be_miss[b0 + k] == (int)j + 1) continue;
            const int ex = a.be_off[2 * (b0 + k)], ey = a.be_off[2 * (b0 + k) + 1];
            const unsigned p = wrap360(j + (unsigned)ex);
            const float tv = (tot + ey)[p];
            const float mo =
                first ? 0.0f : M::ld((msg + (size_t)(b0 + k) * QC_ROW) + p, step);
            const float xv = tv - mo;
            float val = clamp_mag(xv) == m1 ? a2 : a1;
            if (par ^ (unsigned)(xv < 0.0f)) val = -val;
            M::st((msg + (size_t)(b0 + k) * QC_ROW) + p, val, inv_step);
        }
    }
    return syn;
}

// G variable groups of degree D at position p: all loads first, then the
// ordered sums.
template <int FMT, int D, int G, bool TA>
__device__ __forceinline__ void var_multi(const QcArgs &a, const float *__restrict__ lam,
                                          float *__restrict__ tot,
                                          const typename Msg<FMT>::T *__restrict__ msg,
                                          unsigned p, int start, float step) {
    using M = Msg<FMT>;
    float val[G][D], lv[G];
    int w0[G], w1[G], n0[G], tt[G];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int info = __builtin_amdgcn_readfirstlane(a.vg_info[start + gi]);
        const int e0 = (info >> 12) & 2047;
        n0[gi] = (info >> 6) & 63;
        tt[gi] = (info >> 24) & 255;
        lv[gi] = at(lam + tt[gi] * ATSC3_Z, p);
        w0[gi] = 0;
        w1[gi] = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const int ex = a.ve_off[2 * (e0 + k)], ey = a.ve_off[2 * (e0 + k) + 1];
            val[gi][k] = M::ld(&at(msg + ex, p), step);
            const int wr = (int)(p + (unsigned)ey >= 360u);
            if (TA && k < n0[gi])
                w0[gi] += wr;
            else
                w1[gi] += wr;
        }
    }
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        float s = 0.0f;
        if (TA) {
            const int t0 = n0[gi] - w0[gi];
#pragma unroll
            for (int k = 0; k < D; ++k)
                if (k < n0[gi] && k >= t0) s += val[gi][k];
#pragma unroll
            for (int k = 0; k < D; ++k)
                if (k < n0[gi] && k < t0) s += val[gi][k];
        }
        const int t1 = D - w1[gi];
#pragma unroll
        for (int k = 0; k < D; ++k)
            if ((!TA || k >= n0[gi]) && k >= t1) s += val[gi][k];
#pragma unroll
        for (int k = 0; k < D; ++k)
            if ((!TA || k >= n0[gi]) && k < t1) s += val[gi][k];
        at(tot + tt[gi] * QC_ROW, p) = lv[gi] + s;
    }
}

// Variable group holding the missing corner entry (or degree > 32): runtime
// degree up to 32, per-entry validity.
template <int FMT, int VMAX>
__device__ void var_generic(const QcArgs &a, const float *__restrict__ lam_t,
                            float *__restrict__ tot_t,
                            const typename Msg<FMT>::T *__restrict__ msg, unsigned p, int e0,
                            int dv, int n0, float step) {
    using M = Msg<FMT>;
    float val[VMAX];
    unsigned wrapped = 0u, valid = 0u;
#pragma unroll
    for (int k = 0; k < VMAX; ++k) {
        val[k] = 0.0f;
        if (k < dv) {
            const int ex = a.ve_off[2 * (e0 + k)], ey = a.ve_off[2 * (e0 + k) + 1];
            val[k] = M::ld((msg + ex) + p, step);
            wrapped |= (unsigned)(p + (unsigned)ey >= 360u) << k;
            valid |= (unsigned)(a.ve_miss[e0 + k] != (int)p + 1) << k;
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int w = 1; w >= 0; --w) {
#pragma unroll
            for (int k = 0; k < VMAX; ++k) {
                const bool in_r = r == 0 ? k < n0 : k >= n0;
                if (((valid >> k) & 1u) && in_r && ((wrapped >> k) & 1u) == (unsigned)w)
                    s += val[k];
            }
        }
    }
    tot_t[p] = lam_t[p] + s;
}


#define QC_CASES4(X) X(1) X(2) X(3) X(4)
#define QC_CASES8(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8)
#define QC_CASES16(X) \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
#define QC_CASES(X) \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
    X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) \
    X(31) X(32)

}  // namespace atsc3
