// Normalized min-sum over the 360-lifted QC base graph, one workgroup per
// codeword, all iterations in one launch.
//
// Semantics (plain FP32 mode) are those of lab/ldpc_kernel.c:82-143:
//   for it = 0..iters:
//     unsat = #checks whose XOR of (tot[v] < 0) is 1       (syndrome first)
//     unsat == 0 -> bits = tot < 0, conv = 1, its = it, bad = 0, stop
//     it == iters -> bits = tot < 0, conv = 0, its = iters, bad = unsat, stop
//     per check: x = tot[v] - m; min1/min2 of |x|; the first edge (lowest
//       variable index) attaining min1 takes min2, the others min1; times
//       alpha; negated iff parity(x < 0) ^ (x < 0)
//     per variable: s = 0; s += m in ascending check index; tot = lam + s
//
// Layout.  Check (g, j) -- group g, position j in 0..359 -- and variable
// (t, p).  Base edge b joins check group g to variable group t with shift s:
// check (g, j) meets variable t*360 + (j+s)%360.  The message of that edge
// lives at msg[b*360 + j], so lane j of the workgroup reads 360 consecutive
// words per base edge in the check phase, and lane p reads a rotated run of
// 360 consecutive words in the variable phase.  HBM sees only full lines.
//
// Bit-exactness notes.
//  * The first-minimum rule is evaluated as a lexicographic minimum of
//    (|x|, variable index); NaN and +inf magnitudes rank as +inf, which is
//    exactly what the strict '<' scan from min1 = min2 = +inf does.
//  * The per-variable sum must be sequential in ascending check index.  The
//    host sorts each variable group's entries by (region, u, g); at position
//    p the entries whose check position wrapped past 359 come first, so the
//    lane adds (region 0: wrapped, unwrapped), then (region 1: wrapped,
//    unwrapped) with predicated adds -- one FP32 add per message, in order.
//  * Check degrees up to 32 keep x in registers; larger degrees (64800
//    r5/15 47, r7/15 86) recompute x in a second pass from the same operands.
//  * Compiled with -ffp-contract=off: no FMA contraction anywhere.
//  * The syndrome of iteration `it` is computed from the same tot loads as
//    the check update; when it reads zero the (speculative) message writes
//    are simply never consumed -- tot is not touched until the variable
//    phase, so the reported bits are those the syndrome saw.
#include <hip/hip_fp16.h>

#include "fec_common.h"

namespace atsc3 {

struct QcArgs {
    const float *llr;
    int n, ncg, nvg, nbe;
    const int *cg_ptr, *be_word, *vg_ptr, *ve_word;
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
template <>
struct Msg<0> {
    using T = float;
    __device__ static float ld(const T *p, float) { return *p; }
    __device__ static void st(T *p, float v, float) { *p = v; }
};
template <>
struct Msg<1> {
    using T = __half;
    __device__ static float ld(const T *p, float) { return __half2float(*p); }
    __device__ static void st(T *p, float v, float) {
        v = fminf(fmaxf(v, -65504.0f), 65504.0f);
        *p = __float2half_rn(v);
    }
};
template <>
struct Msg<2> {
    using T = int8_t;
    __device__ static float ld(const T *p, float step) { return (float)(*p) * step; }
    __device__ static void st(T *p, float v, float inv_step) {
        float q = rintf(v * inv_step);
        q = fminf(fmaxf(q, -127.0f), 127.0f);
        *p = (int8_t)q;
    }
};

// INT8 step when the caller gives none: clip level = 8 * mean|llr| of the
// block, 127 steps.  Deterministic fixed-order reduction.
__device__ float int8_step_of_block(const float *lam, int n, float *fslots) {
    float acc = 0.0f;
    for (int v = threadIdx.x; v < n; v += ATSC3_WG) acc += fabsf(lam[v]);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) fslots[threadIdx.x >> 6] = acc;
    __syncthreads();
    float t = 0.0f;
#pragma unroll
    for (int w = 0; w < ATSC3_WG / 64; ++w) t += fslots[w];
    __syncthreads();
    const float mean = t / (float)n;
    return fmaxf(8.0f * mean / 127.0f, 1e-30f);
}

template <int FMT, int MAXDC, int MAXDV>
__global__ __launch_bounds__(ATSC3_WG) void minsum_qc_kernel(QcArgs a) {
    using M = Msg<FMT>;
    using T = typename M::T;
    const int blk = blockIdx.x;
    const int tid = threadIdx.x;
    const int n = a.n;
    const float *lam = a.llr + (size_t)blk * n;
    float *tot = a.tot_ws + (size_t)blk * n;
    T *msg = reinterpret_cast<T *>(a.msg_ws) + (size_t)blk * a.nbe * ATSC3_Z;
    __shared__ int slots[ATSC3_WG / 64];
    __shared__ float fslots[ATSC3_WG / 64];

    float step = 1.0f, inv_step = 1.0f;
    if (FMT == 2) {
        step = a.msg_scale ? a.msg_scale[blk] : int8_step_of_block(lam, n, fslots);
        inv_step = 1.0f / step;
    }

    for (int it = 0;; ++it) {
        const float *src = it == 0 ? lam : tot;
        const bool update = it < a.iters;
        int unsat = 0;
        if (tid < ATSC3_Z) {
            const int j = tid;
            for (int g = 0; g < a.ncg; ++g) {
                const int b0 = a.cg_ptr[g];
                const int dc = a.cg_ptr[g + 1] - b0;
                unsigned syn = 0u, par = 0u;
                float lo1 = INFINITY, lo2 = INFINITY;
                int first = 0x7fffffff;
                // x of edge k of this check, and its variable (-1 if the
                // edge is the missing corner entry)
                auto edge = [&](int k, float &xv, float &tv) -> int {
                    const int w = a.be_word[b0 + k];
                    const int s = w & 511;
                    const int t = (w >> 9) & 511;
                    const int miss = (w >> 18) & 1023;
                    if (miss == j + 1) return -1;
                    int p = j + s;
                    if (p >= ATSC3_Z) p -= ATSC3_Z;
                    const int v = t * ATSC3_Z + p;
                    tv = src[v];
                    const float mo =
                        it == 0 ? 0.0f : M::ld(msg + (size_t)(b0 + k) * ATSC3_Z + j, step);
                    xv = tv - mo;
                    return v;
                };
                auto fold = [&](int v, float xv, float tv) {
                    syn ^= (unsigned)(tv < 0.0f);
                    par ^= (unsigned)(xv < 0.0f);
                    float mag = fabsf(xv);
                    mag = mag < INFINITY ? mag : INFINITY;
                    if (mag < lo1 || (mag == lo1 && v < first)) {
                        lo2 = lo1;
                        lo1 = mag;
                        first = v;
                    } else if (mag < lo2) {
                        lo2 = mag;
                    }
                };
                auto emit = [&](int k, int v, float xv) {
                    float val = (v == first ? lo2 : lo1) * a.alpha;
                    if (par ^ (unsigned)(xv < 0.0f)) val = -val;
                    M::st(msg + (size_t)(b0 + k) * ATSC3_Z + j, val, inv_step);
                };
                if constexpr (MAXDC > 0) {
                    // degree bounded at compile time: keep x in registers
                    float x[MAXDC > 0 ? MAXDC : 1];
                    int vv[MAXDC > 0 ? MAXDC : 1];
#pragma unroll
                    for (int k = 0; k < MAXDC; ++k) {
                        vv[k] = -1;
                        x[k] = 0.0f;
                        if (k < dc) {
                            float tv = 0.0f;
                            vv[k] = edge(k, x[k], tv);
                            if (vv[k] >= 0) fold(vv[k], x[k], tv);
                        }
                    }
                    if (update) {
#pragma unroll
                        for (int k = 0; k < MAXDC; ++k)
                            if (k < dc && vv[k] >= 0) emit(k, vv[k], x[k]);
                    }
                } else {
                    // any degree: second pass recomputes x from the same
                    // operands (identical FP32 result) instead of holding it
                    for (int k = 0; k < dc; ++k) {
                        float xv = 0.0f, tv = 0.0f;
                        const int v = edge(k, xv, tv);
                        if (v >= 0) fold(v, xv, tv);
                    }
                    if (update) {
                        for (int k = 0; k < dc; ++k) {
                            float xv = 0.0f, tv = 0.0f;
                            const int v = edge(k, xv, tv);
                            if (v >= 0) emit(k, v, xv);
                        }
                    }
                }
                unsat += (int)syn;
            }
        }
        const int total = wg_sum_i32(unsat, slots);
        if (total == 0 || it == a.iters) {
            uint8_t *bout = a.bits + (size_t)blk * n;
            for (int v = tid; v < n; v += ATSC3_WG) bout[v] = (uint8_t)(src[v] < 0.0f);
            if (tid == 0) {
                a.conv[blk] = (uint8_t)(total == 0);
                a.its[blk] = total == 0 ? it : a.iters;
                a.bad[blk] = total;
            }
            return;
        }
        if (tid < ATSC3_Z) {
            const int p = tid;
            for (int t = 0; t < a.nvg; ++t) {
                const int e0 = a.vg_ptr[t];
                const int dv = a.vg_ptr[t + 1] - e0;
                float val[MAXDV];
                unsigned wrapped = 0u, valid = 0u, reg1 = 0u;
#pragma unroll
                for (int k = 0; k < MAXDV; ++k) {
                    val[k] = 0.0f;
                    if (k < dv) {
                        const int w = a.ve_word[e0 + k];
                        const int b = w & 2047;
                        const int u = (w >> 11) & 511;
                        const int region = (w >> 20) & 1;
                        const int vmiss = (w >> 21) & 1023;
                        int jj = p + u;
                        const bool wr = jj >= ATSC3_Z;
                        if (wr) jj -= ATSC3_Z;
                        val[k] = M::ld(msg + (size_t)b * ATSC3_Z + jj, step);
                        wrapped |= (unsigned)wr << k;
                        valid |= (unsigned)(vmiss != p + 1) << k;
                        reg1 |= (unsigned)region << k;
                    }
                }
                float s = 0.0f;
#pragma unroll
                for (int r = 0; r < 2; ++r) {
#pragma unroll
                    for (int w = 1; w >= 0; --w) {
#pragma unroll
                        for (int k = 0; k < MAXDV; ++k) {
                            const bool take = ((valid >> k) & 1u) &&
                                              ((reg1 >> k) & 1u) == (unsigned)r &&
                                              ((wrapped >> k) & 1u) == (unsigned)w;
                            if (take) s += val[k];
                        }
                    }
                }
                const int v = t * ATSC3_Z + p;
                tot[v] = lam[v] + s;
            }
        }
        __syncthreads();
    }
}

template <int FMT, int DC, int DV>
static int32_t launch_qc(const QcArgs &a, int nblk, hipStream_t st) {
    hipLaunchKernelGGL((minsum_qc_kernel<FMT, DC, DV>), dim3(nblk), dim3(ATSC3_WG), 0, st, a);
    return launch_status();
}

template <int FMT>
static int32_t dispatch_qc(const QcArgs &a, int nblk, int max_cdeg, int max_vdeg,
                           hipStream_t st) {
    if (max_cdeg <= 16 && max_vdeg <= 16) return launch_qc<FMT, 16, 16>(a, nblk, st);
    if (max_cdeg <= 32 && max_vdeg <= 16) return launch_qc<FMT, 32, 16>(a, nblk, st);
    if (max_cdeg <= 16 && max_vdeg <= 32) return launch_qc<FMT, 16, 32>(a, nblk, st);
    if (max_cdeg <= 32) return launch_qc<FMT, 32, 32>(a, nblk, st);
    if (max_vdeg <= 16) return launch_qc<FMT, 0, 16>(a, nblk, st);
    return launch_qc<FMT, 0, 32>(a, nblk, st);
}

}  // namespace atsc3

extern "C" int32_t atsc3_minsum_qc_dev(const float *llr, int32_t nblk, int32_t n,
                                       int32_t ncg, int32_t nvg, int32_t nbe,
                                       const int32_t *cg_ptr, const int32_t *be_word,
                                       const int32_t *vg_ptr, const int32_t *ve_word,
                                       int32_t max_cdeg, int32_t max_vdeg, int32_t iters,
                                       float alpha, int32_t msg_fmt,
                                       const float *msg_scale, float *tot_ws,
                                       void *msg_ws, uint8_t *bits, uint8_t *conv,
                                       int32_t *its, int32_t *bad, void *stream) {
    if (nblk < 0 || iters < 0) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !cg_ptr || !be_word || !vg_ptr || !ve_word || !tot_ws || !msg_ws ||
        !bits || !conv || !its || !bad)
        return ATSC3_E_ARG;
    if (n != nvg * ATSC3_Z || ncg <= 0 || nvg <= 0 || nvg > 511 || nbe <= 0 ||
        nbe >= 2048)
        return ATSC3_E_ARG;
    if (max_cdeg < 1 || max_vdeg < 1) return ATSC3_E_ARG;
    if (max_vdeg > 32) return ATSC3_E_DEGREE;   // check degree is unbounded
    if (msg_fmt < 0 || msg_fmt > 2) return ATSC3_E_ARG;
    atsc3::QcArgs a{llr,   n,        ncg,    nvg,      nbe,  cg_ptr, be_word,
                    vg_ptr, ve_word, iters,  alpha,    msg_scale, tot_ws, msg_ws,
                    bits,  conv,     its,    bad};
    hipStream_t st = atsc3::as_stream(stream);
    switch (msg_fmt) {
        case 0: return atsc3::dispatch_qc<0>(a, nblk, max_cdeg, max_vdeg, st);
        case 1: return atsc3::dispatch_qc<1>(a, nblk, max_cdeg, max_vdeg, st);
        default: return atsc3::dispatch_qc<2>(a, nblk, max_cdeg, max_vdeg, st);
    }
}
