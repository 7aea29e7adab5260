// Normalized min-sum over the 360-lifted QC base graph, one workgroup per
// codeword, all iterations in one launch.
//
// Semantics (plain FP32 mode) are those of lab/ldpc_kernel.c:82-143:
//   for it = 0..iters:
//     unsat = #checks whose XOR of (tot[v] < 0) is 1       (syndrome first)
//     unsat == 0 -> bits = tot < 0, conv = 1, its = it, bad = 0, stop
//     it == iters -> bits = tot < 0, conv = 0, its = iters, bad = unsat, stop
//     per check: x = tot[v] - m; min1/min2 of |x|; the first edge attaining
//       min1 takes min2, the others min1; times alpha; negated iff
//       parity(x < 0) ^ (x < 0)
//     per variable: s = 0; s += m in ascending check index; tot = lam + s
//
// Layout.  Check (g, j) -- group g, position j in 0..359 -- and variable
// (t, p).  Base edge b joins check group g to variable group t with shift s:
// check (g, j) meets variable position p = (j + s) % 360 of group t.  The
// message of that edge is stored at the VARIABLE position, msg[b*384 + p],
// and the posterior at tot[t*384 + p] (rows padded to 384 words).  Lane j of
// the check phase computes p once per edge and uses it for the tot load,
// the message load and the message store; lane p of the variable phase
// reads msg[b*384 + p] -- a fixed, 128-byte aligned row slice -- with no
// address arithmetic at all.
//
// Bit-exactness notes.
//  * min1/min2 are a running minimum and a running median-of-three
//    (fmed3(min1, min2, |x|) is the new second minimum): the multiset
//    semantics of the strict '<' scan.  NaN magnitudes are mapped to +inf
//    first, which is what the scan does with them (NaN < m is false).
//  * "The first edge attaining min1 takes min2" is applied as "an edge with
//    |x| == min1 takes min2": when min1 < min2 exactly one edge attains it;
//    when two or more tie, min2 == min1 and every edge gets the same value,
//    whichever is called first.  Products min*alpha are formed once per
//    check from the same operands: identical values.
//  * The per-variable sum must be sequential in ascending check index.  The
//    host sorts each variable group's entries by (region, u, g) with u =
//    (360 - s) % 360; at position p the entries whose check position
//    wrapped past 359 (p + u >= 360) are the last ones of their region and
//    come first in check order.  The lane adds, per region, the wrapped
//    suffix then the unwrapped prefix -- one FP32 add per message, in order.
//  * Compiled with -ffp-contract=off: no FMA contraction anywhere.
//  * The syndrome of iteration `it` comes from the same tot loads as the
//    check update; when it reads zero the speculative message writes are
//    never consumed -- tot is untouched until the variable phase.
#include "ldpc_qc_common.h"

namespace atsc3 {

template <int FMT, bool TA, int CMAX, int VMAX>
__global__ __launch_bounds__(ATSC3_WG) void minsum_qc_kernel(QcArgs a) {
    using M = Msg<FMT>;
    using T = typename M::T;
    const int blk = blockIdx.x;
    const unsigned tid = threadIdx.x;
    const int n = a.n;
    const float *lam = a.llr + (size_t)blk * n;
    float *tot = a.tot_ws + (size_t)blk * a.nvg * QC_ROW;
    T *msg = reinterpret_cast<T *>(a.msg_ws) + (size_t)blk * a.nbe * QC_ROW;
    __shared__ int slots[ATSC3_WG / 64];
    __shared__ int s_hist[256];

    float step = 1.0f, inv_step = 1.0f;
    if (FMT == 2) {
        step = a.msg_scale ? a.msg_scale[blk] : int8_step_of_block(lam, n, s_hist);
        inv_step = 1.0f / step;
    }
    if (tid < ATSC3_Z)
        for (int t = 0; t < a.nvg; ++t) tot[t * QC_ROW + tid] = lam[t * ATSC3_Z + tid];
    __syncthreads();

    for (int it = 0;; ++it) {
        const bool first = it == 0;
        const bool update = it < a.iters;
        unsigned unsat = 0u;
        if (tid < ATSC3_Z) {
            const unsigned j = tid;
            for (int si = 0; si < a.nsc; ++si) {
                const int sw = __builtin_amdgcn_readfirstlane(a.c_steps[si]);
                const int D = sw & 63;
                const int G = (sw >> 6) & 7;
                const int start = (sw >> 9) & 4095;
                if ((sw >> 21) & 1) {
                    // generic step: the degree comes from the group word
                    // (the step's 6-bit degree field saturates)
                    const int info = a.cg_info[start];
                    unsat += check_generic<FMT>(a, tot, msg, j, (info >> 9) & 2047, info & 511,
                                                first, update, step, inv_step);
                    continue;
                }
                switch (G * 64 + D) {
#define QC_C1(D)                                                                      \
    case 64 + D:                                                                      \
        if constexpr (D <= CMAX)                                                      \
            unsat += check_multi<FMT, D, 1>(a, tot, msg, j, start, first, update, step, \
                                            inv_step);                                \
        break;
#define QC_C2(D)                                                                       \
    case 128 + D:                                                                      \
        unsat += check_multi<FMT, D, 2>(a, tot, msg, j, start, first, update, step,    \
                                        inv_step);                                     \
        break;
                    QC_CASES16(QC_C1)
                    QC_CASES8(QC_C2)
#undef QC_C1
#undef QC_C2
                    default: break;
                }
            }
        }
        const int total = wg_sum_i32((int)unsat, slots);
        if (total == 0 || it == a.iters) {
            uint8_t *bout = a.bits + (size_t)blk * n;
            if (tid < ATSC3_Z)
                for (int t = 0; t < a.nvg; ++t)
                    bout[t * ATSC3_Z + tid] = (uint8_t)(tot[t * QC_ROW + tid] < 0.0f);
            if (tid == 0) {
                a.conv[blk] = (uint8_t)(total == 0);
                a.its[blk] = total == 0 ? it : a.iters;
                a.bad[blk] = total;
            }
            return;
        }
        if (tid < ATSC3_Z) {
            const unsigned p = tid;
            for (int si = 0; si < a.nsv; ++si) {
                const int sw = __builtin_amdgcn_readfirstlane(a.v_steps[si]);
                const int D = sw & 63;
                const int G = (sw >> 6) & 7;
                const int start = (sw >> 9) & 4095;
                if ((sw >> 21) & 1) {
                    const int info = a.vg_info[start];
                    const int t = (info >> 24) & 255;
                    var_generic<FMT, VMAX>(a, lam + t * ATSC3_Z, tot + t * QC_ROW, msg, p,
                                           (info >> 12) & 2047, info & 63, (info >> 6) & 63,
                                           step);
                    continue;
                }
                switch (G * 64 + D) {
#define QC_V1(D)                                                                      \
    case 64 + D:                                                                      \
        if constexpr (D <= VMAX) var_multi<FMT, D, 1, TA>(a, lam, tot, msg, p, start, step); \
        break;
#define QC_V2(D)                                                                      \
    case 128 + D:                                                                     \
        var_multi<FMT, D, 2, TA>(a, lam, tot, msg, p, start, step);                   \
        break;
#define QC_V4(D)                                                                      \
    case 256 + D:                                                                     \
        var_multi<FMT, D, 4, TA>(a, lam, tot, msg, p, start, step);                   \
        break;
                    QC_CASES(QC_V1)
                    QC_CASES8(QC_V2)
                    QC_CASES4(QC_V4)
#undef QC_V1
#undef QC_V2
#undef QC_V4
                    default: break;
                }
            }
        }
        __syncthreads();
    }
}

template <int FMT, bool TA, int CMAX, int VMAX>
static int32_t launch_one(const QcArgs &a, int nblk, hipStream_t st) {
    hipLaunchKernelGGL((minsum_qc_kernel<FMT, TA, CMAX, VMAX>), dim3(nblk), dim3(ATSC3_WG), 0,
                       st, a);
    return launch_status();
}

// Degree tiers: the kernel is sized for the largest fast-path degree it
// contains.  Check degrees above 16 (at most 4% of the checks of any code
// here) take the generic two-pass path, which holds no per-edge registers;
// variable degrees up to 16 or 32.
template <int FMT, bool TA>
static int32_t launch_tier(const QcArgs &a, int nblk, int max_cdeg, int max_vdeg,
                           hipStream_t st) {
    (void)max_cdeg;   // check degrees above 16 take the register-free path
    if (max_vdeg <= 16) return launch_one<FMT, TA, 16, 16>(a, nblk, st);
    return launch_one<FMT, TA, 16, 32>(a, nblk, st);
}

template <int FMT>
static int32_t launch_qc(const QcArgs &a, int nblk, int type_a, int max_cdeg, int max_vdeg,
                         hipStream_t st) {
    return type_a ? launch_tier<FMT, true>(a, nblk, max_cdeg, max_vdeg, st)
                  : launch_tier<FMT, false>(a, nblk, max_cdeg, max_vdeg, st);
}

}  // namespace atsc3

extern "C" int32_t atsc3_minsum_qc_dev(const float *llr, int32_t nblk, int32_t n, int32_t ncg,
                                       int32_t nvg, int32_t nbe, int32_t type_a,
                                       const int32_t *cg_info, const int32_t *be_off,
                                       const int32_t *be_miss, const int32_t *vg_info,
                                       const int32_t *ve_off, const int32_t *ve_miss,
                                       const int32_t *c_steps, int32_t nsc,
                                       const int32_t *v_steps, int32_t nsv,
                                       int32_t max_cdeg, int32_t max_vdeg, int32_t iters,
                                       float alpha,
                                       int32_t msg_fmt, const float *msg_scale,
                                       float *tot_ws, void *msg_ws, uint8_t *bits,
                                       uint8_t *conv, int32_t *its, int32_t *bad,
                                       void *stream) {
    if (nblk < 0 || iters < 0) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !cg_info || !be_off || !be_miss || !vg_info || !ve_off || !ve_miss ||
        !c_steps || !v_steps || !tot_ws || !msg_ws || !bits || !conv || !its || !bad)
        return ATSC3_E_ARG;
    if (nsc <= 0 || nsc > ncg || nsv <= 0 || nsv > nvg || nvg > 255) return ATSC3_E_ARG;
    if (n != nvg * ATSC3_Z || ncg <= 0 || nvg <= 0 || nbe <= 0 || nbe >= 2048)
        return ATSC3_E_ARG;
    if (max_vdeg < 1 || max_cdeg < 1) return ATSC3_E_ARG;
    if (max_vdeg > 32) return ATSC3_E_DEGREE;
    if (msg_fmt < 0 || msg_fmt > 2) return ATSC3_E_ARG;
    using CI = const QC_CONST int *;
    
    atsc3::QcArgs a{llr,
                    n,
                    ncg,
                    nvg,
                    nbe,
                    (CI)cg_info,
                    (CI)be_off,
                    (CI)be_miss,
                    (CI)vg_info,
                    (CI)ve_off,
                    (CI)ve_miss,
                    (CI)c_steps,
                    (CI)v_steps,
                    nsc,
                    nsv,
                    iters,
                    alpha,
                    msg_scale,
                    tot_ws,
                    msg_ws,
                    bits,
                    conv,
                    its,
                    bad};
    hipStream_t st = atsc3::as_stream(stream);
    switch (msg_fmt) {
        case 0: return atsc3::launch_qc<0>(a, nblk, type_a, max_cdeg, max_vdeg, st);
        case 1: return atsc3::launch_qc<1>(a, nblk, type_a, max_cdeg, max_vdeg, st);
        default: return atsc3::launch_qc<2>(a, nblk, type_a, max_cdeg, max_vdeg, st);
    }
}
