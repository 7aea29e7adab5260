// Normalized min-sum over the QC base graph as per-iteration phase kernels:
// the throughput engine for large batches of codewords.
//
// One decode call enqueues, for it = 0..iters,
//   check kernel   every (active block, check step) work item: syndrome of
//                  the step's checks (atomically summed per block) and the
//                  speculative message update (skipped at it == iters)
//   var kernel     every (active block, variable step) item: if the block's
//                  syndrome read zero or it == iters, write its hard bits
//                  (and conv/its/bad once); otherwise the ordered variable
//                  sums tot = lam + s
//   compact kernel the next active list (stable order), counters reset
// Kernel boundaries are the phase barriers, so no workgroup ever waits for
// another.  Work items are walked in block-major contiguous chunks per
// workgroup, so one workgroup touches few blocks and the rows of tot it
// gathers stay hot in its XCD's L2.  When every block has retired, the
// remaining launches exit at their first load.
//
// Results are those of the fused kernel (ldpc_qc.hip) and therefore of
// lab/ldpc_kernel.c: the per-check and per-variable arithmetic is the same
// code (ldpc_qc_common.h); only the schedule of independent work changes.
#include "ldpc_qc_common.h"

#define PHASE_GRID 2048

namespace atsc3 {

struct PhaseCtl {
    int *active0, *active1;   // active block lists, alternating by iteration
    int *n_active;            // [2] list lengths
    unsigned *unsat;          // [nblk] unsatisfied checks of the current iteration
    float *step;              // [nblk] INT8 message step (FMT 2)
};

__device__ __forceinline__ int *active_list(const PhaseCtl &c, int sel) {
    return sel ? c.active1 : c.active0;
}

// tot = lam (padded rows), active list = all blocks, counters zeroed.
template <int FMT>
__global__ __launch_bounds__(ATSC3_WG) void phase_init(QcArgs a, PhaseCtl c, int nblk) {
    __shared__ int s_hist[256];
    const int blk = blockIdx.x;
    const unsigned tid = threadIdx.x;
    const float *lam = a.llr + (size_t)blk * a.n;
    float *tot = a.tot_ws + (size_t)blk * a.nvg * QC_ROW;
    if (tid < ATSC3_Z)
        for (int t = 0; t < a.nvg; ++t) tot[t * QC_ROW + tid] = lam[t * ATSC3_Z + tid];
    if (FMT == 2) {
        const float s = a.msg_scale ? a.msg_scale[blk] : int8_step_of_block(lam, a.n, s_hist);
        if (tid == 0) c.step[blk] = s;
    }
    if (tid == 0) {
        c.active0[blk] = blk;
        c.unsat[blk] = 0u;
        if (blk == 0) {
            c.n_active[0] = nblk;
            c.n_active[1] = 0;
        }
    }
}

__device__ __forceinline__ void flush_unsat(unsigned *unsat, int blk, unsigned cnt) {
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(unsat + blk, cnt);
}

template <int FMT, int CMAX>
__global__ __launch_bounds__(ATSC3_WG) void phase_check(QcArgs a, PhaseCtl c, int it) {
    using T = typename Msg<FMT>::T;
    const int sel = it & 1;
    const int nact = __builtin_amdgcn_readfirstlane(c.n_active[sel]);
    if (nact == 0) return;
    const int *act = active_list(c, sel);
    const unsigned tid = threadIdx.x;
    // Contiguous block-major chunk per workgroup.  (An XCD-affine round-robin
    // order raises the L2 hit rate of the tot gathers from ~44% to ~62% but
    // measured slower: per-item counter flushes and less sequential message
    // streaming cost more than the saved HBM reads.)
    const long long items = (long long)nact * a.nsc;
    const long long chunk = (items + gridDim.x - 1) / gridDim.x;
    const long long i0 = (long long)blockIdx.x * chunk;
    const long long i1 = i0 + chunk < items ? i0 + chunk : items;
    const bool first = it == 0, update = it < a.iters;
    unsigned cnt = 0u;
    int cur = -1;
    for (long long i = i0; i < i1; ++i) {
        const int blk = __builtin_amdgcn_readfirstlane(act[i / a.nsc]);
        const int si = (int)(i % a.nsc);
        if (blk != cur) {
            if (cur >= 0) flush_unsat(c.unsat, cur, cnt);
            cnt = 0u;
            cur = blk;
        }
        if (tid >= ATSC3_Z) continue;
        const float *tot = a.tot_ws + (size_t)blk * a.nvg * QC_ROW;
        T *msg = reinterpret_cast<T *>(a.msg_ws) + (size_t)blk * a.nbe * QC_ROW;
        const float step = FMT == 2 ? c.step[blk] : 1.0f;
        const float inv_step = FMT == 2 ? 1.0f / step : 1.0f;
        const unsigned j = tid;
        const int sw = __builtin_amdgcn_readfirstlane(a.c_steps[si]);
        const int D = sw & 63;
        const int G = (sw >> 6) & 7;
        const int start = (sw >> 9) & 4095;
        if ((sw >> 21) & 1) {
            const int info = a.cg_info[start];
            cnt += check_generic<FMT>(a, tot, msg, j, (info >> 9) & 2047, info & 511, first,
                                      update, step, inv_step);
            continue;
        }
        switch (G * 64 + D) {
#define PH_C1(D)                                                                        \
    case 64 + D:                                                                        \
        if constexpr (D <= CMAX)                                                        \
            cnt += check_multi<FMT, D, 1>(a, tot, msg, j, start, first, update, step,   \
                                          inv_step);                                    \
        break;
#define PH_C2(D)                                                                          \
    case 128 + D:                                                                         \
        cnt += check_multi<FMT, D, 2>(a, tot, msg, j, start, first, update, step, inv_step); \
        break;
            QC_CASES16(PH_C1)
            QC_CASES8(PH_C2)
#undef PH_C1
#undef PH_C2
            default: break;
        }
    }
    if (cur >= 0) flush_unsat(c.unsat, cur, cnt);
}

template <int FMT, bool TA, int VMAX>
__global__ __launch_bounds__(ATSC3_WG) void phase_var(QcArgs a, PhaseCtl c, int it) {
    using T = typename Msg<FMT>::T;
    const int sel = it & 1;
    const int nact = __builtin_amdgcn_readfirstlane(c.n_active[sel]);
    if (nact == 0) return;
    const int *act = active_list(c, sel);
    const unsigned tid = threadIdx.x;
    const long long items = (long long)nact * a.nsv;
    const long long chunk = (items + gridDim.x - 1) / gridDim.x;
    const long long i0 = (long long)blockIdx.x * chunk;
    const long long i1 = i0 + chunk < items ? i0 + chunk : items;
    for (long long i = i0; i < i1; ++i) {
        const int blk = __builtin_amdgcn_readfirstlane(act[i / a.nsv]);
        const int si = (int)(i % a.nsv);
        const unsigned u = __builtin_amdgcn_readfirstlane(c.unsat[blk]);
        const float *lam = a.llr + (size_t)blk * a.n;
        float *tot = a.tot_ws + (size_t)blk * a.nvg * QC_ROW;
        const int sw = __builtin_amdgcn_readfirstlane(a.v_steps[si]);
        const int D = sw & 63;
        const int G = (sw >> 6) & 7;
        const int start = (sw >> 9) & 4095;
        if (u == 0u || it == a.iters) {
            // the block retires at this iteration: bits of this step's groups
            if (tid < ATSC3_Z) {
                uint8_t *bout = a.bits + (size_t)blk * a.n;
                for (int gi = 0; gi < G; ++gi) {
                    const int t = (a.vg_info[start + gi] >> 24) & 255;
                    bout[t * ATSC3_Z + tid] = (uint8_t)(tot[t * QC_ROW + tid] < 0.0f);
                }
            }
            if (si == 0 && tid == 0) {
                a.conv[blk] = (uint8_t)(u == 0u);
                a.its[blk] = u == 0u ? it : a.iters;
                a.bad[blk] = (int)u;
            }
            continue;
        }
        if (tid >= ATSC3_Z) continue;
        const T *msg = reinterpret_cast<const T *>(a.msg_ws) + (size_t)blk * a.nbe * QC_ROW;
        const float step = FMT == 2 ? c.step[blk] : 1.0f;
        const unsigned p = tid;
        if ((sw >> 21) & 1) {
            const int info = a.vg_info[start];
            const int t = (info >> 24) & 255;
            var_generic<FMT, VMAX>(a, lam + t * ATSC3_Z, tot + t * QC_ROW, msg, p,
                                   (info >> 12) & 2047, info & 63, (info >> 6) & 63, step);
            continue;
        }
        switch (G * 64 + D) {
#define PH_V1(D)                                                                       \
    case 64 + D:                                                                       \
        if constexpr (D <= VMAX) var_multi<FMT, D, 1, TA>(a, lam, tot, msg, p, start, step); \
        break;
#define PH_V2(D)                                                   \
    case 128 + D:                                                  \
        var_multi<FMT, D, 2, TA>(a, lam, tot, msg, p, start, step); \
        break;
#define PH_V4(D)                                                   \
    case 256 + D:                                                  \
        var_multi<FMT, D, 4, TA>(a, lam, tot, msg, p, start, step); \
        break;
            QC_CASES(PH_V1)
            QC_CASES8(PH_V2)
            QC_CASES4(PH_V4)
#undef PH_V1
#undef PH_V2
#undef PH_V4
            default: break;
        }
    }
}

// Next active list: the blocks of this iteration's list that did not retire
// (syndrome non-zero and it < iters), in the same order; their counters are
// reset for the next iteration.  One workgroup, 1024 threads.
__global__ __launch_bounds__(1024) void phase_compact(PhaseCtl c, int it, int iters) {
    __shared__ int wsum[16];
    __shared__ int base_s;
    const int sel = it & 1;
    const int nact = c.n_active[sel];
    const int *act = active_list(c, sel);
    int *nxt = active_list(c, sel ^ 1);
    const int tid = threadIdx.x;
    if (tid == 0) base_s = 0;
    __syncthreads();
    for (int k0 = 0; k0 < nact; k0 += 1024) {
        const int k = k0 + tid;
        int blk = -1, keep = 0;
        if (k < nact) {
            blk = act[k];
            keep = (c.unsat[blk] != 0u && it < iters) ? 1 : 0;
        }
        // exclusive prefix sum of keep over the workgroup
        int x = keep;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if ((tid & 63) >= off) x += y;
        }
        if ((tid & 63) == 63) wsum[tid >> 6] = x;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < (tid >> 6); ++w) before += wsum[w];
        const int pos = base_s + before + x - keep;
        if (keep) {
            nxt[pos] = blk;
            c.unsat[blk] = 0u;
        }
        __syncthreads();
        if (tid == 1023) base_s = pos + keep;
        __syncthreads();
    }
    if (tid == 0) c.n_active[sel ^ 1] = base_s;
}

template <int FMT, bool TA, int VMAX>
static int32_t run_phases(const QcArgs &a, const PhaseCtl &c, int nblk, hipStream_t st) {
    hipLaunchKernelGGL((phase_init<FMT>), dim3(nblk), dim3(ATSC3_WG), 0, st, a, c, nblk);
    for (int it = 0; it <= a.iters; ++it) {
        hipLaunchKernelGGL((phase_check<FMT, 16>), dim3(PHASE_GRID), dim3(ATSC3_WG), 0, st, a,
                           c, it);
        hipLaunchKernelGGL((phase_var<FMT, TA, VMAX>), dim3(PHASE_GRID), dim3(ATSC3_WG), 0, st,
                           a, c, it);
        hipLaunchKernelGGL(phase_compact, dim3(1), dim3(1024), 0, st, c, it, a.iters);
    }
    return launch_status();
}

template <int FMT>
static int32_t run_fmt(const QcArgs &a, const PhaseCtl &c, int nblk, int type_a, int max_vdeg,
                       hipStream_t st) {
    if (type_a)
        return max_vdeg <= 16 ? run_phases<FMT, true, 16>(a, c, nblk, st)
                              : run_phases<FMT, true, 32>(a, c, nblk, st);
    return max_vdeg <= 16 ? run_phases<FMT, false, 16>(a, c, nblk, st)
                          : run_phases<FMT, false, 32>(a, c, nblk, st);
}

}  // namespace atsc3

extern "C" int32_t atsc3_minsum_qc_ctl_words(int32_t nblk) {
    if (nblk < 0) return ATSC3_E_ARG;
    return 4 * nblk + 2;
}

extern "C" int32_t atsc3_minsum_qc_phased_dev(
    const float *llr, int32_t nblk, int32_t n, int32_t ncg, int32_t nvg, int32_t nbe,
    int32_t type_a, const int32_t *cg_info, const int32_t *be_off, const int32_t *be_miss,
    const int32_t *vg_info, const int32_t *ve_off, const int32_t *ve_miss,
    const int32_t *c_steps, int32_t nsc, const int32_t *v_steps, int32_t nsv, int32_t max_cdeg,
    int32_t max_vdeg, int32_t iters, float alpha, int32_t msg_fmt, const float *msg_scale,
    float *tot_ws, void *msg_ws, int32_t *ctl_ws, uint8_t *bits, uint8_t *conv, int32_t *its,
    int32_t *bad, void *stream) {
    if (nblk < 0 || iters < 0) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !cg_info || !be_off || !be_miss || !vg_info || !ve_off || !ve_miss ||
        !c_steps || !v_steps || !tot_ws || !msg_ws || !ctl_ws || !bits || !conv || !its || !bad)
        return ATSC3_E_ARG;
    if (n != nvg * ATSC3_Z || ncg <= 0 || nvg <= 0 || nvg > 255 || nbe <= 0 || nbe >= 2048)
        return ATSC3_E_ARG;
    if (nsc <= 0 || nsc > ncg || nsv <= 0 || nsv > nvg) return ATSC3_E_ARG;
    if (max_vdeg < 1 || max_cdeg < 1) return ATSC3_E_ARG;
    if (max_vdeg > 32) return ATSC3_E_DEGREE;
    if (msg_fmt < 0 || msg_fmt > 2) return ATSC3_E_ARG;
    using CI = const QC_CONST int *;
    
    atsc3::QcArgs a{llr,         n,        ncg,         nvg,       nbe,          (CI)cg_info,
                    (CI)be_off, (CI)be_miss, (CI)vg_info, (CI)ve_off, (CI)ve_miss, (CI)c_steps,
                    (CI)v_steps, nsc,      nsv,         iters,     alpha,        msg_scale,
                    tot_ws,      msg_ws,   bits,        conv,      its,          bad};
    atsc3::PhaseCtl c{ctl_ws, ctl_ws + nblk, ctl_ws + 2 * nblk,
                      reinterpret_cast<unsigned *>(ctl_ws + 2 * nblk + 2),
                      reinterpret_cast<float *>(ctl_ws + 3 * nblk + 2)};
    hipStream_t st = atsc3::as_stream(stream);
    switch (msg_fmt) {
        case 0: return atsc3::run_fmt<0>(a, c, nblk, type_a, max_vdeg, st);
        case 1: return atsc3::run_fmt<1>(a, c, nblk, type_a, max_vdeg, st);
        default: return atsc3::run_fmt<2>(a, c, nblk, type_a, max_vdeg, st);
    }
}
