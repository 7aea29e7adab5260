// Batched normalized min-sum: ONE persistent launch decodes a queue of
// codewords of mixed A/322 codes (64800- and 16200-bit, any rate, any mix).
//
// Slots.  `slots` decode slots each hold one codeword's working state: its
// tot rows (nvg x 384 floats) and its check-to-variable messages (nbe x 384
// in the storage format).  The number of slots bounds the working set (the
// autotuner sizes it against the last-level cache); codewords wait in a
// device queue (the input order) until a slot frees up.
//
// Rounds.  The kernel alternates two grid-wide phases:
//   VAR    every (entry, variable step) item of the var list.  An entry is
//          INIT   (a newly admitted codeword: tot = lam),
//          CONT   (ordered variable sums tot = lam + s), or
//          RETIRE (hard bits and conv/its/bad of a codeword whose syndrome
//                  read zero or that reached `iters`);
//   CHECK  every (entry, check step) item of the check list: the syndrome of
//          the step's checks (summed per slot with atomics) and, unless the
//          codeword is at iteration `iters`, the message update.
// Grid barriers separate the phases.  The last workgroup to arrive at the
// barrier after CHECK builds the next lists (compaction): entries whose
// syndrome is zero or that reached `iters` retire, the others continue one
// iteration further; slots freed by the previous round's retirements (whose
// RETIRE pass has completed) are refilled from the queue.  Every codeword thus
// runs exactly the iteration sequence of the per-codeword kernels
// (ldpc_qc.hip) -- the per-check and per-variable arithmetic is the same code
// (ldpc_qc_common.h) -- so results are bit-identical to them and to
// lab/ldpc_kernel.c in FP32 mode.
//
// Co-residency.  Grid barriers need every workgroup resident at once: the
// host entry rejects a grid larger than the occupancy-derived capacity.  Each
// wait is bounded: a workgroup that waits beyond the limit sets the abort
// word and every workgroup leaves at its next barrier, so the kernel always
// ends (the host reads the abort word to report the failure).
#include <cstdlib>

#include "ldpc_batch.h"

extern "C" int32_t atsc3_minsum_batch_ctl_words(int32_t nblk, int32_t slots) {
    if (nblk < 0 || slots < 1) return ATSC3_E_ARG;
    return atsc3::CtlLayout(nblk, slots).words;
}

extern "C" int32_t atsc3_minsum_batch_max_grid(int32_t msg_fmt) {
    switch (msg_fmt) {
        case 0: return atsc3::batch_capacity_f32();
        case 1: return atsc3::batch_capacity_f16();
        case 2: return atsc3::batch_capacity_i8();
        default: return ATSC3_E_ARG;
    }
}

extern "C" int32_t atsc3_minsum_batch_status_word(void) { return atsc3::H_ABORT; }

extern "C" int32_t atsc3_minsum_batch_dev(
    const float *llr, int32_t nblk, const int64_t *blk_off, const int32_t *blk_code,
    const void *codes, int32_t ncode, int32_t max_nvg, int32_t max_nbe, int32_t max_vdeg,
    int32_t iters, float alpha, const float *ladder, int32_t nrungs, int32_t msg_fmt,
    const float *msg_scale, int32_t slots, int32_t grid,
    float *tot_ws, void *msg_ws, int32_t *ctl_ws, uint8_t *bits, uint8_t *conv, int32_t *its,
    int32_t *bad, void *stream) {
    if (nblk < 0 || iters < 0 || iters >= (1 << 19)) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !blk_off || !blk_code || !codes || !tot_ws || !msg_ws || !ctl_ws || !bits ||
        !conv || !its || !bad)
        return ATSC3_E_ARG;
    if (ncode < 1 || ncode > 256 || max_nvg < 1 || max_nvg > 255 || max_nbe < 1 ||
        max_nbe >= 2048 || slots < 1 || max_vdeg < 1 || max_vdeg > 32)
        return ATSC3_E_ARG;
    if (msg_fmt < 0 || msg_fmt > 2) return ATSC3_E_ARG;
    if (ladder && (nrungs < 1 || nrungs > BATCH_MAX_RUNGS)) return ATSC3_E_ARG;
    const int cap = atsc3_minsum_batch_max_grid(msg_fmt);
    if (cap <= 0) return ATSC3_E_HIP;
    if (grid < 1 || grid > cap) return ATSC3_E_SIZE;
    atsc3::CtlLayout L(nblk, slots);
    atsc3::BatchArgs b{};
    b.llr = llr;
    b.blk_off = (const QC_CONST long long *)blk_off;
    b.blk_code = (const QC_CONST int *)blk_code;
    b.codes = (const QC_CONST atsc3::CodeTab *)codes;
    b.nblk = nblk;
    b.slots = slots;
    b.iters = iters;
    b.nrungs = ladder ? nrungs : 1;
    for (int r = 0; r < BATCH_MAX_RUNGS; ++r) b.ladder[r] = ladder && r < nrungs ? ladder[r] : alpha;
    {
        const long long per = (long long)b.nrungs * (iters + 2) + 1;
        const long long mr = ((long long)nblk + 1) * per + 8;
        b.max_rounds = (int)(mr < (1LL << 30) ? mr : (1LL << 30));
        // test hook: a lower round limit forces the abort path from the host
        // (the kernel then leaves through its normal abort exit)
        if (const char *e = getenv("ATSC3_BATCH_MAX_ROUNDS")) {
            const int v = atoi(e);
            if (v > 0 && v < b.max_rounds) b.max_rounds = v;
        }
    }
    b.alpha = b.ladder[0];
    b.msg_scale = msg_scale;
    b.tot_ws = tot_ws;
    b.msg_ws = msg_ws;
    b.tot_stride = (long long)max_nvg * QC_ROW;
    b.msg_stride = (long long)max_nbe * QC_ROW;
    b.ctl = ctl_ws;
    b.var_list = L.var_list;
    b.var_pre = L.var_pre;
    b.chk_list = L.chk_list;
    b.chk_pre = L.chk_pre;
    b.unsat = L.unsat;
    b.free_list = L.free_list;
    b.pend_list = L.pend_list;
    b.steps = L.steps;
    b.best = L.best;
    b.bar = L.bar;
    b.bits = bits;
    b.conv = conv;
    b.its = its;
    b.bad = bad;
    hipStream_t st = atsc3::as_stream(stream);
    const bool v16 = max_vdeg <= 16;
    switch (msg_fmt) {
        case 0: return atsc3::batch_launch_f32(b, grid, st, v16);
        case 1: return atsc3::batch_launch_f16(b, grid, st, v16);
        default: return atsc3::batch_launch_i8(b, grid, st, v16);
    }
}
