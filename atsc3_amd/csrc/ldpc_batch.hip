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
//
// XCD-partitioned mode (atsc3_minsum_batch_xcd_dev).  A grid barrier makes
// every XCD write its L2 back and waits for the slowest workgroup of the
// chip, ~15 us each, two per round: for a frame whose slowest codeword needs
// 50 iterations that is most of the frame's time.  In this mode the host
// cuts the queue into parts and the workgroups of each XCD decode their own
// parts in their own slots, with their own lists and barriers.  The
// workgroups of one XCD share its L2, so a barrier among them needs no L2
// write-back: stores complete in the L2, one agent-scope atomic per
// workgroup arrives, the last one compacts and bumps the group's generation,
// and each waiter drops its CU's L1 before it reads.  The per-codeword
// arithmetic is the same code, so results are bit-identical to the
// grid-wide mode.
#include <cstdlib>

#include "ldpc_batch.h"

namespace {

// Arguments common to both modes; ATSC3_OK or the error to return.
int32_t fill_args(atsc3::BatchArgs &b, const atsc3::CtlLayout &L, const float *llr,
                  int32_t nblk, const int64_t *blk_off, const int32_t *blk_code,
                  const void *codes, int32_t ncode, int32_t max_nvg, int32_t max_nbe,
                  int32_t max_vdeg, int32_t iters, float alpha, const float *ladder,
                  int32_t nrungs, int32_t msg_fmt, const float *msg_scale, int32_t slots,
                  int32_t grid, int cap, float *tot_ws, void *msg_ws, int32_t *ctl_ws,
                  uint8_t *bits, uint8_t *conv, int32_t *its, int32_t *bad) {
    if (!llr || !blk_off || !blk_code || !codes || !tot_ws || !msg_ws || !ctl_ws || !bits ||
        !conv || !its || !bad)
        return ATSC3_E_ARG;
    if (ncode < 1 || ncode > 256 || max_nvg < 1 || max_nvg > 255 || max_nbe < 1 ||
        max_nbe >= 2048 || slots < 1 || max_vdeg < 1 || max_vdeg > 32)
        return ATSC3_E_ARG;
    if (msg_fmt < 0 || msg_fmt > 2) return ATSC3_E_ARG;
    if (ladder && (nrungs < 1 || nrungs > BATCH_MAX_RUNGS)) return ATSC3_E_ARG;
    if (cap <= 0) return ATSC3_E_HIP;
    if (grid < 1 || grid > cap) return ATSC3_E_SIZE;
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
    b.region = L.region;
    b.region_words = L.region_words;
    b.order = nullptr;
    b.part_off = nullptr;
    b.nparts = 0;
    b.bits = bits;
    b.conv = conv;
    b.its = its;
    b.bad = bad;
    return ATSC3_OK;
}

int32_t launch(const atsc3::BatchArgs &b, int32_t msg_fmt, int32_t max_vdeg, int32_t grid,
               void *stream, bool xcd) {
    hipStream_t st = atsc3::as_stream(stream);
    const bool v16 = max_vdeg <= 16;
    switch (msg_fmt) {
        case 0: return atsc3::batch_launch_f32(b, grid, st, v16, xcd);
        case 1: return atsc3::batch_launch_f16(b, grid, st, v16, xcd);
        default: return atsc3::batch_launch_i8(b, grid, st, v16, xcd);
    }
}

int capacity(int32_t msg_fmt, bool xcd) {
    switch (msg_fmt) {
        case 0: return atsc3::batch_capacity_f32(xcd);
        case 1: return atsc3::batch_capacity_f16(xcd);
        case 2: return atsc3::batch_capacity_i8(xcd);
        default: return ATSC3_E_ARG;
    }
}

}  // namespace

extern "C" int32_t atsc3_minsum_batch_ctl_words(int32_t nblk, int32_t slots) {
    if (nblk < 0 || slots < 1) return ATSC3_E_ARG;
    return atsc3::CtlLayout(nblk, slots).words;
}

extern "C" int32_t atsc3_minsum_batch_max_grid(int32_t msg_fmt) { return capacity(msg_fmt, false); }

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
    atsc3::BatchArgs b{};
    const int32_t rc = fill_args(b, atsc3::CtlLayout(nblk, slots > 0 ? slots : 1), llr, nblk,
                                 blk_off, blk_code, codes, ncode, max_nvg, max_nbe, max_vdeg,
                                 iters, alpha, ladder, nrungs, msg_fmt, msg_scale, slots, grid,
                                 capacity(msg_fmt, false), tot_ws, msg_ws, ctl_ws, bits, conv,
                                 its, bad);
    if (rc != ATSC3_OK) return rc;
    return launch(b, msg_fmt, max_vdeg, grid, stream, false);
}

extern "C" int32_t atsc3_minsum_batch_xcd_ctl_words(int32_t nblk, int32_t slots) {
    if (nblk < 0 || slots < 1) return ATSC3_E_ARG;
    return atsc3::CtlLayout(nblk, slots, true).words;
}

extern "C" int32_t atsc3_minsum_batch_xcd_max_grid(int32_t msg_fmt) {
    return capacity(msg_fmt, true);
}

extern "C" int32_t atsc3_minsum_batch_xcd_dev(
    const float *llr, int32_t nblk, const int64_t *blk_off, const int32_t *blk_code,
    const void *codes, int32_t ncode, int32_t max_nvg, int32_t max_nbe, int32_t max_vdeg,
    int32_t iters, float alpha, const float *ladder, int32_t nrungs, int32_t msg_fmt,
    const float *msg_scale, const int32_t *order, const int32_t *part_off, int32_t nparts,
    int32_t slots, int32_t grid, float *tot_ws, void *msg_ws, int32_t *ctl_ws, uint8_t *bits,
    uint8_t *conv, int32_t *its, int32_t *bad, void *stream) {
    if (nblk < 0 || iters < 0 || iters >= (1 << 19)) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!order || !part_off || nparts < 1 || nparts > 4096 || slots < BATCH_BAR_GROUPS)
        return ATSC3_E_ARG;
    atsc3::BatchArgs b{};
    const int32_t rc = fill_args(b, atsc3::CtlLayout(nblk, slots, true), llr, nblk, blk_off,
                                 blk_code, codes, ncode, max_nvg, max_nbe, max_vdeg, iters, alpha,
                                 ladder, nrungs, msg_fmt, msg_scale, slots, grid,
                                 capacity(msg_fmt, true), tot_ws, msg_ws, ctl_ws, bits, conv,
                                 its, bad);
    if (rc != ATSC3_OK) return rc;
    b.order = (const QC_CONST int *)order;
    b.part_off = (const QC_CONST int *)part_off;
    b.nparts = nparts;
    return launch(b, msg_fmt, max_vdeg, grid, stream, true);
}
