// Shared definitions of the batched persistent min-sum engine (control
// layout and kernel arguments); see ldpc_batch.hip for the schedule.
#pragma once
#include "ldpc_qc_common.h"

namespace atsc3 {

// Mirrors struct atsc3_code_tab of the public header (same field order and
// sizes); pointers are device addresses of the code's kernel tables.
struct CodeTab {
    int n, ncg, nvg, nbe, nsc, nsv, type_a, max_vdeg;
    const QC_CONST int *cg_info, *be_off, *be_miss, *vg_info, *ve_off, *ve_miss, *c_steps,
        *v_steps;
};
static_assert(sizeof(CodeTab) == 96, "code table descriptor layout");

enum { H_ABORT = 2, H_QNEXT, H_NVAR, H_NCHK, H_NFREE, H_NPEND, H_DONE, H_ROUNDS, H_WORDS = 16 };
// Grid-barrier words, each on its own 128-byte line (32 ints) of the barrier
// region: the global arrival count, the generation, one arrival count per
// XCD (a workgroup's group is the XCD it runs on, read from the hardware),
// then the launch-time census: its arrival count, its release flag and the
// number of workgroups found on each XCD.
// The XCD-partitioned mode adds one generation word per XCD group.
#define BATCH_BAR_GROUPS 16            // XCC_ID is a 4-bit field
enum { BAR_ARRIVE = 0, BAR_GEN = 32, BAR_GROUP = 64,
       BAR_CENSUS_ARRIVE = 64 + 32 * BATCH_BAR_GROUPS, BAR_CENSUS_GO = BAR_CENSUS_ARRIVE + 32,
       BAR_CENSUS = BAR_CENSUS_GO + 32, BAR_XGEN = BAR_CENSUS + 32,
       BAR_WORDS = BAR_XGEN + 32 * BATCH_BAR_GROUPS };
// Var-list entry modes.  RETIRE_INIT: the codeword failed at its ladder rung
// and starts the next one in the same slot (result kept first if taken).
enum { M_CONT = 0, M_RETIRE = 1, M_INIT = 2, M_RETIRE_INIT = 3 };
#define BATCH_MAX_RUNGS 4

__host__ __device__ inline int round4(int x) { return (x + 3) & ~3; }

// Control words.  unsat, steps, best and bar are offsets from the start of
// the control buffer; the round's lists (var_list .. pend_list) and the
// H_NVAR .. H_ROUNDS words are offsets from a group's region.  The grid-wide
// mode has one region, the start of the buffer (region_words 0); the
// XCD-partitioned mode one region of region_words per XCD group, from
// `region` on, and its H_ABORT / H_ROUNDS words at the start of the buffer.
struct CtlLayout {
    int var_list, var_pre, chk_list, chk_pre, unsat, free_list, pend_list, steps, best, bar,
        region, region_words, words;
    __host__ __device__ CtlLayout(int nblk, int slots, bool xcd = false) {
        if (!xcd) {
            var_list = H_WORDS;
            var_pre = var_list + 4 * slots;
            chk_list = var_pre + round4(slots + 1);
            chk_pre = chk_list + 4 * slots;
            unsat = chk_pre + round4(slots + 1);
            free_list = unsat + round4(slots);
            pend_list = free_list + round4(slots);
            steps = pend_list + round4(slots);
            best = steps + round4(nblk);
            bar = (best + round4(nblk) + 31) & ~31;
            words = bar + BAR_WORDS;
            region = 0;
            region_words = 0;
            return;
        }
        unsat = H_WORDS;
        steps = unsat + round4(slots);
        best = steps + round4(nblk);
        bar = (best + round4(nblk) + 31) & ~31;
        region = bar + BAR_WORDS;
        var_list = H_WORDS;
        var_pre = var_list + 4 * slots;
        chk_list = var_pre + round4(slots + 1);
        chk_pre = chk_list + 4 * slots;
        free_list = chk_pre + round4(slots + 1);
        pend_list = free_list + round4(slots);
        region_words = (pend_list + round4(slots) + 31) & ~31;
        words = region + BATCH_BAR_GROUPS * region_words;
    }
};

struct BatchArgs {
    const float *llr;
    const QC_CONST long long *blk_off;
    const QC_CONST int *blk_code;
    const QC_CONST CodeTab *codes;
    int nblk, slots, iters, max_rounds;
    float alpha;
    float ladder[BATCH_MAX_RUNGS];   // normalization per ladder rung
    int nrungs;
    const float *msg_scale;
    float *tot_ws;
    void *msg_ws;
    long long tot_stride, msg_stride;   // elements per slot
    int *ctl;
    int var_list, var_pre, chk_list, chk_pre, unsat, free_list, pend_list, steps, best, bar;
    int region, region_words;
    // XCD-partitioned mode: the queue is cut into nparts parts, part c being
    // the codewords order[part_off[c] .. part_off[c + 1]); XCD group k of ng
    // decodes parts k, k + ng, ... in that order.  order == nullptr: one
    // queue, the input order.
    const QC_CONST int *order, *part_off;
    int nparts;
    uint8_t *bits, *conv;
    int *its, *bad;
};


// Per-format launchers and capacities, one translation unit per message
// format (ldpc_batch_f32/_f16/_i8.hip) so the variants build in parallel.
// xcd: the XCD-partitioned kernel (b.order set), else the grid-wide one.
int32_t batch_launch_f32(const BatchArgs &b, int grid, hipStream_t st, bool v16, bool xcd);
int32_t batch_launch_f16(const BatchArgs &b, int grid, hipStream_t st, bool v16, bool xcd);
int32_t batch_launch_i8(const BatchArgs &b, int grid, hipStream_t st, bool v16, bool xcd);
int batch_capacity_f32(bool xcd);
int batch_capacity_f16(bool xcd);
int batch_capacity_i8(bool xcd);

}  // namespace atsc3
