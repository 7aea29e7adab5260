// Device code of the batched persistent min-sum engine, instantiated once
// per message format by ldpc_batch_f32/_f16/_i8.hip (BATCH_FMT, BATCH_NAME).
#pragma once
#include "ldpc_batch.h"

#define BATCH_SPIN_LIMIT (1u << 22)

namespace atsc3 {
namespace {

// Loads of words the kernel itself writes: relaxed atomic loads are never
// turned into scalar-cache loads, and after the barrier's acquire fence they
// read what the other workgroups published.
__device__ __forceinline__ int ldw(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ QcArgs qc_of(const QC_CONST CodeTab *ct, float alpha) {
    QcArgs a{};
    a.n = ct->n;
    a.ncg = ct->ncg;
    a.nvg = ct->nvg;
    a.nbe = ct->nbe;
    a.cg_info = ct->cg_info;
    a.be_off = ct->be_off;
    a.be_miss = ct->be_miss;
    a.vg_info = ct->vg_info;
    a.ve_off = ct->ve_off;
    a.ve_miss = ct->ve_miss;
    a.c_steps = ct->c_steps;
    a.v_steps = ct->v_steps;
    a.nsc = ct->nsc;
    a.nsv = ct->nsv;
    a.alpha = alpha;
    return a;
}

// Exclusive workgroup scan of two ints; totals returned to every thread.
// Uses 2 * (ATSC3_WG / 64) ints of LDS and two barriers.
__device__ inline void wg_scan2(int x0, int x1, int &e0, int &e1, int &t0, int &t1, int *lds) {
    int s0 = x0, s1 = x1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const int y0 = __shfl_up(s0, off, 64), y1 = __shfl_up(s1, off, 64);
        if (lane >= off) {
            s0 += y0;
            s1 += y1;
        }
    }
    if (lane == 63) {
        lds[wave] = s0;
        lds[ATSC3_WG / 64 + wave] = s1;
    }
    __syncthreads();
    int b0 = 0, b1 = 0;
    t0 = 0;
    t1 = 0;
#pragma unroll
    for (int w = 0; w < ATSC3_WG / 64; ++w) {
        const int v0 = lds[w], v1 = lds[ATSC3_WG / 64 + w];
        if (w < wave) {
            b0 += v0;
            b1 += v1;
        }
        t0 += v0;
        t1 += v1;
    }
    e0 = b0 + s0 - x0;
    e1 = b1 + s1 - x1;
    __syncthreads();
}

// The workgroups that share one round's lists: the whole grid (grid-wide
// mode) or the workgroups of one XCD (XCD-partitioned mode).  h: the group's
// control region; rank / size: this workgroup's place in the group; k / ng:
// the group's index and the number of groups (its queue: parts k, k + ng, ...).
struct Grp {
    int *h;
    int rank, size, k, ng;
};

// Unsatisfied-check counts are summed with agent-scope atomics; every other
// access to them is agent-scope too, so all of them see one order per word
// whichever L2 the workgroup's XCD has.
__device__ __forceinline__ int unsat_load(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void unsat_clear(int *p) {
    __hip_atomic_store(p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void set_abort(const BatchArgs &b) {
    __hip_atomic_store(b.ctl + H_ABORT, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The group's queue: its length, and the codeword at a position of it (-1
// when the host's order or part table points outside the batch).
__device__ int queue_len(const BatchArgs &b, const Grp &G) {
    if (!b.order) return b.nblk;
    int len = 0;
    for (int c = G.k; c < b.nparts; c += G.ng) len += b.part_off[c + 1] - b.part_off[c];
    return len;
}
__device__ int queue_blk(const BatchArgs &b, const Grp &G, int pos) {
    if (!b.order) return pos;
    for (int c = G.k; c < b.nparts; c += G.ng) {
        const int lo = b.part_off[c], len = b.part_off[c + 1] - lo;
        if (pos < len) {
            if (lo + pos < 0 || lo + pos >= b.nblk) return -1;
            const int blk = b.order[lo + pos];
            return blk >= 0 && blk < b.nblk ? blk : -1;
        }
        pos -= len;
    }
    return -1;
}

__device__ __forceinline__ void put4(int *p, int x, int y, int z, int w) {
    p[0] = x;
    p[1] = y;
    p[2] = z;
    p[3] = w;
}

// Next round's lists of group G (run by one whole workgroup).  With n_chk =
// 0 it only admits codewords (the initial fill).
__device__ void compact(const BatchArgs &b, const Grp &G) {
    __shared__ int lds[4 * (ATSC3_WG / 64)];
    int *h = G.h;
    int *var_list = h + b.var_list, *var_pre = h + b.var_pre;
    int *chk_list = h + b.chk_list, *chk_pre = h + b.chk_pre;
    int *unsat = b.ctl + b.unsat, *fl = h + b.free_list, *pl = h + b.pend_list;
    const int tid = threadIdx.x;
    const int n = ldw(h + H_NCHK), n_pend = ldw(h + H_NPEND), qn = ldw(h + H_QNEXT);
    int n_free = ldw(h + H_NFREE);
    // slots retired last round have been read by their RETIRE pass: free them
    for (int k = tid; k < n_pend; k += ATSC3_WG) fl[n_free + k] = ldw(pl + k);
    n_free += n_pend;
    __syncthreads();
    int n_keep = 0, n_ret = 0, bv = 0, bc = 0;
    int *best = b.ctl + b.best;
    for (int k0 = 0; k0 < n; k0 += ATSC3_WG) {
        const int k = k0 + tid;
        const bool valid = k < n;
        int slot = 0, blk = 0, cz = 0, u = 0, code = 0, bit = 0, rung = 0;
        bool ret = false, more = false, take = false;
        if (valid) {
            slot = ldw(chk_list + 4 * k);
            blk = ldw(chk_list + 4 * k + 1);
            cz = ldw(chk_list + 4 * k + 2);
            code = cz & 255;
            rung = (cz >> 10) & 3;
            bit = cz >> 12;
            u = unsat_load(unsat + slot);
            ret = u == 0 || bit >= b.iters;
            if (ret) {
                // a rung's result is taken when it converged or leaves fewer
                // unsatisfied checks than every earlier rung; a failed rung
                // hands the codeword to the next one in the same slot
                take = u == 0 || u < ldw(best + blk);
                if (take) best[blk] = u;
                more = u != 0 && rung + 1 < b.nrungs;
            }
        }
        const bool keep = valid && (!ret || more);
        const bool fin = valid && ret && !more;
        int ek, er, tk, tr, ev, ec, tv, tc;
        wg_scan2(keep ? 1 : 0, fin ? 1 : 0, ek, er, tk, tr, lds);
        wg_scan2(valid ? b.codes[code].nsv : 0, keep ? b.codes[code].nsc : 0, ev, ec, tv, tc,
                 lds);
        if (valid) {
            const int mode = !ret ? M_CONT : (more ? M_RETIRE_INIT : M_RETIRE);
            put4(var_list + 4 * k, slot, blk,
                 code | (mode << 8) | (rung << 10) | (bit << 12), u | ((take ? 1 : 0) << 30));
            var_pre[k] = bv + ev;
            if (keep) {
                const int q = n_keep + ek;
                const int nz = ret ? code | ((rung + 1) << 10) : code | (rung << 10) | ((bit + 1) << 12);
                put4(chk_list + 4 * q, slot, blk, nz, 0);
                chk_pre[q] = bc + ec;
                unsat_clear(unsat + slot);
            } else {
                pl[n_ret + er] = slot;
            }
        }
        n_keep += tk;
        n_ret += tr;
        bv += tv;
        bc += tc;
        __syncthreads();
    }
    // admit queued codewords into free slots (most recently freed first)
    const int qlen = queue_len(b, G);
    const int rem = qlen - qn;
    const int adm = n_free < rem ? n_free : rem;
    for (int i0 = 0; i0 < adm; i0 += ATSC3_WG) {
        const int i = i0 + tid;
        bool valid = i < adm;
        int slot = 0, blk = 0, code = 0;
        if (valid) {
            slot = ldw(fl + n_free - 1 - i);
            blk = queue_blk(b, G, qn + i);
            if (blk < 0) {
                set_abort(b);
                valid = false;
                blk = 0;
            }
            code = b.blk_code[blk];
        }
        int ev, ec, tv, tc;
        wg_scan2(valid ? b.codes[code].nsv : 0, valid ? b.codes[code].nsc : 0, ev, ec, tv, tc,
                 lds);
        if (valid) {
            put4(var_list + 4 * (n + i), slot, blk, code | (M_INIT << 8), 0);
            var_pre[n + i] = bv + ev;
            put4(chk_list + 4 * (n_keep + i), slot, blk, code, 0);
            chk_pre[n_keep + i] = bc + ec;
            unsat_clear(unsat + slot);
        }
        bv += tv;
        bc += tc;
        __syncthreads();
    }
    if (tid == 0) {
        var_pre[n + adm] = bv;
        chk_pre[n_keep + adm] = bc;
        h[H_NVAR] = n + adm;
        h[H_NCHK] = n_keep + adm;
        h[H_NFREE] = n_free - adm;
        h[H_NPEND] = n_ret;
        h[H_QNEXT] = qn + adm;
        h[H_DONE] = (n_keep + adm == 0 && qn + adm == qlen) ? 1 : 0;
        const int rounds = ldw(h + H_ROUNDS) + 1;
        h[H_ROUNDS] = rounds;
        // the most rounds of any group, where the host reads them
        if (h != b.ctl)
            __hip_atomic_fetch_max(b.ctl + H_ROUNDS, rounds, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        // every codeword leaves its slot within iters + 3 rounds of admission,
        // so this bound is never reached by a correct schedule: it only
        // guarantees that the kernel ends
        if (rounds > b.max_rounds) set_abort(b);
    }
}

// Wait until this wave's vector memory operations (stores and atomics
// included) have completed at the L2 (gfx9 counts stores in vmcnt):
// s_waitcnt vmcnt(0), expcnt and lgkmcnt left unconstrained.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// The XCD (accelerator complex die) this wave runs on: HW_REG_XCC_ID
// (hardware register 20), bits [3:0].  Each XCD has its own L2.
__device__ __forceinline__ unsigned xcc_id() {
    return __builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11)) & (BATCH_BAR_GROUPS - 1);
}

// Where this workgroup arrives in the grid barrier: its XCD's group, the
// group's size and the number of non-empty groups, from a census taken once
// at launch (every workgroup counts itself on its XCD, then a counted
// one-time barrier).  False when the census wait timed out.
struct BarGroup {
    unsigned g, gsize, ng, rank, k;   // rank: order of arrival in the XCD's count;
                                      // k: the XCD's index among the non-empty ones
};

__device__ bool bar_census(const BatchArgs &b, BarGroup &bg, int *s_flag) {
    unsigned *h = reinterpret_cast<unsigned *>(b.ctl);
    unsigned *bar = reinterpret_cast<unsigned *>(b.ctl + b.bar);
    __shared__ unsigned s_cnt[BATCH_BAR_GROUPS];
    __shared__ unsigned s_rank;
    const unsigned g = xcc_id();
    if (threadIdx.x == 0) {
        s_rank = __hip_atomic_fetch_add(bar + BAR_CENSUS + g, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned old = __hip_atomic_fetch_add(bar + BAR_CENSUS_ARRIVE, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1u == gridDim.x) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_store(bar + BAR_CENSUS_GO, 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(bar + BAR_CENSUS_GO, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                if (__hip_atomic_load(h + H_ABORT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
                if (++spins >= BATCH_SPIN_LIMIT) {
                    __hip_atomic_store(h + H_ABORT, 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        for (int q = 0; q < BATCH_BAR_GROUPS; ++q)
            s_cnt[q] = __hip_atomic_load(bar + BAR_CENSUS + q, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        s_flag[1] = __hip_atomic_load(h + H_ABORT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    unsigned ng = 0, k = 0;
    for (int q = 0; q < BATCH_BAR_GROUPS; ++q) {
        ng += s_cnt[q] ? 1u : 0u;
        k += s_cnt[q] && (unsigned)q < g ? 1u : 0u;
    }
    bg.g = (unsigned)uni((int)g);
    bg.gsize = (unsigned)uni((int)s_cnt[g]);
    bg.ng = (unsigned)uni((int)ng);
    bg.rank = (unsigned)uni((int)s_rank);
    bg.k = (unsigned)uni((int)k);
    const bool ok = s_flag[1] == 0;
    __syncthreads();
    return ok;
}

// Grid barrier number `epoch` (1, 2, ...); the last workgroup to arrive runs
// the compaction first when asked.  Returns false when the abort word is set.
//
// Visibility across XCDs (each has its own L2): every wave first waits for
// its own stores to complete at its XCD's L2, then one wave per workgroup
// arrives in the XCD's group count.  The last workgroup of an XCD to arrive
// issues the agent-scope release -- an L2 write-back, which covers the stores
// of every workgroup of that XCD, all of them complete in that L2 before
// their arrivals -- and counts the XCD in the global word; the last XCD's
// last workgroup acquires, runs the compaction, releases and bumps the
// generation.  After the generation changes ONE wave per workgroup issues the
// agent-scope acquire (L2/L1 invalidation, the workgroup's waves share the
// CU's L1).  One write-back per XCD and one invalidation per workgroup: a
// write-back per workgroup made a one-codeword round 20% slower.  Counts on
// separate 128-byte lines (a flat counter serialised 512 same-line atomics);
// waiters poll the generation word on a line of its own.
__device__ bool grid_barrier(const BatchArgs &b, const Grp &G, const BarGroup &bg,
                             unsigned epoch, bool do_compact, int *s_flag) {
    unsigned *h = reinterpret_cast<unsigned *>(b.ctl);
    unsigned *bar = reinterpret_cast<unsigned *>(b.ctl + b.bar);
    wait_vmem();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned old = __hip_atomic_fetch_add(bar + BAR_GROUP + 32 * bg.g, 1u,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int last = 0;
        if (old + 1u == epoch * bg.gsize) {
            // every workgroup of this XCD has arrived: write its L2 back once
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            const unsigned o2 = __hip_atomic_fetch_add(bar + BAR_ARRIVE, 1u, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
            last = o2 + 1u == epoch * bg.ng;
            if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        s_flag[0] = last;
    }
    __syncthreads();
    if (s_flag[0]) {
        if (do_compact) compact(b, G);
        wait_vmem();
        __syncthreads();
        if (threadIdx.x == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(bar + BAR_GEN, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (threadIdx.x == 0) {
        unsigned spins = 0;
        while (__hip_atomic_load(bar + BAR_GEN, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            if (__hip_atomic_load(h + H_ABORT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (++spins >= BATCH_SPIN_LIMIT) {
                __hip_atomic_store(h + H_ABORT, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(8);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (threadIdx.x == 0)
        s_flag[1] = __hip_atomic_load(h + H_ABORT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return s_flag[1] == 0;
}

// Barrier number `epoch` of one XCD group (XCD-partitioned mode): the
// group's workgroups share one L2, so what they publish to each other needs
// no L2 write-back.  Every wave waits for its stores to complete in that L2,
// one wave per workgroup arrives in the group's count; the last to arrive
// drops its CU's L1 (agent-scope acquire), runs the compaction when asked,
// waits for its stores and bumps the group's generation; every other
// workgroup polls the generation and then drops its L1.  A bounded wait, as
// in the grid barrier.  Returns false when the abort word is set.
__device__ bool xcd_barrier(const BatchArgs &b, const Grp &G, const BarGroup &bg,
                            unsigned epoch, bool do_compact, int *s_flag) {
    unsigned *ab = reinterpret_cast<unsigned *>(b.ctl + H_ABORT);
    unsigned *bar = reinterpret_cast<unsigned *>(b.ctl + b.bar);
    unsigned *cnt = bar + BAR_GROUP + 32 * bg.g, *gen = bar + BAR_XGEN + 32 * bg.g;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    wait_vmem();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned old = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        const int last = old + 1u == epoch * bg.gsize;
        if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_flag[0] = last;
    }
    __syncthreads();
    if (s_flag[0]) {
        if (do_compact) compact(b, G);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        wait_vmem();
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(gen, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
        unsigned spins = 0;
        while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            if (__hip_atomic_load(ab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (++spins >= BATCH_SPIN_LIMIT) {
                __hip_atomic_store(ab, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (threadIdx.x == 0)
        s_flag[1] = __hip_atomic_load(ab, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return s_flag[1] == 0;
}

// This workgroup's contiguous item range of a list; k0 = the entry holding
// item i0 (found by the whole workgroup).  Returns false when empty.
__device__ bool my_range(const Grp &G, const int *pre, int n, int &i0, int &i1, int &k0,
                         int *s_k) {
    const int total = uni(ldw(pre + n));
    const int chunk = (total + G.size - 1) / G.size;
    i0 = G.rank * chunk;
    i1 = i0 + chunk < total ? i0 + chunk : total;
    if (i0 >= i1) return false;
    if (threadIdx.x == 0) s_k[0] = 0;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += ATSC3_WG)
        if (ldw(pre + k) <= i0 && i0 < ldw(pre + k + 1)) s_k[0] = k;
    __syncthreads();
    k0 = uni(s_k[0]);
    __syncthreads();
    return true;
}

__device__ __forceinline__ void flush_unsat_slot(int *unsat, int slot, unsigned cnt) {
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(reinterpret_cast<unsigned *>(unsat + slot), cnt);
}

// check_phase and var_phase are inlined into each kernel variant (a shared
// out-of-line copy would be a real call and multiply the register use).
template <int FMT, int CMAX>
__device__ __forceinline__ void check_phase(const BatchArgs &b, const Grp &G, int *s_k) {
    using T = typename Msg<FMT>::T;
    const int n = uni(ldw(G.h + H_NCHK));
    const int *list = G.h + b.chk_list, *pre = G.h + b.chk_pre;
    int i0, i1, k;
    if (!my_range(G, pre, n, i0, i1, k, s_k)) return;
    const unsigned j = threadIdx.x;
    for (int i = i0; i < i1; ++k) {
        const int slot = uni(ldw(list + 4 * k)), blk = uni(ldw(list + 4 * k + 1));
        const int cz = uni(ldw(list + 4 * k + 2));
        const int p0 = uni(ldw(pre + k)), p1 = uni(ldw(pre + k + 1));
        const int end = p1 < i1 ? p1 : i1;
        const QC_CONST CodeTab *ct = b.codes + (cz & 255);
        const QcArgs a = qc_of(ct, b.ladder[(cz >> 10) & 3]);
        const int bit = cz >> 12;
        const bool first = bit == 0, update = bit < b.iters;
        const float *tot = b.tot_ws + (size_t)slot * b.tot_stride;
        T *msg = reinterpret_cast<T *>(b.msg_ws) + (size_t)slot * b.msg_stride;
        const float step =
            FMT == 2 ? __builtin_bit_cast(float, uni(ldw(b.ctl + b.steps + blk))) : 1.0f;
        const float inv_step = FMT == 2 ? 1.0f / step : 1.0f;
        unsigned cnt = 0u;
        for (; i < end; ++i) {
            if (j >= ATSC3_Z) continue;
            const int sw = uni(a.c_steps[i - p0]);
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
#define BT_C1(D)                                                                        \
    case 64 + D:                                                                        \
        if constexpr (D <= CMAX)                                                        \
            cnt += check_multi<FMT, D, 1>(a, tot, msg, j, start, first, update, step, \
                                             inv_step);                                 \
        break;
#define BT_C2(D)                                                                          \
    case 128 + D:                                                                         \
        cnt += check_multi<FMT, D, 2>(a, tot, msg, j, start, first, update, step, inv_step); \
        break;
                QC_CASES16(BT_C1)
                QC_CASES8(BT_C2)
#undef BT_C1
#undef BT_C2
                default: break;
            }
        }
        flush_unsat_slot(b.ctl + b.unsat, slot, cnt);
    }
}

template <int FMT, bool TA, int VMAX>
__device__ __forceinline__ void var_step(const QcArgs &a, const float *lam, float *tot,
                                         const typename Msg<FMT>::T *msg, unsigned p, int sw,
                                         float step) {
    const int D = sw & 63;
    const int G = (sw >> 6) & 7;
    const int start = (sw >> 9) & 4095;
    if ((sw >> 21) & 1) {
        const int info = a.vg_info[start];
        const int t = (info >> 24) & 255;
        var_generic<FMT, VMAX>(a, lam + t * ATSC3_Z, tot + t * QC_ROW, msg, p,
                               (info >> 12) & 2047, info & 63, (info >> 6) & 63, step);
        return;
    }
    switch (G * 64 + D) {
#define BT_V1(D)                                                                       \
    case 64 + D:                                                                       \
        if constexpr (D <= VMAX) var_multi<FMT, D, 1, TA>(a, lam, tot, msg, p, start, step); \
        break;
#define BT_V2(D)                                                   \
    case 128 + D:                                                  \
        var_multi<FMT, D, 2, TA>(a, lam, tot, msg, p, start, step); \
        break;
#define BT_V4(D)                                                   \
    case 256 + D:                                                  \
        var_multi<FMT, D, 4, TA>(a, lam, tot, msg, p, start, step); \
        break;
        QC_CASES(BT_V1)
        QC_CASES8(BT_V2)
        QC_CASES4(BT_V4)
#undef BT_V1
#undef BT_V2
#undef BT_V4
        default: break;
    }
}

template <int FMT, int VMAX>
__device__ __forceinline__ void var_phase(const BatchArgs &b, const Grp &G, int *s_k) {
    using T = typename Msg<FMT>::T;
    const int n = uni(ldw(G.h + H_NVAR));
    const int *list = G.h + b.var_list, *pre = G.h + b.var_pre;
    int i0, i1, k;
    if (!my_range(G, pre, n, i0, i1, k, s_k)) return;
    const unsigned p = threadIdx.x;
    for (int i = i0; i < i1; ++k) {
        const int slot = uni(ldw(list + 4 * k)), blk = uni(ldw(list + 4 * k + 1));
        const int cz = uni(ldw(list + 4 * k + 2)), u = uni(ldw(list + 4 * k + 3));
        const int p0 = uni(ldw(pre + k)), p1 = uni(ldw(pre + k + 1));
        const int end = p1 < i1 ? p1 : i1;
        const QC_CONST CodeTab *ct = b.codes + (cz & 255);
        const QcArgs a = qc_of(ct, b.alpha);
        const int mode = (cz >> 8) & 3, bit = cz >> 12;
        const long long off = b.blk_off[blk];
        const float *lam = b.llr + off;
        float *tot = b.tot_ws + (size_t)slot * b.tot_stride;
        if (mode == M_RETIRE || mode == M_RETIRE_INIT) {
            // the rung's result, when taken: bits, conv, its, bad; a failed
            // rung with a next one then restarts the slot (tot = lam) -- each
            // lane reads its position's tot before overwriting it
            const bool take = (u >> 30) & 1;
            const int ucnt = u & ((1 << 30) - 1);
            if (take && i == p0 && p == 0) {
                b.conv[blk] = (uint8_t)(ucnt == 0);
                b.its[blk] = bit;
                b.bad[blk] = ucnt;
            }
            uint8_t *bout = b.bits + off;
            for (; i < end; ++i) {
                const int sw = uni(a.v_steps[i - p0]);
                const int G = (sw >> 6) & 7, start = (sw >> 9) & 4095;
                if (p < ATSC3_Z)
                    for (int gi = 0; gi < G; ++gi) {
                        const int t = (a.vg_info[start + gi] >> 24) & 255;
                        if (take) bout[t * ATSC3_Z + p] = (uint8_t)(tot[t * QC_ROW + p] < 0.0f);
                        if (mode == M_RETIRE_INIT) tot[t * QC_ROW + p] = lam[t * ATSC3_Z + p];
                    }
            }
            continue;
        }
        if (mode == M_INIT) {
            for (; i < end; ++i) {
                const int sw = uni(a.v_steps[i - p0]);
                const int G = (sw >> 6) & 7, start = (sw >> 9) & 4095;
                if (p < ATSC3_Z)
                    for (int gi = 0; gi < G; ++gi) {
                        const int t = (a.vg_info[start + gi] >> 24) & 255;
                        tot[t * QC_ROW + p] = lam[t * ATSC3_Z + p];
                    }
            }
            continue;
        }
        const T *msg = reinterpret_cast<const T *>(b.msg_ws) + (size_t)slot * b.msg_stride;
        const float step =
            FMT == 2 ? __builtin_bit_cast(float, uni(ldw(b.ctl + b.steps + blk))) : 1.0f;
        if (ct->type_a) {
            for (; i < end; ++i)
                if (p < ATSC3_Z) var_step<FMT, true, VMAX>(a, lam, tot, msg, p, uni(a.v_steps[i - p0]), step);
        } else {
            for (; i < end; ++i)
                if (p < ATSC3_Z) var_step<FMT, false, VMAX>(a, lam, tot, msg, p, uni(a.v_steps[i - p0]), step);
        }
    }
}

// VMAX: largest variable degree served by the unrolled variable steps (16
// when every code of the batch allows it: fewer registers; else 32).  At most
// 128 VGPRs (4 waves per SIMD): two workgroups per CU stay co-resident.
template <int FMT, int VMAX>
__global__ __launch_bounds__(ATSC3_WG) __attribute__((amdgpu_waves_per_eu(4)))
void batch_kernel(BatchArgs b) {
    __shared__ int s_k[1];
    __shared__ int s_flag[2];
    BarGroup bg;
    if (!bar_census(b, bg, s_flag)) return;
    const Grp G{b.ctl, (int)blockIdx.x, (int)gridDim.x, 0, 1};
    unsigned epoch = 0;
    for (;;) {
        var_phase<FMT, VMAX>(b, G, s_k);
        if (!grid_barrier(b, G, bg, ++epoch, false, s_flag)) return;
        if (uni(ldw(G.h + H_DONE))) return;
        check_phase<FMT, 16>(b, G, s_k);
        if (!grid_barrier(b, G, bg, ++epoch, true, s_flag)) return;
    }
}

// XCD-partitioned mode: the workgroups of each XCD decode their own queue
// (parts k, k + ng, ... of the host's order) in their own slots
// (slots / ng of them) with their own lists and barriers; no wait ever
// spans two XCDs after the launch census.  The group's first workgroup
// fills its region (every slot free, first admissions) before the group's
// first barrier.
template <int FMT, int VMAX>
__global__ __launch_bounds__(ATSC3_WG) __attribute__((amdgpu_waves_per_eu(4)))
void batch_kernel_xcd(BatchArgs b) {
    __shared__ int s_k[1];
    __shared__ int s_flag[2];
    BarGroup bg;
    if (!bar_census(b, bg, s_flag)) return;
    const Grp G{b.ctl + b.region + (int)bg.k * b.region_words, (int)bg.rank, (int)bg.gsize,
                (int)bg.k, (int)bg.ng};
    if (G.rank == 0) {
        const int nslot = b.slots / G.ng, slot0 = G.k * nslot;
        for (int k = threadIdx.x; k < H_WORDS; k += ATSC3_WG) G.h[k] = 0;
        for (int k = threadIdx.x; k < nslot; k += ATSC3_WG)
            G.h[b.free_list + k] = slot0 + nslot - 1 - k;
        if (threadIdx.x == 0) {
            G.h[H_NFREE] = nslot;
            // a group with codewords but no slot could never finish
            if (nslot == 0 && queue_len(b, G) > 0) set_abort(b);
        }
        __syncthreads();
        compact(b, G);
    }
    unsigned epoch = 0;
    if (!xcd_barrier(b, G, bg, ++epoch, false, s_flag)) return;
    for (;;) {
        var_phase<FMT, VMAX>(b, G, s_k);
        if (!xcd_barrier(b, G, bg, ++epoch, false, s_flag)) return;
        if (uni(ldw(G.h + H_DONE))) return;
        check_phase<FMT, 16>(b, G, s_k);
        if (!xcd_barrier(b, G, bg, ++epoch, true, s_flag)) return;
    }
}

// Control words reset, every slot free, first admissions (grid-wide mode);
// the XCD-partitioned mode resets the shared words only (its groups fill
// their regions themselves).
__global__ __launch_bounds__(ATSC3_WG) void batch_init(BatchArgs b) {
    int *h = b.ctl;
    const bool xcd = b.order != nullptr;
    for (int k = threadIdx.x; k < H_WORDS; k += ATSC3_WG) h[k] = 0;
    for (int k = threadIdx.x; k < BAR_WORDS; k += ATSC3_WG) b.ctl[b.bar + k] = 0;
    if (!xcd)
        for (int k = threadIdx.x; k < b.slots; k += ATSC3_WG)
            b.ctl[b.free_list + k] = b.slots - 1 - k;
    for (int k = threadIdx.x; k < b.nblk; k += ATSC3_WG) {
        b.ctl[b.best + k] = 0x7fffffff;
        // "not decoded" until the codeword's result is written: a launch
        // that ends on the abort word leaves its = -1 on every codeword it
        // did not finish, so the outputs themselves carry the failure
        b.conv[k] = 0;
        b.its[k] = -1;
        b.bad[k] = -1;
    }
    if (xcd) return;
    if (threadIdx.x == 0) h[H_NFREE] = b.slots;
    __syncthreads();
    compact(b, Grp{b.ctl, 0, 1, 0, 1});
}

// INT8 message step per codeword (given, or 8 * mean|llr| / 127).
__global__ __launch_bounds__(ATSC3_WG) void batch_steps(BatchArgs b) {
    __shared__ int s_hist[256];
    const int blk = blockIdx.x;
    const QC_CONST CodeTab *ct = b.codes + b.blk_code[blk];
    const float s = b.msg_scale ? b.msg_scale[blk]
                                : int8_step_of_block(b.llr + b.blk_off[blk], ct->n, s_hist);
    if (threadIdx.x == 0) b.ctl[b.steps + blk] = __builtin_bit_cast(int, s);
}

// Co-resident workgroups of one kernel variant on the current device.  The
// occupancy query assumes the waves of the resident workgroups pack perfectly
// over the CU's 4 SIMDs; a 6-wave workgroup can place 2 waves on one SIMD, so
// only floor(waves_per_simd / 2) workgroups per CU are certain to be resident
// together (with 3 waves per SIMD the query reports 2 but 1 is guaranteed;
// a grid that is not co-resident makes every barrier wait time out).
template <int FMT, int VMAX, bool XCD>
static int capacity_of() {
    static int cap[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    int &c = cap[dev];
    if (c == 0) {
        int per_cu = 0, cus = 0;
        const void *fn = XCD ? reinterpret_cast<const void *>(batch_kernel_xcd<FMT, VMAX>)
                             : reinterpret_cast<const void *>(batch_kernel<FMT, VMAX>);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, ATSC3_WG, 0) !=
            hipSuccess)
            return 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess)
            return 0;
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, fn) != hipSuccess) return 0;
        const int vgprs = fa.numRegs > 0 ? ((fa.numRegs + 7) / 8) * 8 : 512;
        int waves_per_simd = 512 / vgprs;
        if (waves_per_simd > 8) waves_per_simd = 8;
        const int wg_waves = ATSC3_WG / 64;
        const int per_simd_per_wg = (wg_waves + 3) / 4;
        int safe = waves_per_simd / per_simd_per_wg;
        if (safe < 1) safe = 1;
        if (safe < per_cu) per_cu = safe;
        c = per_cu * cus;
    }
    return c;
}

template <int FMT, int VMAX, bool XCD>
static int32_t run_batch(const BatchArgs &b, int grid, hipStream_t st) {
    if (FMT == 2)
        hipLaunchKernelGGL(batch_steps, dim3(b.nblk), dim3(ATSC3_WG), 0, st, b);
    hipLaunchKernelGGL(batch_init, dim3(1), dim3(ATSC3_WG), 0, st, b);
    if (XCD)
        hipLaunchKernelGGL((batch_kernel_xcd<FMT, VMAX>), dim3(grid), dim3(ATSC3_WG), 0, st, b);
    else
        hipLaunchKernelGGL((batch_kernel<FMT, VMAX>), dim3(grid), dim3(ATSC3_WG), 0, st, b);
    return launch_status();
}


}  // namespace

int32_t BATCH_LAUNCH(const BatchArgs &b, int grid, hipStream_t st, bool v16, bool xcd) {
    if (xcd)
        return v16 ? run_batch<BATCH_FMT, 16, true>(b, grid, st)
                   : run_batch<BATCH_FMT, 32, true>(b, grid, st);
    return v16 ? run_batch<BATCH_FMT, 16, false>(b, grid, st)
               : run_batch<BATCH_FMT, 32, false>(b, grid, st);
}

// The smaller of the two register variants' capacities, so a grid valid for
// a format is valid whatever the codes' degrees.
int BATCH_CAPACITY(bool xcd) {
    const int a = xcd ? capacity_of<BATCH_FMT, 16, true>() : capacity_of<BATCH_FMT, 16, false>();
    const int c = xcd ? capacity_of<BATCH_FMT, 32, true>() : capacity_of<BATCH_FMT, 32, false>();
    return a < c ? a : c;
}

}  // namespace atsc3
