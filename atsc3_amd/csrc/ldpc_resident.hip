// Normalized min-sum with the whole codeword resident on one compute unit:
// FP16 check-to-variable messages, the variable posteriors held in LDS.
//
// One workgroup of 1024 threads decodes one codeword at a time and pulls the
// next from a device queue when it finishes (persistent, no grid barrier:
// workgroups never wait for each other).  Its working state:
//
//   tot   the posterior of every variable, FP16, in LDS (64800 x 2 bytes =
//         129,600 bytes of the 160 KB a gfx950 workgroup may hold).  The
//         check phase reads it with ds_read_u16 -- 64 consecutive positions
//         of one variable group per wave instruction, conflict-free -- so
//         the E-sized posterior gather of the flooding schedule never
//         reaches the memory hierarchy.
//   state one compressed record per check (min-sum messages of a check take
//         only two magnitudes): word 0 the FP16 pair (a1, a2) = alpha *
//         (min1, min2); word 1 the signs of edges 0..24 and the index of the
//         edge that takes a2 (bits 25..31); a 64-bit third word the signs of
//         edges 25..88, touched only by checks of degree > 25.  8 bytes per
//         check in this workgroup's slot in global memory (21,600 checks =
//         173 KB for rate 10/15), against 2 bytes per EDGE for stored FP16
//         messages.  Each record is written twice, at positions j and j + 360
//         of its group's 720-record row, so the variable side reads check
//         position p + u without reducing it mod 360.
//
// Per iteration (plain semantics, lab/ldpc_kernel.c:82-143):
//   CHECK  lane j of a wave takes check (g, j): reads its old record and the
//          posteriors of its edges, counts the syndrome from tot < 0, forms
//          x = tot - m_old, min1/min2/first argmin/signs and writes the new
//          record in place (only its own thread ever reads it back here).
//   (barrier; the syndrome total decides: converged or at `iters` -> bits,
//    conv, its, bad of the codeword, next codeword)
//   VAR    lane p of a wave takes variable (t, p): gathers the new record of
//          each of its checks (coalesced: 64 consecutive checks of one group
//          per wave load), rebuilds the message, sums in FP32 in ascending
//          check index (the reference's order: per region, the entries whose
//          check position wrapped past 359 first, then the rest, each in
//          table order) and stores tot = fp16(lam + s) in LDS.
//
// Arithmetic: every message is fp16_rn(sat(alpha * min)) with the sign of
// the plain rule, exactly as the FP16 message format stores it; every sum is
// FP32 in the reference's order; the posterior is rounded to FP16 (saturated
// to +-65504) where it is stored.  The checker (oracle/resident_ref.py) runs
// the same steps in numpy; with FP32 storage in place of FP16 it reproduces
// the plain FP32 min-sum bit for bit.
#include <hip/hip_fp16.h>

#include "ldpc_qc_common.h"

#define RES_WG 1024
#define RES_WAVES (RES_WG / 64)
#define RES_NMAX 64800
#define RES_MAX_DC 89
#define RES_MAX_DV 32
#define RES_SIGN_LO 25         // sign bits held in the second record word
// Records per check-group row: 720 = each record stored twice (positions j
// and j + 360), so variable-side reads need no mod 360; 360 = stored once.
#ifndef RES_RROW
#define RES_RROW 360
#endif

namespace atsc3 {
namespace {

// Mirrors struct atsc3_res_code of the public header.
struct ResTab {
    int n, ncg, nvg, nbe, max_dc, max_dv, nsc, nsv;
    const QC_CONST int *grp_c;    // per check group: dc | b0 << 7 | has_missing << 18
    const QC_CONST int *edge_c;   // per base edge: s | t << 9 | (missing j + 1) << 18
    const QC_CONST int *grp_v;    // per variable group: dv | e0 << 6 | n0 << 17
    const QC_CONST int *edge_v;   // per entry, two words: u | g << 9 | k << 17; missing p + 1
    const QC_CONST int *c_steps;  // check steps: D | G << 7 | start << 10 | generic << 22
    const QC_CONST int *c_tasks;  // task: group | c6 << 8 (rows c6*64 .. c6*64+63)
    const QC_CONST int *v_steps;  // variable steps, same packing (+ Type A region << 23)
    const QC_CONST int *v_tasks;
};
static_assert(sizeof(ResTab) == 96, "resident code descriptor layout");

struct ResArgs {
    const float *llr;
    const QC_CONST long long *blk_off;
    const QC_CONST int *blk_code;
    const QC_CONST ResTab *codes;
    int nblk, iters;
    float alpha;
    long long stride;            // records per workgroup slot (2 per check)
    unsigned *p0, *p1;
    unsigned long long *p2;
    int *queue;
    uint8_t *bits, *conv;
    int *its, *bad;
};

// FP16 storage of a posterior or a message magnitude: saturated to the
// largest finite half (+-65504), NaN kept (as the FP32 path keeps it).
__device__ __forceinline__ unsigned short h_bits_sat(float v) {
    v = v > 65504.0f ? 65504.0f : (v < -65504.0f ? -65504.0f : v);
    return __half_as_ushort(__float2half_rn(v));
}
__device__ __forceinline__ float h_val(unsigned short h) {
    return __half2float(__ushort_as_half(h));
}

// Bit 31 set iff edge kk's message is negative (record words w1, w2).
__device__ __forceinline__ unsigned rec_sign(unsigned w1, unsigned long long w2, unsigned kk) {
    return (kk < RES_SIGN_LO ? w1 << (31u - kk) : (unsigned)(w2 >> (kk - RES_SIGN_LO)) << 31) &
           0x80000000u;
}

// Message of edge kk of a check from its record words w0 (a1 | a2 << 16),
// w1 (signs 0..24 | a2 edge << 25) and w2 (signs 25..88).  a1, a2 are
// non-negative (alpha >= 0), so setting the sign bit negates.
__device__ __forceinline__ float rec_msg(unsigned w0, unsigned w1, unsigned long long w2,
                                         unsigned kk) {
    const unsigned h = kk == (w1 >> RES_SIGN_LO) ? (w0 >> 16) : w0;
    const float v = h_val((unsigned short)h);
    return __uint_as_float(__float_as_uint(v) | rec_sign(w1, w2, kk));
}

// Record store at row position j and its copy at j + 360.
template <typename T>
__device__ __forceinline__ void rec_put(T *plane, unsigned r, T v) {
    at(plane, r) = v;
    if (RES_RROW == 2 * ATSC3_Z) at(plane, r + (unsigned)ATSC3_Z) = v;
}

// Row position of check position p + u (0 <= p, u < 360) whose wrap flag is
// `wr` (p + u >= 360).
__device__ __forceinline__ unsigned rec_pos(unsigned pu, bool wr) {
    return RES_RROW == 2 * ATSC3_Z ? pu : pu - (wr ? (unsigned)ATSC3_Z : 0u);
}

// G check tasks of degree D <= 16 with every circulant full: the syndrome
// bits of the current posteriors and, unless `update` is false, the new
// compressed records.  Every load of the G tasks (old records, posteriors)
// is issued before the first use.
template <int D, int G>
__device__ __forceinline__ unsigned check_multi(const QC_CONST ResTab *ct, const __half *tot,
                                                unsigned *p0, unsigned *p1, int start,
                                                unsigned lane, bool first, bool update,
                                                float alpha) {
    unsigned rix[G], w0[G], w1[G];
    bool valid[G];
    float tv[G][D];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int task = ct->c_tasks[start + gi];
        const int g = task & 255;
        const unsigned j = (unsigned)(task >> 8) * 64u + lane;
        valid[gi] = j < (unsigned)ATSC3_Z;
        const unsigned jj = valid[gi] ? j : 0u;
        rix[gi] = (unsigned)g * RES_RROW + jj;
        w0[gi] = 0u;
        w1[gi] = 0u;
        if (!first) {
            w0[gi] = at(p0, rix[gi]);
            w1[gi] = at(p1, rix[gi]);
        }
        const int b0 = (ct->grp_c[g] >> 7) & 2047;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const unsigned ew = (unsigned)ct->edge_c[b0 + k];
            const unsigned p = wrap360(jj + (ew & 511u));
            tv[gi][k] = __half2float(tot[((ew >> 9) & 511u) * ATSC3_Z + p]);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    unsigned syn_total = 0u;
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const float a1o = h_val((unsigned short)(w0[gi] & 0xffffu));
        const float a2o = h_val((unsigned short)(w0[gi] >> 16));
        const unsigned idxo = w1[gi] >> RES_SIGN_LO;
        bool syn = false, par = false;
        unsigned idx = 0u, neg = 0u;
        float m1 = INFINITY, m2 = INFINITY;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float ao = (unsigned)k == idxo ? a2o : a1o;
            const float mo =
                __uint_as_float(__float_as_uint(ao) | ((w1[gi] << (31 - k)) & 0x80000000u));
            const float x = tv[gi][k] - mo;
            syn ^= tv[gi][k] < 0.0f;
            const bool ng = x < 0.0f;
            par ^= ng;
            neg |= ng ? 1u << k : 0u;
            const float mag = clamp_mag(x);
            idx = mag < m1 ? (unsigned)k : idx;
            m2 = __builtin_amdgcn_fmed3f(m1, m2, mag);
            m1 = fminf(m1, mag);
        }
        if (update && valid[gi]) {
            const unsigned sg = par ? ~neg : neg;
            rec_put(p0, rix[gi],
                    (unsigned)h_bits_sat(m1 * alpha) | ((unsigned)h_bits_sat(m2 * alpha) << 16));
            rec_put(p1, rix[gi], (sg & ((1u << RES_SIGN_LO) - 1u)) | (idx << RES_SIGN_LO));
        }
        syn_total += (valid[gi] && syn) ? 1u : 0u;
    }
    return syn_total;
}

// Any check degree up to RES_MAX_DC (records of degree > 25 use the third
// word): the same steps as check_multi with a runtime edge loop.
__device__ unsigned check_generic(const QC_CONST ResTab *ct, const __half *tot, unsigned *p0,
                                  unsigned *p1, unsigned long long *p2, int g, int b0, int dc,
                                  bool hasmiss, unsigned j, bool valid, bool first, bool update,
                                  float alpha) {
    const unsigned jj = valid ? j : 0u;
    const unsigned rix = (unsigned)g * RES_RROW + jj;
    unsigned w0 = 0u, w1 = 0u;
    unsigned long long w2 = 0ull;
    if (!first) {
        w0 = at(p0, rix);
        w1 = at(p1, rix);
        if (dc > RES_SIGN_LO) w2 = at(p2, rix);
    }
    const float a1o = h_val((unsigned short)(w0 & 0xffffu));
    const float a2o = h_val((unsigned short)(w0 >> 16));
    const unsigned idxo = w1 >> RES_SIGN_LO;
    unsigned syn = 0u, par = 0u, idx = 0u, neg_lo = 0u;
    unsigned long long neg_hi = 0ull;
    float m1 = INFINITY, m2 = INFINITY;
    for (int k = 0; k < dc; ++k) {
        const unsigned ew = (unsigned)ct->edge_c[b0 + k];
        if (hasmiss && ((ew >> 18) & 1023u) == jj + 1u) continue;
        const unsigned p = wrap360(jj + (ew & 511u));
        const float tv = __half2float(tot[((ew >> 9) & 511u) * ATSC3_Z + p]);
        const float ao = (unsigned)k == idxo ? a2o : a1o;
        const float mo = __uint_as_float(__float_as_uint(ao) | rec_sign(w1, w2, (unsigned)k));
        const float x = tv - mo;
        syn ^= (unsigned)(tv < 0.0f);
        const unsigned ng = (unsigned)(x < 0.0f);
        par ^= ng;
        if (k < RES_SIGN_LO)
            neg_lo |= ng << k;
        else
            neg_hi |= (unsigned long long)ng << (k - RES_SIGN_LO);
        const float mag = clamp_mag(x);
        if (mag < m1) {
            m2 = m1;
            m1 = mag;
            idx = (unsigned)k;
        } else {
            m2 = fminf(m2, mag);
        }
    }
    if (update && valid) {
        const unsigned sg_lo = par ? ~neg_lo : neg_lo;
        rec_put(p0, rix,
                (unsigned)h_bits_sat(m1 * alpha) | ((unsigned)h_bits_sat(m2 * alpha) << 16));
        rec_put(p1, rix, (sg_lo & ((1u << RES_SIGN_LO) - 1u)) | (idx << RES_SIGN_LO));
        if (dc > RES_SIGN_LO) rec_put(p2, rix, par ? ~neg_hi : neg_hi);
    }
    return valid ? syn : 0u;
}

// G variable tasks of degree D <= 16 whose entries are all full circulants
// and all name check edges below 25 (record word 1 holds their signs): every
// record and LLR load of the G tasks first, then the sums in ascending check
// index.  Within a region the entries are sorted by u, so the ones whose
// check position wrapped past 359 (p + u >= 360) are a suffix of the
// region's list; the order is that suffix, then the prefix.  A skipped
// message is added as +0.0, which leaves the sum unchanged (it starts at
// +0.0 and can never become -0.0).  TA: the group has a first region
// (Type A codes, n0 > 0 entries); else one region.  With one region of two
// entries the order does not matter: (0 + a) + b == (0 + b) + a in IEEE
// arithmetic, signed zeros included.
template <int D, int G, bool TA>
__device__ __forceinline__ void var_multi(const QC_CONST ResTab *ct, __half *tot,
                                          const float *lam, const unsigned *p0,
                                          const unsigned *p1, int start, unsigned lane) {
    unsigned w0[G][D], w1[G][D], dst[G];
    bool wr[G][D];
    float lv[G];
    bool valid[G];
    int e0[G], n0[G];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int task = ct->v_tasks[start + gi];
        const int t = task & 255;
        const unsigned p = (unsigned)(task >> 8) * 64u + lane;
        valid[gi] = p < (unsigned)ATSC3_Z;
        const unsigned pp = valid[gi] ? p : 0u;
        dst[gi] = (unsigned)t * ATSC3_Z + pp;
        const int vw = ct->grp_v[t];
        e0[gi] = (vw >> 6) & 2047;
        n0[gi] = TA ? (vw >> 17) & 63 : 0;
        lv[gi] = at(lam, dst[gi]);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const unsigned ew = (unsigned)ct->edge_v[2 * (e0[gi] + k)];
            const unsigned u = ew & 511u, g = (ew >> 9) & 255u;
            wr[gi][k] = pp + u >= (unsigned)ATSC3_Z;
            const unsigned r = g * RES_RROW + rec_pos(pp + u, wr[gi][k]);
            w0[gi][k] = at(p0, r);
            w1[gi][k] = at(p1, r);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        float m[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const unsigned kk = ((unsigned)ct->edge_v[2 * (e0[gi] + k)] >> 17) & 127u;
            const unsigned h = kk == (w1[gi][k] >> RES_SIGN_LO) ? (w0[gi][k] >> 16) : w0[gi][k];
            m[k] = __uint_as_float(__float_as_uint(h_val((unsigned short)h)) |
                                   ((w1[gi][k] << (31u - kk)) & 0x80000000u));
        }
        float s = 0.0f;
        if (!TA && D == 2) {
            s = (s + m[0]) + m[1];
        } else {
            if (TA) {
#pragma unroll
                for (int k = 0; k < D; ++k)
                    if (k < n0[gi]) s += wr[gi][k] ? m[k] : 0.0f;
#pragma unroll
                for (int k = 0; k < D; ++k)
                    if (k < n0[gi]) s += wr[gi][k] ? 0.0f : m[k];
            }
#pragma unroll
            for (int k = 0; k < D; ++k)
                if (k >= n0[gi]) s += wr[gi][k] ? m[k] : 0.0f;
#pragma unroll
            for (int k = 0; k < D; ++k)
                if (k >= n0[gi]) s += wr[gi][k] ? 0.0f : m[k];
        }
        if (valid[gi]) tot[dst[gi]] = __ushort_as_half(h_bits_sat(lv[gi] + s));
    }
}

// Any degree: one record load per entry, visited in the summation order
// directly (suffix then prefix of each region).
__device__ void var_generic(const QC_CONST ResTab *ct, __half *tot, const float *lam,
                            const unsigned *p0, const unsigned *p1,
                            const unsigned long long *p2, int t, int e0, int dv, int n0,
                            unsigned p, bool valid) {
    const unsigned pp = valid ? p : 0u;
    const float lv = lam[(unsigned)t * ATSC3_Z + pp];
    float s = 0.0f;
    for (int r = 0; r < 2; ++r) {
        const int lo = r == 0 ? 0 : n0, hi = r == 0 ? n0 : dv;
        // first wrapped entry of the region (u ascending: a suffix)
        int tr = hi;
        while (tr > lo &&
               pp + ((unsigned)ct->edge_v[2 * (e0 + tr - 1)] & 511u) >= (unsigned)ATSC3_Z)
            --tr;
        for (int q = 0; q < hi - lo; ++q) {
            int k = tr + q;
            if (k >= hi) k -= hi - lo;
            const unsigned ew = (unsigned)ct->edge_v[2 * (e0 + k)];
            const unsigned vm = (unsigned)ct->edge_v[2 * (e0 + k) + 1];
            const unsigned u = ew & 511u, g = (ew >> 9) & 255u, kk = (ew >> 17) & 127u;
            const unsigned rix = g * RES_RROW + rec_pos(pp + u, pp + u >= (unsigned)ATSC3_Z);
            const unsigned long long w2 = kk >= RES_SIGN_LO ? at(p2, rix) : 0ull;
            const float mv = rec_msg(at(p0, rix), at(p1, rix), w2, kk);
            if (vm != pp + 1u) s += mv;
        }
    }
    if (valid) tot[(unsigned)t * ATSC3_Z + p] = __ushort_as_half(h_bits_sat(lv + s));
}

__global__ void resident_init(int *queue) {
    if (threadIdx.x == 0) *queue = 0;
}

__global__ __launch_bounds__(RES_WG) void resident_kernel(ResArgs a) {
    __shared__ __half tot[RES_NMAX];
    __shared__ int s_blk;
    __shared__ int s_unsat[2];
    const unsigned tid = threadIdx.x, lane = tid & 63u;
    const int wave = __builtin_amdgcn_readfirstlane((int)(tid >> 6));
    unsigned *p0 = a.p0 + (size_t)blockIdx.x * a.stride;
    unsigned *p1 = a.p1 + (size_t)blockIdx.x * a.stride;
    unsigned long long *p2 = a.p2 + (size_t)blockIdx.x * a.stride;
    for (;;) {
        if (tid == 0) {
            s_blk = atomicAdd(a.queue, 1);
            s_unsat[0] = 0;
            s_unsat[1] = 0;
        }
        __syncthreads();
        const int blk = __builtin_amdgcn_readfirstlane(s_blk);
        if (blk >= a.nblk) return;
        const QC_CONST ResTab *ct = &a.codes[a.blk_code[blk]];
        const int n = ct->n;
        const long long off = a.blk_off[blk];
        const float *lam = a.llr + off;
        for (int v = (int)tid; v < n; v += RES_WG)
            tot[v] = __ushort_as_half(h_bits_sat(lam[v]));
        __syncthreads();
        for (int it = 0;; ++it) {
            if (tid == 0) s_unsat[(it + 1) & 1] = 0;
            const bool first = it == 0, update = it < a.iters;
            unsigned unsat = 0u;
            for (int si = wave; si < ct->nsc; si += RES_WAVES) {
                const int sw = ct->c_steps[si];
                const int D = sw & 127, G = (sw >> 7) & 7, start = (sw >> 10) & 4095;
                if ((sw >> 22) & 1) {
                    const int task = ct->c_tasks[start];
                    const int g = task & 255;
                    const unsigned j = (unsigned)(task >> 8) * 64u + lane;
                    const int gw = ct->grp_c[g];
                    unsat += check_generic(ct, tot, p0, p1, p2, g, (gw >> 7) & 2047, gw & 127,
                                           (gw >> 18) & 1, j, j < (unsigned)ATSC3_Z, first,
                                           update, a.alpha);
                    continue;
                }
                switch (G * 128 + D) {
#define RES_C1(DD)                                                                        \
    case 128 + DD:                                                                        \
        unsat += check_multi<DD, 1>(ct, tot, p0, p1, start, lane, first, update, a.alpha); \
        break;
#define RES_C2(DD)                                                                        \
    case 256 + DD:                                                                        \
        unsat += check_multi<DD, 2>(ct, tot, p0, p1, start, lane, first, update, a.alpha); \
        break;
                    QC_CASES16(RES_C1)
                    QC_CASES16(RES_C2)
#undef RES_C1
#undef RES_C2
                    default: break;
                }
            }
            if (unsat) atomicAdd(&s_unsat[it & 1], (int)unsat);
            __syncthreads();
            const int total = s_unsat[it & 1];
            if (total == 0 || it == a.iters) {
                uint8_t *bout = a.bits + off;
                for (int v = (int)tid; v < n; v += RES_WG)
                    bout[v] = (uint8_t)(__half2float(tot[v]) < 0.0f);
                if (tid == 0) {
                    a.conv[blk] = (uint8_t)(total == 0);
                    a.its[blk] = total == 0 ? it : a.iters;
                    a.bad[blk] = total;
                }
                break;
            }
            for (int si = wave; si < ct->nsv; si += RES_WAVES) {
                const int sw = ct->v_steps[si];
                const int D = sw & 127, G = (sw >> 7) & 7, start = (sw >> 10) & 4095;
                if ((sw >> 22) & 1) {
                    const int task = ct->v_tasks[start];
                    const int t = task & 255;
                    const unsigned p = (unsigned)(task >> 8) * 64u + lane;
                    const int vw = ct->grp_v[t];
                    var_generic(ct, tot, lam, p0, p1, p2, t, (vw >> 6) & 2047, vw & 63,
                                (vw >> 17) & 63, p, p < (unsigned)ATSC3_Z);
                    continue;
                }
                const bool ta = (sw >> 23) & 1;
                switch ((ta ? 1024 : 0) + G * 128 + D) {
#define RES_V1(DD)                                                      \
    case 128 + DD:                                                      \
        var_multi<DD, 1, false>(ct, tot, lam, p0, p1, start, lane);     \
        break;                                                          \
    case 1024 + 128 + DD:                                               \
        var_multi<DD, 1, true>(ct, tot, lam, p0, p1, start, lane);      \
        break;
#define RES_V2(DD)                                                      \
    case 256 + DD:                                                      \
        var_multi<DD, 2, false>(ct, tot, lam, p0, p1, start, lane);     \
        break;                                                          \
    case 1024 + 256 + DD:                                               \
        var_multi<DD, 2, true>(ct, tot, lam, p0, p1, start, lane);      \
        break;
#define RES_V4(DD)                                                      \
    case 512 + DD:                                                      \
        var_multi<DD, 4, false>(ct, tot, lam, p0, p1, start, lane);     \
        break;                                                          \
    case 1024 + 512 + DD:                                               \
        var_multi<DD, 4, true>(ct, tot, lam, p0, p1, start, lane);      \
        break;
                    QC_CASES16(RES_V1)
                    QC_CASES8(RES_V2)
                    QC_CASES4(RES_V4)
#undef RES_V1
#undef RES_V2
#undef RES_V4
                    default: break;
                }
            }
            __syncthreads();
        }
        // every thread has read s_unsat and tot before thread 0 starts the
        // next codeword
        __syncthreads();
    }
}

}  // namespace
}  // namespace atsc3

extern "C" int64_t atsc3_minsum_resident_ws_bytes(int32_t grid, int32_t max_checks) {
    if (grid < 1 || max_checks < 1) return ATSC3_E_ARG;
    const int64_t stride = ((int64_t)max_checks * (RES_RROW / ATSC3_Z) + 63) & ~(int64_t)63;
    return 256 + (int64_t)grid * stride * 16;
}

extern "C" int32_t atsc3_minsum_resident_grid(void) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ATSC3_E_HIP;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return ATSC3_E_HIP;
    return cus > 0 ? cus : ATSC3_E_HIP;
}

extern "C" int32_t atsc3_minsum_resident_dev(const float *llr, int32_t nblk,
                                             const int64_t *blk_off, const int32_t *blk_code,
                                             const void *codes, int32_t ncode,
                                             int32_t max_checks, int32_t max_n, int32_t max_dc,
                                             int32_t max_dv, int32_t iters, float alpha,
                                             int32_t grid, void *ws, int64_t ws_bytes,
                                             uint8_t *bits, uint8_t *conv, int32_t *its,
                                             int32_t *bad, void *stream) {
    if (nblk < 0 || iters < 0 || iters >= (1 << 24) || !(alpha >= 0.0f)) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !blk_off || !blk_code || !codes || !ws || !bits || !conv || !its || !bad)
        return ATSC3_E_ARG;
    if (ncode < 1 || max_checks < 1 || grid < 1) return ATSC3_E_ARG;
    if (max_n < 1 || max_n > RES_NMAX) return ATSC3_E_SIZE;
    if (max_dc < 1 || max_dc > RES_MAX_DC || max_dv < 1 || max_dv > RES_MAX_DV)
        return ATSC3_E_DEGREE;
    const int64_t need = atsc3_minsum_resident_ws_bytes(grid, max_checks);
    if (need < 0 || ws_bytes < need) return ATSC3_E_SIZE;
    const int64_t stride = ((int64_t)max_checks * (RES_RROW / ATSC3_Z) + 63) & ~(int64_t)63;
    char *base = static_cast<char *>(ws);
    atsc3::ResArgs a{};
    a.llr = llr;
    a.blk_off = (const QC_CONST long long *)blk_off;
    a.blk_code = (const QC_CONST int *)blk_code;
    a.codes = (const QC_CONST atsc3::ResTab *)codes;
    a.nblk = nblk;
    a.iters = iters;
    a.alpha = alpha;
    a.stride = stride;
    a.queue = reinterpret_cast<int *>(base);
    a.p0 = reinterpret_cast<unsigned *>(base + 256);
    a.p1 = a.p0 + (size_t)grid * stride;
    a.p2 = reinterpret_cast<unsigned long long *>(a.p1 + (size_t)grid * stride);
    a.bits = bits;
    a.conv = conv;
    a.its = its;
    a.bad = bad;
    hipStream_t st = atsc3::as_stream(stream);
    hipLaunchKernelGGL(atsc3::resident_init, dim3(1), dim3(64), 0, st, a.queue);
    if (hipGetLastError() != hipSuccess) return ATSC3_E_HIP;
    hipLaunchKernelGGL(atsc3::resident_kernel, dim3(grid), dim3(RES_WG), 0, st, a);
    return atsc3::launch_status();
}
