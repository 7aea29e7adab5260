// Normalized min-sum with the whole codeword resident on one compute unit.
//
// One workgroup of 1024 threads decodes one codeword at a time and pulls the
// next from a device queue when it finishes (persistent, no grid barrier:
// workgroups never wait for each other).  Two storage modes share the code:
//
//   FP16   every posterior in LDS as FP16 (180 rows of 424 halves, the last
//          64 a copy of the first 64: 152,640 bytes of the 160 KB a gfx950
//          workgroup may hold); check records hold FP16 magnitudes.  The
//          FP16 message form.
//   FP32   the plain decoder, bit-exact with the reference: posteriors FP32,
//          the rows of the highest-degree variable groups in LDS (up to 113
//          rows of 360 floats = 162,720 bytes) and the remaining rows in this
//          workgroup's slot in global memory; records hold FP32 magnitudes.
//          The host chooses the LDS rows (for 64800 r10/15: 82% of the edge
//          reads) and marks each row's place in the tables.
//
// The check phase reads posteriors with ds_read (64 consecutive positions of
// one variable group per wave instruction, conflict-free), so most of the
// E-sized posterior gather of the flooding schedule never leaves the CU.
//
// Check records: min-sum messages of a check take only two magnitudes, so
// each check keeps one record: a1 = alpha * min1 and a2 = alpha * min2 (FP16:
// word 0 = a1 | a2 << 16; FP32: word 0 = a1 and a separate word a2), word 1 =
// the signs of edges 0..23 and the index of the edge that takes a2 (bits
// 24..31, one byte), and a 64-bit word with the signs of edges 24..88,
// touched only by checks of degree > 24.  Words 0 and 1 sit side by side (one
// 8-byte access per record).
// Checks of degree <= 16 without missing entries (check_multi) build their
// sign word by shifting in one bit per edge, so edge k's sign sits at bit
// D - 1 - k; the tables give the variable side each entry's bit.
// Records live in this workgroup's slot in global memory (21,600 checks x 8
// or 12 bytes for rate 10/15, against 2 or 4 bytes per EDGE for stored
// messages); the check phase rewrites them in place (only the owning lane
// reads a record back there) and the variable phase gathers them, 64
// consecutive checks of one group per wave load.  The FP16 variable phase
// rebuilds each message as an FP16 value (the stored half with the edge's
// sign bit rotated in) and adds it with one mixed-precision fma: the FP32
// sum of the FP16 value, rounded once, as a conversion and an add give.
//
// Per iteration (plain semantics, lab/ldpc_kernel.c:82-143):
//   CHECK  lane j of a wave takes check (g, j): reads its old record and the
//          posteriors of its edges, counts the syndrome from tot < 0, forms
//          x = tot - m_old, min1/min2/first argmin/signs and writes the new
//          record.
//   (barrier; the syndrome total decides: converged or at `iters` -> bits,
//    conv, its, bad of the codeword, next codeword)
//   VAR    lane p of a wave takes variable (t, p): gathers the new record of
//          each of its checks, rebuilds the message, sums in FP32 and
//          stores tot = lam + s.  FP32 mode sums in ascending check index
//          (the reference's order: per region, the entries whose check
//          position wrapped past 359 first, then the rest, each in table
//          order); FP16 and INT8 modes sum in table order (no per-lane
//          order selects: they are not bound to the reference's rounding).
//
// Arithmetic: every message is alpha * min with the sign of the plain rule
// (FP16 mode: fp16_rn(sat(alpha * min)), exactly as the FP16 message format
// stores it); every sum is FP32 in the order above; FP16 mode rounds
// the posterior to FP16 (saturated to +-65504) where it is stored.  The
// checker (oracle/resident_ref.py) runs the same steps in numpy; its FP32
// mode reproduces the reference kernel's golden outputs bit for bit.
#include <hip/hip_fp16.h>

#include <mutex>

#include "ldpc_qc_common.h"

#define RES_WG 1024
#define RES_WAVES (RES_WG / 64)
#define RES_NMAX 64800
#define RES_MAX_DC 89
#define RES_MAX_DV 32
#define RES_SIGN_LO 24                  // sign bits held in record word 1
#define RES_F32_LDS_ROWS 113            // FP32 posterior rows that fit in LDS
// FP16 posterior rows in LDS are padded: positions 0..359, then 0..63 again,
// so the 64 positions a wave reads for one edge are contiguous (no mod 360).
#define RES_PROW 424
#define RES_F16_LDS_BYTES ((RES_NMAX / ATSC3_Z) * RES_PROW * 2)
#define RES_F32_LDS_BYTES (RES_F32_LDS_ROWS * ATSC3_Z * 4)

namespace atsc3 {
namespace {

// Mirrors struct atsc3_res_code of the public header.
struct ResTab {
    int n, ncg, nvg, nbe, max_dc, max_dv, nsc, nsv;
    const QC_CONST int *grp_c;    // per check group: dc | b0 << 7 | has_missing << 18
    const QC_CONST int *edge_c;   // per base edge: s | row << 9 | (missing j + 1) << 18
                                  //   | global << 28
    const QC_CONST int *grp_v;    // per variable group: dv | e0 << 6 | n0 << 17 | row << 23
                                  //   | global << 31
    const QC_CONST int *edge_v;   // per entry, two words: u | g << 9 | k << 17 | sign bit << 24;
                                  //   missing p + 1
    const QC_CONST int *c_steps;  // check steps: D | log2 G << 7 | start << 10 | generic << 22
    const QC_CONST int *c_tasks;  // task: group | c6 << 8 (rows c6*64 .. c6*64+63)
    const QC_CONST int *v_steps;  // variable steps, same packing (+ Type A region << 23)
    const QC_CONST int *v_tasks;
    // FP16 / INT8 modes: at b * 6 + c6, the byte offset in LDS of the padded
    // posterior row position that check (g, c6*64) meets on base edge b
    const QC_CONST int *c_off;
    // per variable step: offset into v_sdata (fast steps), where each of the
    // step's G tasks has a head word t | c6 << 8 | n0 << 11 | row << 17 |
    // global << 25 followed by its D entry words (edge_v word 0): the whole
    // step in one scalar fetch after the step word
    const QC_CONST int *v_soff;
    const QC_CONST int *v_sdata;
    // the same for check steps: per task g | c6 << 8, then its D words
    // (FP32: edge_c; FP16 / INT8: the c_off byte offsets for this c6)
    const QC_CONST int *c_soff;
    const QC_CONST int *c_sdata;
};
static_assert(sizeof(ResTab) == 136, "resident code descriptor layout");

struct ResArgs {
    const float *llr;
    const QC_CONST long long *blk_off;
    const QC_CONST int *blk_code;
    const QC_CONST ResTab *codes;
    int nblk, iters;
    float alpha;
    const float *msg_scale;       // INT8 mode: step per codeword, or NULL (median rule)
    long long stride;             // records per workgroup slot
    long long tstride;            // global posterior floats per workgroup slot
    uint2 *p01;                   // record words 0 and 1
    unsigned *p0b;                // FP32 mode's a2 word
    unsigned long long *p2;
    float *totg;
    int *queue;
    uint8_t *bits, *conv;
    int *its, *bad;
    float *post;                  // NULL, or the final posterior of every bit
};

// FP16 storage of a posterior or a message magnitude: saturated to the
// largest finite half (+-65504), NaN kept (as the FP32 path keeps it).
// Rounded first: a value that rounds past 65504 becomes +-inf, which steps
// back one code to +-65504 -- the same half as clamping to +-65504 first.
// v is first made an FP32 value in a register: the compiler would otherwise
// fuse the product or sum that forms it with the conversion into one
// mixed-precision op (one rounding to FP16 instead of FP32 then FP16).
__device__ __forceinline__ unsigned short h_bits_sat(float v) {
    asm volatile("" : "+v"(v));
    const unsigned short h = __half_as_ushort(__float2half_rn(v));
    return (h & 0x7fffu) == 0x7c00u ? (unsigned short)(h - 1u) : h;
}
__device__ __forceinline__ float h_val(unsigned short h) {
    return __half2float(__ushort_as_half(h));
}

// Posterior rows: FP16 mode, row r of the LDS array; FP32 mode, row r of the
// LDS array or (global) of this workgroup's global slot.  Team mode (several
// workgroups on one codeword): the variable phase publishes the rows it
// computes to `pub` (row r at r * 360 elements: FP16 halves, FP32 floats;
// FP32 global rows stay in `glob`), and every member copies the published
// rows into its LDS after the team barrier.
struct Tot {
    unsigned char *lds;
    float *glob;
    unsigned char *pub;
};

template <int M>
__device__ __forceinline__ float tot_ld(const Tot &T, unsigned row, bool glob, unsigned p) {
    if constexpr (M != 0) {
        return __half2float(reinterpret_cast<const __half *>(T.lds)[row * RES_PROW + p]);
    } else {
        // compiled as one generic (flat) load of the selected pointer; a
        // branch per edge instead makes every join wait for all loads in
        // flight, which serialises the step's gathers
        return glob ? at(T.glob, row * ATSC3_Z + p)
                    : reinterpret_cast<const float *>(T.lds)[row * ATSC3_Z + p];
    }
}

template <int M, bool PUB = false>
__device__ __forceinline__ void tot_st(const Tot &T, unsigned row, bool glob, unsigned p,
                                       float v) {
    if constexpr (M != 0) {
        const __half hv = __ushort_as_half(h_bits_sat(v));
        if constexpr (PUB) {
            at(reinterpret_cast<__half *>(T.pub), row * ATSC3_Z + p) = hv;
        } else {
            __half *h = reinterpret_cast<__half *>(T.lds) + row * RES_PROW;
            h[p] = hv;
            if (p < 64u) h[p + ATSC3_Z] = hv;
        }
    } else {
        if (glob)
            at(T.glob, row * ATSC3_Z + p) = v;
        else if constexpr (PUB)
            at(reinterpret_cast<float *>(T.pub), row * ATSC3_Z + p) = v;
        else
            reinterpret_cast<float *>(T.lds)[row * ATSC3_Z + p] = v;
    }
}

// Record planes of this workgroup.
struct Rec {
    uint2 *p01;
    unsigned *p0b;
    unsigned long long *p2;
    float step, inv;              // INT8 mode: the codeword's step and 1 / step
};

// Magnitudes (a1, a2) of a record from its magnitude words: FP32 words
// (w0, w0b), an FP16 pair, or two INT8 levels times the codeword's step.
template <int M>
__device__ __forceinline__ float rec_a1(unsigned w0, unsigned, const Rec &R) {
    if constexpr (M == 0) return __uint_as_float(w0);
    else if constexpr (M == 1) return h_val((unsigned short)(w0 & 0xffffu));
    else return (float)(w0 & 0xffu) * R.step;
}
template <int M>
__device__ __forceinline__ float rec_a2(unsigned w0, unsigned w0b, const Rec &R) {
    if constexpr (M == 0) return __uint_as_float(w0b);
    else if constexpr (M == 1) return h_val((unsigned short)(w0 >> 16));
    else return (float)((w0 >> 8) & 0xffu) * R.step;
}
// INT8 level of a magnitude: rint(a / step) saturated at 127 (the INT8
// message format's codec).
__device__ __forceinline__ unsigned i8_level(float a, float inv) {
    return (unsigned)fminf(rintf(a * inv), 127.0f);
}
// Record r: magnitudes (a1, a2) = alpha * (m1, m2) and word 1.
template <int M>
__device__ __forceinline__ void rec_put(const Rec &R, unsigned r, float a1, float a2,
                                        unsigned w1) {
    unsigned w0;
    if constexpr (M == 0) {
        w0 = __float_as_uint(a1);
        at(R.p0b, r) = __float_as_uint(a2);
    } else if constexpr (M == 1) {
        w0 = (unsigned)h_bits_sat(a1) | ((unsigned)h_bits_sat(a2) << 16);
    } else {
        w0 = i8_level(a1, R.inv) | (i8_level(a2, R.inv) << 8);
    }
    at(R.p01, r) = make_uint2(w0, w1);
}

// neg * 2 + this lane's `bit`, as one v_addc_co_u32 whose carry-in is the
// wave's lane mask of `bit` (the compiler otherwise spends a select, a shift
// and an or per bit).
__device__ __forceinline__ unsigned shl1_add_lane_bit(unsigned neg, bool bit) {
    const unsigned long long m = __builtin_amdgcn_ballot_w64(bit);
    unsigned long long co;
    unsigned r;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(co) : "v"(neg), "s"(m));
    return r;
}

// The wave's count of lanes where `bit` holds (the same in every lane).
__device__ __forceinline__ unsigned wave_count(bool bit) {
    return (unsigned)__builtin_popcountll(__builtin_amdgcn_ballot_w64(bit));
}

// Bit 31 set iff the message whose sign sits at bit sp is negative (record
// words w1, w2; bits 0..24 of w1, then w2).
__device__ __forceinline__ unsigned rec_sign(unsigned w1, unsigned long long w2, unsigned sp) {
    return (sp < RES_SIGN_LO ? w1 << (31u - sp) : (unsigned)(w2 >> (sp - RES_SIGN_LO)) << 31) &
           0x80000000u;
}

// Message of edge kk of a check from its record, its sign at bit sp.  a1,
// a2 are non-negative (alpha >= 0), so setting the sign bit negates.
template <int M>
__device__ __forceinline__ float rec_msg(unsigned w0, unsigned w0b, unsigned w1,
                                         unsigned long long w2, unsigned kk, unsigned sp,
                                         const Rec &R) {
    const bool a2 = kk == (w1 >> RES_SIGN_LO);
    float v;
    if constexpr (M == 0) v = __uint_as_float(a2 ? w0b : w0);
    else if constexpr (M == 1) v = h_val((unsigned short)(a2 ? (w0 >> 16) : w0));
    else v = (float)((a2 ? (w0 >> 8) : w0) & 0xffu) * R.step;
    return __uint_as_float(__float_as_uint(v) | rec_sign(w1, w2, sp));
}

// G check tasks of degree D <= 16 with every circulant full: the syndrome
// bits of the current posteriors and, unless `update` is false, the new
// compressed records.  Every load of the G tasks (old records, posteriors)
// is issued before the first use.
template <int M, int D, int G>
__device__ __forceinline__ unsigned check_multi(const QC_CONST int *sd, const Tot &T,
                                                const Rec &R, unsigned lane, bool first,
                                                bool update, float alpha) {
    unsigned rix[G], w0[G], w0b[G], w1[G];
    bool valid[G];
    float tv[G][D];
    __half th[G][D];                // FP16 / INT8: the posteriors as loaded
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int task = sd[gi * (D + 1)];
        const int g = task & 255;
        const unsigned j = (unsigned)(task >> 8) * 64u + lane;
        valid[gi] = j < (unsigned)ATSC3_Z;
        const unsigned jj = valid[gi] ? j : 0u;
        rix[gi] = (unsigned)g * ATSC3_Z + jj;
        w0[gi] = 0u;
        w0b[gi] = 0u;
        w1[gi] = 0u;
        if (!first) {
            const uint2 w = at(R.p01, rix[gi]);
            w0[gi] = w.x;
            w1[gi] = w.y;
            if (M == 0) w0b[gi] = at(R.p0b, rix[gi]);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            // FP32: the edge_c word; FP16 / INT8: the c_off byte offset
            const unsigned ew = (unsigned)sd[gi * (D + 1) + 1 + k];
            if constexpr (M == 0) {
                const unsigned p = wrap360(jj + (ew & 511u));
                tv[gi][k] = tot_ld<M>(T, (ew >> 9) & 511u, (ew >> 28) & 1u, p);
            } else {
                th[gi][k] = *reinterpret_cast<const __half *>(T.lds + ew + 2u * lane);
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    // FP16 / INT8: x = tot - m as fma(tot, 1, -m), which is exactly the FP32
    // difference (tot * 1 is exact; one rounding), taking tot as FP16: the
    // compiler then emits one mixed-precision fma per edge instead of a
    // conversion and a subtraction (the 1.0 is hidden from constant folding)
    float one = 1.0f, ninf = -INFINITY;     // -inf hidden too: med3(a, b, -inf)
    asm volatile("" : "+s"(one), "+s"(ninf)); // would fold back to a quieting fminf
    unsigned syn_total = 0u;
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        float a1o = rec_a1<M>(w0[gi], w0b[gi], R);
        float a2o = rec_a2<M>(w0[gi], w0b[gi], R);
        // converted once per check: keeps the compiler from converting the
        // selected FP16/INT8 magnitude again for every edge (no instruction)
        asm volatile("" : "+v"(a1o), "+v"(a2o));
        // shifted out once per check: compared against inline constants
        // (a byte-select compare per edge would need each constant in a VGPR)
        unsigned idxo = w1[gi] >> RES_SIGN_LO;
        asm volatile("" : "+v"(idxo));
        bool syn = false, par = false;
        unsigned idx = 0u, neg = 0u;
        float m1 = INFINITY, m2 = INFINITY;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float ao = (unsigned)k == idxo ? a2o : a1o;
            // fast checks keep edge k's sign at bit D - 1 - k
            const float mo = __uint_as_float(__float_as_uint(ao) |
                                             ((w1[gi] << (32 - D + k)) & 0x80000000u));
            float x;
            if constexpr (M == 0) {
                x = tv[gi][k] - mo;
                syn ^= tv[gi][k] < 0.0f;
            } else {
                x = __builtin_fmaf(__half2float(th[gi][k]), one, -mo);
                syn ^= __half2float(th[gi][k]) < 0.0f;
            }
            const bool ng = x < 0.0f;
            par ^= ng;
            neg = shl1_add_lane_bit(neg, ng);       // edge k ends at bit D - 1 - k
            // mag is never NaN (clamp_mag takes NaN to +inf), so min1 is the
            // median of (m1, mag, -inf): med3 needs no NaN-quieting max on
            // its operands, which fminf gets once the compiler sinks this
            // chain into the record-update branch
            const float mag = clamp_mag(x);
            asm volatile("" ::"v"(mag));           // clamp beside x, where x is known canonical
            idx = mag < m1 ? (unsigned)k : idx;
            m2 = __builtin_amdgcn_fmed3f(m1, m2, mag);
            m1 = __builtin_amdgcn_fmed3f(m1, mag, ninf);
        }
        if (update && valid[gi]) {
            const unsigned sg = par ? ~neg : neg;
            // the edge index as it is: folding the shift into the selects
            // would turn their inline constants into literals
            unsigned ix = idx;
            asm volatile("" : "+v"(ix));
            rec_put<M>(R, rix[gi], m1 * alpha, m2 * alpha,
                       (sg & ((1u << RES_SIGN_LO) - 1u)) | (ix << RES_SIGN_LO));
        }
        syn_total += wave_count(valid[gi] && syn);
    }
    return syn_total;
}

// Any check degree up to RES_MAX_DC (records of degree > 24 use the third
// word): the same steps as check_multi with a runtime edge loop.
template <int M>
__device__ unsigned check_generic(const QC_CONST ResTab *ct, const Tot &T, const Rec &R, int g,
                                  int b0, int dc, bool hasmiss, unsigned j, bool valid,
                                  bool first, bool update, float alpha) {
    const unsigned jj = valid ? j : 0u;
    const unsigned rix = (unsigned)g * ATSC3_Z + jj;
    unsigned w0 = 0u, w0b = 0u, w1 = 0u;
    unsigned long long w2 = 0ull;
    if (!first) {
        const uint2 w = at(R.p01, rix);
        w0 = w.x;
        w1 = w.y;
        if (M == 0) w0b = at(R.p0b, rix);
        if (dc > RES_SIGN_LO) w2 = at(R.p2, rix);
    }
    const float a1o = rec_a1<M>(w0, w0b, R);
    const float a2o = rec_a2<M>(w0, w0b, R);
    const unsigned idxo = w1 >> RES_SIGN_LO;
    unsigned syn = 0u, par = 0u, idx = 0u, neg_lo = 0u;
    unsigned long long neg_hi = 0ull;
    float m1 = INFINITY, m2 = INFINITY;
    for (int k = 0; k < dc; ++k) {
        const unsigned ew = (unsigned)ct->edge_c[b0 + k];
        if (hasmiss && ((ew >> 18) & 1023u) == jj + 1u) continue;
        const unsigned p = wrap360(jj + (ew & 511u));
        const float tv = tot_ld<M>(T, (ew >> 9) & 511u, (ew >> 28) & 1u, p);
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
        rec_put<M>(R, rix, m1 * alpha, m2 * alpha,
                   (sg_lo & ((1u << RES_SIGN_LO) - 1u)) | (idx << RES_SIGN_LO));
        if (dc > RES_SIGN_LO) at(R.p2, rix) = par ? ~neg_hi : neg_hi;
    }
    return wave_count(valid && syn != 0u);
}

// G variable tasks of degree D <= 16 whose entries are all full circulants
// and all name check edges below 24 (record word 1 holds their signs): every
// record and LLR load of the G tasks first, then the sums: FP16 / INT8 in
// table order; FP32 in ascending check index.  Within a region the entries are sorted by u, so the ones whose
// check position wrapped past 359 (p + u >= 360) are a suffix of the
// region's list; the order is that suffix, then the prefix.  A skipped
// message is added as +0.0, which leaves the sum unchanged (it starts at
// +0.0 and can never become -0.0).  TA: the group has a first region
// (Type A codes, n0 > 0 entries); else one region.  With one region of two
// entries the order does not matter: (0 + a) + b == (0 + b) + a in IEEE
// arithmetic, signed zeros included.
template <int M, int D, int G, bool TA, bool PUB>
__device__ __forceinline__ void var_multi(const QC_CONST int *sd, const Tot &T,
                                          const float *lam, const Rec &R, unsigned lane) {
    unsigned w0[G][D], w0b[G][D], w1[G][D], dst[G], row[G], pp8[G];
    bool wr[G][D], glob[G];
    float lv[G];
    bool valid[G];
    int n0[G];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const unsigned hw = (unsigned)sd[gi * (D + 1)];
        const int t = hw & 255;
        const unsigned p = ((hw >> 8) & 7u) * 64u + lane;
        valid[gi] = p < (unsigned)ATSC3_Z;
        const unsigned pp = valid[gi] ? p : 0u;
        dst[gi] = pp;
        pp8[gi] = 8u * pp;
        n0[gi] = TA ? (hw >> 11) & 63 : 0;
        row[gi] = (hw >> 17) & 255u;
        glob[gi] = (hw >> 25) & 1u;
        lv[gi] = at(lam, (unsigned)t * ATSC3_Z + pp);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const unsigned ew = (unsigned)sd[gi * (D + 1) + 1 + k];
            const unsigned u = ew & 511u, g = (ew >> 9) & 255u;
            if constexpr (M == 0) {
                wr[gi][k] = pp + u >= (unsigned)ATSC3_Z;
                const unsigned r = g * ATSC3_Z + pp + u - (wr[gi][k] ? (unsigned)ATSC3_Z : 0u);
                const uint2 w = at(R.p01, r);
                w0[gi][k] = w.x;
                w1[gi][k] = w.y;
                w0b[gi][k] = at(R.p0b, r);
            } else {
                // byte offset within group g's records, wrapped (min(q, q - 2880)),
                // on a wave-uniform base: add, subtract, min per entry
                const unsigned q = pp8[gi] + 8u * u;
                const unsigned ob = min(q, q - 8u * ATSC3_Z);
                const uint2 w = at(R.p01 + g * ATSC3_Z, 0u, ob);
                w0[gi][k] = w.x;
                w1[gi][k] = w.y;
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    // FP16: the 1.0 of the mixed-precision fma, hidden from constant folding
    float one = 1.0f;
    asm volatile("" : "+s"(one));
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        float m[D];
        float s = 0.0f;
        if constexpr (M == 1) {
            // message = the selected FP16 magnitude with the edge's sign
            // (record bit sp) rotated into bit 15, added as s = fma(h, 1, s)
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const unsigned ew = (unsigned)sd[gi * (D + 1) + 1 + k];
                const unsigned kk = (ew >> 17) & 127u, sp = ew >> 24;
                const bool a2 = kk == (w1[gi][k] >> RES_SIGN_LO);
                const unsigned hm = a2 ? (w0[gi][k] >> 16) : w0[gi][k];
                const unsigned sg = __builtin_amdgcn_alignbit(w1[gi][k], w1[gi][k], (sp + 17u) & 31u);
                // magnitudes are non-negative (bit 15 clear), so the sign is
                // OR-ed in, as the FP32 forms do (bits 16..31 are not read)
                const unsigned h = hm | (sg & 0x8000u);
                s = __builtin_fmaf(h_val((unsigned short)h), one, s);
            }
            if (valid[gi]) tot_st<M, PUB>(T, row[gi], glob[gi], dst[gi], lv[gi] + s);
            continue;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const unsigned ew = (unsigned)sd[gi * (D + 1) + 1 + k];
            const unsigned kk = (ew >> 17) & 127u, sp = ew >> 24;
            const bool a2 = kk == (w1[gi][k] >> RES_SIGN_LO);
            float v;
            if constexpr (M == 0) {
                v = __uint_as_float(a2 ? w0b[gi][k] : w0[gi][k]);
            } else {
                v = (float)((a2 ? (w0[gi][k] >> 8) : w0[gi][k]) & 0xffu) * R.step;
            }
            m[k] = __uint_as_float(__float_as_uint(v) | ((w1[gi][k] << (31u - sp)) & 0x80000000u));
        }
        if constexpr (M != 0 || (!TA && D == 2)) {
#pragma unroll
            for (int k = 0; k < D; ++k) s += m[k];
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
        if (valid[gi]) tot_st<M, PUB>(T, row[gi], glob[gi], dst[gi], lv[gi] + s);
    }
}

// Any degree: one record load per entry, visited in the summation order
// directly (FP32: suffix then prefix of each region; FP16 / INT8: table
// order).
template <int M>
__device__ float var_term(const QC_CONST ResTab *ct, const Rec &R, int e, unsigned pp) {
    const unsigned ew = (unsigned)ct->edge_v[2 * e];
    const unsigned vm = (unsigned)ct->edge_v[2 * e + 1];
    const unsigned u = ew & 511u, g = (ew >> 9) & 255u, kk = (ew >> 17) & 127u, sp = ew >> 24;
    const unsigned rix = g * ATSC3_Z + wrap360(pp + u);
    const unsigned long long w2 = sp >= RES_SIGN_LO ? at(R.p2, rix) : 0ull;
    const uint2 w = at(R.p01, rix);
    const float mv = rec_msg<M>(w.x, M == 0 ? at(R.p0b, rix) : 0u, w.y, w2, kk, sp, R);
    // a missing circulant entry adds +0.0: the sum (started at +0.0) is unchanged
    return vm != pp + 1u ? mv : 0.0f;
}

template <int M, bool PUB>
__device__ void var_generic(const QC_CONST ResTab *ct, const Tot &T, const float *lam,
                            const Rec &R, int t, unsigned p, bool valid) {
    const unsigned vw = (unsigned)ct->grp_v[t];
    const int dv = vw & 63, e0 = (vw >> 6) & 2047, n0 = (vw >> 17) & 63;
    const unsigned pp = valid ? p : 0u;
    const float lv = lam[(unsigned)t * ATSC3_Z + pp];
    float s = 0.0f;
    if constexpr (M != 0) {
        for (int k = 0; k < dv; ++k) s += var_term<M>(ct, R, e0 + k, pp);
    } else {
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
                s += var_term<M>(ct, R, e0 + k, pp);
            }
        }
    }
    if (valid) tot_st<M, PUB>(T, (vw >> 23) & 255u, vw >> 31, p, lv + s);
}

// Step dealing.  Static (ctr == nullptr): this wave's steps are w0, w0 +
// wstep, ...  Dynamic: the waves of a workgroup take their steps one at a
// time from an LDS counter, so a wave that finds its steps slow (misses, a
// busy SIMD) takes fewer of them and the phase ends when the work does, not
// when the slowest share does.  Which wave computes a step does not change
// any value.
struct StepDeal {
    int *ctr;
    int w0, wstep;
};
__device__ __forceinline__ int step_next(const StepDeal &d, unsigned lane, int si) {
    if (!d.ctr) return si + d.wstep;
    int v = 0;
    if (lane == 0u) v = atomicAdd(d.ctr, 1);
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int step_first(const StepDeal &d, unsigned lane) {
    return d.ctr ? step_next(d, lane, 0) : d.w0;
}

// One CHECK phase over the check steps that d deals this wave; returns the
// wave's count of unsatisfied checks (the same in every lane).
template <int M>
__device__ __forceinline__ unsigned check_phase(const QC_CONST ResTab *ct, const Tot &T,
                                                const Rec &R, const StepDeal &d, unsigned lane,
                                                bool first, bool update, float alpha) {
    unsigned unsat = 0u;
    for (int si = step_first(d, lane); si < ct->nsc; si = step_next(d, lane, si)) {
        const int sw = ct->c_steps[si];
        const int D = sw & 127, G = 1 << ((sw >> 7) & 7), start = (sw >> 10) & 4095;
        if ((sw >> 22) & 1) {
            const int task = ct->c_tasks[start];
            const int g = task & 255;
            const unsigned j = (unsigned)(task >> 8) * 64u + lane;
            const int gw = ct->grp_c[g];
            unsat += check_generic<M>(ct, T, R, g, (gw >> 7) & 2047, gw & 127,
                                        (gw >> 18) & 1, j, j < (unsigned)ATSC3_Z, first,
                                        update, alpha);
            continue;
        }
        const QC_CONST int *sd = ct->c_sdata + ct->c_soff[si];
        switch (G * 128 + D) {
#define RES_C1(DD)                                                                           \
    case 128 + DD:                                                                           \
        unsat += check_multi<M, DD, 1>(sd, T, R, lane, first, update, alpha);           \
        break;
#define RES_C2(DD)                                                                           \
    case 256 + DD:                                                                           \
        unsat += check_multi<M, DD, 2>(sd, T, R, lane, first, update, alpha);           \
        break;
            QC_CASES16(RES_C1)
            QC_CASES16(RES_C2)
#undef RES_C1
#undef RES_C2
            default: break;
        }
    }
    return unsat;
}

// Syndrome alone: the unsatisfied checks among the G tasks of a fast check
// step (the signs of the current posteriors, as check_multi counts them),
// with no record read or written.
template <int M, int D, int G>
__device__ __forceinline__ unsigned syndrome_multi(const QC_CONST int *sd, const Tot &T,
                                                   unsigned lane) {
    bool valid[G];
    float tv[G][D];
    __half th[G][D];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        const int task = sd[gi * (D + 1)];
        const unsigned j = (unsigned)(task >> 8) * 64u + lane;
        valid[gi] = j < (unsigned)ATSC3_Z;
        const unsigned jj = valid[gi] ? j : 0u;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const unsigned ew = (unsigned)sd[gi * (D + 1) + 1 + k];
            if constexpr (M == 0) {
                const unsigned p = wrap360(jj + (ew & 511u));
                tv[gi][k] = tot_ld<M>(T, (ew >> 9) & 511u, (ew >> 28) & 1u, p);
            } else {
                th[gi][k] = *reinterpret_cast<const __half *>(T.lds + ew + 2u * lane);
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    unsigned cnt = 0u;
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
        bool syn = false;
#pragma unroll
        for (int k = 0; k < D; ++k)
            syn ^= (M == 0 ? tv[gi][k] : __half2float(th[gi][k])) < 0.0f;
        cnt += wave_count(valid[gi] && syn);
    }
    return cnt;
}

// The same for a check of any degree (check_generic's edge loop).
template <int M>
__device__ unsigned syndrome_generic(const QC_CONST ResTab *ct, const Tot &T, int b0, int dc,
                                     bool hasmiss, unsigned j, bool valid) {
    const unsigned jj = valid ? j : 0u;
    bool syn = false;
    for (int k = 0; k < dc; ++k) {
        const unsigned ew = (unsigned)ct->edge_c[b0 + k];
        if (hasmiss && ((ew >> 18) & 1023u) == jj + 1u) continue;
        const unsigned p = wrap360(jj + (ew & 511u));
        syn ^= tot_ld<M>(T, (ew >> 9) & 511u, (ew >> 28) & 1u, p) < 0.0f;
    }
    return wave_count(valid && syn);
}

// The syndrome of the current posteriors over the check steps that d deals
// this wave: the count check_phase would return, at a fraction of its work.
template <int M>
__device__ __forceinline__ unsigned syndrome_phase(const QC_CONST ResTab *ct, const Tot &T,
                                                   const StepDeal &d, unsigned lane) {
    unsigned unsat = 0u;
    for (int si = step_first(d, lane); si < ct->nsc; si = step_next(d, lane, si)) {
        const int sw = ct->c_steps[si];
        const int D = sw & 127, G = 1 << ((sw >> 7) & 7), start = (sw >> 10) & 4095;
        if ((sw >> 22) & 1) {
            const int task = ct->c_tasks[start];
            const int g = task & 255;
            const unsigned j = (unsigned)(task >> 8) * 64u + lane;
            const int gw = ct->grp_c[g];
            unsat += syndrome_generic<M>(ct, T, (gw >> 7) & 2047, gw & 127, (gw >> 18) & 1, j,
                                         j < (unsigned)ATSC3_Z);
            continue;
        }
        const QC_CONST int *sd = ct->c_sdata + ct->c_soff[si];
        switch (G * 128 + D) {
#define RES_S1(DD)                                                                           \
    case 128 + DD:                                                                           \
        unsat += syndrome_multi<M, DD, 1>(sd, T, lane);                                     \
        break;
#define RES_S2(DD)                                                                           \
    case 256 + DD:                                                                           \
        unsat += syndrome_multi<M, DD, 2>(sd, T, lane);                                     \
        break;
            QC_CASES16(RES_S1)
            QC_CASES16(RES_S2)
#undef RES_S1
#undef RES_S2
            default: break;
        }
    }
    return unsat;
}

// Speculative stop test: once a check phase has left at most RES_SYN_FIRST
// unsatisfied checks, the next iteration first counts the unsatisfied
// checks alone (syndrome_phase: one posterior read per edge, no record
// traffic) and stops there when none is left; otherwise the check phase
// runs as always and counts them again.  The results are those of the
// plain schedule: the stop test sees the same syndrome, and the records a
// stopping check phase would have written are never read.  The pass costs
// about a fifth of a check phase; the threshold 4 stops ~3 of 4 codewords
// there at 21 dB on the 256-NUC r10/15 live-frame window, with ~1 wasted
// pass per codeword (thresholds 2, 3, 4, 6, 16: 8.86, 8.82, 8.78, 8.81,
// 8.85 ms per 2566-block window; without the pass 8.99 ms).
#define RES_SYN_FIRST 4

// One VAR phase over the variable steps that d deals this wave; PUB: team
// mode (posteriors published, see Tot).
template <int M, bool PUB>
__device__ __forceinline__ void var_phase(const QC_CONST ResTab *ct, const Tot &T,
                                          const float *lam, const Rec &R, const StepDeal &d,
                                          unsigned lane) {
    for (int si = step_first(d, lane); si < ct->nsv; si = step_next(d, lane, si)) {
        const int sw = ct->v_steps[si];
        const int D = sw & 127, G = 1 << ((sw >> 7) & 7), start = (sw >> 10) & 4095;
        if ((sw >> 22) & 1) {
            const int task = ct->v_tasks[start];
            const unsigned p = (unsigned)(task >> 8) * 64u + lane;
            var_generic<M, PUB>(ct, T, lam, R, task & 255, p, p < (unsigned)ATSC3_Z);
            continue;
        }
        const bool ta = M == 0 && ((sw >> 23) & 1);
        const QC_CONST int *sd = ct->v_sdata + ct->v_soff[si];
        switch ((ta ? 2048 : 0) + G * 128 + D) {
#define RES_V1(DD)                                                      \
    case 128 + DD:                                                      \
        var_multi<M, DD, 1, false, PUB>(sd, T, lam, R, lane);       \
        break;                                                          \
    case 2048 + 128 + DD:                                               \
        var_multi<M, DD, 1, true, PUB>(sd, T, lam, R, lane);        \
        break;
// pairs of degree 9..16 tasks: FP16 / INT8 only (FP32 keeps them single:
// its two record words and ordered sums would not fit the register budget)
#define RES_V2(DD)                                                      \
    case 256 + DD:                                                      \
        if constexpr (M != 0 || DD <= 8)                                \
            var_multi<M, DD, 2, false, PUB>(sd, T, lam, R, lane);   \
        break;                                                          \
    case 2048 + 256 + DD:                                               \
        if constexpr (M == 0 && DD <= 8)                                \
            var_multi<M, DD, 2, true, PUB>(sd, T, lam, R, lane);    \
        break;
#define RES_V4(DD)                                                      \
    case 512 + DD:                                                      \
        var_multi<M, DD, 4, false, PUB>(sd, T, lam, R, lane);       \
        break;                                                          \
    case 2048 + 512 + DD:                                               \
        var_multi<M, DD, 4, true, PUB>(sd, T, lam, R, lane);        \
        break;
// eight tasks of degree <= 3: FP16 / INT8 only
#define RES_V8(DD)                                                      \
    case 1024 + DD:                                                     \
        if constexpr (M != 0)                                           \
            var_multi<M, DD, 8, false, PUB>(sd, T, lam, R, lane);   \
        break;
            QC_CASES16(RES_V1)
            QC_CASES16(RES_V2)
            QC_CASES4(RES_V4)
            RES_V8(1) RES_V8(2) RES_V8(3)
#undef RES_V1
#undef RES_V2
#undef RES_V4
#undef RES_V8
            default: break;
        }
    }
}

// tot = lam for the 64-row tasks w0, w0 + wst, ... (8 loads in flight).
// Four positions per lane (16-byte LLR loads) measured 0.8% faster with the
// write-out below done the same way, but raised the grow kernel from 112 to
// 128 VGPRs, which leaves no room for a demap wave beside a decoding
// workgroup: the receiver pipeline's overlap is worth more.
template <int M>
__device__ __forceinline__ void res_tot_init(const QC_CONST ResTab *ct, const Tot &T,
                                             const float *lam, int w0, int wst, unsigned lane) {
    const int nvg = ct->nvg;
    for (int t8 = w0; t8 < nvg * 6; t8 += wst * 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int task = t8 + k * wst, tg = task / 6;
            const unsigned p = (unsigned)(task - tg * 6) * 64u + lane;
            v[k] = task < nvg * 6 && p < (unsigned)ATSC3_Z ? lam[(unsigned)tg * ATSC3_Z + p] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int task = t8 + k * wst, tg = task / 6;
            const unsigned p = (unsigned)(task - tg * 6) * 64u + lane;
            if (task < nvg * 6 && p < (unsigned)ATSC3_Z) {
                const unsigned vw = (unsigned)ct->grp_v[tg];
                tot_st<M>(T, (vw >> 23) & 255u, vw >> 31, p, v[k]);
            }
        }
    }
}

// Hard decisions (and posteriors when asked) of the 64-row tasks w0, w0 +
// wst, ...
template <int M>
__device__ __forceinline__ void res_write_out(const ResArgs &a, const QC_CONST ResTab *ct,
                                              const Tot &T, long long off, int w0, int wst,
                                              unsigned lane) {
    uint8_t *bout = a.bits + off;
    for (int task = w0; task < ct->nvg * 6; task += wst) {
        const int tg = task / 6;
        const unsigned p = (unsigned)(task - tg * 6) * 64u + lane;
        const unsigned vw = (unsigned)ct->grp_v[tg];
        if (p < (unsigned)ATSC3_Z) {
            const float v = tot_ld<M>(T, (vw >> 23) & 255u, vw >> 31, p);
            bout[(unsigned)tg * ATSC3_Z + p] = (uint8_t)(v < 0.0f);
            if (a.post) a.post[off + (unsigned)tg * ATSC3_Z + p] = v;
        }
    }
}

__global__ void resident_init(int *queue) {
    if (threadIdx.x == 0) *queue = 0;
}

template <int M>
__global__ __launch_bounds__(RES_WG) void resident_kernel(ResArgs a) {
    __shared__ __align__(16) unsigned char lds[M == 0 ? RES_F32_LDS_BYTES : RES_F16_LDS_BYTES];
    __shared__ int s_blk;
    __shared__ int s_unsat[3];       // check phases (by iteration parity), syndrome pass
    __shared__ int s_next[3];        // step counters: syndrome pass, check, variable phase
    __shared__ int s_hist[M == 2 ? 256 : 1];
    const unsigned tid = threadIdx.x, lane = tid & 63u;
    const int wave = __builtin_amdgcn_readfirstlane((int)(tid >> 6));
    Rec R;
    R.p01 = a.p01 + (size_t)blockIdx.x * a.stride;
    R.p0b = M == 0 ? a.p0b + (size_t)blockIdx.x * a.stride : nullptr;
    R.p2 = a.p2 + (size_t)blockIdx.x * a.stride;
    Tot T;
    T.lds = lds;
    T.glob = M == 0 ? a.totg + (size_t)blockIdx.x * a.tstride : nullptr;
    for (;;) {
        if (tid == 0) {
            s_blk = atomicAdd(a.queue, 1);
            s_unsat[0] = 0;
            s_unsat[1] = 0;
            s_unsat[2] = 0;
            s_next[0] = 0;
            s_next[1] = 0;
            s_next[2] = 0;
        }
        __syncthreads();
        const int blk = __builtin_amdgcn_readfirstlane(s_blk);
        if (blk >= a.nblk) return;
        const QC_CONST ResTab *ct = &a.codes[a.blk_code[blk]];
        const long long off = a.blk_off[blk];
        const float *lam = a.llr + off;
        if constexpr (M == 2) {
            // the codeword's INT8 step: the caller's, or 6 * median|llr| / 127
            R.step = a.msg_scale ? a.msg_scale[blk] : int8_step_of_block(lam, ct->n, s_hist);
            R.inv = 1.0f / R.step;
        }
        res_tot_init<M>(ct, T, lam, wave, RES_WAVES, lane);
        __syncthreads();
        int prev = 1 << 30;          // the last check phase's count
        // step counters: each is reset by thread 0 in a phase that no wave
        // can be taking its steps in (between two barriers that enclose none
        // of its phases)
        for (int it = 0;; ++it) {
            if (tid == 0) s_next[2] = 0;
            bool stop = false;
            if (it > 0 && prev <= RES_SYN_FIRST) {
                const unsigned u0 = syndrome_phase<M>(ct, T, StepDeal{&s_next[0], 0, 0}, lane);
                if (lane == 0u && u0) atomicAdd(&s_unsat[2], (int)u0);
                __syncthreads();
                stop = s_unsat[2] == 0;
            }
            int total = 0;
            if (!stop) {
                if (tid == 0) s_unsat[(it + 1) & 1] = 0;
                const bool first = it == 0, update = it < a.iters;
                const unsigned unsat = check_phase<M>(ct, T, R, StepDeal{&s_next[1], 0, 0}, lane,
                                                      first, update, a.alpha);
                if (lane == 0u && unsat) atomicAdd(&s_unsat[it & 1], (int)unsat);
                __syncthreads();
                total = s_unsat[it & 1];
                // every thread read the syndrome pass's count before this
                // barrier
                if (tid == 0) s_unsat[2] = 0;
            }
            if (total == 0 || it == a.iters) {
                res_write_out<M>(a, ct, T, off, wave, RES_WAVES, lane);
                if (tid == 0) {
                    a.conv[blk] = (uint8_t)(total == 0);
                    a.its[blk] = total == 0 ? it : a.iters;
                    a.bad[blk] = total;
                }
                break;
            }
            prev = total;
            if (tid == 0) {
                s_next[0] = 0;
                s_next[1] = 0;
            }
            var_phase<M, false>(ct, T, lam, R, StepDeal{&s_next[2], 0, 0}, lane);
            __syncthreads();
        }
        // every thread has read s_unsat and tot before thread 0 starts the
        // next codeword
        __syncthreads();
    }
}

struct WsLayout {
    int64_t stride, tstride, p01, p0b, p2, totg, bytes;
    WsLayout(int64_t grid, int64_t max_checks, int64_t max_grows, bool f32) {
        stride = (max_checks + 63) & ~(int64_t)63;
        tstride = f32 ? ((max_grows * ATSC3_Z + 63) & ~(int64_t)63) : 0;
        p01 = 256;
        p2 = p01 + grid * stride * 8;
        p0b = p2 + grid * stride * 8;
        totg = p0b + (f32 ? grid * stride * 4 : 0);
        bytes = totg + grid * tstride * 4;
    }
};

// ---------------------------------------------------------------------------
// Team mode: one codeword decoded by a team of workgroups on one XCD.
//
// The resident kernel puts a codeword on one compute unit, so a frame whose
// slowest codeword needs 50 iterations waits for 50 iterations of one CU.
// Here `team` workgroups (one per CU, all on one XCD) share each codeword:
// wave w of member m takes the check steps and variable steps w * team + m,
// stride 16 * team (together the steps the resident kernel's waves take), the team's
// records live in its slot in global memory, and every member keeps a full
// copy of the posteriors in LDS.  Per iteration:
//   CHECK (own share; old records, LDS posteriors) -> new records, the
//     member's unsatisfied-check count added to the team's counter
//   team barrier; total read; converged or at `iters` -> output (each member
//     its share of the bits; member 0 conv/its/bad), next codeword
//   VAR (own share; records of the whole team) -> posteriors published to
//     the team's global rows
//   team barrier; every member copies the published rows into its LDS
// so the results are the resident kernel's, bit for bit (same steps, same
// arithmetic, same order).
//
// Team barriers stay inside one XCD, whose workgroups share one L2: a writer
// waits for its stores to complete in that L2 (s_waitcnt vmcnt(0)) and
// arrives on the team's counter; a reader polls the counter and then drops
// its CU's L1 (agent-scope acquire) before reading what the others wrote --
// no L2 write-back.  Teams are formed from a census taken at launch: every
// workgroup reads its XCD id (HW_REG_XCC_ID) and draws a rank on its XCD's
// counter; after a counted one-time wait for the whole grid (every
// workgroup must be co-resident: grid <= compute units, one workgroup per
// CU), ranks r of XCD x form teams r / team (the last one smaller when the
// XCD's count is not a multiple).  Every wait is bounded: a wait that runs
// out sets the abort word, every workgroup leaves, and codewords never
// finished keep its = -1 (bad = -1, conv 0) from the launch's init kernel.
// ---------------------------------------------------------------------------
#define RES_TEAM_MAX 32
#define RES_XCDS 16                   // XCC_ID is a 4-bit field
#define RES_SPIN_LIMIT (1u << 23)
// header words of the team workspace (each group on its own 128-byte line)
#define TH_QUEUE 0
#define TH_ABORT 32
#define TH_CENSUS 64                  // RES_XCDS counters
#define TH_ARRIVE 96
#define TH_GO 128
#define TH_WORDS 160
// per-team control block (words): barrier counter, then on another line the
// three rotating unsatisfied-check counters and the two codeword words
#define TC_BAR 0
#define TC_UNSAT 32
#define TC_BLK 36
#define TC_WORDS 64

#define RES_MAX_RUNGS 4

struct TeamArgs {
    unsigned *hdr;                    // TH_WORDS header words, then the control blocks
    unsigned char *tpost;             // team posterior slots
    long long pstride;                // bytes per team posterior slot
    const QC_CONST int *order;        // queue position -> codeword, or NULL
    int team;
    int nrungs;                       // 0: one decode at ResArgs::alpha; else the ladder
    float ladder[RES_MAX_RUNGS];
};

// Wait until this wave's vector memory operations (stores and atomics
// included) have completed at the L2: s_waitcnt vmcnt(0) (gfx9 counts stores
// in vmcnt), expcnt and lgkmcnt left unconstrained.
__device__ __forceinline__ void res_wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// The XCD this wave runs on: HW_REG_XCC_ID (hardware register 20), bits [3:0].
__device__ __forceinline__ unsigned res_xcc_id() {
    return __builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11)) & (RES_XCDS - 1);
}

__device__ __forceinline__ unsigned ld_agent(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(unsigned *p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Spin (thread 0) until *w >= target; false (and the abort word set) when
// the abort word is seen or the bound runs out.
__device__ bool res_spin_until(const unsigned *w, unsigned target, unsigned *ab) {
    unsigned spins = 0;
    while (ld_agent(w) < target) {
        if ((++spins & 15u) == 0u) {
            if (ld_agent(ab)) return false;
            if (spins >= RES_SPIN_LIMIT) {
                st_agent(ab, 1u);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// Team barrier number `epoch` of a team of `size` members; when `val` is
// given, thread 0 reads it after the barrier into s_w[1].  Returns false
// when the abort word is set.
__device__ __forceinline__ bool team_barrier(unsigned *ctl, unsigned epoch, unsigned size,
                                             unsigned *ab, const unsigned *val, unsigned *s_w) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    res_wait_vmem();
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(ctl + TC_BAR, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool ok = res_spin_until(ctl + TC_BAR, epoch * size, ab);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_w[0] = ok ? 0u : 1u;
        if (val) s_w[1] = ld_agent(val);
    }
    __syncthreads();
    return s_w[0] == 0u;
}

// Every member's LDS rows from the team's published rows: FP16 rows of 360
// halves into the padded 424-half rows (positions 0..63 written twice), FP32
// rows of 360 floats as they are; 16-byte chunks, 8 loads in flight per
// thread.
template <int M>
__device__ __forceinline__ void team_refresh(unsigned char *lds, const unsigned char *pub,
                                             int nrows) {
    constexpr unsigned RB = M == 0 ? ATSC3_Z * 4 : ATSC3_Z * 2;   // published row bytes
    constexpr unsigned LB = M == 0 ? ATSC3_Z * 4 : RES_PROW * 2;  // LDS row bytes
    constexpr unsigned NC = RB / 16;
    const unsigned total = (unsigned)nrows * NC;
    for (unsigned c0 = threadIdx.x; c0 < total; c0 += 8u * RES_WG) {
        uint4 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            // past the end: chunk 0 again (in bounds, not stored)
            const unsigned c = c0 + (unsigned)q * RES_WG;
            v[q] = at(reinterpret_cast<const uint4 *>(pub), c < total ? c : 0u);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const unsigned c = c0 + (unsigned)q * RES_WG;
            if (c < total) {
                const unsigned r = c / NC, k = c - r * NC;
                *reinterpret_cast<uint4 *>(lds + r * LB + k * 16u) = v[q];
                if (M != 0 && k < 8u) *reinterpret_cast<uint4 *>(lds + r * LB + RB + k * 16u) = v[q];
            }
        }
    }
}

__global__ void resident_team_init(unsigned *hdr, int words, uint8_t *conv, int *its, int *bad,
                                   int nblk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) hdr[i] = 0u;
    if (i < nblk) {
        conv[i] = 0;
        its[i] = -1;
        bad[i] = -1;
    }
}

template <int M, bool DIRECT>
__global__ __launch_bounds__(RES_WG) void resident_team_kernel(ResArgs a, TeamArgs t) {
    __shared__ __align__(16) unsigned char
        lds[DIRECT ? 16 : (M == 0 ? RES_F32_LDS_BYTES : RES_F16_LDS_BYTES)];
    __shared__ int s_unsat[2];
    __shared__ int s_hist[M == 2 ? 256 : 1];
    __shared__ unsigned s_cnt[RES_XCDS];
    __shared__ unsigned s_w[4];
    const unsigned tid = threadIdx.x, lane = tid & 63u;
    const int wave = __builtin_amdgcn_readfirstlane((int)(tid >> 6));
    unsigned *hdr = t.hdr, *ab = t.hdr + TH_ABORT;

    // census: rank on this XCD, then wait for the whole grid
    const unsigned x = res_xcc_id();
    if (tid == 0) {
        s_w[2] = __hip_atomic_fetch_add(hdr + TH_CENSUS + x, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned old = __hip_atomic_fetch_add(hdr + TH_ARRIVE, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        bool ok = true;
        if (old + 1u == gridDim.x) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            st_agent(hdr + TH_GO, 1u);
        } else {
            ok = res_spin_until(hdr + TH_GO, 1u, ab);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        for (int q = 0; q < RES_XCDS; ++q) s_cnt[q] = ld_agent(hdr + TH_CENSUS + q);
        s_w[0] = ok ? 0u : 1u;
    }
    __syncthreads();
    if (s_w[0]) return;
    const unsigned tm = (unsigned)t.team;
    const unsigned rank = (unsigned)__builtin_amdgcn_readfirstlane((int)s_w[2]);
    unsigned slot = 0;
    for (unsigned q = 0; q < x; ++q) slot += (s_cnt[q] + tm - 1u) / tm;
    const unsigned nx = (unsigned)__builtin_amdgcn_readfirstlane((int)s_cnt[x]);
    const unsigned tix = rank / tm, m = rank - tix * tm;
    slot = (unsigned)__builtin_amdgcn_readfirstlane((int)(slot + tix));
    const unsigned S = min(tm, nx - tix * tm);   // members of this team
    unsigned *ctl = hdr + TH_WORDS + (size_t)slot * TC_WORDS;
    const bool leader = m == 0u && tid == 0u;

    Rec R;
    R.p01 = a.p01 + (size_t)slot * a.stride;
    R.p0b = M == 0 ? a.p0b + (size_t)slot * a.stride : nullptr;
    R.p2 = a.p2 + (size_t)slot * a.stride;
    Tot T;
    T.pub = t.tpost + (size_t)slot * t.pstride;
    // DIRECT: no LDS copy; the team's rows in global memory (FP16 padded as
    // the LDS rows are) are read and written where the LDS rows would be
    T.lds = DIRECT ? T.pub : lds;
    // FP32: the LDS rows' published copies first, then the global rows
    T.glob = M == 0 ? reinterpret_cast<float *>(T.pub) + RES_F32_LDS_ROWS * ATSC3_Z : nullptr;
    // steps and tasks dealt member-major (wave w of member m takes w * S + m,
    // then every S * 16-th): a phase with fewer steps than the team has
    // waves then spreads them over every member's SIMDs
    const int w0 = wave * (int)S + (int)m, wst = (int)S * RES_WAVES;
    unsigned epoch = 0;
    for (unsigned cw = 0;; ++cw) {
        if (leader)
            st_agent(ctl + TC_BLK + (cw & 1u),
                     (unsigned)__hip_atomic_fetch_add(hdr + TH_QUEUE, 1u, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT));
        if (!team_barrier(ctl, ++epoch, S, ab, ctl + TC_BLK + (cw & 1u), s_w)) return;
        const int q = __builtin_amdgcn_readfirstlane((int)s_w[1]);
        if (q >= a.nblk) return;
        // every member has read the last codeword's total: clear the counters
        if (leader)
            for (int k = 0; k < 3; ++k) st_agent(ctl + TC_UNSAT + k, 0u);
        const int blk = t.order ? t.order[q] : q;
        const QC_CONST ResTab *ct = &a.codes[a.blk_code[blk]];
        const int nvg = ct->nvg;
        const long long off = a.blk_off[blk];
        const float *lam = a.llr + off;
        if constexpr (M == 2) {
            R.step = a.msg_scale ? a.msg_scale[blk] : int8_step_of_block(lam, ct->n, s_hist);
            R.inv = 1.0f / R.step;
        }
        // the alpha ladder (lab/m9_fast.py:843-891): every rung decodes the
        // codeword afresh from its LLRs; a rung's result is taken when it
        // converged or leaves fewer unsatisfied checks than the best so far
        // (the first rung's always), and the ladder stops at convergence
        const int nr = t.nrungs > 0 ? t.nrungs : 1;
        int best_bad = 0x7fffffff;
        for (int r = 0; r < nr; ++r) {
            const float alpha = t.nrungs > 0 ? t.ladder[r] : a.alpha;
            if (r > 0) {
                // every member has read the last rung's total and written its
                // share of that rung's outputs before the rows are rewritten
                if (!team_barrier(ctl, ++epoch, S, ab, nullptr, s_w)) return;
                if (leader)
                    for (int k = 0; k < 3; ++k) st_agent(ctl + TC_UNSAT + k, 0u);
            }
            bool converged = false;
            // tot = lam for this member's share of the 64-row tasks, published
            for (int t8 = w0; t8 < nvg * 6; t8 += wst * 8) {
                float v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int task = t8 + k * wst, tg = task / 6;
                    const unsigned p = (unsigned)(task - tg * 6) * 64u + lane;
                    v[k] = task < nvg * 6 && p < (unsigned)ATSC3_Z ? lam[(unsigned)tg * ATSC3_Z + p]
                                                                   : 0.0f;
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int task = t8 + k * wst, tg = task / 6;
                    const unsigned p = (unsigned)(task - tg * 6) * 64u + lane;
                    if (task < nvg * 6 && p < (unsigned)ATSC3_Z) {
                        const unsigned vw = (unsigned)ct->grp_v[tg];
                        tot_st<M, !DIRECT>(T, (vw >> 23) & 255u, vw >> 31, p, v[k]);
                    }
                }
            }
            const int nlds = M == 0 ? min(nvg, RES_F32_LDS_ROWS) : nvg;
            if (!team_barrier(ctl, ++epoch, S, ab, nullptr, s_w)) return;
            if (!DIRECT) team_refresh<M>(lds, T.pub, nlds);
            if (tid == 0) s_unsat[0] = 0;
            __syncthreads();
            for (int it = 0;; ++it) {
                if (tid == 0) s_unsat[(it + 1) & 1] = 0;
                const bool first = it == 0, update = it < a.iters;
                const unsigned unsat = check_phase<M>(ct, T, R, StepDeal{nullptr, w0, wst}, lane,
                                                      first, update, alpha);
                if (lane == 0u && unsat) atomicAdd(&s_unsat[it & 1], (int)unsat);
                __syncthreads();
                unsigned *cnt = ctl + TC_UNSAT + it % 3;
                if (tid == 0 && s_unsat[it & 1])
                    __hip_atomic_fetch_add(cnt, (unsigned)s_unsat[it & 1], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                if (!team_barrier(ctl, ++epoch, S, ab, cnt, s_w)) return;
                const int total = __builtin_amdgcn_readfirstlane((int)s_w[1]);
                // counter (it + 1) % 3 was last read in iteration it - 2, before
                // every member's second barrier of that iteration
                if (leader) st_agent(ctl + TC_UNSAT + (it + 1) % 3, 0u);
                if (total == 0 || it == a.iters) {
                    converged = total == 0;
                    if (r == 0 || converged || total < best_bad) {
                        best_bad = total;
                        uint8_t *bout = a.bits + off;
                        for (int task = w0; task < nvg * 6; task += wst) {
                            const int tg = task / 6;
                            const unsigned p = (unsigned)(task - tg * 6) * 64u + lane;
                            const unsigned vw = (unsigned)ct->grp_v[tg];
                            if (p < (unsigned)ATSC3_Z) {
                                const float v = tot_ld<M>(T, (vw >> 23) & 255u, vw >> 31, p);
                                bout[(unsigned)tg * ATSC3_Z + p] = (uint8_t)(v < 0.0f);
                                if (a.post) a.post[off + (unsigned)tg * ATSC3_Z + p] = v;
                            }
                        }
                        if (leader) {
                            a.conv[blk] = (uint8_t)(total == 0);
                            a.its[blk] = total == 0 ? it : a.iters;
                            a.bad[blk] = total;
                        }
                    }
                    break;
                }
                var_phase<M, !DIRECT>(ct, T, lam, R, StepDeal{nullptr, w0, wst}, lane);
                if (!team_barrier(ctl, ++epoch, S, ab, nullptr, s_w)) return;
                if (!DIRECT) {
                    team_refresh<M>(lds, T.pub, nlds);
                    __syncthreads();
                }
            }
            if (converged) break;
        }
    }
}

// ---------------------------------------------------------------------------
// Grow mode: the resident kernel's one workgroup per codeword, whose last
// codewords grow into teams.
//
// Once the queue is empty, a workgroup that finishes its codeword has
// nothing left to start, while the codewords still running (the slow ones
// of the launch's last round) hold one compute unit each: the launch ends
// with most units idle.  Here a workgroup that finds the queue empty
// becomes a helper: it takes a seat on its XCD's seat counter and waits.
// After every variable phase a codeword's owner reads its XCD's seat and
// claim counters; when `team` - 1 unclaimed seats wait, it claims them (one
// compare-and-swap on the claim counter), copies its LDS posterior rows to
// its global slot (the layout the direct team flavor reads) and posts a
// request; the seated helpers read it, and the team (the owner is member 0)
// runs the codeword's remaining iterations as resident_team_kernel's direct
// flavor does: the same steps dealt member-major, the posteriors in the
// slot's global rows, the records in the owner's record slot.  Seat s of an
// XCD serves that XCD's request s / (team - 1) as member 1 + s % (team - 1).
// When the team finishes every member becomes a helper again.  A helper
// leaves when every codeword is finished (the done counter reaches nblk),
// when its seat lies past the request table, or on the abort word.  A
// codeword grows at most once, so each owner's control block serves one
// team.  Results are the resident kernel's bit for bit: every step computes
// the same values whichever wave runs it.  Owners never wait for helpers
// they have not claimed, and claimed helpers are running, so no wait
// depends on workgroups that are not resident.
// ---------------------------------------------------------------------------
#define GH_DONE 0                     // codewords finished (grow header words)
#define GH_XCD 32                     // per XCD, its own line: seats taken, seats claimed
#define GH_REQ (GH_XCD + RES_XCDS * 32)
#define GROW_RMAX 64                  // requests per XCD
#define GR_WORDS 8                    // request: posted, codeword, owner, first iteration, step
#define GH_WORDS (GH_REQ + RES_XCDS * GROW_RMAX * GR_WORDS)

// The owner's LDS posterior rows, byte for byte, to its global slot: 16-byte
// chunks, 8 in flight per thread.
__device__ __forceinline__ void grow_publish(const unsigned char *lds, unsigned char *pub,
                                             unsigned bytes) {
    const unsigned total = bytes / 16u;
    for (unsigned c0 = threadIdx.x; c0 < total; c0 += 8u * RES_WG) {
        uint4 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const unsigned c = c0 + (unsigned)q * RES_WG;
            // past the end: chunk 0 again (in bounds, not stored)
            v[q] = reinterpret_cast<const uint4 *>(lds)[c < total ? c : 0u];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const unsigned c = c0 + (unsigned)q * RES_WG;
            if (c < total) at(reinterpret_cast<uint4 *>(pub), c) = v[q];
        }
    }
}

// Thread 0 of a helper: 1 once the request at rq is posted, 0 once every
// codeword is finished, -1 on the abort word or when the bound runs out.
__device__ int grow_wait(const unsigned *rq, const unsigned *done, unsigned nblk, unsigned *ab) {
    for (unsigned spins = 1;; ++spins) {
        if (ld_agent(rq)) return 1;
        // a posted request's codeword cannot finish without this member, so
        // done == nblk means no request will come for this seat
        if (ld_agent(done) >= nblk) return 0;
        if ((spins & 15u) == 0u) {
            if (ld_agent(ab)) return -1;
            if (spins >= RES_SPIN_LIMIT) {
                st_agent(ab, 1u);
                return -1;
            }
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

// A grown codeword's iterations from `from` on: member m of a team of S,
// the owner's slot `own` (records, global rows, control block).  False on
// the abort word.
template <int M>
__device__ __forceinline__ bool grow_team(const ResArgs &a, const TeamArgs &t, unsigned *ctl,
                                          unsigned *done, unsigned *ab, int blk, unsigned own,
                                          unsigned m, unsigned S, int from, float step,
                                          int *s_unsat, unsigned *s_w) {
    const unsigned tid = threadIdx.x, lane = tid & 63u;
    const int wave = __builtin_amdgcn_readfirstlane((int)(tid >> 6));
    Rec R;
    R.p01 = a.p01 + (size_t)own * a.stride;
    R.p0b = M == 0 ? a.p0b + (size_t)own * a.stride : nullptr;
    R.p2 = a.p2 + (size_t)own * a.stride;
    R.step = step;
    R.inv = 1.0f / step;
    Tot T;
    T.pub = t.tpost + (size_t)own * t.pstride;
    T.lds = T.pub;
    T.glob = M == 0 ? a.totg + (size_t)own * a.tstride : nullptr;
    const QC_CONST ResTab *ct = &a.codes[a.blk_code[blk]];
    const long long off = a.blk_off[blk];
    const float *lam = a.llr + off;
    const int w0 = wave * (int)S + (int)m, wst = (int)S * RES_WAVES;
    const bool leader = m == 0u && tid == 0u;
    unsigned epoch = 0;
    if (tid == 0) {
        s_unsat[0] = 0;
        s_unsat[1] = 0;
    }
    __syncthreads();
    for (int it = from;; ++it) {
        if (tid == 0) s_unsat[(it + 1) & 1] = 0;
        const bool first = it == 0, update = it < a.iters;
        const unsigned unsat = check_phase<M>(ct, T, R, StepDeal{nullptr, w0, wst}, lane, first,
                                              update, a.alpha);
        if (lane == 0u && unsat) atomicAdd(&s_unsat[it & 1], (int)unsat);
        __syncthreads();
        unsigned *cnt = ctl + TC_UNSAT + it % 3;
        if (tid == 0 && s_unsat[it & 1])
            __hip_atomic_fetch_add(cnt, (unsigned)s_unsat[it & 1], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        if (!team_barrier(ctl, ++epoch, S, ab, cnt, s_w)) return false;
        const int total = __builtin_amdgcn_readfirstlane((int)s_w[1]);
        // counter (it + 1) % 3 was last read in iteration it - 2, before
        // every member's second barrier of that iteration
        if (leader) st_agent(ctl + TC_UNSAT + (it + 1) % 3, 0u);
        if (total == 0 || it == a.iters) {
            res_write_out<M>(a, ct, T, off, w0, wst, lane);
            if (leader) {
                a.conv[blk] = (uint8_t)(total == 0);
                a.its[blk] = total == 0 ? it : a.iters;
                a.bad[blk] = total;
                __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return true;
        }
        var_phase<M, false>(ct, T, lam, R, StepDeal{nullptr, w0, wst}, lane);
        if (!team_barrier(ctl, ++epoch, S, ab, nullptr, s_w)) return false;
    }
}

template <int M>
__global__ __launch_bounds__(RES_WG) void resident_grow_kernel(ResArgs a, TeamArgs t) {
    __shared__ __align__(16) unsigned char lds[M == 0 ? RES_F32_LDS_BYTES : RES_F16_LDS_BYTES];
    __shared__ int s_blk;
    __shared__ int s_unsat[3];       // check phases (by iteration parity), syndrome pass
    __shared__ int s_next[3];        // step counters: syndrome pass, check, variable phase
    __shared__ int s_hist[M == 2 ? 256 : 1];
    __shared__ unsigned s_w[8];
    const unsigned tid = threadIdx.x, lane = tid & 63u;
    const int wave = __builtin_amdgcn_readfirstlane((int)(tid >> 6));
    unsigned *hdr = t.hdr, *ab = hdr + TH_ABORT, *gh = hdr + TH_WORDS;
    unsigned *done = gh + GH_DONE;
    unsigned *ctl0 = gh + GH_WORDS;
    const unsigned x = res_xcc_id();
    unsigned *xw = gh + GH_XCD + x * 32u;
    unsigned *reqs = gh + GH_REQ + x * (GROW_RMAX * GR_WORDS);
    const unsigned S = (unsigned)t.team, need = S - 1u;
    const unsigned me = blockIdx.x;
    bool solo = true;                 // still taking codewords from the queue
    for (;;) {
        // the next team this workgroup joins: (codeword, owner, member,
        // first iteration, INT8 step)
        int tblk = -1, tfrom = 0;
        unsigned town = 0, tm = 0;
        float tstep = 1.0f;
        if (solo) {
            Rec R;
            R.p01 = a.p01 + (size_t)me * a.stride;
            R.p0b = M == 0 ? a.p0b + (size_t)me * a.stride : nullptr;
            R.p2 = a.p2 + (size_t)me * a.stride;
            R.step = 1.0f;
            R.inv = 1.0f;
            Tot T;
            T.pub = t.tpost + (size_t)me * t.pstride;
            T.lds = lds;
            // FP32: the rows beyond LDS, where the team reads them after
            // growing
            T.glob = M == 0 ? a.totg + (size_t)me * a.tstride : nullptr;
            for (;;) {
                if (tid == 0) {
                    s_blk = (int)__hip_atomic_fetch_add(hdr + TH_QUEUE, 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
                    s_unsat[0] = 0;
                    s_unsat[1] = 0;
                    s_unsat[2] = 0;
                    s_next[0] = 0;
                    s_next[1] = 0;
                    s_next[2] = 0;
                }
                __syncthreads();
                const int q = __builtin_amdgcn_readfirstlane(s_blk);
                if (q >= a.nblk) {
                    solo = false;
                    break;
                }
                const int blk = t.order ? t.order[q] : q;
                const QC_CONST ResTab *ct = &a.codes[a.blk_code[blk]];
                const int nvg = ct->nvg;
                const long long off = a.blk_off[blk];
                const float *lam = a.llr + off;
                if constexpr (M == 2) {
                    R.step = a.msg_scale ? a.msg_scale[blk]
                                         : int8_step_of_block(lam, ct->n, s_hist);
                    R.inv = 1.0f / R.step;
                }
                res_tot_init<M>(ct, T, lam, wave, RES_WAVES, lane);
                __syncthreads();
                int from = -1;        // grown: the team's first iteration
                unsigned rix = 0;
                int prev = 1 << 30;   // the last check phase's count
                // step counters reset as in resident_kernel
                for (int it = 0;; ++it) {
                    if (tid == 0) s_next[2] = 0;
                    bool stop = false;
                    if (it > 0 && prev <= RES_SYN_FIRST) {
                        const unsigned u0 = syndrome_phase<M>(ct, T, StepDeal{&s_next[0], 0, 0},
                                                              lane);
                        if (lane == 0u && u0) atomicAdd(&s_unsat[2], (int)u0);
                        __syncthreads();
                        stop = s_unsat[2] == 0;
                    }
                    int total = 0;
                    if (!stop) {
                        if (tid == 0) s_unsat[(it + 1) & 1] = 0;
                        const bool first = it == 0, update = it < a.iters;
                        const unsigned unsat = check_phase<M>(ct, T, R,
                                                              StepDeal{&s_next[1], 0, 0}, lane,
                                                              first, update, a.alpha);
                        if (lane == 0u && unsat) atomicAdd(&s_unsat[it & 1], (int)unsat);
                        __syncthreads();
                        total = s_unsat[it & 1];
                        // every thread read the syndrome pass's count before
                        // this barrier
                        if (tid == 0) s_unsat[2] = 0;
                    }
                    if (total == 0 || it == a.iters) {
                        res_write_out<M>(a, ct, T, off, wave, RES_WAVES, lane);
                        if (tid == 0) {
                            a.conv[blk] = (uint8_t)(total == 0);
                            a.its[blk] = total == 0 ? it : a.iters;
                            a.bad[blk] = total;
                            __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                        }
                        break;
                    }
                    // the XCD's counters, loaded before the variable phase and
                    // looked at after it (a codeword in its last iteration
                    // does not grow)
                    prev = total;
                    const bool look = need != 0u && tid == 0 && it + 1 < a.iters;
                    unsigned sv = 0u, cv = 0u;
                    if (look) {
                        sv = ld_agent(xw);
                        cv = ld_agent(xw + 1);
                    }
                    if (tid == 0) {
                        s_next[0] = 0;
                        s_next[1] = 0;
                    }
                    var_phase<M, false>(ct, T, lam, R, StepDeal{&s_next[2], 0, 0}, lane);
                    if (tid == 0) {
                        s_w[0] = 0u;
                        // seats only grow and the swap re-reads the claims, so
                        // sv >= cv + need holds at the swap in whichever order
                        // the two loads saw the counters
                        if (look && sv >= cv + need && cv / need < GROW_RMAX) {
                            unsigned exp = cv;
                            if (__hip_atomic_compare_exchange_strong(
                                    xw + 1, &exp, cv + need, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT)) {
                                s_w[0] = 1u;
                                s_w[1] = cv / need;
                            }
                        }
                    }
                    __syncthreads();
                    if (s_w[0]) {
                        from = it + 1;
                        rix = (unsigned)__builtin_amdgcn_readfirstlane((int)s_w[1]);
                        break;
                    }
                }
                if (from >= 0) {
                    // the rows (and the records of the phases before) complete
                    // in L2 before the request is posted
                    const unsigned nrows =
                        M == 0 ? (unsigned)min(nvg, RES_F32_LDS_ROWS) : (unsigned)nvg;
                    grow_publish(lds, T.pub, nrows * (M == 0 ? ATSC3_Z * 4u : RES_PROW * 2u));
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    res_wait_vmem();
                    __syncthreads();
                    if (tid == 0) {
                        unsigned *rq = reqs + rix * GR_WORDS;
                        st_agent(rq + 1, (unsigned)blk);
                        st_agent(rq + 2, me);
                        st_agent(rq + 3, (unsigned)from);
                        st_agent(rq + 4, __float_as_uint(R.step));
                        res_wait_vmem();
                        st_agent(rq, 1u);
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    }
                    tblk = blk;
                    town = me;
                    tm = 0u;
                    tfrom = from;
                    tstep = R.step;
                    // after growing, the workgroup is a helper (the queue is
                    // empty)
                    solo = false;
                    break;
                }
                // every thread has read s_unsat and the rows before thread 0
                // starts the next codeword
                __syncthreads();
            }
        }
        if (tblk < 0) {
            // helper: a seat on this XCD, then its request or the end
            if (tid == 0) {
                s_w[0] = 0u;
                if (need) {
                    const unsigned s = __hip_atomic_fetch_add(xw, 1u, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
                    if (s / need < GROW_RMAX) {
                        const unsigned *rq = reqs + (s / need) * GR_WORDS;
                        if (grow_wait(rq, done, (unsigned)a.nblk, ab) > 0) {
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                            s_w[0] = 1u;
                            s_w[1] = ld_agent(rq + 1);
                            s_w[2] = ld_agent(rq + 2);
                            s_w[3] = ld_agent(rq + 3);
                            s_w[4] = ld_agent(rq + 4);
                            s_w[5] = 1u + s % need;
                        }
                    }
                }
            }
            __syncthreads();
            if (!s_w[0]) return;
            tblk = __builtin_amdgcn_readfirstlane((int)s_w[1]);
            town = (unsigned)__builtin_amdgcn_readfirstlane((int)s_w[2]);
            tfrom = __builtin_amdgcn_readfirstlane((int)s_w[3]);
            tstep = __uint_as_float((unsigned)__builtin_amdgcn_readfirstlane((int)s_w[4]));
            tm = (unsigned)__builtin_amdgcn_readfirstlane((int)s_w[5]);
        }
        // every thread has read s_w before the team run reuses it
        __syncthreads();
        if (!grow_team<M>(a, t, ctl0 + (size_t)town * TC_WORDS, done, ab, tblk, town, tm, S, tfrom,
                          tstep, s_unsat, s_w))
            return;
        // every thread has read the team's last total (s_w) before thread 0
        // takes the next seat
        __syncthreads();
    }
}

// Queue order for team mode: the codewords that will need the most
// iterations first, so that no slow codeword starts late.  The estimate is
// the fraction of unsatisfied checks of the channel's hard decisions over a
// sample of every check group's rows (positions 0..63; noise is independent
// of the position): per code it ranks iteration counts closely, and the
// codewords that exhaust the iteration limit sit near one half.  One
// workgroup per codeword.  FP32 tables name a variable group's posterior
// row, not the group: the inverse of grp_v's row map is built in LDS.
#define RES_KEY_ROWS 64
#define RES_KEY_WG 1024
__global__ __launch_bounds__(RES_KEY_WG) void team_difficulty(const float *llr,
                                                        const QC_CONST long long *blk_off,
                                                        const QC_CONST int *blk_code,
                                                        const QC_CONST ResTab *codes, int f32,
                                                        float *key) {
    __shared__ unsigned char inv[256];
    __shared__ int s_cnt;
    const int b = blockIdx.x;
    const QC_CONST ResTab *ct = &codes[blk_code[b]];
    const float *lam = llr + blk_off[b];
    const int nvg = ct->nvg, ncg = ct->ncg;
    if (threadIdx.x == 0) s_cnt = 0;
    if (f32)
        for (int t = threadIdx.x; t < nvg; t += blockDim.x) {
            const unsigned vw = (unsigned)ct->grp_v[t];
            inv[((vw >> 31) << 7) | ((vw >> 23) & 127u)] = (unsigned char)t;
        }
    __syncthreads();
    unsigned cnt = 0;
    for (int c = threadIdx.x; c < ncg * RES_KEY_ROWS; c += blockDim.x) {
        const int g = c / RES_KEY_ROWS;
        const unsigned j = (unsigned)(c - g * RES_KEY_ROWS);
        const unsigned gw = (unsigned)ct->grp_c[g];
        const int dc = gw & 127, b0 = (gw >> 7) & 2047;
        const bool hasmiss = (gw >> 18) & 1u;
        unsigned par = 0;
        // 16 loads in flight (one round for the common degrees <= 16)
        for (int k0 = 0; k0 < dc; k0 += 16) {
            float v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                v[q] = 0.0f;                       // past the degree or a missing entry: +0
                if (k0 + q < dc) {
                    const unsigned ew = (unsigned)ct->edge_c[b0 + k0 + q];
                    if (!(hasmiss && ((ew >> 18) & 1023u) == j + 1u)) {
                        const unsigned row = (ew >> 9) & 511u;
                        const unsigned t = f32 ? inv[(((ew >> 28) & 1u) << 7) | (row & 127u)] : row;
                        v[q] = lam[t * ATSC3_Z + wrap360(j + (ew & 511u))];
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) par ^= v[q] < 0.0f ? 1u : 0u;
        }
        cnt += par;
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63u) == 0u && cnt) atomicAdd(&s_cnt, (int)cnt);
    __syncthreads();
    if (threadIdx.x == 0) key[b] = (float)s_cnt / (float)(ncg * RES_KEY_ROWS);
}

// order[r] = the codeword of rank r by decreasing key (ties: lower index
// first): each codeword counts the ones ranked before it.
__global__ void team_rank(const float *key, int nblk, int *order) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nblk) return;
    const float ki = key[i];
    int r = 0;
    for (int j = 0; j < nblk; ++j) {
        const float kj = key[j];
        r += (kj > ki || (kj == ki && j < i)) ? 1 : 0;
    }
    order[r] = i;
}

struct TeamLayout {
    int64_t nslots, stride, pstride, tstride, hdr, p01, p0b, p2, totg, tpost, bytes, gwords;
    // grow: grow mode (one slot per workgroup, the grow header after the
    // team header words; FP32 global rows in a region of their own, laid out
    // as the resident kernel's, the slots holding only the LDS rows' copies)
    TeamLayout(int64_t grid, int64_t team, int64_t max_checks, int64_t max_grows, bool f32,
               bool grow = false) {
        // at most ceil(n_x / team) teams on each XCD x: grid / team + one per XCD
        nslots = grow ? grid : grid / team + RES_XCDS;
        gwords = grow ? GH_WORDS : 0;
        stride = (max_checks + 63) & ~(int64_t)63;
        tstride = f32 && grow ? ((max_grows * ATSC3_Z + 63) & ~(int64_t)63) : 0;
        // FP16: padded rows (the direct flavor reads them as the LDS rows)
        pstride = f32 ? ((RES_F32_LDS_ROWS + (grow ? 0 : max_grows)) * ATSC3_Z * 4 + 255) &
                            ~(int64_t)255
                      : ((int64_t)RES_F16_LDS_BYTES + 255) & ~(int64_t)255;
        hdr = 0;
        const int64_t words = TH_WORDS + gwords + nslots * TC_WORDS;
        p01 = (words * 4 + 255) & ~(int64_t)255;
        p2 = p01 + nslots * stride * 8;
        p0b = p2 + nslots * stride * 8;
        totg = p0b + (f32 ? nslots * stride * 4 : 0);
        tpost = totg + nslots * tstride * 4;
        tpost = (tpost + 255) & ~(int64_t)255;
        bytes = tpost + nslots * pstride;
    }
    int64_t words() const { return TH_WORDS + gwords + nslots * TC_WORDS; }
};

}  // namespace
}  // namespace atsc3

extern "C" int64_t atsc3_minsum_resident_ws_bytes(int32_t grid, int32_t max_checks,
                                                  int32_t max_global_rows, int32_t msg_fmt) {
    if (grid < 1 || max_checks < 1 || max_global_rows < 0 || msg_fmt < 0 || msg_fmt > 2)
        return ATSC3_E_ARG;
    return atsc3::WsLayout(grid, max_checks, max_global_rows, msg_fmt == 0).bytes;
}

extern "C" int32_t atsc3_minsum_resident_lds_rows(int32_t msg_fmt) {
    if (msg_fmt == 0) return RES_F32_LDS_ROWS;
    if (msg_fmt == 1 || msg_fmt == 2) return RES_NMAX / ATSC3_Z;
    return ATSC3_E_ARG;
}

extern "C" int32_t atsc3_minsum_resident_grid(void) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ATSC3_E_HIP;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return ATSC3_E_HIP;
    return cus > 0 ? cus : ATSC3_E_HIP;
}

namespace atsc3 {
namespace {
// Validates the arguments and enqueues the queue reset and the decode with
// the workspace `ws` of `ws_bytes` bytes.
int32_t resident_launch(const float *llr, int32_t nblk, const int64_t *blk_off,
                        const int32_t *blk_code, const void *codes, int32_t ncode,
                        int32_t max_checks, int32_t max_global_rows, int32_t max_n,
                        int32_t max_dc, int32_t max_dv, int32_t iters, float alpha,
                        int32_t msg_fmt, const float *msg_scale, int32_t grid, void *ws,
                        int64_t ws_bytes, uint8_t *bits, uint8_t *conv, int32_t *its,
                        int32_t *bad, float *post, hipStream_t st) {
    const bool f32 = msg_fmt == 0;
    const WsLayout L(grid, max_checks, max_global_rows, f32);
    if (ws_bytes < L.bytes) return ATSC3_E_SIZE;
    char *base = static_cast<char *>(ws);
    ResArgs a{};
    a.llr = llr;
    a.blk_off = (const QC_CONST long long *)blk_off;
    a.blk_code = (const QC_CONST int *)blk_code;
    a.codes = (const QC_CONST ResTab *)codes;
    a.nblk = nblk;
    a.iters = iters;
    a.alpha = alpha;
    a.stride = L.stride;
    a.tstride = L.tstride;
    a.queue = reinterpret_cast<int *>(base);
    a.p01 = reinterpret_cast<uint2 *>(base + L.p01);
    a.p2 = reinterpret_cast<unsigned long long *>(base + L.p2);
    a.p0b = f32 ? reinterpret_cast<unsigned *>(base + L.p0b) : nullptr;
    a.totg = f32 ? reinterpret_cast<float *>(base + L.totg) : nullptr;
    a.bits = bits;
    a.conv = conv;
    a.its = its;
    a.bad = bad;
    a.post = post;
    hipLaunchKernelGGL(resident_init, dim3(1), dim3(64), 0, st, a.queue);
    if (hipGetLastError() != hipSuccess) return ATSC3_E_HIP;
    a.msg_scale = msg_fmt == 2 ? msg_scale : nullptr;
    if (f32)
        hipLaunchKernelGGL(resident_kernel<0>, dim3(grid), dim3(RES_WG), 0, st, a);
    else if (msg_fmt == 1)
        hipLaunchKernelGGL(resident_kernel<1>, dim3(grid), dim3(RES_WG), 0, st, a);
    else
        hipLaunchKernelGGL(resident_kernel<2>, dim3(grid), dim3(RES_WG), 0, st, a);
    return launch_status();
}

int32_t resident_check_args(const float *llr, int32_t nblk, const int64_t *blk_off,
                            const int32_t *blk_code, const void *codes, int32_t ncode,
                            int32_t max_checks, int32_t max_global_rows, int32_t max_n,
                            int32_t max_dc, int32_t max_dv, int32_t iters, float alpha,
                            int32_t msg_fmt, int32_t grid, const uint8_t *bits,
                            const uint8_t *conv, const int32_t *its, const int32_t *bad) {
    if (nblk < 0 || iters < 0 || iters >= (1 << 24) || !(alpha >= 0.0f)) return ATSC3_E_ARG;
    if (msg_fmt < 0 || msg_fmt > 2) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !blk_off || !blk_code || !codes || !bits || !conv || !its || !bad)
        return ATSC3_E_ARG;
    if (ncode < 1 || max_checks < 1 || grid < 1 || max_global_rows < 0) return ATSC3_E_ARG;
    if (max_n < 1 || max_n > RES_NMAX) return ATSC3_E_SIZE;
    if (max_dc < 1 || max_dc > RES_MAX_DC || max_dv < 1 || max_dv > RES_MAX_DV)
        return ATSC3_E_DEGREE;
    return ATSC3_OK;
}
}  // namespace
}  // namespace atsc3

extern "C" int32_t atsc3_minsum_resident_dev(
    const float *llr, int32_t nblk, const int64_t *blk_off, const int32_t *blk_code,
    const void *codes, int32_t ncode, int32_t max_checks, int32_t max_global_rows,
    int32_t max_n, int32_t max_dc, int32_t max_dv, int32_t iters, float alpha, int32_t msg_fmt,
    const float *msg_scale, int32_t grid, void *ws, int64_t ws_bytes, uint8_t *bits,
    uint8_t *conv, int32_t *its,
    int32_t *bad, void *stream) {
    const int32_t rc = atsc3::resident_check_args(llr, nblk, blk_off, blk_code, codes, ncode,
                                                  max_checks, max_global_rows, max_n, max_dc,
                                                  max_dv, iters, alpha, msg_fmt, grid, bits,
                                                  conv, its, bad);
    if (rc != ATSC3_OK || nblk == 0) return rc;
    if (!ws) return ATSC3_E_ARG;
    return atsc3::resident_launch(llr, nblk, blk_off, blk_code, codes, ncode, max_checks,
                                  max_global_rows, max_n, max_dc, max_dv, iters, alpha, msg_fmt,
                                  msg_scale, grid, ws, ws_bytes, bits, conv, its, bad,
                                  nullptr, atsc3::as_stream(stream));
}

// The same decode also writing each codeword's final posterior LLRs
// (llr_out, laid out like bits): FP32 mode the FP32 posteriors, FP16 / INT8
// modes the FP16 posteriors widened to FP32.
extern "C" int32_t atsc3_minsum_resident_llr_dev(
    const float *llr, int32_t nblk, const int64_t *blk_off, const int32_t *blk_code,
    const void *codes, int32_t ncode, int32_t max_checks, int32_t max_global_rows,
    int32_t max_n, int32_t max_dc, int32_t max_dv, int32_t iters, float alpha, int32_t msg_fmt,
    const float *msg_scale, int32_t grid, void *ws, int64_t ws_bytes, uint8_t *bits,
    uint8_t *conv, int32_t *its, int32_t *bad, float *llr_out, void *stream) {
    const int32_t rc = atsc3::resident_check_args(llr, nblk, blk_off, blk_code, codes, ncode,
                                                  max_checks, max_global_rows, max_n, max_dc,
                                                  max_dv, iters, alpha, msg_fmt, grid, bits,
                                                  conv, its, bad);
    if (rc != ATSC3_OK || nblk == 0) return rc;
    if (!ws || !llr_out) return ATSC3_E_ARG;
    return atsc3::resident_launch(llr, nblk, blk_off, blk_code, codes, ncode, max_checks,
                                  max_global_rows, max_n, max_dc, max_dv, iters, alpha, msg_fmt,
                                  msg_scale, grid, ws, ws_bytes, bits, conv, its, bad, llr_out,
                                  atsc3::as_stream(stream));
}

// Team mode (see resident_team_kernel): the same decode, each codeword on a
// team of `team` workgroups of one XCD.
extern "C" int64_t atsc3_minsum_resident_team_ws_bytes(int32_t grid, int32_t team,
                                                       int32_t max_checks,
                                                       int32_t max_global_rows,
                                                       int32_t msg_fmt) {
    const bool grow = (team & ATSC3_TEAM_GROW) != 0;
    team &= ~(ATSC3_TEAM_DIRECT | ATSC3_TEAM_GROW);
    if (grid < 1 || team < 1 || team > RES_TEAM_MAX || max_checks < 1 || max_global_rows < 0 ||
        msg_fmt < 0 || msg_fmt > 2)
        return ATSC3_E_ARG;
    return atsc3::TeamLayout(grid, team, max_checks, max_global_rows, msg_fmt == 0, grow).bytes;
}

// Index of the int32 abort word in the team workspace: nonzero after a
// launch whose bounded waits ran out (its = -1 marks the codewords it left).
extern "C" int32_t atsc3_minsum_resident_team_status_word(void) { return TH_ABORT; }

// The team queue order (see team_difficulty): order[0..nblk) = the codewords
// by decreasing sampled fraction of unsatisfied checks; key: nblk floats of
// device scratch.
extern "C" int32_t atsc3_minsum_resident_team_order_dev(const float *llr, int32_t nblk,
                                                        const int64_t *blk_off,
                                                        const int32_t *blk_code,
                                                        const void *codes, int32_t msg_fmt,
                                                        float *key, int32_t *order,
                                                        void *stream) {
    if (nblk < 0 || msg_fmt < 0 || msg_fmt > 2) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !blk_off || !blk_code || !codes || !key || !order) return ATSC3_E_ARG;
    hipStream_t st = atsc3::as_stream(stream);
    hipLaunchKernelGGL(atsc3::team_difficulty, dim3(nblk), dim3(RES_KEY_WG), 0, st, llr,
                       (const QC_CONST long long *)blk_off, (const QC_CONST int *)blk_code,
                       (const QC_CONST atsc3::ResTab *)codes, msg_fmt == 0 ? 1 : 0, key);
    if (hipGetLastError() != hipSuccess) return ATSC3_E_HIP;
    hipLaunchKernelGGL(atsc3::team_rank, dim3((nblk + 255) / 256), dim3(256), 0, st, key, nblk,
                       order);
    return atsc3::launch_status();
}

extern "C" int32_t atsc3_minsum_resident_team_dev(
    const float *llr, int32_t nblk, const int64_t *blk_off, const int32_t *blk_code,
    const void *codes, int32_t ncode, int32_t max_checks, int32_t max_global_rows,
    int32_t max_n, int32_t max_dc, int32_t max_dv, int32_t iters, float alpha, int32_t msg_fmt,
    const float *msg_scale, const int32_t *order, const float *ladder, int32_t nrungs,
    int32_t grid, int32_t team, void *ws, int64_t ws_bytes, uint8_t *bits, uint8_t *conv,
    int32_t *its, int32_t *bad, void *stream) {
    const int32_t rc = atsc3::resident_check_args(llr, nblk, blk_off, blk_code, codes, ncode,
                                                  max_checks, max_global_rows, max_n, max_dc,
                                                  max_dv, iters, alpha, msg_fmt, grid, bits,
                                                  conv, its, bad);
    if (rc != ATSC3_OK) return rc;
    const bool direct = (team & ATSC3_TEAM_DIRECT) != 0;
    const bool grow = (team & ATSC3_TEAM_GROW) != 0;
    team &= ~(ATSC3_TEAM_DIRECT | ATSC3_TEAM_GROW);
    if (team < 1 || team > RES_TEAM_MAX) return ATSC3_E_ARG;
    if (nrungs < 0 || nrungs > RES_MAX_RUNGS || (nrungs > 0 && !ladder)) return ATSC3_E_ARG;
    // grow mode runs one decode per codeword (no ladder)
    if (grow && nrungs > 0) return ATSC3_E_ARG;
    for (int r = 0; r < nrungs; ++r)
        if (!(ladder[r] >= 0.0f)) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!ws) return ATSC3_E_ARG;
    const bool f32 = msg_fmt == 0;
    const atsc3::TeamLayout L(grid, team, max_checks, max_global_rows, f32, grow);
    if (ws_bytes < L.bytes) return ATSC3_E_SIZE;
    // every workgroup must be co-resident (one per compute unit): the census
    // waits for the whole grid
    const int32_t cus = atsc3_minsum_resident_grid();
    if (cus < 0) return cus;
    if (grid > cus) return ATSC3_E_SIZE;
    char *base = static_cast<char *>(ws);
    atsc3::ResArgs a{};
    a.llr = llr;
    a.blk_off = (const QC_CONST long long *)blk_off;
    a.blk_code = (const QC_CONST int *)blk_code;
    a.codes = (const QC_CONST atsc3::ResTab *)codes;
    a.nblk = nblk;
    a.iters = iters;
    a.alpha = alpha;
    a.msg_scale = msg_fmt == 2 ? msg_scale : nullptr;
    a.stride = L.stride;
    a.p01 = reinterpret_cast<uint2 *>(base + L.p01);
    a.p2 = reinterpret_cast<unsigned long long *>(base + L.p2);
    a.p0b = f32 ? reinterpret_cast<unsigned *>(base + L.p0b) : nullptr;
    a.tstride = L.tstride;
    a.totg = L.tstride ? reinterpret_cast<float *>(base + L.totg) : nullptr;
    a.bits = bits;
    a.conv = conv;
    a.its = its;
    a.bad = bad;
    atsc3::TeamArgs t{};
    t.hdr = reinterpret_cast<unsigned *>(base + L.hdr);
    t.tpost = reinterpret_cast<unsigned char *>(base + L.tpost);
    t.pstride = L.pstride;
    t.order = (const QC_CONST int *)order;
    t.team = team;
    t.nrungs = nrungs;
    for (int r = 0; r < nrungs; ++r) t.ladder[r] = ladder[r];
    hipStream_t st = atsc3::as_stream(stream);
    const int words = (int)L.words();
    const int n0 = words > nblk ? words : nblk;
    hipLaunchKernelGGL(atsc3::resident_team_init, dim3((n0 + 255) / 256), dim3(256), 0, st,
                       t.hdr, words, conv, its, bad, nblk);
    if (hipGetLastError() != hipSuccess) return ATSC3_E_HIP;
    const dim3 g(grid), b(RES_WG);
    if (grow) {
        if (f32)
            hipLaunchKernelGGL(atsc3::resident_grow_kernel<0>, g, b, 0, st, a, t);
        else if (msg_fmt == 1)
            hipLaunchKernelGGL(atsc3::resident_grow_kernel<1>, g, b, 0, st, a, t);
        else
            hipLaunchKernelGGL(atsc3::resident_grow_kernel<2>, g, b, 0, st, a, t);
        return atsc3::launch_status();
    }
    if (f32 && direct)
        hipLaunchKernelGGL((atsc3::resident_team_kernel<0, true>), g, b, 0, st, a, t);
    else if (f32)
        hipLaunchKernelGGL((atsc3::resident_team_kernel<0, false>), g, b, 0, st, a, t);
    else if (msg_fmt == 1 && direct)
        hipLaunchKernelGGL((atsc3::resident_team_kernel<1, true>), g, b, 0, st, a, t);
    else if (msg_fmt == 1)
        hipLaunchKernelGGL((atsc3::resident_team_kernel<1, false>), g, b, 0, st, a, t);
    else if (direct)
        hipLaunchKernelGGL((atsc3::resident_team_kernel<2, true>), g, b, 0, st, a, t);
    else
        hipLaunchKernelGGL((atsc3::resident_team_kernel<2, false>), g, b, 0, st, a, t);
    return atsc3::launch_status();
}

// ---------------------------------------------------------------------------
// The same decode with no allocation at all: the workspace is `scratch`, a
// file-scope device array of RES_SCRATCH_BYTES that every GPU the library
// runs on holds until the process ends.  In host code the array's name is
// not a device address, so its address is taken once per device with
// hipGetSymbolAddress (with that device current) and only that address is
// handed to the kernels.  All calls on one GPU share the array: calls that
// overlap on one GPU race (enqueue them on one stream).  The kernels need
// no zeroed memory beyond the queue word, which a kernel resets each call.
// ---------------------------------------------------------------------------
#define RES_SCRATCH_BYTES (128ll << 20)
#define RES_SCRATCH_DEVICES 64

static __device__ __align__(256) unsigned char scratch[RES_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[RES_SCRATCH_DEVICES];

int32_t scratch_of_current_device(void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return ATSC3_E_HIP;
    if (dev < 0 || dev >= RES_SCRATCH_DEVICES) return ATSC3_E_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return ATSC3_E_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return ATSC3_OK;
}
}  // namespace

extern "C" int64_t atsc3_minsum_resident_noalloc_max_bytes(void) { return RES_SCRATCH_BYTES; }

extern "C" int32_t atsc3_minsum_resident_noalloc_dev(
    const float *llr, int32_t nblk, const int64_t *blk_off, const int32_t *blk_code,
    const void *codes, int32_t ncode, int32_t max_checks, int32_t max_global_rows,
    int32_t max_n, int32_t max_dc, int32_t max_dv, int32_t iters, float alpha, int32_t msg_fmt,
    const float *msg_scale, int32_t grid, uint8_t *bits, uint8_t *conv, int32_t *its,
    int32_t *bad, void *stream) {
    const int32_t rc = atsc3::resident_check_args(llr, nblk, blk_off, blk_code, codes, ncode,
                                                  max_checks, max_global_rows, max_n, max_dc,
                                                  max_dv, iters, alpha, msg_fmt, grid, bits,
                                                  conv, its, bad);
    if (rc != ATSC3_OK || nblk == 0) return rc;
    // the workspace this launch needs must fit the scratch array
    if (atsc3::WsLayout(grid, max_checks, max_global_rows, msg_fmt == 0).bytes >
        RES_SCRATCH_BYTES)
        return ATSC3_E_SIZE;
    void *ws = nullptr;
    const int32_t rs = scratch_of_current_device(&ws);
    if (rs != ATSC3_OK) return rs;
    return atsc3::resident_launch(llr, nblk, blk_off, blk_code, codes, ncode, max_checks,
                                  max_global_rows, max_n, max_dc, max_dv, iters, alpha, msg_fmt,
                                  msg_scale, grid, ws, RES_SCRATCH_BYTES, bits, conv, its, bad,
                                  nullptr, atsc3::as_stream(stream));
}
