/* minsum_oracle.c -- CPU oracle for the FP32 normalized min-sum decoder.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this
 * file; it is the checker the HIP decoder is compared against (tests/,
 * the smoke check and the CPU-baseline leg of bench.py).
 *
 * It restates the semantics of the upstream compiled kernel
 * lab/ldpc_kernel.c:65-146 (ldpc_minsum_f32) -- same C ABI, same results --
 * organised as one function per phase so each rule is visible on its own:
 *
 *   syndrome_count   lab/ldpc_kernel.c:94-101   XOR of (tot[v] < 0) per check,
 *                                               taken BEFORE each update, so a
 *                                               clean block retires at its = 0
 *   check_update     lab/ldpc_kernel.c:115-134  v = tot - m (FP32); min1/min2
 *                                               by strict '<' in ascending
 *                                               column order; only the FIRST
 *                                               edge attaining min1 takes min2;
 *                                               m = +-(min * alpha), negated
 *                                               iff parity ^ (v < 0)
 *   variable_update  lab/ldpc_kernel.c:136-142  s = 0.0f; s += m in ascending
 *                                               check order (vedge); tot = lam+s
 *   decode_one       lab/ldpc_kernel.c:82-143   its = iters, conv = 0 for a
 *                                               block that never converges,
 *                                               bad = its last unsatisfied
 *                                               count
 *
 * Build with -O3 -ffp-contract=off and no fast-math, as the upstream kernel
 * is built (lab/build_ldpc_kernel.py:36): every FP32 operation rounds once,
 * in source order.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define ORACLE_MAX_CHECK_DEGREE 128   /* lab/ldpc_kernel.c:48 */

/* ORACLE_PLANT != 0 builds a deliberately wrong variant used only as a
 * negative control by the tests (make plant). */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

typedef struct {
    int32_t n, nc, nedges;
    const int32_t *cptr, *cidx, *vptr, *vedge;
} tanner_t;

int32_t oracle_minsum_abi(void) { return 1; }

static int32_t syndrome_count(const tanner_t *g, const float *tot)
{
    int32_t bad = 0;
    for (int32_t c = 0; c < g->nc; ++c) {
        unsigned par = 0;
        for (int32_t e = g->cptr[c]; e < g->cptr[c + 1]; ++e)
            par ^= (unsigned)(tot[g->cidx[e]] < 0.0f);
        bad += (int32_t)par;
    }
    return bad;
}

static void check_update(const tanner_t *g, const float *tot, float *m,
                         float alpha)
{
    float v[ORACLE_MAX_CHECK_DEGREE];
    for (int32_t c = 0; c < g->nc; ++c) {
        const int32_t lo = g->cptr[c], hi = g->cptr[c + 1];
        float lo1 = INFINITY, lo2 = INFINITY;
        int32_t first = lo;
        unsigned parity = 0;
        for (int32_t e = lo; e < hi; ++e) {
            const float x = tot[g->cidx[e]] - m[e];
            const float a = fabsf(x);
            v[e - lo] = x;
            parity ^= (unsigned)(x < 0.0f);
            if (a < lo1) {
                lo2 = lo1;
                lo1 = a;
                first = e;
            } else if (a < lo2) {
                lo2 = a;
            }
        }
        for (int32_t e = lo; e < hi; ++e) {
            float out = (e == first ? lo2 : lo1) * alpha;
            if (parity ^ (unsigned)(v[e - lo] < 0.0f))
                out = -out;
            m[e] = out;
        }
    }
}

static void variable_update(const tanner_t *g, const float *lam,
                            const float *m, float *tot)
{
    for (int32_t i = 0; i < g->n; ++i) {
        float s = 0.0f;
#if ORACLE_PLANT == 1   /* negative control: descending check order */
        for (int32_t k = g->vptr[i + 1] - 1; k >= g->vptr[i]; --k)
            s += m[g->vedge[k]];
#else
        for (int32_t k = g->vptr[i]; k < g->vptr[i + 1]; ++k)
            s += m[g->vedge[k]];
#endif
        tot[i] = lam[i] + s;
    }
}

static void hard_bits(const float *tot, int32_t n, uint8_t *bits)
{
    for (int32_t i = 0; i < n; ++i)
        bits[i] = (uint8_t)(tot[i] < 0.0f);
}

static void decode_one(const tanner_t *g, const float *lam, int32_t iters,
                       float alpha, float *tot, float *m, uint8_t *bits,
                       uint8_t *conv, int32_t *its, int32_t *bad)
{
    memcpy(tot, lam, sizeof(float) * (size_t)g->n);
    for (int32_t e = 0; e < g->nedges; ++e)
        m[e] = 0.0f;
    *conv = 0;
    *its = iters;
    *bad = -1;
    for (int32_t it = 0;; ++it) {
#if ORACLE_PLANT == 2   /* negative control: update before the syndrome */
        if (it == 0 && iters > 0) {
            check_update(g, tot, m, alpha);
            variable_update(g, lam, m, tot);
        }
#endif
        const int32_t unsat = syndrome_count(g, tot);
        if (unsat == 0 || it == iters) {
            hard_bits(tot, g->n, bits);
            if (unsat == 0) {
                *conv = 1;
                *its = it;
                *bad = 0;
            } else {
                *bad = unsat;
            }
            return;
        }
        check_update(g, tot, m, alpha);
        variable_update(g, lam, m, tot);
    }
}

/* Same argument list and meaning as lab/ldpc_kernel.c:65-71. */
int32_t oracle_minsum_f32(const float *llr, int32_t nblk, int32_t n,
                          int32_t nc, int32_t nedges, const int32_t *cptr,
                          const int32_t *cidx, const int32_t *vptr,
                          const int32_t *vedge, int32_t iters, float alpha,
                          uint8_t *bits, uint8_t *conv, int32_t *its,
                          int32_t *bad, float *work)
{
    const tanner_t g = {n, nc, nedges, cptr, cidx, vptr, vedge};
    for (int32_t c = 0; c < nc; ++c)
        if (cptr[c + 1] - cptr[c] > ORACLE_MAX_CHECK_DEGREE)
            return -1;
    for (int32_t b = 0; b < nblk; ++b)
        decode_one(&g, llr + (size_t)b * (size_t)n, iters, alpha, work,
                   work + n, bits + (size_t)b * (size_t)n, conv + b, its + b,
                   bad + b);
    return 0;
}
