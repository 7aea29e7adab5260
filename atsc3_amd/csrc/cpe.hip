// Cell-side steps in front of the demapper, gfx950:
//   * decision-directed common-phase-error correction per OFDM symbol
//     (lab/m6_cells.py:176-209; on the device in lab/m9_gpu.py:283-299);
//   * the channel SNR off the known +-1 dummy cells (lab/m16_margin.py:303-318);
//   (the per-part noise weights of the weighted-LLR path are in demap.hip).
//
// CPE: the pool is normalised by its RMS (one FP64 sum over the pool, split
// into at most 256 per-workgroup partials that every symbol workgroup re-adds
// in the same tree order), then each symbol -- a list of pool cells -- is handled by one
// workgroup that runs all `iters` rounds itself: symbols do not interact, so
// no grid-wide step is needed.  Per round: each cell's value is the
// normalised pool value divided by the gains found so far (in order); the
// cells that estimate the gain (a second per-symbol list: every cell, or the
// upstream fast path's every 4th pool cell, lab/m9_fast.py:667-678) take their
// hard decision, the nearest point of its region's alphabet (FP64 squared
// distances to the FP64 alphabet, first point on ties) or its known dummy
// value, and the round's
// gain c = sum(conj(h) z) / max(sum |h|^2, 1e-12) with FP64 sums reduced in a
// fixed tree order.  Output: the corrected cells (complex64) and each
// symbol's total gain (complex128).
#include "fec_common.h"

#define CPE_WG 256
#define CPE_SYM_WG 1024     // one symbol per workgroup: wide for single-frame latency
#define CPE_MAX_ITERS 8
#define CPE_MAX_POINTS 4096

namespace atsc3 {
namespace {

struct C64 {
    double re, im;
};

__device__ __forceinline__ C64 cdiv(C64 a, C64 b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

__device__ __forceinline__ C64 cmul(C64 a, C64 b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// Sum of a double over the workgroup (fixed order); every thread gets it.
__device__ double wg_sum_f64(double x, double *red) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)blockDim.x / 64; ++w) t += red[w];
    return t;
}

__global__ __launch_bounds__(CPE_WG) void cpe_power(const float2 *__restrict__ pool, int n,
                                                    double *__restrict__ partial) {
    __shared__ double red[CPE_WG / 64];
    double s = 0.0;
    for (int i = blockIdx.x * CPE_WG + threadIdx.x; i < n; i += gridDim.x * CPE_WG) {
        const float2 z = pool[i];
        s += (double)z.x * z.x + (double)z.y * z.y;
    }
    s = wg_sum_f64(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(CPE_SYM_WG) void cpe_symbols(
    const float2 *__restrict__ pool, int npool, const double *__restrict__ partial, int nparts,
    const int32_t *__restrict__ cells, const int32_t *__restrict__ seg,
    const int32_t *__restrict__ ecells, const int32_t *__restrict__ eseg,
    const uint8_t *__restrict__ region, const double2 *__restrict__ pts0, int m0,
    const double2 *__restrict__ pts1, int m1, const float2 *__restrict__ dummy, int iters,
    float2 *__restrict__ out, double *__restrict__ gains) {
    __shared__ double red[CPE_SYM_WG / 64];
    __shared__ C64 cs[CPE_MAX_ITERS];
    const int sym = blockIdx.x;
    // the pool power: one partial per thread (nparts <= CPE_WG <= CPE_SYM_WG), summed in
    // the same fixed tree order by every symbol workgroup
    const double tot =
        wg_sum_f64((int)threadIdx.x < nparts ? partial[threadIdx.x] : 0.0, red);
    const double rms = sqrt(tot / (double)npool);
    C64 ctot = {1.0, 0.0};
    for (int it = 0; it <= iters; ++it) {
        double nr = 0.0, ni = 0.0, dn = 0.0;
        // decision rounds walk the estimating cells, the last pass every cell
        const int32_t *list = it == iters ? cells : ecells;
        const int lo = it == iters ? seg[sym] : eseg[sym];
        const int hi = it == iters ? seg[sym + 1] : eseg[sym + 1];
        for (int k = lo + (int)threadIdx.x; k < hi; k += CPE_SYM_WG) {
            const int c = list[k];
            const float2 p = pool[c];
            C64 z = {(double)p.x / rms, (double)p.y / rms};
            for (int q = 0; q < it; ++q) z = cdiv(z, cs[q]);
            if (it == iters) {
                out[c] = make_float2((float)z.re, (float)z.im);
                continue;
            }
            const int r = region[c];
            double hr, hi_;
            if (r == 2 || r > 2) {
                // known dummy value (no dummy table: no contribution)
                const float2 d = dummy && r == 2 ? dummy[c] : make_float2(0.0f, 0.0f);
                hr = d.x;
                hi_ = d.y;
            } else {
                const double2 *pts = r == 0 ? pts0 : pts1;
                const int m = r == 0 ? m0 : m1;
                if (m == 0) continue;              // no alphabet for the region
                double best = INFINITY;
                int arg = 0;
                if (m > 0) {
                    for (int j = 0; j < m; ++j) {
                        const double2 w = pts[j];
                        const double dr = z.re - w.x, di = z.im - w.y;
                        const double d2 = dr * dr + di * di;
                        if (d2 < best) {
                            best = d2;
                            arg = j;
                        }
                    }
                    hr = pts[arg].x;
                    hi_ = pts[arg].y;
                } else {
                    // quadrant-rule alphabet given by its -m first-quadrant
                    // base points: the nearest point lies in z's quadrant
                    // (|z - p| >= |z~ - w| for the mirror images), and there
                    // (|zr| - wr)^2 is bit-identical to (zr - (+-wr))^2
                    const double ar = fabs(z.re), ai = fabs(z.im);
                    for (int j = 0; j < -m; ++j) {
                        const double2 w = pts[j];
                        const double dr = ar - w.x, di = ai - w.y;
                        const double d2 = dr * dr + di * di;
                        if (d2 < best) {
                            best = d2;
                            arg = j;
                        }
                    }
                    hr = z.re >= 0.0 ? pts[arg].x : -pts[arg].x;
                    hi_ = z.im >= 0.0 ? pts[arg].y : -pts[arg].y;
                }
            }
            // conj(h) * z
            nr += hr * z.re + hi_ * z.im;
            ni += hr * z.im - hi_ * z.re;
            dn += hr * hr + hi_ * hi_;
        }
        if (it == iters) break;
        nr = wg_sum_f64(nr, red);
        ni = wg_sum_f64(ni, red);
        dn = wg_sum_f64(dn, red);
        const double den = dn > 1e-12 ? dn : 1e-12;
        const C64 c = {nr / den, ni / den};
        ctot = cmul(ctot, c);
        __syncthreads();
        if (threadIdx.x == 0) cs[it] = c;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        gains[2 * sym] = ctot.re;
        gains[2 * sym + 1] = ctot.im;
    }
}

// One workgroup: a = sum(d.re s) / sum(s^2), noise = mean |d - a s|^2,
// out = {10 log10(a^2 / noise), valid}.
__global__ __launch_bounds__(1024) void dummy_snr(const float2 *__restrict__ cells, int n,
                                                  const uint8_t *__restrict__ bits,
                                                  double *__restrict__ out) {
    __shared__ double red[1024 / 64];
    double ds = 0.0, ss = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const double s = 1.0 - 2.0 * (double)(bits[i] & 1);
        ds += (double)cells[i].x * s;
        ss += s * s;
    }
    ds = wg_sum_f64(ds, red);
    ss = wg_sum_f64(ss, red);
    const double a = ds / ss;
    double e = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const double s = 1.0 - 2.0 * (double)(bits[i] & 1);
        const double er = (double)cells[i].x - a * s, ei = (double)cells[i].y;
        e += er * er + ei * ei;
    }
    e = wg_sum_f64(e, red);
    if (threadIdx.x == 0) {
        const double noise = e / (double)n;
        const bool ok = noise > 0.0 && a != 0.0;
        out[0] = ok ? 10.0 * log10(a * a / noise) : 0.0;
        out[1] = ok ? 1.0 : 0.0;
    }
}

}  // namespace
}  // namespace atsc3

using namespace atsc3;

extern "C" int32_t atsc3_cpe_ws_doubles(int32_t npool) {
    if (npool < 0) return ATSC3_E_ARG;
    const int g = (npool + CPE_WG - 1) / CPE_WG;
    return g < 256 ? (g > 0 ? g : 1) : 256;
}

extern "C" int32_t atsc3_cpe_dev(const float *pool, int32_t npool, const int32_t *cells,
                                 const int32_t *seg, const int32_t *est_cells,
                                 const int32_t *est_seg, int32_t nsym, const uint8_t *region,
                                 const double *pts0, int32_t m0, const double *pts1, int32_t m1,
                                 const float *dummy, int32_t iters, float *out, double *gains,
                                 double *ws, void *stream) {
    if (npool < 0 || nsym < 0 || iters < 0 || iters > CPE_MAX_ITERS) return ATSC3_E_ARG;
    if (m0 > CPE_MAX_POINTS || m1 > CPE_MAX_POINTS || m0 < -CPE_MAX_POINTS / 4 ||
        m1 < -CPE_MAX_POINTS / 4)
        return ATSC3_E_SIZE;
    if (npool == 0 || nsym == 0) return ATSC3_OK;
    if ((est_cells == nullptr) != (est_seg == nullptr)) return ATSC3_E_ARG;
    if (!est_cells) {
        est_cells = cells;
        est_seg = seg;
    }
    if (!pool || !cells || !seg || !region || !out || !gains || !ws) return ATSC3_E_ARG;
    if ((m0 != 0 && !pts0) || (m1 != 0 && !pts1)) return ATSC3_E_ARG;
    hipStream_t st = as_stream(stream);
    const int np = atsc3_cpe_ws_doubles(npool);
    hipLaunchKernelGGL(cpe_power, dim3(np), dim3(CPE_WG), 0, st,
                       reinterpret_cast<const float2 *>(pool), npool, ws);
    hipLaunchKernelGGL(cpe_symbols, dim3(nsym), dim3(CPE_SYM_WG), 0, st,
                       reinterpret_cast<const float2 *>(pool), npool, ws, np, cells, seg,
                       est_cells, est_seg, region,
                       reinterpret_cast<const double2 *>(pts0), m0,
                       reinterpret_cast<const double2 *>(pts1), m1,
                       reinterpret_cast<const float2 *>(dummy), iters,
                       reinterpret_cast<float2 *>(out), gains);
    return launch_status();
}

extern "C" int32_t atsc3_dummy_snr_dev(const float *cells, int32_t n, const uint8_t *bits,
                                       double *out, void *stream) {
    if (n < 0 || !out) return ATSC3_E_ARG;
    if (n > 0 && (!cells || !bits)) return ATSC3_E_ARG;
    hipLaunchKernelGGL(dummy_snr, dim3(1), dim3(1024), 0, as_stream(stream),
                       reinterpret_cast<const float2 *>(cells), n, bits, out);
    return launch_status();
}

