// LDM successive interference cancellation: re-modulate decoded core-layer
// codewords and subtract them from the received cells.
//
// Follows the cancellation step of lab/e69_enh.py:86-121: for every core
// block that decoded (conv = 1) the transmitted symbols s are rebuilt from the
// codeword (bit interleave q = cw[lam_of_q], label k = sum_i q_i << (nb-1-i),
// s = points[k], lab/m6_bicm.py:228-236), the complex gain
//     g = vdot(s, y) / vdot(s, s) = sum conj(s) y / sum |s|^2
// is estimated per block in FP64, and r = y - g s.  Blocks that did not
// decode are passed through (r = y).  Every output cell is multiplied by
// out_scale (the enhanced layer's amplitude normalisation).
//
// Pass 1 (one workgroup per (chunk, block)) writes per-workgroup FP64
// partial sums; pass 2 reduces them in a fixed order (deterministic) and
// writes the residual.
#include "fec_common.h"

#define LDM_WG 256

namespace atsc3 {

__device__ __forceinline__ float2 remap_symbol(const uint8_t *cw, const int *lam_of_q,
                                               const float2 *pts, int s, int nb) {
    int k = 0;
    for (int i = 0; i < nb; ++i) k = (k << 1) | (cw[lam_of_q[s * nb + i]] & 1);
    return pts[k];
}

__global__ __launch_bounds__(LDM_WG) void ldm_dot_kernel(const float2 *__restrict__ y,
                                                         const uint8_t *__restrict__ cw,
                                                         const int *__restrict__ lam_of_q,
                                                         const float2 *__restrict__ pts, int nb,
                                                         int ncell, int n,
                                                         double *__restrict__ partial) {
    __shared__ double red[3][LDM_WG / 64];
    const int blk = blockIdx.y;
    const int s = blockIdx.x * LDM_WG + threadIdx.x;
    double re = 0.0, im = 0.0, den = 0.0;
    if (s < ncell) {
        const float2 sym = remap_symbol(cw + (size_t)blk * n, lam_of_q, pts, s, nb);
        const float2 v = y[(size_t)blk * ncell + s];
        // conj(sym) * v
        re = (double)sym.x * v.x + (double)sym.y * v.y;
        im = (double)sym.x * v.y - (double)sym.y * v.x;
        den = (double)sym.x * sym.x + (double)sym.y * sym.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
        den += __shfl_xor(den, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = re;
        red[1][threadIdx.x >> 6] = im;
        red[2][threadIdx.x >> 6] = den;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int w = 0; w < LDM_WG / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
            c += red[2][w];
        }
        double *p = partial + ((size_t)blk * gridDim.x + blockIdx.x) * 3;
        p[0] = a;
        p[1] = b;
        p[2] = c;
    }
}

__global__ __launch_bounds__(LDM_WG) void ldm_cancel_kernel(
    const float2 *__restrict__ y, const uint8_t *__restrict__ cw,
    const uint8_t *__restrict__ conv, const int *__restrict__ lam_of_q,
    const float2 *__restrict__ pts, int nb, int ncell, int n, const double *__restrict__ partial,
    int nwg, float out_scale, float2 *__restrict__ r, double *__restrict__ gain_out) {
    __shared__ double g_s[2];
    const int blk = blockIdx.y;
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int i = 0; i < nwg; ++i) {
            const double *p = partial + ((size_t)blk * nwg + i) * 3;
            a += p[0];
            b += p[1];
            c += p[2];
        }
        const bool use = conv[blk] != 0 && c > 0.0;
        g_s[0] = use ? a / c : 0.0;
        g_s[1] = use ? b / c : 0.0;
        if (gain_out && blockIdx.x == 0) {
            gain_out[2 * blk] = g_s[0];
            gain_out[2 * blk + 1] = g_s[1];
        }
    }
    __syncthreads();
    const int s = blockIdx.x * LDM_WG + threadIdx.x;
    if (s >= ncell) return;
    const float2 v = y[(size_t)blk * ncell + s];
    float2 out = v;
    if (conv[blk]) {
        const float2 sym = remap_symbol(cw + (size_t)blk * n, lam_of_q, pts, s, nb);
        const double gr = g_s[0], gi = g_s[1];
        out.x = (float)((double)v.x - (gr * sym.x - gi * sym.y));
        out.y = (float)((double)v.y - (gr * sym.y + gi * sym.x));
    }
    out.x *= out_scale;
    out.y *= out_scale;
    r[(size_t)blk * ncell + s] = out;
}

}  // namespace atsc3

extern "C" int32_t atsc3_ldm_ws_doubles(int32_t nblk, int32_t ncell) {
    if (nblk < 0 || ncell < 0) return ATSC3_E_ARG;
    return 3 * nblk * ((ncell + LDM_WG - 1) / LDM_WG);
}

extern "C" int32_t atsc3_ldm_cancel_dev(const float *cells, const uint8_t *cw,
                                        const uint8_t *conv, int32_t nblk, int32_t ncell,
                                        int32_t n, const int32_t *lam_of_q,
                                        const float *points, int32_t nbits, float out_scale,
                                        float *resid_out, double *gain_out, double *ws,
                                        void *stream) {
    if (nblk < 0 || ncell < 0 || n <= 0) return ATSC3_E_ARG;
    if (nblk == 0 || ncell == 0) return ATSC3_OK;
    if (!cells || !cw || !conv || !lam_of_q || !points || !resid_out || !ws) return ATSC3_E_ARG;
    if (nbits < 1 || nbits > 12 || ncell * nbits != n) return ATSC3_E_ARG;
    hipStream_t st = atsc3::as_stream(stream);
    const int nwg = (ncell + LDM_WG - 1) / LDM_WG;
    const float2 *y = reinterpret_cast<const float2 *>(cells);
    const float2 *p = reinterpret_cast<const float2 *>(points);
    hipLaunchKernelGGL(atsc3::ldm_dot_kernel, dim3(nwg, nblk), dim3(LDM_WG), 0, st, y, cw,
                       lam_of_q, p, nbits, ncell, n, ws);
    hipLaunchKernelGGL(atsc3::ldm_cancel_kernel, dim3(nwg, nblk), dim3(LDM_WG), 0, st, y, cw,
                       conv, lam_of_q, p, nbits, ncell, n, ws, nwg, out_scale,
                       reinterpret_cast<float2 *>(resid_out), gain_out);
    return atsc3::launch_status();
}
