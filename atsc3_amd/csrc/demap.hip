// Max-log NUC demapper with fused bit de-interleave, gfx950.
//
// Maths (lab/m6_bicm.py:213-225, max-log, llr > 0 means bit 0):
//   LLR_i(z) = w * (min_{p: y_i = 1} |z - p|^2 - min_{p: y_i = 0} |z - p|^2) / sigma2
//
// Every A/322 alphabet handled here (QPSK and the 2D NUCs of Annex C) is
// built by the quadrant rule x = [w, -conj w, conj w, -w] from M/4 base
// points w in the closed first quadrant, with label bit y0 selecting the sign
// of the imaginary part and y1 the sign of the real part (the host checks
// this before it hands the base points over).  Hence, with z~ = (|Re z|,
// |Im z|) and per base point
//     A = (|zr| - wr)^2,  B = (|zr| + wr)^2,  C = (|zi| - wi)^2,  D = (|zi| + wi)^2,
//   * the nearest point with a given index k of w is at distance A + C, so for
//     label bits y2..y_{nb-1} (the bits of k) the subset minima are minima of
//     A + C over k;
//   * for y0 the subset with the same imaginary sign as z has minimum
//     min_k (A + C) and the other min_k (A + D); likewise y1 with B + C.
// That is the exhaustive max-log over all M points, evaluated exactly (in
// real arithmetic) with M/4 distance evaluations: A + C by direct
// differences (no |p|^2 - 2 Re(z p*) expansion, which loses relative
// accuracy on the closely spaced 256-NUC points), A + D = A + C + 4|zi|wi and
// B + C = A + C + 4|zr|wr as one fma each (sums of non-negative terms).
//
// Output position of q-stream bit q = cell*nb + i is lam_of_q[q]
// (lab/m6_bicm.py:331-335), so the LLRs land in codeword order.
//
// Cell gather (atsc3_demap_gather_dev): cell s of block b is read from a
// pool through an index map instead of cells[b * ncell + s], which fuses the
// time de-interleaver's FEC-block gather into the demap read:
//   * TABLE: pool[idx[b * ncell + s]] -- the hybrid time interleaver's
//     twisted-block map (A/322 7.1.5.4, virtual FEC Blocks included) as the
//     upstream gathers it (lab/m9_fast.py:1087-1093);
//   * CTI: output cell i = first + b * ncell + s of the convolutional time
//     de-interleaver reads received cell q = i + Nrows * ((start_row + i) mod
//     Nrows) (A/322 9.3.9.1; lab/m10_cti.py:108-146), held in a ring buffer
//     of `ring` cells at q mod ring (lab/m44_ldm.py:1289-1304).
// The CSI weight of a cell is read through the same index.
//
// sigma2 == NULL: per block sigma2 = max(mean(w * d2min), 1e-9) with d2min
// = min_p |z - p|^2 (lab/m6_bicm.py:219-220), reduced in FP64 in a fixed
// order (per-workgroup tree, then sequential over workgroups), then a second
// pass divides the block's LLRs by it.
#include "fec_common.h"

#define DEMAP_WG 256
// At least 4 waves per SIMD (<= 128 VGPRs): a demap wave then fits beside
// the 4 waves of a resident decode workgroup (<= 96 VGPRs each, 512 per
// SIMD), so a demap overlapped with a decode on another stream runs on
// every compute unit in the decode's idle issue slots instead of waiting
// for its tail.
#define DEMAP_MIN_WAVES 4

namespace atsc3 {

enum { GATHER_NONE = 0, GATHER_TABLE = 1, GATHER_CTI = 2 };

struct Gather {
    const int32_t *idx;
    long long first, ring;
    int nrows, start_row;
};

template <int GM>
__device__ __forceinline__ size_t cell_of(const Gather &g, int blk, int ncell, int s) {
    if constexpr (GM == GATHER_TABLE) {
        return (size_t)(uint32_t)g.idx[(size_t)blk * ncell + s];
    } else if constexpr (GM == GATHER_CTI) {
        const long long i = g.first + (long long)blk * ncell + s;
        const long long q = i + (long long)g.nrows * ((g.start_row + i) % g.nrows);
        return (size_t)(q % g.ring);
    } else {
        return (size_t)blk * ncell + s;
    }
}

// Max-log differences of one cell z against the NW = 2^(NB-2) base points w
// (in LDS): diff[i] = min over y_i = 1 minus min over y_i = 0 (before the
// CSI weight and sigma2), dmin = the distance to the nearest point.
//
// Per base point p = (px, py) >= 0 and z~ = (ar, ai) = (|Re z|, |Im z|):
//   ac = (ar - px)^2 + (ai - py)^2          the point of z's quadrant
//   ad = (ar - px)^2 + (ai + py)^2 = ac + 4 ai py   (imaginary sign flipped)
//   bc = (ar + px)^2 + (ai - py)^2 = ac + 4 ar px   (real sign flipped)
// -- (a + b)^2 = (a - b)^2 + 4ab, and every term is >= 0, so each distance
// keeps its relative accuracy (no cancellation).  The minima of ac over the
// points whose index bit j is 0 or 1 (label bits y2..y_{nb-1}) come from the
// row and column minima of the index split as k = h * L + l: bits of h from
// the rows R[h] = min_l ac, bits of l from the columns C[l] = min_h ac.
template <int NB>
__device__ __forceinline__ void demap_cell(float2 z, const float2 *w, float (&diff)[NB],
                                           float &dmin_out) {
    constexpr int NW = 1 << (NB - 2);
    constexpr int LB = (NB - 2) / 2;             // low index bits (columns)
    constexpr int L = 1 << LB, H = NW / L;
    const float ar = fabsf(z.x), ai = fabsf(z.y);
    const float ar4 = 4.0f * ar, ai4 = 4.0f * ai;  // exact
    float opp_im = INFINITY, opp_re = INFINITY;
    float R[H], C[L];
#pragma unroll
    for (int h = 0; h < H; ++h) R[h] = INFINITY;
#pragma unroll
    for (int l = 0; l < L; ++l) C[l] = INFINITY;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const float2 p = w[k];
        const float dr = ar - p.x, di = ai - p.y;
        const float ac = __builtin_fmaf(dr, dr, di * di);
        opp_im = fminf(opp_im, __builtin_fmaf(ai4, p.y, ac));
        opp_re = fminf(opp_re, __builtin_fmaf(ar4, p.x, ac));
        R[k / L] = fminf(R[k / L], ac);
        C[k % L] = fminf(C[k % L], ac);
        // two rows of base points at a time: the compiler would otherwise
        // load every point into registers up front (128 VGPRs for 256-NUC);
        // the rows' results are inputs of an empty statement that also
        // orders memory, so their arithmetic stays above it and the next
        // rows' loads below (inputs only: the values stay known to the
        // compiler, which would otherwise re-quiet them before every fminf)
        if (k % (2 * L) == 2 * L - 1) {
            asm volatile("" : : "v"(opp_im), "v"(opp_re), "v"(R[k / L]), "v"(R[k / L - 1])
                         : "memory");
#pragma unroll
            for (int l = 0; l < L; ++l) asm volatile("" : : "v"(C[l]));
        }
    }
    float dmin = R[0];
#pragma unroll
    for (int h = 1; h < H; ++h) dmin = fminf(dmin, R[h]);
    // y0: imaginary sign; y1: real sign (see header)
    diff[0] = z.y >= 0.0f ? opp_im - dmin : dmin - opp_im;
    diff[1] = z.x >= 0.0f ? opp_re - dmin : dmin - opp_re;
    // label bit i (2..NB-1) is index bit j = NB - 1 - i
#pragma unroll
    for (int i = 2; i < NB; ++i) {
        const int j = NB - 1 - i;
        float m0 = INFINITY, m1 = INFINITY;
        if (j >= LB) {
#pragma unroll
            for (int h = 0; h < H; ++h) {
                if ((h >> (j - LB)) & 1) m1 = fminf(m1, R[h]);
                else m0 = fminf(m0, R[h]);
            }
        } else {
#pragma unroll
            for (int l = 0; l < L; ++l) {
                if ((l >> j) & 1) m1 = fminf(m1, C[l]);
                else m0 = fminf(m0, C[l]);
            }
        }
        diff[i] = m1 - m0;
    }
    dmin_out = dmin;
}

template <int NB, int GM>
__global__ __launch_bounds__(DEMAP_WG, DEMAP_MIN_WAVES) void demap_kernel(
    const float2 *__restrict__ cells, const float *__restrict__ csi,
    const double *__restrict__ sigma2, int ncell, const float2 *__restrict__ base,
    const int *__restrict__ lam_of_q, float *__restrict__ llr, double *__restrict__ partial,
    Gather g, float *__restrict__ d2_out) {
    constexpr int NW = 1 << (NB - 2);
    __shared__ float2 w[NW];
    __shared__ double red[DEMAP_WG / 64];
    const int blk = blockIdx.y;
    const int tid = threadIdx.x;
    for (int k = tid; k < NW; k += DEMAP_WG) w[k] = base[k];
    __syncthreads();
    const int s = blockIdx.x * DEMAP_WG + tid;
    const int N = ncell * NB;
    double d2w = 0.0;
    if (s < ncell) {
        const size_t ci = cell_of<GM>(g, blk, ncell, s);
        const float2 z = cells[ci];
        float diff[NB], dmin;
        demap_cell<NB>(z, w, diff, dmin);
        const float wgt = csi ? csi[ci] : 1.0f;
        const int *lq = lam_of_q + (size_t)s * NB;
        float *out = llr + (size_t)blk * N;
        if (sigma2) {
            const float s2 = (float)sigma2[blk];
#pragma unroll
            for (int i = 0; i < NB; ++i) out[lq[i]] = (diff[i] * wgt) / s2;
        } else {
#pragma unroll
            for (int i = 0; i < NB; ++i) out[lq[i]] = diff[i] * wgt;
            d2w = (double)wgt * (double)dmin;
            if (d2_out) d2_out[(size_t)blk * ncell + s] = dmin;
        }
    }
    if (!sigma2) {
        for (int off = 32; off > 0; off >>= 1) d2w += __shfl_xor(d2w, off, 64);
        if ((tid & 63) == 0) red[tid >> 6] = d2w;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int i = 0; i < DEMAP_WG / 64; ++i) t += red[i];
            partial[(size_t)blk * gridDim.x + blockIdx.x] = t;
        }
    }
}

// One workgroup per block: sigma2 = max(sum(partials) / ncell, 1e-9), then
// llr /= sigma2 over the block's N values.
__global__ __launch_bounds__(DEMAP_WG) void demap_scale_kernel(
    float *__restrict__ llr, int N, int ncell, const double *__restrict__ partial, int nwg,
    double *__restrict__ sigma2_out) {
    __shared__ float s2f;
    const int blk = blockIdx.x;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < nwg; ++i) t += partial[(size_t)blk * nwg + i];
        double s2 = t / (double)ncell;
        s2 = s2 > 1e-9 ? s2 : 1e-9;
        if (sigma2_out) sigma2_out[blk] = s2;
        s2f = (float)s2;
    }
    __syncthreads();
    const float d = s2f;
    float *out = llr + (size_t)blk * N;
    for (int i = threadIdx.x; i < N; i += DEMAP_WG) out[i] = out[i] / d;
}

// Every (code, constellation) group of a frame in one launch: workgroup
// (x, y) takes cells x*256.. of the frame's block y, found in the group
// table by its first-block index (groups in increasing blk0); sigma2 given
// per block, no gather.  The arithmetic is demap_kernel's.
struct DemapGroup {                 // mirrors struct atsc3_demap_group
    const float2 *cells;
    const float *csi;
    const double *sigma2;
    const float2 *points;
    const int32_t *lam_of_q;
    float *llr;
    int32_t ncell, nbits, nblk, blk0;
};
static_assert(sizeof(DemapGroup) == 64, "demap group descriptor layout");

template <int NB>
__device__ __forceinline__ void demap_group_block(const DemapGroup &G, int blk, float2 *w) {
    constexpr int NW = 1 << (NB - 2);
    for (int k = threadIdx.x; k < NW; k += DEMAP_WG) w[k] = G.points[k];
    __syncthreads();
    const int s = blockIdx.x * DEMAP_WG + threadIdx.x;
    if (s >= G.ncell) return;
    const size_t ci = (size_t)blk * G.ncell + s;
    float diff[NB], dmin;
    demap_cell<NB>(G.cells[ci], w, diff, dmin);
    const float wgt = G.csi ? G.csi[ci] : 1.0f;
    const int *lq = G.lam_of_q + (size_t)s * NB;
    float *out = G.llr + (size_t)blk * G.ncell * NB;
    const float s2 = (float)G.sigma2[blk];
#pragma unroll
    for (int i = 0; i < NB; ++i) out[lq[i]] = (diff[i] * wgt) / s2;
}

__global__ __launch_bounds__(DEMAP_WG, DEMAP_MIN_WAVES) void demap_groups_kernel(
    const DemapGroup *__restrict__ grp, int ngroups) {
    __shared__ float2 w[64];
    const int gb = blockIdx.y;
    int gi = 0;
    while (gi + 1 < ngroups && grp[gi + 1].blk0 <= gb) ++gi;
    const DemapGroup G = grp[gi];
    const int blk = gb - G.blk0;
    if (blk >= G.nblk || (int)blockIdx.x * DEMAP_WG >= G.ncell) return;
    switch (G.nbits) {
        case 2: demap_group_block<2>(G, blk, w); break;
        case 4: demap_group_block<4>(G, blk, w); break;
        case 6: demap_group_block<6>(G, blk, w); break;
        case 8: demap_group_block<8>(G, blk, w); break;
        default: break;
    }
}

__global__ void copy_sigma2_kernel(const double *src, double *dst, int nblk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nblk) dst[i] = src[i];
}

template <int NB, int GM>
static void launch_demap(const float *cells, const float *csi, const double *sigma2,
                         int nblk, int ncell, const float *points,
                         const int32_t *lam_of_q, float *llr, double *ws, const Gather &g,
                         hipStream_t st, float *d2_out = nullptr) {
    const int nwg = (ncell + DEMAP_WG - 1) / DEMAP_WG;
    hipLaunchKernelGGL((demap_kernel<NB, GM>), dim3(nwg, nblk), dim3(DEMAP_WG), 0, st,
                       reinterpret_cast<const float2 *>(cells), csi, sigma2, ncell,
                       reinterpret_cast<const float2 *>(points), lam_of_q, llr, ws, g, d2_out);
}

// s2[p] = max(mean of d2 over the cells of part p, 1e-6) as float; the cells
// of part p are order[seg[p] .. seg[p+1]).  One workgroup per part.
__global__ __launch_bounds__(DEMAP_WG) void part_noise(const float *__restrict__ d2,
                                                     const int32_t *__restrict__ order,
                                                     const int32_t *__restrict__ seg,
                                                     float *__restrict__ s2) {
    __shared__ double red[DEMAP_WG / 64];
    const int p = blockIdx.x;
    const int lo = seg[p], hi = seg[p + 1];
    double s = 0.0;
    for (int k = lo + (int)threadIdx.x; k < hi; k += DEMAP_WG) s += (double)d2[order[k]];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        s = 0.0;
        for (int w = 0; w < DEMAP_WG / 64; ++w) s += red[w];
        const int cnt = hi - lo > 1 ? hi - lo : 1;
        const double v = s / (double)cnt;
        s2[p] = (float)(v > 1e-6 ? v : 1e-6);
    }
}

// LLRs of block cell (b, s) (codeword positions lam_of_q[s * nb + i]) divided
// by the noise variance of the cell's part.
__global__ __launch_bounds__(DEMAP_WG) void part_rescale(float *__restrict__ llr, int nblk,
                                                       int ncell, int nb,
                                                       const int32_t *__restrict__ lam_of_q,
                                                       const int32_t *__restrict__ part,
                                                       const float *__restrict__ s2) {
    const int b = blockIdx.y;
    const int s = blockIdx.x * DEMAP_WG + threadIdx.x;
    if (b >= nblk || s >= ncell) return;
    const float v = s2[part[(size_t)b * ncell + s]];
    float *row = llr + (size_t)b * ncell * nb;
    for (int i = 0; i < nb; ++i) {
        float &x = row[lam_of_q[s * nb + i]];
        x = x / v;
    }
}



template <int GM>
static int32_t demap_any(const float *cells, const float *csi, const double *sigma2,
                         int nblk, int ncell, const float *points, int nbits,
                         const int32_t *lam_of_q, float *llr_out, double *sigma2_out,
                         double *ws, const Gather &g, hipStream_t st) {
    // `points` holds the M/4 first-quadrant base points (host-checked).
    switch (nbits) {
        case 2: launch_demap<2, GM>(cells, csi, sigma2, nblk, ncell, points, lam_of_q, llr_out, ws, g, st); break;
        case 4: launch_demap<4, GM>(cells, csi, sigma2, nblk, ncell, points, lam_of_q, llr_out, ws, g, st); break;
        case 6: launch_demap<6, GM>(cells, csi, sigma2, nblk, ncell, points, lam_of_q, llr_out, ws, g, st); break;
        case 8: launch_demap<8, GM>(cells, csi, sigma2, nblk, ncell, points, lam_of_q, llr_out, ws, g, st); break;
        default: return ATSC3_E_ARG;
    }
    if (hipGetLastError() != hipSuccess) return ATSC3_E_HIP;
    if (!sigma2) {
        const int nwg = (ncell + DEMAP_WG - 1) / DEMAP_WG;
        hipLaunchKernelGGL(demap_scale_kernel, dim3(nblk), dim3(DEMAP_WG), 0, st, llr_out,
                           ncell * nbits, ncell, ws, nwg, sigma2_out);
    } else if (sigma2_out) {
        hipLaunchKernelGGL(copy_sigma2_kernel, dim3((nblk + 255) / 256), dim3(256), 0, st,
                           sigma2, sigma2_out, nblk);
    }
    return launch_status();
}

}  // namespace atsc3

extern "C" int32_t atsc3_demap_groups_dev(const void *groups, int32_t ngroups,
                                          int32_t nblk_total, int32_t max_ncell, void *stream) {
    if (ngroups < 0 || nblk_total < 0 || max_ncell < 0) return ATSC3_E_ARG;
    if (ngroups == 0 || nblk_total == 0 || max_ncell == 0) return ATSC3_OK;
    if (!groups || nblk_total > 65535) return ATSC3_E_ARG;
    hipLaunchKernelGGL(atsc3::demap_groups_kernel,
                       dim3((max_ncell + DEMAP_WG - 1) / DEMAP_WG, nblk_total), dim3(DEMAP_WG),
                       0, atsc3::as_stream(stream),
                       static_cast<const atsc3::DemapGroup *>(groups), ngroups);
    return atsc3::launch_status();
}

extern "C" int32_t atsc3_demap_ws_doubles(int32_t nblk, int32_t ncell) {
    if (nblk < 0 || ncell < 0) return ATSC3_E_ARG;
    return nblk * ((ncell + DEMAP_WG - 1) / DEMAP_WG);
}

extern "C" int32_t atsc3_demap_dev(const float *cells, const float *csi,
                                   const double *sigma2, int32_t nblk, int32_t ncell,
                                   const float *points, int32_t nbits,
                                   const int32_t *lam_of_q, float *llr_out,
                                   double *sigma2_out, double *ws, void *stream) {
    if (nblk < 0 || ncell < 0) return ATSC3_E_ARG;
    if (nblk == 0 || ncell == 0) return ATSC3_OK;
    if (!cells || !points || !lam_of_q || !llr_out) return ATSC3_E_ARG;
    if (!sigma2 && !ws) return ATSC3_E_ARG;
    const atsc3::Gather g{nullptr, 0, 0, 0, 0};
    return atsc3::demap_any<atsc3::GATHER_NONE>(cells, csi, sigma2, nblk, ncell, points, nbits,
                                                lam_of_q, llr_out, sigma2_out, ws, g,
                                                atsc3::as_stream(stream));
}

extern "C" int32_t atsc3_demap_raw_dev(const float *cells, const float *csi, int32_t nblk,
                                       int32_t ncell, const float *points, int32_t nbits,
                                       const int32_t *lam_of_q, float *llr_out, float *d2_out,
                                       double *ws, void *stream) {
    if (nblk < 0 || ncell < 0) return ATSC3_E_ARG;
    if (nblk == 0 || ncell == 0) return ATSC3_OK;
    if (!cells || !points || !lam_of_q || !llr_out || !d2_out || !ws) return ATSC3_E_ARG;
    const atsc3::Gather g{nullptr, 0, 0, 0, 0};
    hipStream_t st = atsc3::as_stream(stream);
    // sigma2 NULL: the kernel writes the unscaled differences, d2min per cell
    // and its FP64 partial sums (unused here); no scale pass follows.
    switch (nbits) {
        case 2: atsc3::launch_demap<2, atsc3::GATHER_NONE>(cells, csi, nullptr, nblk, ncell, points, lam_of_q, llr_out, ws, g, st, d2_out); break;
        case 4: atsc3::launch_demap<4, atsc3::GATHER_NONE>(cells, csi, nullptr, nblk, ncell, points, lam_of_q, llr_out, ws, g, st, d2_out); break;
        case 6: atsc3::launch_demap<6, atsc3::GATHER_NONE>(cells, csi, nullptr, nblk, ncell, points, lam_of_q, llr_out, ws, g, st, d2_out); break;
        case 8: atsc3::launch_demap<8, atsc3::GATHER_NONE>(cells, csi, nullptr, nblk, ncell, points, lam_of_q, llr_out, ws, g, st, d2_out); break;
        default: return ATSC3_E_ARG;
    }
    return atsc3::launch_status();
}

extern "C" int32_t atsc3_demap_gather_dev(const float *pool, const float *pool_csi,
                                          const atsc3_gather *gather, const double *sigma2,
                                          int32_t nblk, int32_t ncell, const float *points,
                                          int32_t nbits, const int32_t *lam_of_q,
                                          float *llr_out, double *sigma2_out, double *ws,
                                          void *stream) {
    if (nblk < 0 || ncell < 0 || !gather) return ATSC3_E_ARG;
    if (nblk == 0 || ncell == 0) return ATSC3_OK;
    if (!pool || !points || !lam_of_q || !llr_out) return ATSC3_E_ARG;
    if (!sigma2 && !ws) return ATSC3_E_ARG;
    atsc3::Gather g{gather->idx, gather->first, gather->ring, gather->nrows, gather->start_row};
    hipStream_t st = atsc3::as_stream(stream);
    if (gather->mode == ATSC3_GATHER_TABLE) {
        if (!g.idx) return ATSC3_E_ARG;
        return atsc3::demap_any<atsc3::GATHER_TABLE>(pool, pool_csi, sigma2, nblk, ncell, points,
                                                     nbits, lam_of_q, llr_out, sigma2_out, ws, g,
                                                     st);
    }
    if (gather->mode == ATSC3_GATHER_CTI) {
        if (g.nrows <= 0 || g.ring <= 0 || g.first < 0 || g.start_row < 0 ||
            g.start_row >= g.nrows)
            return ATSC3_E_ARG;
        return atsc3::demap_any<atsc3::GATHER_CTI>(pool, pool_csi, sigma2, nblk, ncell, points,
                                                   nbits, lam_of_q, llr_out, sigma2_out, ws, g,
                                                   st);
    }
    return ATSC3_E_ARG;
}

extern "C" int32_t atsc3_demap_weighted_dev(
    const float *pool, const float *pool_csi, const atsc3_gather *gather, int32_t nblk,
    int32_t ncell, const float *points, int32_t nbits, const int32_t *lam_of_q,
    const int32_t *part, const int32_t *order, const int32_t *seg, int32_t nparts,
    float *llr_out, float *s2_out, float *d2_ws, double *ws, void *stream) {
    if (nblk < 0 || ncell < 0 || nparts < 0) return ATSC3_E_ARG;
    if (nblk == 0 || ncell == 0) return ATSC3_OK;
    if (!pool || !points || !lam_of_q || !part || !order || !seg || !llr_out || !s2_out ||
        !d2_ws || !ws || nparts == 0)
        return ATSC3_E_ARG;
    atsc3::Gather g{nullptr, 0, 0, 0, 0};
    int mode = atsc3::GATHER_NONE;
    if (gather) {
        g = atsc3::Gather{gather->idx, gather->first, gather->ring, gather->nrows,
                          gather->start_row};
        mode = gather->mode;
        if (mode == ATSC3_GATHER_TABLE && !g.idx) return ATSC3_E_ARG;
        if (mode == ATSC3_GATHER_CTI && (g.nrows <= 0 || g.ring <= 0 || g.first < 0 ||
                                         g.start_row < 0 || g.start_row >= g.nrows))
            return ATSC3_E_ARG;
        if (mode != ATSC3_GATHER_TABLE && mode != ATSC3_GATHER_CTI) return ATSC3_E_ARG;
    }
    hipStream_t st = atsc3::as_stream(stream);
#define W_LAUNCH(NB, GM)                                                                      \
    atsc3::launch_demap<NB, GM>(pool, pool_csi, nullptr, nblk, ncell, points, lam_of_q,        \
                                llr_out, ws, g, st, d2_ws)
#define W_MODES(NB)                                                                           \
    if (mode == atsc3::GATHER_TABLE)                                                          \
        W_LAUNCH(NB, atsc3::GATHER_TABLE);                                                    \
    else if (mode == atsc3::GATHER_CTI)                                                       \
        W_LAUNCH(NB, atsc3::GATHER_CTI);                                                      \
    else                                                                                      \
        W_LAUNCH(NB, atsc3::GATHER_NONE);
    switch (nbits) {
        case 2: W_MODES(2) break;
        case 4: W_MODES(4) break;
        case 6: W_MODES(6) break;
        case 8: W_MODES(8) break;
        default: return ATSC3_E_ARG;
    }
#undef W_MODES
#undef W_LAUNCH
    if (hipGetLastError() != hipSuccess) return ATSC3_E_HIP;
    hipLaunchKernelGGL(atsc3::part_noise, dim3(nparts), dim3(DEMAP_WG), 0, st, d2_ws, order, seg,
                       s2_out);
    hipLaunchKernelGGL(atsc3::part_rescale, dim3((ncell + DEMAP_WG - 1) / DEMAP_WG, nblk),
                       dim3(DEMAP_WG), 0, st, llr_out, nblk, ncell, nbits, lam_of_q, part,
                       s2_out);
    return atsc3::launch_status();
}

namespace atsc3 {
// out[b * ncell + s] = pool[cell(b, s)] through the gather map.
template <int GM>
__global__ __launch_bounds__(DEMAP_WG) void gather_cells(const float2 *__restrict__ pool,
                                                         Gather g, int ncell,
                                                         float2 *__restrict__ out) {
    const int b = blockIdx.y;
    const int s = blockIdx.x * DEMAP_WG + threadIdx.x;
    if (s < ncell) out[(size_t)b * ncell + s] = pool[cell_of<GM>(g, b, ncell, s)];
}
}  // namespace atsc3

extern "C" int32_t atsc3_gather_cells_dev(const float *pool, const atsc3_gather *gather,
                                          int32_t nblk, int32_t ncell, float *out,
                                          void *stream) {
    if (nblk < 0 || ncell < 0 || !gather) return ATSC3_E_ARG;
    if (nblk == 0 || ncell == 0) return ATSC3_OK;
    if (!pool || !out) return ATSC3_E_ARG;
    atsc3::Gather g{gather->idx, gather->first, gather->ring, gather->nrows, gather->start_row};
    const dim3 grid((ncell + DEMAP_WG - 1) / DEMAP_WG, nblk);
    hipStream_t st = atsc3::as_stream(stream);
    const float2 *p = reinterpret_cast<const float2 *>(pool);
    float2 *o = reinterpret_cast<float2 *>(out);
    if (gather->mode == ATSC3_GATHER_TABLE) {
        if (!g.idx) return ATSC3_E_ARG;
        hipLaunchKernelGGL(atsc3::gather_cells<atsc3::GATHER_TABLE>, grid, dim3(DEMAP_WG), 0, st,
                           p, g, ncell, o);
    } else if (gather->mode == ATSC3_GATHER_CTI) {
        if (g.nrows <= 0 || g.ring <= 0 || g.first < 0 || g.start_row < 0 ||
            g.start_row >= g.nrows)
            return ATSC3_E_ARG;
        hipLaunchKernelGGL(atsc3::gather_cells<atsc3::GATHER_CTI>, grid, dim3(DEMAP_WG), 0, st,
                           p, g, ncell, o);
    } else {
        return ATSC3_E_ARG;
    }
    return atsc3::launch_status();
}
