// Differentiable decoding: forward and hand-written backward of
//   * the unrolled normalized min-sum (flooding, T iterations, a learnable
//     normalization factor per (iteration, check-degree class)), and
//   * the exhaustive max-log soft demapper (learnable constellation points),
// in FP32 or FP64 (template T).  Any parity-check matrix in CSR form
// (check-major edges, cptr/cidx; per-variable edge lists vptr/vedge in
// ascending check order, as lab/m3_ldpc.py:459-482 builds them).
//
// Unrolled min-sum (iteration t = 0..T-1, m_0 = 0, tot_0 = lam):
//   x_e      = tot_t[v(e)] - m_t[e]
//   m_{t+1,e} = alpha[t][cls(c)] * s_e * mu_e,
//              s_e  = prod_{e' in c, e' != e} sign(x_e')   (sign(0) = +1)
//              mu_e = min_{e' in c, e' != e} |x_e'|  = (e == i1 ? min2 : min1)
//              (i1 = first index attaining min1, i2 = that of min2)
//   tot_{t+1,v} = lam_v + sum_{e in v} m_{t+1,e}   (ascending check order)
// The check kernel can record a compact per-check state (min1, min2, i1, i2,
// parity, sign bits of x) from which the backward recomputes everything it
// needs; the decode therefore checkpoints only (tot, m) every few
// iterations and recomputes a segment's states during the backward.
//
// Backward of one iteration, given G = dL/dtot_{t+1} and C = dL/dm_{t+1}
// carried from iteration t+1 (C = -G_x of that iteration):
//   g_e = G[v(e)] + C[e]                             (dL/dm_{t+1,e})
//   dL/dalpha[t][cls] += sum_e g_e s_e mu_e
//   G_x[i1] = alpha sign(x_i1) sum_{e != i1} g_e s_e  (the subgradient of
//   G_x[i2] = alpha sign(x_i2) g_i1 s_i1               min1/min2 selection)
//   dL/dtot_t[v] = sum_{e in v} G_x[e],  dL/dm_t = -G_x,  dL/dlam += G.
//
// Degrees: check degree <= 32 (sign bits in one word).
#include "fec_common.h"

#define UN_WG 256

namespace atsc3 {

template <typename T>
__device__ __forceinline__ T un_inf();
template <>
__device__ __forceinline__ float un_inf<float>() { return INFINITY; }
template <>
__device__ __forceinline__ double un_inf<double>() { return __builtin_huge_val(); }

template <typename T>
__global__ __launch_bounds__(UN_WG) void un_check_fwd(
    const T *__restrict__ tot, const T *__restrict__ m_in, T *__restrict__ m_out,
    const int *__restrict__ cptr, const int *__restrict__ cidx, const int *__restrict__ ccls,
    const T *__restrict__ alpha_t, int nblk, int n, int nc, int E, T *__restrict__ st_mn,
    int *__restrict__ st_idx, unsigned *__restrict__ st_sgn) {
    const long long gid = (long long)blockIdx.x * UN_WG + threadIdx.x;
    if (gid >= (long long)nblk * nc) return;
    const int b = (int)(gid / nc), c = (int)(gid % nc);
    const int e0 = cptr[c], d = cptr[c + 1] - e0;
    const T *tb = tot + (size_t)b * n;
    const T *mb = m_in ? m_in + (size_t)b * E + e0 : nullptr;
    T m1 = un_inf<T>(), m2 = un_inf<T>();
    int i1 = 0, i2 = 0;
    unsigned sg = 0u;
    for (int k = 0; k < d; ++k) {
        const T x = tb[cidx[e0 + k]] - (mb ? mb[k] : T(0));
        sg |= (unsigned)(x < T(0)) << k;
        const T a = x < T(0) ? -x : x;
        if (a < m1) {
            m2 = m1;
            i2 = i1;
            m1 = a;
            i1 = k;
        } else if (a < m2) {
            m2 = a;
            i2 = k;
        }
    }
    const unsigned par = __builtin_popcount(sg) & 1u;
    const T al = alpha_t[ccls[c]];
    T *ob = m_out + (size_t)b * E + e0;
    for (int k = 0; k < d; ++k) {
        const T mu = k == i1 ? m2 : m1;
        const T v = al * mu;
        ob[k] = ((par ^ (sg >> k)) & 1u) ? -v : v;
    }
    if (st_mn) {
        const size_t s = (size_t)b * nc + c;
        st_mn[2 * s] = m1;
        st_mn[2 * s + 1] = m2;
        st_idx[s] = i1 | (i2 << 8) | (int)(par << 16);
        st_sgn[s] = sg;
    }
}

template <typename T>
__global__ __launch_bounds__(UN_WG) void un_var_fwd(const T *__restrict__ lam,
                                                    const T *__restrict__ m,
                                                    const int *__restrict__ vptr,
                                                    const int *__restrict__ vedge,
                                                    T *__restrict__ tot, int nblk, int n, int E) {
    const long long gid = (long long)blockIdx.x * UN_WG + threadIdx.x;
    if (gid >= (long long)nblk * n) return;
    const int b = (int)(gid / n), v = (int)(gid % n);
    const T *mb = m + (size_t)b * E;
    T s = T(0);
    for (int k = vptr[v]; k < vptr[v + 1]; ++k) s += mb[vedge[k]];
    tot[gid] = lam[gid] + s;
}

// Backward of the check update of one iteration (see the file header).
// g_carry may be NULL (last iteration) and must not alias g_x.  Writes G_x
// for every edge and the per-check alpha partial  sum_e g_e s_e mu_e.
template <typename T>
__global__ __launch_bounds__(UN_WG) void un_check_bwd(
    const T *__restrict__ g_tot, const T *__restrict__ g_carry, const int *__restrict__ cptr,
    const int *__restrict__ cidx, const int *__restrict__ ccls, const T *__restrict__ alpha_t,
    const T *__restrict__ st_mn, const int *__restrict__ st_idx,
    const unsigned *__restrict__ st_sgn, int nblk, int n, int nc, int E, T *__restrict__ g_x,
    T *__restrict__ g_alpha_part) {
    const long long gid = (long long)blockIdx.x * UN_WG + threadIdx.x;
    if (gid >= (long long)nblk * nc) return;
    const int b = (int)(gid / nc), c = (int)(gid % nc);
    const int e0 = cptr[c], d = cptr[c + 1] - e0;
    const size_t s = (size_t)b * nc + c;
    const T m1 = st_mn[2 * s], m2 = st_mn[2 * s + 1];
    const int id = st_idx[s];
    const int i1 = id & 255, i2 = (id >> 8) & 255;
    const unsigned par = (unsigned)(id >> 16) & 1u, sg = st_sgn[s];
    const T al = alpha_t[ccls[c]];
    const T *gt = g_tot + (size_t)b * n;
    const T *gc = g_carry ? g_carry + (size_t)b * E + e0 : nullptr;
    T *gx = g_x + (size_t)b * E + e0;
    T sum_other = T(0), g_first = T(0), ga = T(0);
    for (int k = 0; k < d; ++k) {
        T g = gt[cidx[e0 + k]];
        if (gc) g -= gc[k];   // carried term: dL/dm = -G_x of the next iteration
        const T gs = ((par ^ (sg >> k)) & 1u) ? -g : g;
        ga += gs * (k == i1 ? m2 : m1);
        if (k == i1)
            g_first = gs;
        else
            sum_other += gs;
        gx[k] = T(0);
    }
    const T sx1 = ((sg >> i1) & 1u) ? T(-1) : T(1);
    const T sx2 = ((sg >> i2) & 1u) ? T(-1) : T(1);
    // min1 is +inf only for a degree-0 check; min2 for degree <= 1
    if (d >= 1) gx[i1] = al * sx1 * sum_other;
    if (d >= 2) gx[i2] = al * sx2 * g_first;
    g_alpha_part[s] = ga;
}

// dL/dtot_t[v] = sum_{e in v} G_x[e];  dL/dlam += dL/dtot_{t+1}.
template <typename T>
__global__ __launch_bounds__(UN_WG) void un_var_bwd(const T *__restrict__ g_x,
                                                    const T *__restrict__ g_tot_next,
                                                    const int *__restrict__ vptr,
                                                    const int *__restrict__ vedge,
                                                    T *__restrict__ g_tot, T *__restrict__ g_lam,
                                                    int nblk, int n, int E) {
    const long long gid = (long long)blockIdx.x * UN_WG + threadIdx.x;
    if (gid >= (long long)nblk * n) return;
    const int b = (int)(gid / n), v = (int)(gid % n);
    const T *gb = g_x + (size_t)b * E;
    T s = T(0);
    for (int k = vptr[v]; k < vptr[v + 1]; ++k) s += gb[vedge[k]];
    g_lam[gid] += g_tot_next[gid];
    g_tot[gid] = s;
}

// ---------------------------------------------------------------------------
// Exhaustive max-log soft demapper (any constellation of M = 2^nb points).
// LLR_i = w (min_{p: y_i=1} |z-p|^2 - min_{p: y_i=0} |z-p|^2) / sigma2, y_0 =
// label MSB; output in codeword order out[b][lam_of_q[s*nb + i]].
// ---------------------------------------------------------------------------
#define SD_MAXB 8

template <typename T>
__global__ __launch_bounds__(UN_WG) void soft_demap_fwd(
    const T *__restrict__ z, const T *__restrict__ csi, const T *__restrict__ sigma2,
    const T *__restrict__ pts, int M, int nb, const int *__restrict__ lam_of_q, int nblk,
    int ncell, T *__restrict__ out) {
    const long long gid = (long long)blockIdx.x * UN_WG + threadIdx.x;
    if (gid >= (long long)nblk * ncell) return;
    const int b = (int)(gid / ncell), s = (int)(gid % ncell);
    const T zr = z[2 * gid], zi = z[2 * gid + 1];
    T d0[SD_MAXB], d1[SD_MAXB];
    for (int i = 0; i < nb; ++i) d0[i] = d1[i] = un_inf<T>();
    for (int k = 0; k < M; ++k) {
        const T dr = zr - pts[2 * k], di = zi - pts[2 * k + 1];
        const T dd = dr * dr + di * di;
        for (int i = 0; i < nb; ++i) {
            if ((k >> (nb - 1 - i)) & 1) {
                if (dd < d1[i]) d1[i] = dd;
            } else if (dd < d0[i]) {
                d0[i] = dd;
            }
        }
    }
    const T w = csi ? csi[gid] : T(1);
    const T sc = w / sigma2[b];
    const size_t n = (size_t)ncell * nb;
    for (int i = 0; i < nb; ++i) out[(size_t)b * n + lam_of_q[s * nb + i]] = (d1[i] - d0[i]) * sc;
}

// g_z[b][s] = sum_i G_i sc 2 (p0*_i - p1*_i);  point partials per workgroup:
// g_pts_part[wg][k] += G_i sc (-2)(z - p1*) for p1*, += G_i sc 2 (z - p0*) for
// p0*.  g_sigma2_part[b][wg] = sum -G_i LLR_i / sigma2.
template <typename T>
__global__ __launch_bounds__(UN_WG) void soft_demap_bwd(
    const T *__restrict__ z, const T *__restrict__ csi, const T *__restrict__ sigma2,
    const T *__restrict__ pts, int M, int nb, const int *__restrict__ lam_of_q, int nblk,
    int ncell, const T *__restrict__ g_llr, T *__restrict__ g_z, T *__restrict__ g_pts_part,
    T *__restrict__ g_s2_part) {
    __shared__ T gp[2 * 256];
    __shared__ T gs2[UN_WG];
    for (int k = threadIdx.x; k < 2 * M; k += UN_WG) gp[k] = T(0);
    __syncthreads();
    const long long gid = (long long)blockIdx.x * UN_WG + threadIdx.x;
    T s2c = T(0);
    int bb = -1;
    if (gid < (long long)nblk * ncell) {
        const int b = (int)(gid / ncell), s = (int)(gid % ncell);
        bb = b;
        const T zr = z[2 * gid], zi = z[2 * gid + 1];
        T d0[SD_MAXB], d1[SD_MAXB];
        int k0[SD_MAXB], k1[SD_MAXB];
        for (int i = 0; i < nb; ++i) {
            d0[i] = d1[i] = un_inf<T>();
            k0[i] = k1[i] = 0;
        }
        for (int k = 0; k < M; ++k) {
            const T dr = zr - pts[2 * k], di = zi - pts[2 * k + 1];
            const T dd = dr * dr + di * di;
            for (int i = 0; i < nb; ++i) {
                if ((k >> (nb - 1 - i)) & 1) {
                    if (dd < d1[i]) {
                        d1[i] = dd;
                        k1[i] = k;
                    }
                } else if (dd < d0[i]) {
                    d0[i] = dd;
                    k0[i] = k;
                }
            }
        }
        const T w = csi ? csi[gid] : T(1);
        const T sc = w / sigma2[b];
        const size_t n = (size_t)ncell * nb;
        T gr = T(0), gi = T(0);
        for (int i = 0; i < nb; ++i) {
            const T g = g_llr[(size_t)b * n + lam_of_q[s * nb + i]];
            const T gsc = g * sc;
            const T p1r = pts[2 * k1[i]], p1i = pts[2 * k1[i] + 1];
            const T p0r = pts[2 * k0[i]], p0i = pts[2 * k0[i] + 1];
            gr += gsc * T(2) * (p0r - p1r);
            gi += gsc * T(2) * (p0i - p1i);
            atomicAdd(&gp[2 * k1[i]], -gsc * T(2) * (zr - p1r));
            atomicAdd(&gp[2 * k1[i] + 1], -gsc * T(2) * (zi - p1i));
            atomicAdd(&gp[2 * k0[i]], gsc * T(2) * (zr - p0r));
            atomicAdd(&gp[2 * k0[i] + 1], gsc * T(2) * (zi - p0i));
            s2c -= g * (d1[i] - d0[i]) * sc / sigma2[b];
        }
        g_z[2 * gid] = gr;
        g_z[2 * gid + 1] = gi;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * M; k += UN_WG)
        g_pts_part[(size_t)blockIdx.x * 2 * M + k] = gp[k];
    // sigma2 partials: a workgroup spans at most two blocks when ncell >= UN_WG
    gs2[threadIdx.x] = s2c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const long long base = (long long)blockIdx.x * UN_WG;
        for (int t = 0; t < UN_WG; ++t) {
            const long long g = base + t;
            if (g >= (long long)nblk * ncell) break;
            const int b = (int)(g / ncell);
            g_s2_part[(size_t)b * gridDim.x + blockIdx.x] += gs2[t];
        }
    }
    (void)bb;
}

static inline unsigned grid_of(long long items) {
    return (unsigned)((items + UN_WG - 1) / UN_WG);
}

template <typename T>
static int32_t check_fwd(const T *tot, const T *m_in, T *m_out, const int *cptr, const int *cidx,
                         const int *ccls, const T *alpha_t, int nblk, int n, int nc, int E,
                         T *st_mn, int *st_idx, unsigned *st_sgn, hipStream_t st) {
    hipLaunchKernelGGL(un_check_fwd<T>, dim3(grid_of((long long)nblk * nc)), dim3(UN_WG), 0, st,
                       tot, m_in, m_out, cptr, cidx, ccls, alpha_t, nblk, n, nc, E, st_mn, st_idx,
                       st_sgn);
    return launch_status();
}

}  // namespace atsc3

using atsc3::as_stream;
using atsc3::grid_of;

#define UN_CHECK_ARGS(T)                                                                      \
    const T *tot, const T *m_in, T *m_out, const int32_t *cptr, const int32_t *cidx,          \
        const int32_t *ccls, const T *alpha_t, int32_t nblk, int32_t n, int32_t nc, int32_t E, \
        T *st_mn, int32_t *st_idx, uint32_t *st_sgn, void *stream

#define UN_VALID_CHECK                                                                       \
    if (nblk < 0 || n <= 0 || nc <= 0 || E <= 0) return ATSC3_E_ARG;                         \
    if (nblk == 0) return ATSC3_OK;                                                          \
    if (!tot || !m_out || !cptr || !cidx || !ccls || !alpha_t) return ATSC3_E_ARG;           \
    if ((st_mn != nullptr) != (st_idx != nullptr) || (st_mn != nullptr) != (st_sgn != nullptr)) \
        return ATSC3_E_ARG;

extern "C" int32_t atsc3_unrolled_check_fwd_f32(UN_CHECK_ARGS(float)) {
    UN_VALID_CHECK
    return atsc3::check_fwd<float>(tot, m_in, m_out, cptr, cidx, ccls, alpha_t, nblk, n, nc, E,
                                   st_mn, st_idx, st_sgn, as_stream(stream));
}
extern "C" int32_t atsc3_unrolled_check_fwd_f64(UN_CHECK_ARGS(double)) {
    UN_VALID_CHECK
    return atsc3::check_fwd<double>(tot, m_in, m_out, cptr, cidx, ccls, alpha_t, nblk, n, nc, E,
                                    st_mn, st_idx, st_sgn, as_stream(stream));
}

#define UN_VAR_FWD(T, NAME)                                                                     \
    extern "C" int32_t NAME(const T *lam, const T *m, const int32_t *vptr, const int32_t *vedge,  \
                            T *tot, int32_t nblk, int32_t n, int32_t E, void *stream) {           \
        if (nblk < 0 || n <= 0 || E <= 0) return ATSC3_E_ARG;                                   \
        if (nblk == 0) return ATSC3_OK;                                                         \
        if (!lam || !m || !vptr || !vedge || !tot) return ATSC3_E_ARG;                          \
        hipLaunchKernelGGL(atsc3::un_var_fwd<T>, dim3(grid_of((long long)nblk * n)),              \
                           dim3(UN_WG), 0, as_stream(stream), lam, m, vptr, vedge, tot, nblk, n,  \
                           E);                                                                  \
        return atsc3::launch_status();                                                          \
    }
UN_VAR_FWD(float, atsc3_unrolled_var_fwd_f32)
UN_VAR_FWD(double, atsc3_unrolled_var_fwd_f64)

#define UN_CHECK_BWD(T, NAME)                                                                   \
    extern "C" int32_t NAME(const T *g_tot, const T *g_carry, const int32_t *cptr,              \
                            const int32_t *cidx, const int32_t *ccls, const T *alpha_t,         \
                            const T *st_mn, const int32_t *st_idx, const uint32_t *st_sgn,      \
                            int32_t nblk, int32_t n, int32_t nc, int32_t E, T *g_x,             \
                            T *g_alpha_part, void *stream) {                                    \
        if (nblk < 0 || n <= 0 || nc <= 0 || E <= 0) return ATSC3_E_ARG;                        \
        if (nblk == 0) return ATSC3_OK;                                                         \
        if (!g_tot || !cptr || !cidx || !ccls || !alpha_t || !st_mn || !st_idx || !st_sgn ||    \
            !g_x || !g_alpha_part)                                                              \
            return ATSC3_E_ARG;                                                                 \
        hipLaunchKernelGGL(atsc3::un_check_bwd<T>, dim3(grid_of((long long)nblk * nc)),           \
                           dim3(UN_WG), 0, as_stream(stream), g_tot, g_carry, cptr, cidx, ccls,   \
                           alpha_t, st_mn, st_idx, st_sgn, nblk, n, nc, E, g_x, g_alpha_part);   \
        return atsc3::launch_status();                                                          \
    }
UN_CHECK_BWD(float, atsc3_unrolled_check_bwd_f32)
UN_CHECK_BWD(double, atsc3_unrolled_check_bwd_f64)

#define UN_VAR_BWD(T, NAME)                                                                      \
    extern "C" int32_t NAME(const T *g_x, const T *g_tot_next, const int32_t *vptr,              \
                            const int32_t *vedge, T *g_tot, T *g_lam, int32_t nblk, int32_t n,   \
                            int32_t E, void *stream) {                                           \
        if (nblk < 0 || n <= 0 || E <= 0) return ATSC3_E_ARG;                                    \
        if (nblk == 0) return ATSC3_OK;                                                          \
        if (!g_x || !g_tot_next || !vptr || !vedge || !g_tot || !g_lam) return ATSC3_E_ARG;      \
        hipLaunchKernelGGL(atsc3::un_var_bwd<T>, dim3(grid_of((long long)nblk * n)),               \
                           dim3(UN_WG), 0, as_stream(stream), g_x, g_tot_next, vptr, vedge,        \
                           g_tot, g_lam, nblk, n, E);                                            \
        return atsc3::launch_status();                                                           \
    }
UN_VAR_BWD(float, atsc3_unrolled_var_bwd_f32)
UN_VAR_BWD(double, atsc3_unrolled_var_bwd_f64)

#define SD_FWD(T, NAME)                                                                          \
    extern "C" int32_t NAME(const T *z, const T *csi, const T *sigma2, const T *pts, int32_t M,  \
                            int32_t nb, const int32_t *lam_of_q, int32_t nblk, int32_t ncell,    \
                            T *out, void *stream) {                                              \
        if (nblk < 0 || ncell < 0 || nb < 1 || nb > SD_MAXB || M != (1 << nb)) return ATSC3_E_ARG; \
        if (nblk == 0 || ncell == 0) return ATSC3_OK;                                            \
        if (!z || !sigma2 || !pts || !lam_of_q || !out) return ATSC3_E_ARG;                      \
        hipLaunchKernelGGL(atsc3::soft_demap_fwd<T>, dim3(grid_of((long long)nblk * ncell)),       \
                           dim3(UN_WG), 0, as_stream(stream), z, csi, sigma2, pts, M, nb,          \
                           lam_of_q, nblk, ncell, out);                                          \
        return atsc3::launch_status();                                                           \
    }
SD_FWD(float, atsc3_soft_demap_fwd_f32)
SD_FWD(double, atsc3_soft_demap_fwd_f64)

extern "C" int32_t atsc3_soft_demap_bwd_grid(int32_t nblk, int32_t ncell) {
    if (nblk < 0 || ncell < 0) return ATSC3_E_ARG;
    return (int32_t)grid_of((long long)nblk * ncell);
}

#define SD_BWD(T, NAME)                                                                          \
    extern "C" int32_t NAME(const T *z, const T *csi, const T *sigma2, const T *pts, int32_t M,  \
                            int32_t nb, const int32_t *lam_of_q, int32_t nblk, int32_t ncell,    \
                            const T *g_llr, T *g_z, T *g_pts_part, T *g_s2_part, void *stream) { \
        if (nblk < 0 || ncell < 0 || nb < 1 || nb > SD_MAXB || M != (1 << nb)) return ATSC3_E_ARG; \
        if (nblk == 0 || ncell == 0) return ATSC3_OK;                                            \
        if (!z || !sigma2 || !pts || !lam_of_q || !g_llr || !g_z || !g_pts_part || !g_s2_part)   \
            return ATSC3_E_ARG;                                                                  \
        hipLaunchKernelGGL(atsc3::soft_demap_bwd<T>, dim3(grid_of((long long)nblk * ncell)),       \
                           dim3(UN_WG), 0, as_stream(stream), z, csi, sigma2, pts, M, nb,          \
                           lam_of_q, nblk, ncell, g_llr, g_z, g_pts_part, g_s2_part);            \
        return atsc3::launch_status();                                                           \
    }
SD_BWD(float, atsc3_soft_demap_bwd_f32)
SD_BWD(double, atsc3_soft_demap_bwd_f64)
