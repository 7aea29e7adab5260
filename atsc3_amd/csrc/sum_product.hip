// Exact (float64, log-tanh) sum-product decode of near-miss codewords and
// the device-side rescue policy around it.
//
// Restates lab/m16_margin.py:325-390 (sum_product_decode_batch) on the GPU:
// flooding BP over the check rows padded to the widest check of the code, in
// the arithmetic order of the upstream's numpy statement wherever it is
// defined by the program rather than by a math library:
//   * the padded row's sum of log|tanh| terms in numpy's pairwise order
//     (eight running partial sums over the whole multiple-of-8 prefix,
//     combined ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the tail added in
//     order; plain left-to-right below 8 slots; padding adds 0.0);
//   * every variable's total = LLR + its messages in ascending check order
//     (np.add.at over the row-major index array);
//   * clips, comparisons and the sign rule exactly as written there.
// tanh, log, exp and atanh are the device's double-precision library
// functions, so totals agree with the CPU statement to a few ulps per
// operation, not bit for bit.
//
// Layout (device memory, per code):
//   cvar [dmax][nc]  variable of slot k of check thread i (n = padding); the
//                    thread order is the QC order (group, position), so
//                    neighbouring threads read neighbouring variables
//   vptr [n + 1], vedge [E]  per variable, the message slots k * nc + i of
//                    its edges in ascending check index
// Per row r of a call: tot [r][n] and msg [r][dmax][nc] doubles.
//
// One call runs every iteration as two kernels (check, var) per iteration
// on the caller's stream: no host synchronisation.  A row leaves when its
// syndrome is zero (its = the iteration), at the last iteration, or when the
// time budget is spent (no new iteration starts after the budget, as the
// upstream's deadline; `shed` marks those rows).
#include <climits>

#include "fec_common.h"

#define SP_WG 256
#define SP_MAX_DMAX 128

namespace atsc3 {
namespace {

// Control words per call.
enum { SP_STOP = 0, SP_T0 = 2, SP_HDR = 4 };   // stop[2], t0 (64-bit), then fin[rows], cnt[rows][2]

struct SpArgs {
    const float *llr;
    const int64_t *row_off;   // NULL: row s starts at s * n
    const int32_t *sel;       // NULL: call row r decodes llr row r
    const int32_t *nsel;      // NULL: all `rows` rows are live
    int rows, n, nc, dmax, iters;
    long long budget;         // wall-clock ticks; <= 0: none
    const int32_t *cvar, *vptr, *vedge;
    double *tot, *msg;
    int32_t *ctl;
    uint8_t *bits, *conv, *shed;
    int32_t *its, *bad;
};

__device__ __forceinline__ int *fin_of(const SpArgs &a) { return a.ctl + SP_HDR; }
__device__ __forceinline__ int *cnt_of(const SpArgs &a) { return a.ctl + SP_HDR + a.rows; }

__device__ __forceinline__ const float *lam_row(const SpArgs &a, int r) {
    const int s = a.sel ? a.sel[r] : r;
    return a.llr + (a.row_off ? a.row_off[s] : (int64_t)s * a.n);
}

__device__ __forceinline__ double clip(double x, double lo, double hi) {
    // np.clip: minimum(maximum(x, lo), hi), NaN passes through
    return x < lo ? lo : (x > hi ? hi : x);
}

__global__ __launch_bounds__(SP_WG) void sp_init(SpArgs a) {
    const int r = blockIdx.y;
    const int live = a.nsel ? *a.nsel : a.rows;
    const bool on = r < live;
    const float *lam = on ? lam_row(a, r) : nullptr;
    double *tot = a.tot + (size_t)r * a.n;
    for (int v = blockIdx.x * SP_WG + threadIdx.x; v < a.n; v += gridDim.x * SP_WG)
        tot[v] = on ? (double)lam[v] : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        fin_of(a)[r] = on ? INT_MAX : -1;
        cnt_of(a)[2 * r] = 0;
        cnt_of(a)[2 * r + 1] = 0;
        if (!on) {
            a.conv[r] = 0;
            a.its[r] = 0;
            a.bad[r] = -1;
            a.shed[r] = 0;
        }
        if (r == 0) {
            a.ctl[SP_STOP] = 0;
            a.ctl[SP_STOP + 1] = 0;
            const long long t0 = wall_clock64();
            a.ctl[SP_T0] = (int)(t0 & 0xffffffffLL);
            a.ctl[SP_T0 + 1] = (int)(t0 >> 32);
        }
    }
}

// Syndrome of the current totals (counted per row), then, unless this is
// the last iteration or the budget is spent, the new messages.  The slot
// loop runs twice: the first pass stores each slot's log|tanh| term in the
// message slot it replaces (the thread's own words) and forms the row sum,
// the second turns terms into messages.
__global__ __launch_bounds__(SP_WG) void sp_check(SpArgs a, int it) {
    const int r = blockIdx.y;
    if (fin_of(a)[r] < it) return;
    const int i = blockIdx.x * SP_WG + threadIdx.x;
    const bool on = i < a.nc;
    const double *tot = a.tot + (size_t)r * a.n;
    double *msg = a.msg + (size_t)r * a.dmax * a.nc;
    const bool update = it < a.iters && a.ctl[SP_STOP + (it & 1)] == 0;
    unsigned hard = 0u;
    if (on && !update) {
        for (int k = 0; k < a.dmax; ++k) {
            const int v = a.cvar[(size_t)k * a.nc + i];
            if (v != a.n) hard ^= tot[v] < 0.0 ? 1u : 0u;
        }
    } else if (on) {
        unsigned long long neg0 = 0ull, neg1 = 0ull;   // sign of tanh per slot
        // slot k's log|tanh| term, stored in the message slot it replaces
        auto term = [&](int k) -> double {
            const size_t s = (size_t)k * a.nc + i;
            const int v = a.cvar[s];
            if (v == a.n) return 0.0;
            const double tv = tot[v];
            hard ^= tv < 0.0 ? 1u : 0u;
            const double x = it == 0 ? tv : tv - msg[s];
            const double t = tanh(clip(x, -38.0, 38.0) * 0.5);
            if (t < 0.0) {
                if (k < 64)
                    neg0 |= 1ull << k;
                else
                    neg1 |= 1ull << (k - 64);
            }
            const double lt = log(clip(fabs(t), 1e-300, 1.0));
            msg[s] = lt;
            return lt;
        };
        double lsum = 0.0;
        if (a.dmax < 8) {
            for (int k = 0; k < a.dmax; ++k) lsum += term(k);
        } else {
            const int nmain = a.dmax - a.dmax % 8;
            double rp[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < nmain; k0 += 8) {
#pragma unroll
                for (int j = 0; j < 8; ++j) rp[j] += term(k0 + j);
            }
            lsum = ((rp[0] + rp[1]) + (rp[2] + rp[3])) + ((rp[4] + rp[5]) + (rp[6] + rp[7]));
            for (int k = nmain; k < a.dmax; ++k) lsum += term(k);
        }
        const unsigned par = (unsigned)(__popcll(neg0) + __popcll(neg1)) & 1u;
        for (int k = 0; k < a.dmax; ++k) {
            const size_t s = (size_t)k * a.nc + i;
            if (a.cvar[s] == a.n) continue;
            const double lt = msg[s];
            const double le = clip(lsum - lt, -700.0, -1e-16);
            const double mag = 2.0 * atanh(clip(exp(le), 0.0, 1.0 - 1e-16));
            const unsigned nk = (unsigned)(((k < 64 ? neg0 : neg1) >> (k & 63)) & 1ull);
            msg[s] = (par ^ nk) ? -mag : mag;
        }
    }
    // unsatisfied checks of this wave, one atomic per wave
    const unsigned long long b = __ballot(on && hard);
    if ((threadIdx.x & 63) == 0 && b)
        atomicAdd(cnt_of(a) + 2 * r + (it & 1), (int)__popcll(b));
}

// Rows whose syndrome is zero, at the last iteration or out of budget leave
// with the current decisions; the others rebuild their totals.  Block (0, 0)
// also decides whether iteration it + 1 may start.
__global__ __launch_bounds__(SP_WG) void sp_var(SpArgs a, int it) {
    const int r = blockIdx.y;
    if (blockIdx.x == 0 && r == 0 && threadIdx.x == 0) {
        int stop = 0;
        if (a.budget > 0) {
            const long long t0 = (long long)(unsigned)a.ctl[SP_T0] |
                                 ((long long)a.ctl[SP_T0 + 1] << 32);
            stop = wall_clock64() - t0 > a.budget ? 1 : 0;
        }
        a.ctl[SP_STOP + ((it + 1) & 1)] = stop;
    }
    // fin[r] is only ever lowered to `it` in this kernel, so every
    // workgroup of the row sees it as live here
    if (fin_of(a)[r] < it) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt_of(a)[2 * r + ((it + 1) & 1)] = 0;
    const int u = cnt_of(a)[2 * r + (it & 1)];
    const bool stopped = a.ctl[SP_STOP + (it & 1)] != 0;
    const int v = blockIdx.x * SP_WG + threadIdx.x;
    double *tot = a.tot + (size_t)r * a.n;
    if (u == 0 || it == a.iters || stopped) {
        if (v < a.n) a.bits[(size_t)r * a.n + v] = tot[v] < 0.0 ? 1 : 0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.conv[r] = u == 0 ? 1 : 0;
            a.its[r] = it;
            a.bad[r] = u;
            a.shed[r] = (u != 0 && it < a.iters) ? 1 : 0;
            fin_of(a)[r] = it;
        }
        return;
    }
    if (v >= a.n) return;
    const double *msg = a.msg + (size_t)r * a.dmax * a.nc;
    double acc = (double)lam_row(a, r)[v];
    for (int e = a.vptr[v]; e < a.vptr[v + 1]; ++e) acc += msg[a.vedge[e]];
    tot[v] = acc;
}

// Candidate rows: not converged, bad <= near_max (and of code code_id when
// blk_code is given), ordered by bad then row (a stable sort), the first
// max_rows of them.  One workgroup.
__global__ __launch_bounds__(1024) void sp_select(const uint8_t *conv, const int32_t *bad,
                                                  const int32_t *blk_code, int code_id,
                                                  int nblk, int near_max, int max_rows,
                                                  int32_t *rows_out, int32_t *nrows_out) {
    __shared__ int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const int kb = bad[b];
        if (conv[b] || kb > near_max || (blk_code && blk_code[b] != code_id)) continue;
        atomicAdd(&s_cnt, 1);
        int rank = 0;
        for (int o = 0; o < nblk && rank < max_rows; ++o) {
            if (conv[o] || (blk_code && blk_code[o] != code_id)) continue;
            const int ko = bad[o];
            rank += (ko <= near_max && (ko < kb || (ko == kb && o < b))) ? 1 : 0;
        }
        if (rank < max_rows) rows_out[rank] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) *nrows_out = s_cnt < max_rows ? s_cnt : max_rows;
}

// Converged rescues replace the batch's results of their rows.
__global__ __launch_bounds__(SP_WG) void sp_merge(const int32_t *sel, const int32_t *nsel, int n,
                                                  const int64_t *row_off, const uint8_t *sp_bits,
                                                  const uint8_t *sp_conv, const int32_t *sp_its,
                                                  uint8_t *bits, uint8_t *conv, int32_t *its,
                                                  int32_t *bad) {
    const int r = blockIdx.y;
    if (r >= *nsel || !sp_conv[r]) return;
    const int s = sel[r];
    uint8_t *dst = bits + (row_off ? row_off[s] : (int64_t)s * n);
    const uint8_t *src = sp_bits + (size_t)r * n;
    for (int v = blockIdx.x * SP_WG + threadIdx.x; v < n; v += gridDim.x * SP_WG) dst[v] = src[v];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        conv[s] = 1;
        its[s] = sp_its[r];
        bad[s] = 0;
    }
}

}  // namespace
}  // namespace atsc3

using namespace atsc3;

extern "C" int32_t atsc3_sp_ctl_words(int32_t rows) {
    return rows < 0 ? ATSC3_E_ARG : SP_HDR + 3 * rows;
}

extern "C" int32_t atsc3_sum_product_dev(
    const float *llr, const int64_t *row_off, const int32_t *sel, const int32_t *nsel,
    int32_t rows, int32_t n, int32_t nc, int32_t dmax, const int32_t *cvar, const int32_t *vptr,
    const int32_t *vedge, int32_t iters, double budget_s, double *tot_ws, double *msg_ws,
    int32_t *ctl_ws, uint8_t *bits, uint8_t *conv, int32_t *its, int32_t *bad, uint8_t *shed,
    void *stream) {
    if (rows < 0 || n <= 0 || nc <= 0 || dmax <= 0 || iters < 0) return ATSC3_E_ARG;
    if (dmax > SP_MAX_DMAX) return ATSC3_E_DEGREE;
    if (rows == 0) return ATSC3_OK;
    if (!llr || !cvar || !vptr || !vedge || !tot_ws || !msg_ws || !ctl_ws || !bits || !conv ||
        !its || !bad || !shed)
        return ATSC3_E_ARG;
    if ((long long)rows * dmax * nc > (1LL << 40)) return ATSC3_E_SIZE;
    SpArgs a;
    a.llr = llr;
    a.row_off = row_off;
    a.sel = sel;
    a.nsel = nsel;
    a.rows = rows;
    a.n = n;
    a.nc = nc;
    a.dmax = dmax;
    a.iters = iters;
    a.budget = 0;
    if (budget_s > 0.0) {
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
            khz <= 0)
            return ATSC3_E_HIP;
        const double ticks = budget_s * 1000.0 * khz;
        a.budget = ticks < 1.0 ? 1 : (ticks > 9.0e18 ? (long long)9e18 : (long long)ticks);
    }
    a.cvar = cvar;
    a.vptr = vptr;
    a.vedge = vedge;
    a.tot = tot_ws;
    a.msg = msg_ws;
    a.ctl = ctl_ws;
    a.bits = bits;
    a.conv = conv;
    a.shed = shed;
    a.its = its;
    a.bad = bad;
    hipStream_t st = as_stream(stream);
    const int gv = (n + SP_WG - 1) / SP_WG, gc = (nc + SP_WG - 1) / SP_WG;
    hipLaunchKernelGGL(sp_init, dim3(gv < 64 ? gv : 64, rows), dim3(SP_WG), 0, st, a);
    for (int it = 0; it <= iters; ++it) {
        hipLaunchKernelGGL(sp_check, dim3(gc, rows), dim3(SP_WG), 0, st, a, it);
        hipLaunchKernelGGL(sp_var, dim3(gv, rows), dim3(SP_WG), 0, st, a, it);
    }
    return launch_status();
}

extern "C" int32_t atsc3_sp_select_dev(const uint8_t *conv, const int32_t *bad,
                                       const int32_t *blk_code, int32_t code_id, int32_t nblk,
                                       int32_t near_max, int32_t max_rows, int32_t *rows_out,
                                       int32_t *nrows_out, void *stream) {
    if (nblk < 0 || max_rows <= 0 || !nrows_out || !rows_out) return ATSC3_E_ARG;
    if (nblk > 0 && (!conv || !bad)) return ATSC3_E_ARG;
    hipLaunchKernelGGL(sp_select, dim3(1), dim3(1024), 0, as_stream(stream), conv, bad, blk_code,
                       code_id, nblk, near_max, max_rows, rows_out, nrows_out);
    return launch_status();
}

extern "C" int32_t atsc3_sp_merge_dev(const int32_t *sel, const int32_t *nsel, int32_t rows,
                                      int32_t n, const int64_t *row_off, const uint8_t *sp_bits,
                                      const uint8_t *sp_conv, const int32_t *sp_its,
                                      uint8_t *bits, uint8_t *conv, int32_t *its, int32_t *bad,
                                      void *stream) {
    if (rows < 0 || n <= 0) return ATSC3_E_ARG;
    if (rows == 0) return ATSC3_OK;
    if (!sel || !nsel || !sp_bits || !sp_conv || !sp_its || !bits || !conv || !its || !bad)
        return ATSC3_E_ARG;
    const int gv = (n + SP_WG - 1) / SP_WG;
    hipLaunchKernelGGL(sp_merge, dim3(gv < 64 ? gv : 64, rows), dim3(SP_WG), 0,
                       as_stream(stream), sel, nsel, n, row_off, sp_bits, sp_conv, sp_its, bits,
                       conv, its, bad);
    return launch_status();
}
