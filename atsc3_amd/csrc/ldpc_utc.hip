// Update-then-check normalized min-sum in FP32 or FP64 over any Tanner graph
// in CSR form: the contract of the upstream's torch decoder GpuMinSum.decode
// (lab/m9_gpu.py:138-211) at dtype float32, and of its float64 decoders
// min_sum_decode / min_sum_decode_batch(dtype=float64)
// (lab/m3_ldpc.py:245-286, 509-630).
//
// Per codeword, for k = 1 .. iters:
//   * every check e of c: x_e = tot[v_e] - m_e (tot = llr, m = 0 before the
//     first update); min1 <= min2 are the two smallest of the multiset
//     {|x_e|} plus (dmax - deg(c)) pad entries of magnitude 1e9 (the padded
//     rows of the upstream, lab/m9_gpu.py:157-158); parity = XOR of x_e < 0;
//     m_e = +-(alpha * (|x_e| == min1 ? min2 : min1)), negative iff
//     parity ^ (x_e < 0).  (When several entries tie at min1, min2 == min1,
//     so "the first minimal entry takes min2" and "every minimal entry
//     takes min2" give the same values.)
//   * every variable v: tot[v] = ((llr[v] + m_e1) + m_e2) + ... over v's
//     edges in ascending check index (vedge order): the channel-value-first
//     association of lab/m9_gpu.py:184-188 and np.add.at.
//   * syndrome of tot < 0: zero -> bits = tot < 0, conv 1, its k, bad 0.
// After iters updates without a zero syndrome: conv 0, its iters, bad = the
// unsatisfied count, bits = tot < 0.  Each operation is one IEEE operation
// of T (no contraction), so FP32 reproduces the torch decoder's roundings
// and FP64 numpy's.
//
// A check whose magnitudes include NaN (NaN LLRs, inf - inf) follows one of
// two rules, where the upstream's functions differ (NaN-free checks are the
// same under both):
//   * UTC_NAN_TORCH (GpuMinSum, lab/m9_gpu.py:160-166): min1 = amin, which
//     propagates NaN, so every message is NaN except edge 0's, which takes
//     min2 = amin of the other entries: the smallest of them when edge 0
//     holds the only NaN, else NaN;
//   * UTC_NAN_NUMPY (min_sum_decode, lab/m3_ldpc.py:264-270): min1 <= min2
//     are the two smallest non-NaN entries (np.partition orders NaN last;
//     NaN when fewer exist), and np.argmin's choice -- the first NaN edge --
//     takes min2, every other edge min1.
//
// Schedule: one launch pair per iteration, over (codeword, check) and
// (codeword, variable) threads.  The check launch of iteration it gathers
// tot once for both the syndrome of the posteriors after `it` updates (for
// it >= 1) and the next update; the variable launch retires a codeword whose
// syndrome count read zero (its = it) instead of updating it.  Retired
// codewords drop out of every later launch (one load of a per-codeword word).
#include "fec_common.h"

#define UTC_WG 256
#define UTC_NAN_TORCH 0
#define UTC_NAN_NUMPY 1

namespace atsc3 {

template <typename T>
struct UtcArgs {
    const T *llr;
    int n, nc, dmax;
    const int *cptr, *cidx, *vptr, *vedge;
    int nedges, iters;
    T alpha;
    T *tot, *msg;
    int *done_it;   // [nblk] iteration a codeword retired at (INT_MAX: active)
    int *cnt;       // [2][nblk] unsatisfied-check counts, by iteration parity
    int nblk;
    uint8_t *bits, *conv;
    int *its, *bad;
};

__global__ void utc_init(int *done_it, int *cnt, int nblk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nblk) {
        done_it[i] = 0x7fffffff;
        cnt[i] = 0;
        cnt[nblk + i] = 0;
    }
}

template <typename T>
__device__ __forceinline__ void two_min(T a, T &lo1, T &lo2) {
    if (a < lo1) {
        lo2 = lo1;
        lo1 = a;
    } else if (a < lo2) {
        lo2 = a;
    }
}

// Check launch of iteration `it`: syndrome (it >= 1; KC: it >= 0) and
// update (it < iters).  KC = the compiled kernel's contract
// (lab/ldpc_kernel.c:82-143): no row padding, and the edge that first
// attains min1 in the strict-< scan takes min2 (a NaN magnitude never wins).
// Otherwise the rows are padded and NR (UTC_NAN_*) rules NaN.
template <typename T, bool KC, int NR>
__global__ __launch_bounds__(UTC_WG) void utc_check(UtcArgs<T> a, int it) {
    const int blk = blockIdx.y;
    if (a.done_it[blk] < it) return;
    const int c = blockIdx.x * UTC_WG + threadIdx.x;
    unsigned syn = 0u;
    if (c < a.nc) {
        const T *src = (it == 0 ? a.llr : a.tot) + (size_t)blk * a.n;
        T *m = a.msg + (size_t)blk * a.nedges;
        const int e0 = a.cptr[c], e1 = a.cptr[c + 1];
        const T big = (T)1e9;
        const T zero = (T)0;
        if (it < a.iters) {
            unsigned par = 0u;
            T lo1 = (T)INFINITY, lo2 = (T)INFINITY;
            int first = e0;
            int nanpos = -1, nnan = 0;   // first NaN-magnitude edge, NaN count
            for (int e = e0; e < e1; ++e) {
                const T tv = src[a.cidx[e]];
                const T x = tv - (it == 0 ? zero : m[e]);
                syn ^= (unsigned)(tv < zero);
                par ^= (unsigned)(x < zero);
                const T mag = x < zero ? -x : x;
                if constexpr (KC) {
                    if (mag < lo1) {
                        lo2 = lo1;
                        lo1 = mag;
                        first = e;
                    } else if (mag < lo2) {
                        lo2 = mag;
                    }
                } else {
                    two_min(mag, lo1, lo2);   // a NaN compares below nothing
                    if (mag != mag) {
                        if (nanpos < 0) nanpos = e;
                        ++nnan;
                    }
                }
            }
            const int npad = KC ? 0 : a.dmax - (e1 - e0);
            if constexpr (!KC) {
                if (npad >= 1) two_min(big, lo1, lo2);
                if (npad >= 2) two_min(big, lo1, lo2);
            }
            const T s1 = a.alpha * lo1, s2 = a.alpha * lo2;
            const T qnan = (T)__builtin_nan("");
            // NaN rows: the values of the non-minimal and the min2 edge
            T t1 = s1, t2 = s2;
            int take2 = nanpos;
            if (!KC && nanpos >= 0) {
                if constexpr (NR == UTC_NAN_TORCH) {
                    // edge 0 takes amin of the others (their smallest when it
                    // holds the only NaN, so s1), every other edge NaN
                    t1 = qnan;
                    t2 = (nanpos == e0 && nnan == 1) ? s1 : qnan;
                    take2 = e0;
                } else {
                    const int nok = (e1 - e0) - nnan + npad;
                    t1 = nok >= 1 ? s1 : qnan;
                    t2 = nok >= 2 ? s2 : qnan;
                }
            }
            for (int e = e0; e < e1; ++e) {
                const T x = src[a.cidx[e]] - (it == 0 ? zero : m[e]);
                const T mag = x < zero ? -x : x;
                T v;
                if (KC)
                    v = e == first ? s2 : s1;
                else if (nanpos < 0)
                    v = mag == lo1 ? s2 : s1;
                else
                    v = e == take2 ? t2 : t1;
                m[e] = (par ^ (unsigned)(x < zero)) ? -v : v;
            }
        } else {
            for (int e = e0; e < e1; ++e) syn ^= (unsigned)(src[a.cidx[e]] < zero);
        }
    }
    if (KC || it >= 1) {
        // one atomic per wave
        const unsigned long long b = __ballot(syn != 0u);
        if ((threadIdx.x & 63) == 0 && b)
            atomicAdd(&a.cnt[(it & 1) * a.nblk + blk], (int)__popcll(b));
    }
}

// Variable launch of iteration `it`: retire (syndrome read zero, or it ==
// iters) or update the posteriors.  KC: a codeword may retire at it = 0
// (channel decisions already satisfy H) and the sum is lam + (0 + m_1 + ...).
template <typename T, bool KC>
__global__ __launch_bounds__(UTC_WG) void utc_var(UtcArgs<T> a, int it) {
    const int blk = blockIdx.y;
    const int v = blockIdx.x * UTC_WG + threadIdx.x;
    if (v == 0) a.cnt[((it + 1) & 1) * a.nblk + blk] = 0;   // next count
    if (a.done_it[blk] < it) return;
    const int count = a.cnt[(it & 1) * a.nblk + blk];
    const bool retire = ((KC || it >= 1) && count == 0) || it == a.iters;
    T *tot = a.tot + (size_t)blk * a.n;
    if (retire) {
        const T *src = it == 0 ? a.llr + (size_t)blk * a.n : tot;
        if (v < a.n) a.bits[(size_t)blk * a.n + v] = (uint8_t)(src[v] < (T)0);
        if (v == 0) {
            a.conv[blk] = (uint8_t)(count == 0);
            a.its[blk] = count == 0 ? it : a.iters;
            a.bad[blk] = count;
            a.done_it[blk] = it;
        }
        return;
    }
    if (v < a.n) {
        const T *m = a.msg + (size_t)blk * a.nedges;
        const T lam = a.llr[(size_t)blk * a.n + v];
        T acc = KC ? (T)0 : lam;
        for (int k = a.vptr[v]; k < a.vptr[v + 1]; ++k) acc = acc + m[a.vedge[k]];
        tot[v] = KC ? lam + acc : acc;
    }
}

// iters == 0 (min_sum_decode's contract): bits = llr < 0, conv 0, its 0,
// bad -1.
template <typename T>
__global__ __launch_bounds__(UTC_WG) void utc_zero(UtcArgs<T> a) {
    const int blk = blockIdx.y;
    const int v = blockIdx.x * UTC_WG + threadIdx.x;
    if (v < a.n)
        a.bits[(size_t)blk * a.n + v] = (uint8_t)(a.llr[(size_t)blk * a.n + v] < (T)0);
    if (v == 0) {
        a.conv[blk] = 0;
        a.its[blk] = 0;
        a.bad[blk] = -1;
    }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t utc_ws_bytes(int nblk, int n, int nedges, size_t tsize) {
    return align256((size_t)nblk * n * tsize) + align256((size_t)nblk * nedges * tsize) +
           align256((size_t)nblk * 3 * sizeof(int));
}

template <typename T, bool KC, int NR>
static int32_t utc_launch(const T *llr, int nblk, int n, int nc, int nedges, const int *cptr,
                          const int *cidx, const int *vptr, const int *vedge, int dmax,
                          int iters, double alpha, void *ws, uint8_t *bits, uint8_t *conv,
                          int *its, int *bad, hipStream_t st) {
    char *p = static_cast<char *>(ws);
    UtcArgs<T> a;
    a.llr = llr;
    a.n = n;
    a.nc = nc;
    a.dmax = dmax;
    a.cptr = cptr;
    a.cidx = cidx;
    a.vptr = vptr;
    a.vedge = vedge;
    a.nedges = nedges;
    a.iters = iters;
    a.alpha = (T)alpha;
    a.tot = reinterpret_cast<T *>(p);
    p += align256((size_t)nblk * n * sizeof(T));
    a.msg = reinterpret_cast<T *>(p);
    p += align256((size_t)nblk * nedges * sizeof(T));
    a.done_it = reinterpret_cast<int *>(p);
    a.cnt = a.done_it + nblk;
    a.nblk = nblk;
    a.bits = bits;
    a.conv = conv;
    a.its = its;
    a.bad = bad;
    const dim3 vg((n + UTC_WG - 1) / UTC_WG, nblk);
    if (!KC && iters == 0) {
        hipLaunchKernelGGL(utc_zero<T>, vg, dim3(UTC_WG), 0, st, a);
        return launch_status();
    }
    const dim3 cg(nc > 0 ? (nc + UTC_WG - 1) / UTC_WG : 1, nblk);
    hipLaunchKernelGGL(utc_init, dim3((nblk + 255) / 256), dim3(256), 0, st, a.done_it, a.cnt,
                       nblk);
    for (int it = 0; it <= iters; ++it) {
        hipLaunchKernelGGL((utc_check<T, KC, NR>), cg, dim3(UTC_WG), 0, st, a, it);
        hipLaunchKernelGGL((utc_var<T, KC>), vg, dim3(UTC_WG), 0, st, a, it);
    }
    return launch_status();
}

size_t csr_phase_ws_bytes(int nblk, int n, int nedges) { return utc_ws_bytes(nblk, n, nedges, 4); }

int32_t csr_phase_kc_f32(const float *llr, int nblk, int n, int nc, int nedges, const int *cptr,
                         const int *cidx, const int *vptr, const int *vedge, int iters,
                         float alpha, void *ws, uint8_t *bits, uint8_t *conv, int *its,
                         int *bad, hipStream_t st) {
    return utc_launch<float, true, UTC_NAN_TORCH>(llr, nblk, n, nc, nedges, cptr, cidx, vptr,
                                                  vedge, 0, iters, (double)alpha, ws, bits,
                                                  conv, its, bad, st);
}

}  // namespace atsc3

extern "C" int64_t atsc3_minsum_utc_ws_bytes(int32_t nblk, int32_t n, int32_t nedges,
                                             int32_t dtype) {
    if (nblk < 0 || n < 0 || nedges < 0 || (dtype != 0 && dtype != 1)) return ATSC3_E_ARG;
    return (int64_t)atsc3::utc_ws_bytes(nblk, n, nedges, dtype == 1 ? 8 : 4);
}

extern "C" int32_t atsc3_minsum_utc_dev(const void *llr, int32_t dtype, int32_t nan_rule,
                                        int32_t nblk, int32_t n, int32_t nc, int32_t nedges,
                                        const int32_t *cptr, const int32_t *cidx,
                                        const int32_t *vptr, const int32_t *vedge,
                                        int32_t dmax, int32_t iters, double alpha, void *ws,
                                        int64_t ws_bytes, uint8_t *bits, uint8_t *conv,
                                        int32_t *its, int32_t *bad, void *stream) {
    if (nblk < 0 || n <= 0 || nc < 0 || nedges < 0 || iters < 0 || dmax < 0 ||
        (dtype != 0 && dtype != 1) || (nan_rule != UTC_NAN_TORCH && nan_rule != UTC_NAN_NUMPY) ||
        nblk > 65535)
        return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !cptr || !cidx || !vptr || !vedge || !ws || !bits || !conv || !its || !bad)
        return ATSC3_E_ARG;
    if (ws_bytes < (int64_t)atsc3::utc_ws_bytes(nblk, n, nedges, dtype == 1 ? 8 : 4))
        return ATSC3_E_SIZE;
    hipStream_t st = atsc3::as_stream(stream);
#define UTC_GO(T, NR)                                                                          \
    atsc3::utc_launch<T, false, NR>(static_cast<const T *>(llr), nblk, n, nc, nedges, cptr,     \
                                    cidx, vptr, vedge, dmax, iters, alpha, ws, bits, conv, its, \
                                    bad, st)
    if (dtype == 1)
        return nan_rule == UTC_NAN_NUMPY ? UTC_GO(double, UTC_NAN_NUMPY)
                                         : UTC_GO(double, UTC_NAN_TORCH);
    return nan_rule == UTC_NAN_NUMPY ? UTC_GO(float, UTC_NAN_NUMPY) : UTC_GO(float, UTC_NAN_TORCH);
#undef UTC_GO
}
