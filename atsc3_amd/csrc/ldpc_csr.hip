// Normalized min-sum over an arbitrary Tanner graph in CSR form, and the
// host-pointer drop-in for the upstream compiled kernel.
//
// The CSR tables are those of lab/m3_ldpc.py:459-482: cidx check-major with
// ascending columns inside each check, vedge = stable argsort of cidx (each
// variable's edges in ascending check index).  Semantics are
// lab/ldpc_kernel.c:82-143, statement for statement in FP32 with no
// contraction: this path walks the same edge orders the upstream loop does,
// so it needs no re-ordering argument at all.  It serves any code (degrees up
// to 128, the upstream limit) and is the engine of the drop-in
// ldpc_minsum_f32; the QC kernel is the fast path for A/322 codes.
#include "fec_common.h"

#define CSR_WG ATSC3_WG
#define CSR_MAX_DEG 128   // lab/ldpc_kernel.c:48

namespace atsc3 {

struct CsrArgs {
    const float *llr;
    int n, nc, nedges;
    const int *cptr, *cidx, *vptr, *vedge;
    int iters;
    float alpha;
    float *tot_ws, *msg_ws;
    uint8_t *bits, *conv;
    int *its, *bad;
};

__global__ __launch_bounds__(CSR_WG) void minsum_csr_kernel(CsrArgs a) {
    const int blk = blockIdx.x;
    const int tid = threadIdx.x;
    const int n = a.n;
    const float *lam = a.llr + (size_t)blk * n;
    float *tot = a.tot_ws + (size_t)blk * n;
    float *m = a.msg_ws + (size_t)blk * a.nedges;
    __shared__ int slots[CSR_WG / 64];

    for (int it = 0;; ++it) {
        const float *src = it == 0 ? lam : tot;
        const bool update = it < a.iters;
        int unsat = 0;
        for (int c = tid; c < a.nc; c += CSR_WG) {
            const int e0 = a.cptr[c], e1 = a.cptr[c + 1];
            unsigned syn = 0u, par = 0u;
            float lo1 = INFINITY, lo2 = INFINITY;
            int first = e0;
            for (int e = e0; e < e1; ++e) {
                const float tv = src[a.cidx[e]];
                const float xv = tv - (it == 0 ? 0.0f : m[e]);
                const float mag = fabsf(xv);
                syn ^= (unsigned)(tv < 0.0f);
                par ^= (unsigned)(xv < 0.0f);
                if (mag < lo1) {
                    lo2 = lo1;
                    lo1 = mag;
                    first = e;
                } else if (mag < lo2) {
                    lo2 = mag;
                }
            }
            unsat += (int)syn;
            if (update) {
                // second pass recomputes x exactly (same operands, same op)
                for (int e = e0; e < e1; ++e) {
                    const float xv = src[a.cidx[e]] - (it == 0 ? 0.0f : m[e]);
                    float val = (e == first ? lo2 : lo1) * a.alpha;
                    if (par ^ (unsigned)(xv < 0.0f)) val = -val;
                    m[e] = val;
                }
            }
        }
        const int total = wg_sum_i32(unsat, slots);
        if (total == 0 || it == a.iters) {
            uint8_t *bout = a.bits + (size_t)blk * n;
            for (int v = tid; v < n; v += CSR_WG) bout[v] = (uint8_t)(src[v] < 0.0f);
            if (tid == 0) {
                a.conv[blk] = (uint8_t)(total == 0);
                a.its[blk] = total == 0 ? it : a.iters;
                a.bad[blk] = total;
            }
            return;
        }
        for (int i = tid; i < n; i += CSR_WG) {
            float s = 0.0f;
            for (int k = a.vptr[i]; k < a.vptr[i + 1]; ++k) s += m[a.vedge[k]];
            tot[i] = lam[i] + s;
        }
        __syncthreads();
    }
}

static int32_t check_degrees(const int32_t *cptr_host, int32_t nc) {
    for (int32_t c = 0; c < nc; ++c)
        if (cptr_host[c + 1] - cptr_host[c] > CSR_MAX_DEG) return ATSC3_E_DEGREE;
    return ATSC3_OK;
}

}  // namespace atsc3

extern "C" int32_t atsc3_minsum_csr_f32_dev(const float *llr, int32_t nblk, int32_t n,
                                            int32_t nc, int32_t nedges,
                                            const int32_t *cptr, const int32_t *cidx,
                                            const int32_t *vptr, const int32_t *vedge,
                                            int32_t iters, float alpha, float *tot_ws,
                                            float *msg_ws, uint8_t *bits, uint8_t *conv,
                                            int32_t *its, int32_t *bad, void *stream) {
    if (nblk < 0 || n <= 0 || nc < 0 || nedges < 0 || iters < 0) return ATSC3_E_ARG;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !cptr || !cidx || !vptr || !vedge || !tot_ws || !msg_ws || !bits ||
        !conv || !its || !bad)
        return ATSC3_E_ARG;
    atsc3::CsrArgs a{llr,   n,     nc,     nedges, cptr,   cidx, vptr, vedge,
                     iters, alpha, tot_ws, msg_ws, bits,   conv, its,  bad};
    hipLaunchKernelGGL(atsc3::minsum_csr_kernel, dim3(nblk), dim3(CSR_WG), 0,
                       atsc3::as_stream(stream), a);
    return atsc3::launch_status();
}

// ---------------------------------------------------------------------------
// Drop-in: the upstream C ABI (lab/ldpc_kernel.c:52,65-71), host pointers.
// ---------------------------------------------------------------------------
extern "C" int32_t ldpc_kernel_abi(void) { return 1; }

namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes > 0 ? bytes : 1); }
};
}  // namespace

extern "C" int32_t ldpc_minsum_f32(const float *llr, int32_t nblk, int32_t n, int32_t nc,
                                   int32_t nedges, const int32_t *cptr,
                                   const int32_t *cidx, const int32_t *vptr,
                                   const int32_t *vedge, int32_t iters, float alpha,
                                   uint8_t *bits, uint8_t *conv, int32_t *its,
                                   int32_t *bad, float *work) {
    (void)work;
    if (nblk < 0 || n <= 0 || nc < 0 || nedges < 0) return ATSC3_E_ARG;
    if (!cptr) return ATSC3_E_ARG;
    if (atsc3::check_degrees(cptr, nc) != ATSC3_OK) return ATSC3_E_DEGREE;
    if (nblk == 0) return ATSC3_OK;
    const size_t nb = (size_t)nblk, nn = (size_t)n, ne = (size_t)nedges;
    DevBuf d_llr, d_cptr, d_cidx, d_vptr, d_vedge, d_tot, d_msg, d_bits, d_conv, d_its,
        d_bad;
    hipStream_t st = nullptr;
    if (hipStreamCreate(&st) != hipSuccess) return ATSC3_E_HIP;
    int32_t rc = ATSC3_E_HIP;
    do {
        if (d_llr.alloc(nb * nn * 4) || d_cptr.alloc(((size_t)nc + 1) * 4) ||
            d_cidx.alloc(ne * 4) || d_vptr.alloc((nn + 1) * 4) || d_vedge.alloc(ne * 4) ||
            d_tot.alloc(nb * nn * 4) || d_msg.alloc(nb * ne * 4) || d_bits.alloc(nb * nn) ||
            d_conv.alloc(nb) || d_its.alloc(nb * 4) || d_bad.alloc(nb * 4))
            break;
        if (hipMemcpyAsync(d_llr.p, llr, nb * nn * 4, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(d_cptr.p, cptr, ((size_t)nc + 1) * 4, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(d_cidx.p, cidx, ne * 4, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(d_vptr.p, vptr, (nn + 1) * 4, hipMemcpyHostToDevice, st) ||
            hipMemcpyAsync(d_vedge.p, vedge, ne * 4, hipMemcpyHostToDevice, st))
            break;
        rc = atsc3_minsum_csr_f32_dev(
            (const float *)d_llr.p, nblk, n, nc, nedges, (const int32_t *)d_cptr.p,
            (const int32_t *)d_cidx.p, (const int32_t *)d_vptr.p, (const int32_t *)d_vedge.p,
            iters, alpha, (float *)d_tot.p, (float *)d_msg.p, (uint8_t *)d_bits.p,
            (uint8_t *)d_conv.p, (int32_t *)d_its.p, (int32_t *)d_bad.p, st);
        if (rc != ATSC3_OK) break;
        rc = ATSC3_E_HIP;
        if (hipMemcpyAsync(bits, d_bits.p, nb * nn, hipMemcpyDeviceToHost, st) ||
            hipMemcpyAsync(conv, d_conv.p, nb, hipMemcpyDeviceToHost, st) ||
            hipMemcpyAsync(its, d_its.p, nb * 4, hipMemcpyDeviceToHost, st) ||
            hipMemcpyAsync(bad, d_bad.p, nb * 4, hipMemcpyDeviceToHost, st))
            break;
        if (hipStreamSynchronize(st) != hipSuccess) break;
        rc = ATSC3_OK;
    } while (0);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}
