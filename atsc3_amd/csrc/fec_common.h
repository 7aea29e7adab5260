// Shared definitions for the gfx950 demap + min-sum kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/atsc3_fec.h"

#define ATSC3_Z 360            // A/322 lifting size: positions per group
#define ATSC3_WG 384           // 6 wave64s: one lane per group position

namespace atsc3 {

static inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

static inline int32_t launch_status() {
    return hipGetLastError() == hipSuccess ? ATSC3_OK : ATSC3_E_HIP;
}

// Sum of an int over a workgroup of ATSC3_WG threads (6 waves).  `slots`
// needs ATSC3_WG / 64 ints of LDS; every thread gets the total.  Contains
// one barrier; the caller must not reuse `slots` before the next barrier.
__device__ inline int wg_sum_i32(int x, int *slots) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) slots[wave] = x;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int w = 0; w < ATSC3_WG / 64; ++w) t += slots[w];
    return t;
}

// The compiled kernel's FP32 contract (lab/ldpc_kernel.c:82-143) as phase
// kernels over (codeword, check) and (codeword, variable) threads on CSR
// tables (ldpc_utc.hip); workspace csr_phase_ws_bytes(nblk, n, nedges)
// bytes; nblk <= 65535.
size_t csr_phase_ws_bytes(int nblk, int n, int nedges);
int32_t csr_phase_kc_f32(const float *llr, int nblk, int n, int nc, int nedges, const int *cptr,
                         const int *cidx, const int *vptr, const int *vedge, int iters,
                         float alpha, void *ws, uint8_t *bits, uint8_t *conv, int *its,
                         int *bad, hipStream_t st);

}  // namespace atsc3
