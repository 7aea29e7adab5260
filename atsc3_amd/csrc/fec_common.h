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

}  // namespace atsc3
