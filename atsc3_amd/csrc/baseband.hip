// BCH syndrome and descrambling of decoded codewords (Baseband Packet bytes).
//
// Receive-side step after the LDPC decode (lab/m6_bicm.py:357-364,
// lab/m9_fast.py:942-949): of each decoded codeword, the first kldpc bits
// form the BCH codeword.  One workgroup per codeword:
//   1. pack the kldpc bits (one byte each, codeword order) into 32-bit words
//      in LDS, message position t of word w at bit 31 - t;
//   2. syndrome bit j = parity over words of popcount(word & mask[w][j]):
//      a GF(2) matrix-vector product with the code's syndrome matrix (column
//      j = coefficient j of x^(kldpc-1-i) mod g for every position i; the
//      table is read coalesced, [w][j]), i.e. the message polynomial mod g;
//   3. payload byte k = message byte k XOR scrambling byte k, for the first
//      kpayload = kldpc - mouter bits.
// Outputs per codeword: the syndrome as 8 little-endian 32-bit limbs
// (coefficient j at limb j/32, bit j%32), bch_ok = (syndrome == 0), and the
// payload bytes.
#include "fec_common.h"

#define BB_WG 256
#define BB_MAX_WORDS 1824   // 58320-bit BCH codewords (64800 r13/15) / 32, rounded up

namespace atsc3 {

__global__ __launch_bounds__(BB_WG) void baseband_kernel(
    const uint8_t *__restrict__ bits, int n, int kldpc, int mouter,
    const uint32_t *__restrict__ mask, const uint8_t *__restrict__ seq,
    uint32_t *__restrict__ syn_out, uint8_t *__restrict__ ok_out, uint8_t *__restrict__ bytes_out) {
    __shared__ uint32_t msg[BB_MAX_WORDS];
    __shared__ uint32_t sbits[8];
    const int blk = blockIdx.x;
    const int tid = threadIdx.x;
    const int nwords = (kldpc + 31) / 32;
    const uint8_t *b = bits + (size_t)blk * n;
    for (int w = tid; w < nwords; w += BB_WG) {
        uint32_t v = 0u;
        const int base = 32 * w;
        for (int t = 0; t < 32; ++t) {
            const int i = base + t;
            const uint32_t bit = i < kldpc ? (uint32_t)(b[i] & 1u) : 0u;
            v |= bit << (31 - t);
        }
        msg[w] = v;
    }
    if (tid < 8) sbits[tid] = 0u;
    __syncthreads();
    for (int j = tid; j < mouter; j += BB_WG) {
        unsigned acc = 0u;
        for (int w = 0; w < nwords; ++w) acc += __popc(msg[w] & mask[(size_t)w * mouter + j]);
        if (acc & 1u) atomicOr(&sbits[j >> 5], 1u << (j & 31));
    }
    __syncthreads();
    if (tid < 8) syn_out[(size_t)blk * 8 + tid] = sbits[tid];
    if (tid == 0) {
        unsigned any = 0u;
        for (int q = 0; q < 8; ++q) any |= sbits[q];
        ok_out[blk] = (uint8_t)(any == 0u);
    }
    const int kbytes = (kldpc - mouter) / 8;
    uint8_t *out = bytes_out + (size_t)blk * kbytes;
    for (int k = tid; k < kbytes; k += BB_WG) {
        const uint32_t byte = (msg[k >> 2] >> (24 - 8 * (k & 3))) & 0xFFu;
        out[k] = (uint8_t)(byte ^ seq[k]);
    }
}

}  // namespace atsc3

extern "C" int32_t atsc3_baseband_dev(const uint8_t *bits, int32_t nblk, int32_t n,
                                      int32_t kldpc, int32_t mouter, const uint32_t *mask,
                                      const uint8_t *seq, uint32_t *syn_out, uint8_t *ok_out,
                                      uint8_t *bytes_out, void *stream) {
    if (nblk < 0 || n <= 0 || kldpc <= 0 || kldpc > n || mouter <= 0 || mouter > 256 ||
        mouter >= kldpc || (kldpc - mouter) % 8 != 0)
        return ATSC3_E_ARG;
    if ((kldpc + 31) / 32 > BB_MAX_WORDS) return ATSC3_E_SIZE;
    if (nblk == 0) return ATSC3_OK;
    if (!bits || !mask || !seq || !syn_out || !ok_out || !bytes_out) return ATSC3_E_ARG;
    hipLaunchKernelGGL(atsc3::baseband_kernel, dim3(nblk), dim3(BB_WG), 0,
                       atsc3::as_stream(stream), bits, n, kldpc, mouter, mask, seq, syn_out,
                       ok_out, bytes_out);
    return atsc3::launch_status();
}
