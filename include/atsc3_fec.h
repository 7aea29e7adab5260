/* atsc3_fec.h -- C ABI of libatsc3fec.so, the MI355X (gfx950) demap + LDPC
 * min-sum library.
 *
 * Conventions
 *   - Plain pointers and sizes only.  Functions whose name ends in _dev take
 *     DEVICE pointers (memory the caller allocated on the current HIP device)
 *     and enqueue work on `stream` (a hipStream_t passed as void*, NULL = the
 *     null stream); they return as soon as the work is enqueued.
 *   - Functions without the suffix take HOST pointers and are synchronous.
 *   - Every entry point returns 0 on success or a negative ATSC3_E* code; on
 *     error nothing is written to the outputs.
 *   - LLR sign convention: llr > 0 means bit 0 (lab/m6_bicm.py:213).
 *   - Thread safety: concurrent calls on different streams with disjoint
 *     buffers are safe; the host-pointer drop-in keeps a mutex-guarded pool
 *     of per-call contexts and a per-device table cache.
 */
#ifndef ATSC3_FEC_H
#define ATSC3_FEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATSC3_OK 0
#define ATSC3_E_DEGREE (-1)   /* a check degree exceeds the supported maximum */
#define ATSC3_E_ARG (-2)      /* inconsistent sizes or NULL pointers          */
#define ATSC3_E_HIP (-3)      /* a HIP runtime call or kernel launch failed   */
#define ATSC3_E_SIZE (-4)     /* a size exceeds a fixed capacity             */

/* ------------------------------------------------------------------------
 * Drop-in for the upstream compiled kernel.
 * Replaces lab/ldpc_kernel.c:52  int32_t ldpc_kernel_abi(void)
 * ------------------------------------------------------------------------ */
int32_t ldpc_kernel_abi(void);

/* Replaces lab/ldpc_kernel.c:65-71 ldpc_minsum_f32 with identical arguments
 * and results (bits, conv, its, bad bit-exact).  Host pointers; the decode
 * runs on the current HIP device (H2D, kernel, D2H) and `work` is unused.
 * Returns -1 when a check degree exceeds 128, like the upstream kernel,
 * with nothing written.  Thread-safe: concurrent calls take separate
 * pooled contexts (stream, grow-only device and pinned buffers), so after
 * the first calls none allocates.  Calls of an A/322 code that arrive
 * together from several threads (same code, iters, alpha) are decoded in
 * one launch, each caller getting its own blocks' results (a call waits up
 * to 0.5 ms for other recently active threads to join; a lone caller never
 * waits); a failure returns the same error to each, writing nothing.       */
int32_t ldpc_minsum_f32(const float *llr, int32_t nblk, int32_t n, int32_t nc,
                        int32_t nedges, const int32_t *cptr,
                        const int32_t *cidx, const int32_t *vptr,
                        const int32_t *vedge, int32_t iters, float alpha,
                        uint8_t *bits, uint8_t *conv, int32_t *its,
                        int32_t *bad, float *work);

/* Which engine ldpc_minsum_f32 uses for these tables: 1 = the resident
 * engine's FP32 mode (an A/322 code found, by a hash of the four arrays,
 * among the tables the build writes beside the library,
 * libatsc3fec_tables.bin), 0 = the CSR phase kernels (any other graph).
 * No device call; no upstream counterpart (a diagnostic of the drop-in). */
int32_t atsc3_dropin_engine(int32_t n, int32_t nc, int32_t nedges,
                            const int32_t *cptr, const int32_t *cidx,
                            const int32_t *vptr, const int32_t *vedge);

/* ------------------------------------------------------------------------
 * Device-pointer min-sum over the Tanner graph in CSR form.
 * Same semantics as ldpc_minsum_f32 (lab/ldpc_kernel.c:65-146); serves any
 * parity-check matrix.  Workspace: tot_ws [nblk*n] f32, msg_ws [nblk*nedges]
 * f32 (contents on entry are ignored).
 * ------------------------------------------------------------------------ */
int32_t atsc3_minsum_csr_f32_dev(const float *llr, int32_t nblk, int32_t n,
                                 int32_t nc, int32_t nedges,
                                 const int32_t *cptr, const int32_t *cidx,
                                 const int32_t *vptr, const int32_t *vedge,
                                 int32_t iters, float alpha, float *tot_ws,
                                 float *msg_ws, uint8_t *bits, uint8_t *conv,
                                 int32_t *its, int32_t *bad, void *stream);

/* ------------------------------------------------------------------------
 * Update-then-check min-sum in FP32 (dtype 0) or FP64 (dtype 1), the
 * contract of lab/m9_gpu.py:138-211 GpuMinSum.decode (float32 / float64) and
 * lab/m3_ldpc.py:245-286 min_sum_decode, lab/m3_ldpc.py:509-630
 * min_sum_decode_batch(dtype=float64): messages updated first, the syndrome
 * of tot < 0 checked after each update (its >= 1 when iters >= 1);
 * tot[v] = ((llr[v] + m_e1) + m_e2) + ... in ascending check order; each
 * check's two-minimum also sees (dmax - degree) pad entries of magnitude
 * 1e9 (the upstream's rows padded to the largest degree dmax); alpha is
 * rounded to T.  iters == 0: bits = llr < 0, conv 0, its 0, bad -1.
 *   nan_rule how a check whose magnitudes include NaN (NaN LLRs, inf - inf)
 *            updates, where the upstream's functions differ:
 *            0 = GpuMinSum's (lab/m9_gpu.py:160-166: amin propagates NaN, so
 *            every message is NaN but edge 0's, which takes the smallest
 *            other entry when it holds the only NaN);
 *            1 = min_sum_decode's (lab/m3_ldpc.py:264-270: the two smallest
 *            non-NaN entries, the first NaN edge taking the second).
 *            Checks without NaN are the same under both.
 *   llr      [nblk][n] float (dtype 0) or double (dtype 1), device
 *   cptr..vedge  CSR tables as ldpc_minsum_f32 (any degrees)
 *   dmax     the largest check degree of the code (the row padding width)
 *   ws       atsc3_minsum_utc_ws_bytes(nblk, n, nedges, dtype) bytes
 * 2 * (iters + 1) + 1 launches on `stream`; nblk <= 65535.
 * ------------------------------------------------------------------------ */
int64_t atsc3_minsum_utc_ws_bytes(int32_t nblk, int32_t n, int32_t nedges,
                                  int32_t dtype);
int32_t atsc3_minsum_utc_dev(const void *llr, int32_t dtype, int32_t nan_rule,
                             int32_t nblk, int32_t n, int32_t nc, int32_t nedges,
                             const int32_t *cptr, const int32_t *cidx,
                             const int32_t *vptr, const int32_t *vedge,
                             int32_t dmax, int32_t iters, double alpha,
                             void *ws, int64_t ws_bytes, uint8_t *bits,
                             uint8_t *conv, int32_t *its, int32_t *bad,
                             void *stream);

/* ------------------------------------------------------------------------
 * Device-pointer min-sum over the 360-lifted QC base graph.
 * The fast path behind the upstream min-sum calls (lab/m3_ldpc.py:509-630
 * min_sum_decode_batch at float32 -> lab/ldpc_kernel.c:65; lab/m9_gpu.py:138
 * GpuMinSum.decode).  Tables come from atsc3_amd.codes.LdpcCode
 * .kernel_tables(): cg_info[ncg], be_off[2*nbe], be_miss[nbe],
 * vg_info[nvg], ve_off[2*nbe], ve_miss[nbe], c_steps[nsc], v_steps[nsv]
 * (groups in processing order, packed into steps of equal degree);
 * type_a = 1 for A/322 Type A codes (two check regions).
 * msg_fmt selects the check-to-variable message storage:
 *   0 = FP32 (plain mode: bits, conv, its, bad bit-exact with
 *       lab/ldpc_kernel.c on the same LLRs),
 *   1 = FP16 storage, 2 = INT8 storage with a per-block scale
 *       (msg_scale[b] = value of one INT8 step for block b; NULL -> derived
 *       on the device as 6 * median|llr| / 127, the median to 8 mantissa
 *       bits; saturating at +-127 steps).
 * Variable sums always accumulate in FP32.  Any check degree; variable
 * degree <= 32 (every A/322 code), else ATSC3_E_DEGREE.
 * Workspace (contents on entry ignored): tot_ws [nblk * nvg * 384] f32;
 * msg_ws nblk * nbe * 384 elements of the message type (4, 2 or 1 bytes).
 * ------------------------------------------------------------------------ */
int32_t atsc3_minsum_qc_dev(const float *llr, int32_t nblk, int32_t n,
                            int32_t ncg, int32_t nvg, int32_t nbe,
                            int32_t type_a, const int32_t *cg_info,
                            const int32_t *be_off, const int32_t *be_miss,
                            const int32_t *vg_info, const int32_t *ve_off,
                            const int32_t *ve_miss, const int32_t *c_steps,
                            int32_t nsc, const int32_t *v_steps, int32_t nsv,
                            int32_t max_cdeg,
                            int32_t max_vdeg, int32_t iters, float alpha,
                            int32_t msg_fmt,
                            const float *msg_scale, float *tot_ws,
                            void *msg_ws, uint8_t *bits, uint8_t *conv,
                            int32_t *its, int32_t *bad, void *stream);

/* ------------------------------------------------------------------------
 * Same decode, same results, as per-iteration phase kernels: the throughput
 * engine for large batches (hundreds of blocks and more).  Enqueues about
 * 3 * (iters + 1) kernels on `stream`; retired blocks drop out of the work
 * lists, and once all have retired the remaining launches exit at once.
 * Extra workspace: ctl_ws of atsc3_minsum_qc_ctl_words(nblk) int32 words.
 * Other arguments as atsc3_minsum_qc_dev.
 * ------------------------------------------------------------------------ */
int32_t atsc3_minsum_qc_ctl_words(int32_t nblk);
int32_t atsc3_minsum_qc_phased_dev(const float *llr, int32_t nblk, int32_t n,
                                   int32_t ncg, int32_t nvg, int32_t nbe,
                                   int32_t type_a, const int32_t *cg_info,
                                   const int32_t *be_off,
                                   const int32_t *be_miss,
                                   const int32_t *vg_info,
                                   const int32_t *ve_off,
                                   const int32_t *ve_miss,
                                   const int32_t *c_steps, int32_t nsc,
                                   const int32_t *v_steps, int32_t nsv,
                                   int32_t max_cdeg, int32_t max_vdeg,
                                   int32_t iters, float alpha, int32_t msg_fmt,
                                   const float *msg_scale, float *tot_ws,
                                   void *msg_ws, int32_t *ctl_ws, uint8_t *bits,
                                   uint8_t *conv, int32_t *its, int32_t *bad,
                                   void *stream);

/* ------------------------------------------------------------------------
 * Batched min-sum over a queue of codewords of MIXED A/322 codes in one
 * persistent launch (plus one init launch).  Same per-codeword semantics and
 * results as ldpc_minsum_f32 (lab/ldpc_kernel.c:65-146) in FP32 mode; it
 * replaces the upstream's per-code sub-batched calls
 * (lab/m9_fast.py:843-925 FrameDecoder.ldpc_batch, lab/m44_ldm.py:2497-2576
 * ldpc_batch) with one call for a frame's whole mix.
 *
 *   codes     device array of ncode atsc3_code_tab (tables of each code, as
 *             built by the Python layer from the A/322 parity tables)
 *   blk_code  [nblk] int32 index into codes (device)
 *   blk_off   [nblk] int64 offset of block b's LLRs in llr and of its bits
 *             in bits (device; the rows may have different lengths)
 *   max_nvg, max_nbe, max_vdeg  maxima over the codes (variable groups,
 *             base edges, variable degree <= 32)
 *   ladder    NULL: one decode with alpha.  Else host array of nrungs
 *             (1..4) alphas, the alpha ladder of lab/m9_fast.py:843-891 run
 *             inside the kernel: a codeword failing rung r restarts from its
 *             LLRs at rung r+1 in the same slot; a rung's result (bits,
 *             conv, its, bad) is kept when it converged or has fewer
 *             unsatisfied checks than every earlier rung
 *   msg_fmt   0 FP32 messages (bit-exact with the reference), 1 FP16, 2 INT8
 *             (the storage codecs of atsc3_minsum_qc_dev)
 *   slots     codewords decoded concurrently; the rest wait in the queue
 *   grid      workgroups of the persistent kernel, 1 ..
 *             atsc3_minsum_batch_max_grid(msg_fmt) (all must be resident)
 *   tot_ws    [slots * max_nvg * 384] f32
 *   msg_ws    slots * max_nbe * 384 elements of the message format
 *   ctl_ws    atsc3_minsum_batch_ctl_words(nblk, slots) int32; after the
 *             call ctl_ws[atsc3_minsum_batch_status_word()] != 0 means a
 *             workgroup timed out waiting at a grid barrier; every codeword
 *             the launch did not finish then reads its = -1, bad = -1
 *             (its >= 0 for every codeword of a completed launch)
 * ------------------------------------------------------------------------ */
typedef struct atsc3_code_tab {
    int32_t n, ncg, nvg, nbe, nsc, nsv, type_a, max_vdeg;
    const int32_t *cg_info, *be_off, *be_miss, *vg_info, *ve_off, *ve_miss,
        *c_steps, *v_steps;
} atsc3_code_tab;

int32_t atsc3_minsum_batch_ctl_words(int32_t nblk, int32_t slots);
int32_t atsc3_minsum_batch_max_grid(int32_t msg_fmt);
int32_t atsc3_minsum_batch_status_word(void);
int32_t atsc3_minsum_batch_dev(const float *llr, int32_t nblk,
                               const int64_t *blk_off, const int32_t *blk_code,
                               const void *codes, int32_t ncode,
                               int32_t max_nvg, int32_t max_nbe,
                               int32_t max_vdeg, int32_t iters,
                               float alpha, const float *ladder,
                               int32_t nrungs, int32_t msg_fmt,
                               const float *msg_scale, int32_t slots,
                               int32_t grid, float *tot_ws, void *msg_ws,
                               int32_t *ctl_ws, uint8_t *bits, uint8_t *conv,
                               int32_t *its, int32_t *bad, void *stream);

/* ------------------------------------------------------------------------
 * XCD-partitioned batched min-sum: the decode of atsc3_minsum_batch_dev
 * (same arguments, results bit-identical) for frames whose slowest codeword
 * sets their time.  The workgroups of each XCD decode their own share of the
 * queue in their own slots with barriers among themselves only (no L2
 * write-back, no wait for the other XCDs), so a 50-iteration codeword no
 * longer holds every workgroup of the chip to its rounds.
 *   order     [nblk] int32 device: a permutation of 0 .. nblk-1
 *   part_off  [nparts + 1] int32 device, 0 = part_off[0] <= ... <=
 *             part_off[nparts] = nblk: part c is order[part_off[c] ..
 *             part_off[c + 1]); the k-th of the ng XCDs holding workgroups
 *             decodes parts k, k + ng, ... in that order (balance the parts:
 *             a launch takes as long as its slowest XCD)
 *   slots     >= 16; each XCD gets slots / ng of them
 *   grid      1 .. atsc3_minsum_batch_xcd_max_grid(msg_fmt)
 *   ctl_ws    atsc3_minsum_batch_xcd_ctl_words(nblk, slots) int32 (status
 *             word as atsc3_minsum_batch_dev)
 * ------------------------------------------------------------------------ */
int32_t atsc3_minsum_batch_xcd_ctl_words(int32_t nblk, int32_t slots);
int32_t atsc3_minsum_batch_xcd_max_grid(int32_t msg_fmt);
int32_t atsc3_minsum_batch_xcd_dev(const float *llr, int32_t nblk,
                                   const int64_t *blk_off, const int32_t *blk_code,
                                   const void *codes, int32_t ncode,
                                   int32_t max_nvg, int32_t max_nbe,
                                   int32_t max_vdeg, int32_t iters,
                                   float alpha, const float *ladder,
                                   int32_t nrungs, int32_t msg_fmt,
                                   const float *msg_scale, const int32_t *order,
                                   const int32_t *part_off, int32_t nparts,
                                   int32_t slots, int32_t grid, float *tot_ws,
                                   void *msg_ws, int32_t *ctl_ws, uint8_t *bits,
                                   uint8_t *conv, int32_t *its, int32_t *bad,
                                   void *stream);

/* ------------------------------------------------------------------------
 * Resident min-sum: the same decode calls as atsc3_minsum_batch_dev replaces
 * (lab/m9_fast.py:843-925, lab/m44_ldm.py:2497-2576 over
 * lab/ldpc_kernel.c:65-146), one workgroup per codeword holding the
 * codeword's variable posteriors in LDS and each check's messages as one
 * compressed record (alpha*min1, alpha*min2, the edge taking min2, the
 * signs).  Two storage modes:
 *   msg_fmt 1  FP16: posteriors in LDS as FP16, FP16 record magnitudes (the
 *              FP16 message format's values); each variable's messages
 *              summed in FP32 in the order of its entry list (the
 *              atsc3_res_code edge_v order), the posterior rounded to FP16
 *              where it is stored.
 *   msg_fmt 0  FP32, bit-exact with lab/ldpc_kernel.c (bits, conv, its,
 *              bad): FP32 posteriors, the rows of the variable groups whose
 *              grp_v word says so in LDS (at most
 *              atsc3_minsum_resident_lds_rows(0) rows), the other rows in the
 *              workgroup's global slot; FP32 record magnitudes.
 *   msg_fmt 2  INT8: as FP16 with the record magnitudes stored as INT8
 *              levels of the codeword's step (rint(a / step) saturated at
 *              127, the INT8 message format's codec); msg_scale[b] = the
 *              step of codeword b, or NULL: 6 * median|llr| / 127 computed on
 *              the device (the median to 8 mantissa bits), as the other
 *              engines' INT8 format.
 * Codewords of mixed codes (n <= 64800, check degree <= 89, variable degree
 * <= 32: every A/322 code) are taken from a device queue by `grid`
 * workgroups (atsc3_minsum_resident_grid() = one per compute unit; any
 * grid >= 1 is valid: workgroups never wait for each other).
 *
 *   codes      device array of ncode atsc3_res_code built for this msg_fmt
 *   blk_code, blk_off  as atsc3_minsum_batch_dev
 *   max_checks largest ncg * 360 over the codes
 *   max_global_rows  largest number of posterior rows a code keeps in the
 *              global slot (0 for msg_fmt 1)
 *   max_n, max_dc, max_dv  maxima over the codes
 *   alpha      normalization, >= 0 (the records hold magnitudes)
 *   ws         device workspace of atsc3_minsum_resident_ws_bytes(grid,
 *              max_checks, max_global_rows, msg_fmt) bytes (queue word,
 *              per-workgroup records and posterior rows)
 * Outputs bits/conv/its/bad with the plain semantics (its = iteration at
 * which the syndrome read zero, else iters; bad = unsatisfied checks).
 * Two launches on `stream` (queue reset, decode).
 * ------------------------------------------------------------------------ */
typedef struct atsc3_res_code {
    int32_t n, ncg, nvg, nbe, max_dc, max_dv, nsc, nsv;
    const int32_t *grp_c;   /* per check group: dc | b0 << 7 | has_missing << 18 */
    const int32_t *edge_c;  /* per base edge: s | row << 9 | (missing j + 1) << 18
                               | global << 28 (row: the variable group's posterior
                               row in LDS, or in the global slot)                 */
    const int32_t *grp_v;   /* per variable group: dv | e0 << 6 | n0 << 17 |
                               row << 23 | global << 31                           */
    const int32_t *edge_v;  /* per entry, 2 words: u | g << 9 | k << 17 |
                               sp << 24 (sp: the bit of the check's record that
                               holds edge k's sign: dc - 1 - k for checks of
                               the non-generic steps, else k); missing variable
                               position + 1                                      */
    const int32_t *c_steps; /* [nsc] check steps: D | log2(G) << 7 | start << 10 |
                               generic << 22, over c_tasks                       */
    const int32_t *c_tasks; /* group | c6 << 8: rows c6*64 .. c6*64+63         */
    const int32_t *v_steps; /* [nsv] variable steps, same packing, over v_tasks,
                               | Type A region << 23                             */
    const int32_t *v_tasks;
    const int32_t *c_off;   /* FP16/INT8: [nbe * 6] at b*6 + c6, the byte offset
                               in LDS of the posterior that check (g, c6*64)
                               meets on base edge b; rows of 424 halves
                               (positions 0..359, then 0..63 again)           */
    const int32_t *v_soff;  /* [nsv] offset into v_sdata of each variable step
                               (generic steps: unused)                         */
    const int32_t *v_sdata; /* per non-generic variable step, for each of its G
                               tasks: t | c6 << 8 | n0 << 11 | row << 17 |
                               global << 25, then the group's D edge_v word-0
                               entries                                          */
    const int32_t *c_soff;  /* [nsc] offset into c_sdata of each check step    */
    const int32_t *c_sdata; /* per non-generic check step, for each of its G
                               tasks: g | c6 << 8, then D words: msg_fmt 0 the
                               group's edge_c words, else its c_off offsets for
                               c6                                               */
} atsc3_res_code;

int64_t atsc3_minsum_resident_ws_bytes(int32_t grid, int32_t max_checks,
                                       int32_t max_global_rows, int32_t msg_fmt);
int32_t atsc3_minsum_resident_lds_rows(int32_t msg_fmt);
int32_t atsc3_minsum_resident_grid(void);
int32_t atsc3_minsum_resident_dev(const float *llr, int32_t nblk,
                                  const int64_t *blk_off, const int32_t *blk_code,
                                  const void *codes, int32_t ncode,
                                  int32_t max_checks, int32_t max_global_rows,
                                  int32_t max_n, int32_t max_dc, int32_t max_dv,
                                  int32_t iters, float alpha, int32_t msg_fmt,
                                  const float *msg_scale, int32_t grid, void *ws,
                                  int64_t ws_bytes,
                                  uint8_t *bits, uint8_t *conv, int32_t *its,
                                  int32_t *bad, void *stream);

/* The same decode also writing the final posterior LLR of every codeword
 * bit to llr_out (laid out like bits; FP32 mode the FP32 posteriors, FP16 /
 * INT8 modes the stored FP16 posteriors widened to FP32): the soft output
 * the LLR-error measurement of the message forms reads.  No upstream
 * counterpart (the upstream decoders return hard bits only). */
int32_t atsc3_minsum_resident_llr_dev(const float *llr, int32_t nblk,
                                      const int64_t *blk_off,
                                      const int32_t *blk_code,
                                      const void *codes, int32_t ncode,
                                      int32_t max_checks,
                                      int32_t max_global_rows, int32_t max_n,
                                      int32_t max_dc, int32_t max_dv,
                                      int32_t iters, float alpha,
                                      int32_t msg_fmt, const float *msg_scale,
                                      int32_t grid, void *ws, int64_t ws_bytes,
                                      uint8_t *bits, uint8_t *conv,
                                      int32_t *its, int32_t *bad,
                                      float *llr_out, void *stream);

/* ------------------------------------------------------------------------
 * Resident min-sum, team mode: the same decode calls (lab/m9_fast.py:843-925,
 * lab/m44_ldm.py:2497-2576 over lab/ldpc_kernel.c:65-146) for frames whose
 * few slow codewords would otherwise leave one compute unit decoding alone:
 * each codeword is decoded by a team of `team` (1..32) workgroups on one
 * XCD, which split its check and variable steps and synchronise only with
 * each other (two team barriers per iteration, no L2 write-back); every
 * member keeps the whole posterior vector in LDS.  Results are those of
 * atsc3_minsum_resident_dev with the same codes, msg_fmt and msg_scale, bit
 * for bit.  Teams are formed at launch from the workgroups each XCD holds,
 * so every workgroup must be co-resident: grid <= atsc3_minsum_resident_grid()
 * (one workgroup per compute unit; larger grids return ATSC3_E_SIZE).
 *   order  NULL, or nblk codeword indices: the queue's order (a team takes
 *          the codewords order[0], order[1], ... as it becomes free)
 *   ladder, nrungs  nrungs = 0: one decode at `alpha`; 1..4: the alpha
 *          ladder of lab/m9_fast.py:843-891 inside the launch, `ladder` the
 *          host array of the rungs' alphas (each >= 0; `alpha` unused): each
 *          rung decodes the codeword afresh from its LLRs, a rung's result
 *          is taken when it converged or leaves fewer unsatisfied checks
 *          than the best so far (the first rung's always), and a codeword
 *          stops at its first converged rung -- the results of
 *          atsc3_amd.ldpc.ladder_dev / the batched engine's ladder
 *   ws     atsc3_minsum_resident_team_ws_bytes(grid, team, max_checks,
 *          max_global_rows, msg_fmt) bytes: control words, per-team records
 *          and published posterior rows
 * Every wait is bounded: if one runs out (e.g. another kernel holds compute
 * units for seconds), the int32 word atsc3_minsum_resident_team_status_word()
 * of ws is nonzero after the launch and the codewords left undecoded hold
 * its = -1, bad = -1, conv = 0.  Two launches on `stream` (init, decode);
 * runs sharing one ws must not overlap.  No upstream counterpart as an
 * entry point (the upstream's decode calls map onto it as above).
 * ------------------------------------------------------------------------ */
/* A queue order for team mode: order[0..nblk) = the codewords by
 * decreasing fraction of unsatisfied checks of the channel's hard decisions
 * (sign of llr), sampled on positions 0..63 of every check group (ties: lower
 * index first), so that the codewords likely to need the most iterations
 * start first.  codes: the atsc3_res_code array of msg_fmt's tables; key:
 * nblk floats of device scratch (receives the fractions).  Two launches on
 * `stream`; the order changes only scheduling, never results. */
#define ATSC3_TEAM_DIRECT 0x100   /* OR-ed into `team`: no LDS copy of the
                                     posteriors; every member reads the team's
                                     rows in global memory (L2) instead       */
/* ATSC3_TEAM_GROW, OR-ed into `team` (of the decode and of the ws query):
 * grow mode.  Every codeword starts on one workgroup, as
 * atsc3_minsum_resident_dev decodes it; once the queue is empty, workgroups
 * with nothing left to start wait on their XCD, and a codeword still running
 * takes `team` - 1 of them (team 1: never) after a variable phase and runs
 * its remaining iterations as a team (the direct flavor's global rows).  It
 * removes the idle tail of a launch whose last codewords run long.  Same
 * results bit for bit; ladder must be absent (nrungs 0); order as above. */
#define ATSC3_TEAM_GROW 0x200
int32_t atsc3_minsum_resident_team_order_dev(const float *llr, int32_t nblk,
                                             const int64_t *blk_off,
                                             const int32_t *blk_code,
                                             const void *codes, int32_t msg_fmt,
                                             float *key, int32_t *order,
                                             void *stream);
int64_t atsc3_minsum_resident_team_ws_bytes(int32_t grid, int32_t team,
                                            int32_t max_checks,
                                            int32_t max_global_rows,
                                            int32_t msg_fmt);
int32_t atsc3_minsum_resident_team_status_word(void);
int32_t atsc3_minsum_resident_team_dev(
    const float *llr, int32_t nblk, const int64_t *blk_off,
    const int32_t *blk_code, const void *codes, int32_t ncode,
    int32_t max_checks, int32_t max_global_rows, int32_t max_n, int32_t max_dc,
    int32_t max_dv, int32_t iters, float alpha, int32_t msg_fmt,
    const float *msg_scale, const int32_t *order, const float *ladder,
    int32_t nrungs, int32_t grid, int32_t team, void *ws, int64_t ws_bytes,
    uint8_t *bits, uint8_t *conv, int32_t *its, int32_t *bad, void *stream);

/* ------------------------------------------------------------------------
 * Resident min-sum with no allocation.  No counterpart in the upstream: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as atsc3_minsum_resident_dev without ws/ws_bytes:
 * the workspace is a fixed device array inside the library of
 * atsc3_minsum_resident_noalloc_max_bytes() bytes (128 MiB) on each GPU the
 * library runs on, held until the process ends.  Returns ATSC3_E_SIZE
 * (nothing enqueued) when atsc3_minsum_resident_ws_bytes(grid, max_checks,
 * max_global_rows, msg_fmt) exceeds it -- lower `grid` to fit.  Every call
 * on one GPU shares that array: calls that overlap on one GPU race, so
 * enqueue them on one stream.
 * ------------------------------------------------------------------------ */
int64_t atsc3_minsum_resident_noalloc_max_bytes(void);
int32_t atsc3_minsum_resident_noalloc_dev(
    const float *llr, int32_t nblk, const int64_t *blk_off,
    const int32_t *blk_code, const void *codes, int32_t ncode,
    int32_t max_checks, int32_t max_global_rows, int32_t max_n, int32_t max_dc,
    int32_t max_dv, int32_t iters, float alpha, int32_t msg_fmt,
    const float *msg_scale, int32_t grid, uint8_t *bits, uint8_t *conv,
    int32_t *its, int32_t *bad, void *stream);

/* ------------------------------------------------------------------------
 * Max-log NUC demapper with fused bit de-interleave.
 * Replaces the upstream demap calls lab/m6_bicm.py:213-225 (demap_llr),
 * lab/m9_fast.py:756 (FrameDecoder.demap_batch), lab/m44_ldm.py:2454
 * (demap_batch_qam) and lab/m9_gpu.py:301 (GpuBicm.demap), followed by the
 * caller's scatter lam[lam_of_q] = q (lab/m6_bicm.py:331-335).
 *
 *   cells   [nblk*ncell] complex64 (interleaved re, im), equalised
 *   csi     [nblk*ncell] f32 per-cell weight |h|^2, or NULL (weight 1)
 *   sigma2  [nblk] f64 noise variance, or NULL: estimated per block as
 *           max(mean(w * min_p |z - p|^2), 1e-9) in FP64 on the device
 *   points  [M] complex64, M = 2^nbits, label = index (MSB = y0)
 *   lam_of_q[ncell*nbits] int32 codeword position of q-stream bit
 *   llr_out [nblk*ncell*nbits] f32 in codeword order
 *   sigma2_out [nblk] f64 (the value used), may be NULL
 *   ws      f64 scratch of atsc3_demap_ws_doubles(nblk, ncell) elements,
 *           needed only when sigma2 is NULL
 * LLR_i = w * (min_{y_i=1} |z-p|^2 - min_{y_i=0} |z-p|^2) / sigma2.
 * ------------------------------------------------------------------------ */
int32_t atsc3_demap_ws_doubles(int32_t nblk, int32_t ncell);
int32_t atsc3_demap_dev(const float *cells, const float *csi,
                        const double *sigma2, int32_t nblk, int32_t ncell,
                        const float *points, int32_t nbits,
                        const int32_t *lam_of_q, float *llr_out,
                        double *sigma2_out, double *ws, void *stream);

/* Several groups of blocks, each with its own (N, constellation, rate), in
 * ONE launch: the per-(code, constellation) demap calls of a frame's FEC
 * stage (lab/m9_fast.py:1087-1166, one atsc3_demap_dev each) as one kernel,
 * with atsc3_demap_dev's arithmetic (bit-identical LLRs), sigma2 given.
 * groups: device array of ngroups descriptors in increasing blk0, where
 * group g covers the frame's blocks blk0 .. blk0 + nblk - 1; nblk_total =
 * the frame's blocks (<= 65535), max_ncell = the largest ncell.  Every
 * pointer in a descriptor is a device pointer (csi may be NULL). */
typedef struct atsc3_demap_group {
    const float *cells;      /* [nblk*ncell] complex64                        */
    const float *csi;        /* [nblk*ncell] f32 weights, or NULL            */
    const double *sigma2;    /* [nblk] f64                                    */
    const float *points;     /* [2^nbits] complex64 alphabet                  */
    const int32_t *lam_of_q; /* [ncell*nbits]                                 */
    float *llr;              /* [nblk*ncell*nbits] f32, codeword order        */
    int32_t ncell, nbits, nblk, blk0;
} atsc3_demap_group;
int32_t atsc3_demap_groups_dev(const void *groups, int32_t ngroups,
                               int32_t nblk_total, int32_t max_ncell,
                               void *stream);

/* Raw max-log differences and nearest-point distances, no noise scaling:
 * replaces lab/m9_fast.py:756-822 FrameDecoder.demap_batch called with
 * d2min_out (sigma^2 switched off).  llr_out as atsc3_demap_dev with
 * sigma2 = 1 (w * (min_{y_i=1} - min_{y_i=0})), d2_out [nblk*ncell] f32 =
 * min_p |z - p|^2 per cell; ws as atsc3_demap_dev with sigma2 NULL.      */
int32_t atsc3_demap_raw_dev(const float *cells, const float *csi,
                            int32_t nblk, int32_t ncell, const float *points,
                            int32_t nbits, const int32_t *lam_of_q,
                            float *llr_out, float *d2_out, double *ws,
                            void *stream);

/* ------------------------------------------------------------------------
 * Demap with the time de-interleaver's FEC-block gather fused into the read.
 * Replaces the FEC-block gathers in front of the demap calls: the hybrid
 * (twisted block) time de-interleave lab/m6_tbi.py:47-96 as batched in
 * lab/m9_fast.py:1087-1093 (cells = plp0[hidx]), and the convolutional time
 * de-interleave lab/m10_cti.py:108-146 as streamed by CtiStream.take_blocks
 * (lab/m44_ldm.py:1289-1304).  Same outputs as atsc3_demap_dev on the
 * gathered cells; cell s of block b (and its CSI weight) is read from
 *   ATSC3_GATHER_TABLE  pool[idx[b * ncell + s]]
 *   ATSC3_GATHER_CTI    pool[q mod ring], q = i + nrows * ((start_row + i)
 *                       mod nrows), i = first + b * ncell + s
 *   pool, pool_csi      complex64 / f32 cells of the frame, TI block or ring
 * ------------------------------------------------------------------------ */
#define ATSC3_GATHER_TABLE 1
#define ATSC3_GATHER_CTI 2
typedef struct atsc3_gather {
    int32_t mode, nrows, start_row, reserved;
    const int32_t *idx;
    int64_t first, ring;
} atsc3_gather;

/* The gathered cells themselves: out [nblk][ncell] complex64 = the cells
 * atsc3_demap_gather_dev would read (for consumers that need the
 * de-interleaved cells, e.g. the LDM cancellation). */
int32_t atsc3_gather_cells_dev(const float *pool, const atsc3_gather *gather,
                               int32_t nblk, int32_t ncell, float *out,
                               void *stream);
int32_t atsc3_demap_gather_dev(const float *pool, const float *pool_csi,
                               const atsc3_gather *gather,
                               const double *sigma2, int32_t nblk,
                               int32_t ncell, const float *points,
                               int32_t nbits, const int32_t *lam_of_q,
                               float *llr_out, double *sigma2_out, double *ws,
                               void *stream);

/* ------------------------------------------------------------------------
 * Weighted LLRs: per-part noise variance from the demap's own nearest-point
 * distances.  Replaces the weighted-LLR path of lab/m9_fast.py:1096-1115
 * (demap_batch with d2min_out, lab/m9_fast.py:756-822, then per pool part
 * s2 = max(mean d2min, 1e-6) and q /= s2[part]).
 *   pool/gather  as atsc3_demap_gather_dev; gather NULL: cells [nblk][ncell]
 *   part         [nblk*ncell] part id of every block cell (0 .. nparts-1)
 *   order, seg   the block cells of part p are order[seg[p] .. seg[p+1])
 *   llr_out      [nblk][ncell*nbits] codeword order (lam_of_q), each cell's
 *                raw max-log differences divided by its part's s2
 *   s2_out       [nparts] float; d2_ws [nblk*ncell] float;
 *   ws           atsc3_demap_ws_doubles(nblk, ncell) doubles
 * ------------------------------------------------------------------------ */
int32_t atsc3_demap_weighted_dev(const float *pool, const float *pool_csi,
                                 const atsc3_gather *gather, int32_t nblk,
                                 int32_t ncell, const float *points,
                                 int32_t nbits, const int32_t *lam_of_q,
                                 const int32_t *part, const int32_t *order,
                                 const int32_t *seg, int32_t nparts,
                                 float *llr_out, float *s2_out, float *d2_ws,
                                 double *ws, void *stream);

/* ------------------------------------------------------------------------
 * Common-phase-error correction and the dummy-cell SNR (cells in front of
 * the demapper).  atsc3_cpe_dev replaces lab/m6_cells.py:176-209
 * cpe_correct (lab/m9_fast.py:619-634, lab/m9_gpu.py:283-299): pool
 * normalised by its RMS, then per symbol `iters` rounds of
 * c = sum(conj(h) z) / max(sum |h|^2, 1e-12), z /= c, with h the nearest
 * point of the cell's region alphabet (region 0: pts0, 1: pts1) or its
 * known dummy value (region 2).
 *   pool [npool] complex64; cells/seg: symbol s = pool cells
 *   cells[seg[s] .. seg[s+1]); region [npool] u8; dummy [npool] complex64
 *   (may be NULL without region-2 cells); iters <= 8; alphabets pts0/pts1
 *   complex128 (re, im pairs): m > 0 points searched exhaustively, or
 *   m < 0: the -m first-quadrant base points w of a quadrant-rule alphabet
 *   [w, -conj w, conj w, -w] (nearest point searched in z's quadrant); at
 *   most 4096 points; est_cells/est_seg: the cells of symbol s that estimate
 *   its gains, est_cells[est_seg[s] .. est_seg[s+1]) (the upstream's fast
 *   path takes every 4th pool cell, lab/m9_fast.py:667-678), or both NULL:
 *   every cell; every cell of a symbol is corrected; out [npool] complex64 (cells of no symbol are not written);
 *   gains [nsym] complex128 (the product of a symbol's rounds);
 *   ws atsc3_cpe_ws_doubles(npool) doubles.
 * atsc3_dummy_snr_dev replaces lab/m16_margin.py:303-318 frame_snr_db:
 * cells [n] complex64 (the dummies), bits [n] scrambling bits (sign
 * s = 1 - 2 bit); out[0] = 10 log10(a^2 / noise) with a = sum(re d s) /
 * sum(s^2), noise = mean |d - a s|^2; out[1] = 1 if defined, else 0.
 * ------------------------------------------------------------------------ */
int32_t atsc3_cpe_ws_doubles(int32_t npool);
int32_t atsc3_cpe_dev(const float *pool, int32_t npool, const int32_t *cells,
                      const int32_t *seg, const int32_t *est_cells,
                      const int32_t *est_seg, int32_t nsym, const uint8_t *region,
                      const double *pts0, int32_t m0, const double *pts1,
                      int32_t m1, const float *dummy, int32_t iters,
                      float *out, double *gains, double *ws, void *stream);
int32_t atsc3_dummy_snr_dev(const float *cells, int32_t n, const uint8_t *bits,
                            double *out, void *stream);

/* ------------------------------------------------------------------------
 * LDM successive interference cancellation.
 * Replaces the core-layer cancellation of lab/e69_enh.py:86-121: for each
 * block with conv[b] = 1 the symbols s are rebuilt from the decoded codeword
 * cw (q = cw[lam_of_q], label MSB first, s = points[label]), the gain
 * g = vdot(s, y) / vdot(s, s) is estimated in FP64 and resid = y - g s;
 * blocks with conv[b] = 0 pass through.  Every output cell is multiplied by
 * out_scale.
 *   cells, resid_out [nblk*ncell] complex64; cw [nblk*n] u8; conv [nblk] u8
 *   lam_of_q [n] int32, points [2^nbits] complex64, n = ncell * nbits
 *   gain_out [2*nblk] f64 (re, im) or NULL
 *   ws       f64 scratch of atsc3_ldm_ws_doubles(nblk, ncell) elements
 * ------------------------------------------------------------------------ */
int32_t atsc3_ldm_ws_doubles(int32_t nblk, int32_t ncell);
int32_t atsc3_ldm_cancel_dev(const float *cells, const uint8_t *cw,
                             const uint8_t *conv, int32_t nblk, int32_t ncell,
                             int32_t n, const int32_t *lam_of_q,
                             const float *points, int32_t nbits,
                             float out_scale, float *resid_out,
                             double *gain_out, double *ws, void *stream);

/* ------------------------------------------------------------------------
 * Baseband Packet step after the decode: BCH syndrome and descrambling.
 * Replaces lab/m6_bicm.py:357-364 (PlpChain.baseband_packet) and its batched
 * form lab/m9_fast.py:942-949 (bch_syndrome_batch of lab/m9_accel.py:150-236
 * + descramble of lab/m4_scrambler.py:148-165).
 *   bits     [nblk*n] u8 decoded codewords (0/1), first kldpc = BCH codeword
 *   mask     u32 [ceil(kldpc/32)][mouter] syndrome matrix (bit 31-t of
 *            [w][j] = coefficient j of x^(kldpc-1-(32w+t)) mod g)
 *   seq      [(kldpc-mouter)/8] scrambling bytes (A/322 5.2.3, MSB first)
 *   syn_out  [nblk][8] u32: syndrome, coefficient j at limb j/32 bit j%32
 *   ok_out   [nblk] u8: syndrome == 0
 *   bytes_out[nblk][(kldpc-mouter)/8] descrambled payload bytes
 * kldpc <= 58368; mouter <= 256.
 * ------------------------------------------------------------------------ */
int32_t atsc3_baseband_dev(const uint8_t *bits, int32_t nblk, int32_t n,
                           int32_t kldpc, int32_t mouter, const uint32_t *mask,
                           const uint8_t *seq, uint32_t *syn_out,
                           uint8_t *ok_out, uint8_t *bytes_out, void *stream);

/* ------------------------------------------------------------------------
 * Exact sum-product rescue of near misses (float64, log-tanh BP).
 * atsc3_sum_product_dev replaces lab/m16_margin.py:325-390
 * (sum_product_decode_batch; = lab/e58_weighted.py:134-187 with no
 * deadline); atsc3_sp_select_dev + atsc3_sp_merge_dev replace the candidate
 * policy and write-back of lab/m44_ldm.py:2527-2555 and
 * lab/m9_fast.py:1126-1151, on the device (no host synchronisation).
 *
 * atsc3_sum_product_dev decodes call rows r < live (live = *nsel, or rows
 * when nsel is NULL) of one code; row r reads LLR row s = sel[r] (or r when
 * sel is NULL) at llr + row_off[s] (or s * n when row_off is NULL).
 *   cvar     [dmax][nc] int32: variable of slot k of check thread i, n for
 *            padding (slot order = the check's sorted variable list)
 *   vptr     [n + 1], vedge [E] int32: per variable its message slots
 *            k * nc + i, in ascending check index
 *   iters    iteration limit (the last syndrome test is at iteration iters)
 *   budget_s > 0: no iteration starts once this much wall-clock time has
 *            passed since the call's first kernel began (the upstream's
 *            deadline); rows still failing leave with shed = 1
 *   tot_ws   rows * n doubles; msg_ws rows * dmax * nc doubles;
 *   ctl_ws   atsc3_sp_ctl_words(rows) int32
 *   bits     [rows][n] u8, conv/shed [rows] u8, its/bad [rows] int32;
 *            rows past live get conv 0, bad -1
 * dmax <= 128 (ATSC3_E_DEGREE otherwise).
 *
 * atsc3_sp_select_dev: rows_out[0 .. *nrows_out) = the rows b with
 * conv[b] == 0 and bad[b] <= near_max (and blk_code[b] == code_id when
 * blk_code is not NULL: one code of a mixed batch), by ascending bad then
 * b, at most max_rows of them.
 * atsc3_sp_merge_dev: for r < *nsel with sp_conv[r], copy sp_bits row r to
 * bits row sel[r] (at row_off[sel[r]] or sel[r] * n) and set conv = 1,
 * its = sp_its[r], bad = 0 there.
 * ------------------------------------------------------------------------ */
int32_t atsc3_sp_ctl_words(int32_t rows);
int32_t atsc3_sum_product_dev(const float *llr, const int64_t *row_off,
                              const int32_t *sel, const int32_t *nsel,
                              int32_t rows, int32_t n, int32_t nc,
                              int32_t dmax, const int32_t *cvar,
                              const int32_t *vptr, const int32_t *vedge,
                              int32_t iters, double budget_s, double *tot_ws,
                              double *msg_ws, int32_t *ctl_ws, uint8_t *bits,
                              uint8_t *conv, int32_t *its, int32_t *bad,
                              uint8_t *shed, void *stream);
int32_t atsc3_sp_select_dev(const uint8_t *conv, const int32_t *bad,
                            const int32_t *blk_code, int32_t code_id,
                            int32_t nblk, int32_t near_max, int32_t max_rows,
                            int32_t *rows_out, int32_t *nrows_out,
                            void *stream);
int32_t atsc3_sp_merge_dev(const int32_t *sel, const int32_t *nsel,
                           int32_t rows, int32_t n, const int64_t *row_off,
                           const uint8_t *sp_bits, const uint8_t *sp_conv,
                           const int32_t *sp_its, uint8_t *bits,
                           uint8_t *conv, int32_t *its, int32_t *bad,
                           void *stream);

/* ------------------------------------------------------------------------
 * Differentiable decoding (training; no upstream counterpart for the
 * backward -- the forward update is that of lab/ldpc_kernel.c:82-143 without
 * the syndrome stop, on the CSR tables of lab/m3_ldpc.py:459-482).
 *
 * Unrolled normalized min-sum, one iteration per call pair:
 *   check_fwd: m_out[b][e] = alpha_t[ccls[c]] * s_e * mu_e from
 *              x = tot[b][cidx[e]] - m_in[b][e] (m_in NULL = zero messages);
 *              optional state (st_mn [nblk*nc*2], st_idx, st_sgn [nblk*nc])
 *              records min1/min2, argmins | parity << 16, sign bits of x
 *   var_fwd:   tot[b][v] = lam[b][v] + sum over v's edges of m[b][e]
 *   check_bwd: from g_tot = dL/dtot_{t+1}, g_carry = G_x of iteration t+1
 *              (NULL at the last iteration; must not alias g_x) and the
 *              recorded state: g_x [nblk*E] (subgradient of the min1/min2
 *              selection) and g_alpha_part [nblk*nc] (dL/dalpha per check)
 *   var_bwd:   g_tot[b][v] = sum over v's edges of g_x; g_lam += g_tot_next
 * Check degree <= 32.  _f32 and _f64 variants.
 * ------------------------------------------------------------------------ */
int32_t atsc3_unrolled_check_fwd_f32(const float *tot, const float *m_in,
                                     float *m_out, const int32_t *cptr,
                                     const int32_t *cidx, const int32_t *ccls,
                                     const float *alpha_t, int32_t nblk,
                                     int32_t n, int32_t nc, int32_t E,
                                     float *st_mn, int32_t *st_idx,
                                     uint32_t *st_sgn, void *stream);
int32_t atsc3_unrolled_check_fwd_f64(const double *tot, const double *m_in,
                                     double *m_out, const int32_t *cptr,
                                     const int32_t *cidx, const int32_t *ccls,
                                     const double *alpha_t, int32_t nblk,
                                     int32_t n, int32_t nc, int32_t E,
                                     double *st_mn, int32_t *st_idx,
                                     uint32_t *st_sgn, void *stream);
int32_t atsc3_unrolled_var_fwd_f32(const float *lam, const float *m,
                                   const int32_t *vptr, const int32_t *vedge,
                                   float *tot, int32_t nblk, int32_t n,
                                   int32_t E, void *stream);
int32_t atsc3_unrolled_var_fwd_f64(const double *lam, const double *m,
                                   const int32_t *vptr, const int32_t *vedge,
                                   double *tot, int32_t nblk, int32_t n,
                                   int32_t E, void *stream);
int32_t atsc3_unrolled_check_bwd_f32(const float *g_tot, const float *g_carry,
                                     const int32_t *cptr, const int32_t *cidx,
                                     const int32_t *ccls, const float *alpha_t,
                                     const float *st_mn, const int32_t *st_idx,
                                     const uint32_t *st_sgn, int32_t nblk,
                                     int32_t n, int32_t nc, int32_t E,
                                     float *g_x, float *g_alpha_part,
                                     void *stream);
int32_t atsc3_unrolled_check_bwd_f64(const double *g_tot, const double *g_carry,
                                     const int32_t *cptr, const int32_t *cidx,
                                     const int32_t *ccls, const double *alpha_t,
                                     const double *st_mn, const int32_t *st_idx,
                                     const uint32_t *st_sgn, int32_t nblk,
                                     int32_t n, int32_t nc, int32_t E,
                                     double *g_x, double *g_alpha_part,
                                     void *stream);
int32_t atsc3_unrolled_var_bwd_f32(const float *g_x, const float *g_tot_next,
                                   const int32_t *vptr, const int32_t *vedge,
                                   float *g_tot, float *g_lam, int32_t nblk,
                                   int32_t n, int32_t E, void *stream);
int32_t atsc3_unrolled_var_bwd_f64(const double *g_x, const double *g_tot_next,
                                   const int32_t *vptr, const int32_t *vedge,
                                   double *g_tot, double *g_lam, int32_t nblk,
                                   int32_t n, int32_t E, void *stream);

/* Exhaustive max-log soft demapper for any constellation of M = 2^nb points
 * (same LLR as lab/m6_bicm.py:213-225), codeword-order output through
 * lam_of_q, and its backward: g_z [nblk*ncell*2], per-workgroup point
 * partials g_pts_part [grid*M*2], sigma2 partials g_s2_part [nblk*grid]
 * (zeroed by the caller), grid = atsc3_soft_demap_bwd_grid(nblk, ncell).
 * z is the (re, im) view of the cells; csi may be NULL.                     */
int32_t atsc3_soft_demap_fwd_f32(const float *z, const float *csi,
                                 const float *sigma2, const float *pts,
                                 int32_t M, int32_t nb,
                                 const int32_t *lam_of_q, int32_t nblk,
                                 int32_t ncell, float *out, void *stream);
int32_t atsc3_soft_demap_fwd_f64(const double *z, const double *csi,
                                 const double *sigma2, const double *pts,
                                 int32_t M, int32_t nb,
                                 const int32_t *lam_of_q, int32_t nblk,
                                 int32_t ncell, double *out, void *stream);
int32_t atsc3_soft_demap_bwd_grid(int32_t nblk, int32_t ncell);
int32_t atsc3_soft_demap_bwd_f32(const float *z, const float *csi,
                                 const float *sigma2, const float *pts,
                                 int32_t M, int32_t nb,
                                 const int32_t *lam_of_q, int32_t nblk,
                                 int32_t ncell, const float *g_llr, float *g_z,
                                 float *g_pts_part, float *g_s2_part,
                                 void *stream);
int32_t atsc3_soft_demap_bwd_f64(const double *z, const double *csi,
                                 const double *sigma2, const double *pts,
                                 int32_t M, int32_t nb,
                                 const int32_t *lam_of_q, int32_t nblk,
                                 int32_t ncell, const double *g_llr,
                                 double *g_z, double *g_pts_part,
                                 double *g_s2_part, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ATSC3_FEC_H */
