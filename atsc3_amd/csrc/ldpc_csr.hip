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
//
// The upstream drives this call from a ThreadPoolExecutor: many concurrent
// calls of a few blocks each, same tables every time (lab/m3_ldpc.py:485-506,
// lab/m9_fast.py:893-925).  So a call
//   * takes an idle context of the current device from a pool -- its own
//     stream, a grow-only device workspace and grow-only pinned staging --
//     and gives it back at the end: concurrent callers run on separate
//     streams, and once the buffers have grown no call allocates or frees;
//   * finds the code's tables in a per-device cache keyed by a 64-bit hash
//     of (n, nc, nedges, cptr, cidx, vptr, vedge) and uploads them only the
//     first time (up to 64 codes per device; beyond that they travel with
//     the call);
//   * decodes an A/322 code -- recognised by the same hash among the tables
//     the build writes beside this library (libatsc3fec_tables.bin) -- on
//     the resident engine's FP32 mode (bit-exact with the reference kernel),
//     and any other graph on the CSR phase kernels (ldpc_utc.hip, the
//     compiled kernel's contract), which spread the call's checks and
//     variables over the whole GPU.
// Outputs are staged and copied to the caller's arrays only after the whole
// decode succeeded, so on any error nothing is written.
// ---------------------------------------------------------------------------
extern "C" int32_t ldpc_kernel_abi(void) { return 1; }

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

namespace {

struct DropCtx {
    int device = -1;
    bool busy = false;
    hipStream_t st = nullptr;
    char *dev = nullptr;
    size_t dev_cap = 0;
    char *host = nullptr;
    size_t host_cap = 0;
};

struct DropTab {
    int device, n, nc, ne;
    uint64_t hash;
    int32_t *d;   // cptr [nc+1] | cidx [ne] | vptr [n+1] | vedge [ne]
};

std::mutex g_drop_mu;
std::vector<DropCtx *> g_drop_ctx;   // never freed: reused for the process lifetime
std::vector<DropTab> g_drop_tab;
constexpr int kMaxTablesPerDevice = 64;

// Order-sensitive hash of the CSR tables (same value as csr_hash in
// atsc3_amd/resident.py): the sum over the concatenated words w_i of a
// 64-bit finaliser of (w_i, i), plus the sizes.
struct CsrHash {
    uint64_t h = 0, i = 0;
    void add(const int32_t *p, size_t n) {
        for (size_t k = 0; k < n; ++k, ++i) {
            uint64_t x = (uint64_t)(uint32_t)p[k] * 0x9E3779B97F4A7C15ull +
                         i * 0xC2B2AE3D27D4EB4Full;
            x ^= x >> 31;
            x *= 0xBF58476D1CE4E5B9ull;
            x ^= x >> 29;
            h += x;
        }
    }
};

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

DropCtx *acquire_ctx(int device) {
    std::lock_guard<std::mutex> lock(g_drop_mu);
    for (DropCtx *c : g_drop_ctx)
        if (!c->busy && c->device == device) {
            c->busy = true;
            return c;
        }
    DropCtx *c = new DropCtx;
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    c->device = device;
    c->busy = true;
    g_drop_ctx.push_back(c);
    return c;
}

void release_ctx(DropCtx *c) {
    std::lock_guard<std::mutex> lock(g_drop_mu);
    c->busy = false;
}

bool grow(DropCtx *c, size_t dev_need, size_t host_need) {
    if (dev_need > c->dev_cap) {
        if (c->dev) (void)hipFree(c->dev);
        c->dev = nullptr;
        c->dev_cap = 0;
        const size_t cap = dev_need + dev_need / 4;
        if (hipMalloc(reinterpret_cast<void **>(&c->dev), cap) != hipSuccess) return false;
        c->dev_cap = cap;
    }
    if (host_need > c->host_cap) {
        if (c->host) (void)hipHostFree(c->host);
        c->host = nullptr;
        c->host_cap = 0;
        const size_t cap = host_need + host_need / 4;
        if (hipHostMalloc(reinterpret_cast<void **>(&c->host), cap, hipHostMallocDefault) !=
            hipSuccess)
            return false;
        c->host_cap = cap;
    }
    return true;
}

// Device copy of the tables, from the cache or uploaded once; NULL when the
// cache is full for this device (the caller stages them with the call).
const int32_t *cached_tables(int device, int n, int nc, int ne, uint64_t hash,
                             const int32_t *cptr, const int32_t *cidx, const int32_t *vptr,
                             const int32_t *vedge, bool *failed) {
    std::lock_guard<std::mutex> lock(g_drop_mu);
    int count = 0;
    for (const DropTab &t : g_drop_tab) {
        if (t.device != device) continue;
        ++count;
        if (t.n == n && t.nc == nc && t.ne == ne && t.hash == hash) return t.d;
    }
    if (count >= kMaxTablesPerDevice) return nullptr;
    const size_t words = (size_t)nc + 1 + ne + (size_t)n + 1 + ne;
    int32_t *d = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&d), words * 4) != hipSuccess) {
        *failed = true;
        return nullptr;
    }
    size_t o = 0;
    bool ok = hipMemcpy(d + o, cptr, ((size_t)nc + 1) * 4, hipMemcpyHostToDevice) == hipSuccess;
    o += (size_t)nc + 1;
    ok = ok && hipMemcpy(d + o, cidx, (size_t)ne * 4, hipMemcpyHostToDevice) == hipSuccess;
    o += ne;
    ok = ok && hipMemcpy(d + o, vptr, ((size_t)n + 1) * 4, hipMemcpyHostToDevice) == hipSuccess;
    o += (size_t)n + 1;
    ok = ok && hipMemcpy(d + o, vedge, (size_t)ne * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        (void)hipFree(d);
        *failed = true;
        return nullptr;
    }
    g_drop_tab.push_back(DropTab{device, n, nc, ne, hash, d});
    return d;
}

// ---- the resident route -------------------------------------------------
// Entries of the table file written by the build (libatsc3fec_tables.bin
// beside this library); per device, each code's tables and its
// atsc3_res_code descriptor are uploaded at the first call that needs them.
struct ResEntry {
    uint64_t hash;
    int32_t n, nc, ne, ncg, nvg, nbe, max_dc, max_dv, nsc, nsv, grows;
    std::vector<std::vector<int32_t>> tab;   // 13 tables, atsc3_res_code order
};
struct ResCode {
    int device;
    const ResEntry *e;
    void *desc;                              // device atsc3_res_code
};
std::vector<ResEntry> g_res_file;
bool g_res_loaded = false;
std::deque<ResCode> g_res_dev;   // a deque: entries handed out stay valid as it grows

std::string table_file_path() {
    Dl_info info;
    if (!dladdr(reinterpret_cast<void *>(&ldpc_kernel_abi), &info) || !info.dli_fname)
        return std::string();
    std::string p(info.dli_fname);
    const size_t slash = p.rfind('/');
    return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) +
           "/libatsc3fec_tables.bin";
}

// Parses the file once; a missing or malformed file leaves the list empty
// (every code then runs on the CSR phase kernels).
void load_table_file() {
    if (g_res_loaded) return;
    g_res_loaded = true;
    FILE *f = std::fopen(table_file_path().c_str(), "rb");
    if (!f) return;
    std::vector<ResEntry> out;
    char magic[4];
    uint32_t ver = 0, count = 0;
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "A3RT", 4) == 0 &&
              std::fread(&ver, 4, 1, f) == 1 && ver == 1 && std::fread(&count, 4, 1, f) == 1 &&
              count < 1024;
    for (uint32_t c = 0; ok && c < count; ++c) {
        ResEntry e;
        int32_t hdr[11];
        ok = std::fread(&e.hash, 8, 1, f) == 1 && std::fread(hdr, 4, 11, f) == 11;
        if (!ok) break;
        e.n = hdr[0]; e.nc = hdr[1]; e.ne = hdr[2]; e.ncg = hdr[3]; e.nvg = hdr[4];
        e.nbe = hdr[5]; e.max_dc = hdr[6]; e.max_dv = hdr[7]; e.nsc = hdr[8]; e.nsv = hdr[9];
        e.grows = hdr[10];
        for (int t = 0; ok && t < 13; ++t) {
            int32_t len = 0;
            ok = std::fread(&len, 4, 1, f) == 1 && len >= 0 && len < (1 << 26);
            if (!ok) break;
            std::vector<int32_t> v((size_t)len);
            ok = len == 0 || std::fread(v.data(), 4, (size_t)len, f) == (size_t)len;
            e.tab.push_back(std::move(v));
        }
        if (ok) out.push_back(std::move(e));
    }
    std::fclose(f);
    if (ok) g_res_file = std::move(out);
}

// The device copy of code `hash` (uploaded on first use), or NULL when the
// file has no such code.
const ResCode *resident_code(int device, uint64_t hash, int n, int nc, int ne, bool *failed) {
    std::lock_guard<std::mutex> lock(g_drop_mu);
    load_table_file();
    for (const ResCode &r : g_res_dev)
        if (r.device == device && r.e->hash == hash) return &r;
    const ResEntry *e = nullptr;
    for (const ResEntry &x : g_res_file)
        if (x.hash == hash && x.n == n && x.nc == nc && x.ne == ne) e = &x;
    if (!e || e->n > 64800 || e->max_dc > 89 || e->max_dv > 32) return nullptr;
    size_t words = 0;
    for (const auto &v : e->tab) words += v.size() + 64;   // 256-byte aligned tables
    char *buf = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&buf), words * 4 + 256) != hipSuccess) {
        *failed = true;
        return nullptr;
    }
    struct {
        int32_t n, ncg, nvg, nbe, max_dc, max_dv, nsc, nsv;
        const int32_t *t[13];
    } d;
    static_assert(sizeof(d) == sizeof(atsc3_res_code), "descriptor layout");
    d.n = e->n; d.ncg = e->ncg; d.nvg = e->nvg; d.nbe = e->nbe; d.max_dc = e->max_dc;
    d.max_dv = e->max_dv; d.nsc = e->nsc; d.nsv = e->nsv;
    size_t off = 256;                                  // the descriptor first
    for (int t = 0; t < 13; ++t) {
        const auto &v = e->tab[t];
        d.t[t] = reinterpret_cast<const int32_t *>(buf + off);
        if (!v.empty() &&
            hipMemcpy(buf + off, v.data(), v.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
            *failed = true;
            return nullptr;
        }
        off += up256(v.size() * 4);
    }
    if (hipMemcpy(buf, &d, sizeof(d), hipMemcpyHostToDevice) != hipSuccess) {
        *failed = true;
        return nullptr;
    }
    g_res_dev.push_back(ResCode{device, e, buf});
    return &g_res_dev.back();
}

// One caller's share of a combined decode: its host arrays.
struct DropCall {
    const float *llr;
    int32_t nblk;
    uint8_t *bits, *conv;
    int32_t *its, *bad;
};

// Team decodes (each codeword on several compute units of one XCD) need
// every workgroup of their grid co-resident; two running at once could each
// hold compute units the other waits for, so they take turns per device.
constexpr int kTeamDevices = 64;
std::mutex g_team_mu[kTeamDevices];

// The resident decode of every call's blocks in one launch: LLRs staged
// into pinned memory call after call, results copied back to each caller's
// arrays only when the whole decode succeeded (nothing is written on error).
// A launch with at most half as many codewords as compute units runs the
// resident engine's team mode (FP32, no LDS copies: the same bits, conv,
// its, bad), each codeword on 2..16 compute units: a round of the
// upstream's small thread-pool calls then lasts its slowest codeword's
// iterations at the team's rate instead of one compute unit's.
int32_t resident_decode_calls(const ResCode *rc, int device, int32_t n, int32_t iters,
                              float alpha, const std::vector<DropCall> &calls) {
    const ResEntry &e = *rc->e;
    int32_t nblk = 0;
    for (const DropCall &k : calls) nblk += k.nblk;
    const int cus = atsc3_minsum_resident_grid();
    if (cus <= 0) return ATSC3_E_HIP;
    int team = 1;
    while (team < 16 && (int64_t)nblk * team * 2 <= cus) team *= 2;
    if (device < 0 || device >= kTeamDevices) team = 1;
    const int grid = team > 1 ? cus : (nblk < cus ? nblk : cus);
    const int max_checks = e.ncg * ATSC3_Z;
    const int64_t wsb =
        team > 1 ? atsc3_minsum_resident_team_ws_bytes(grid, team, max_checks, e.grows, 0)
                 : atsc3_minsum_resident_ws_bytes(grid, max_checks, e.grows, 0);
    if (wsb <= 0) return ATSC3_E_HIP;
    const size_t nb = (size_t)nblk, nn = (size_t)n;
    // device and pinned host share the first three offsets: llr | blk_off | blk_code
    const size_t o_off = up256(nb * nn * 4);
    const size_t o_code = o_off + up256(nb * 8);
    const size_t o_in_end = o_code + up256(nb * 4);
    const size_t o_ws = o_in_end;
    const size_t o_bits = o_ws + up256((size_t)wsb);
    const size_t o_conv = o_bits + up256(nb * nn);
    const size_t o_its = o_conv + up256(nb);
    const size_t o_bad = o_its + up256(nb * 4);
    const size_t d_need = o_bad + up256(nb * 4);
    // pinned host: blk_off | blk_code at the device offsets, then the outputs
    const size_t h_out = o_in_end;
    const size_t h_need = h_out + nb * nn + nb + 8 * nb;
    DropCtx *c = acquire_ctx(device);
    if (!c) return ATSC3_E_HIP;
    int32_t rc_out = ATSC3_E_HIP;
    std::unique_lock<std::mutex> team_turn;      // held until the decode has finished
    do {
        if (!grow(c, d_need, h_need)) break;
        hipStream_t st = c->st;
        // the LLRs straight from each caller's (pageable) arrays: the
        // runtime's own staged copy beats a memcpy into pinned memory first
        bool ok = true;
        size_t b0 = 0;
        for (const DropCall &k : calls) {
            ok = ok && hipMemcpyAsync(c->dev + b0 * nn * 4, k.llr, (size_t)k.nblk * nn * 4,
                                      hipMemcpyHostToDevice, st) == hipSuccess;
            b0 += (size_t)k.nblk;
        }
        if (!ok) break;
        int64_t *hoff = reinterpret_cast<int64_t *>(c->host + o_off);
        int32_t *hcode = reinterpret_cast<int32_t *>(c->host + o_code);
        for (size_t b = 0; b < nb; ++b) {
            hoff[b] = (int64_t)(b * nn);
            hcode[b] = 0;
        }
        if (hipMemcpyAsync(c->dev + o_off, c->host + o_off, o_in_end - o_off,
                           hipMemcpyHostToDevice, st) != hipSuccess)
            break;
        uint8_t *dbits = reinterpret_cast<uint8_t *>(c->dev + o_bits);
        uint8_t *dconv = reinterpret_cast<uint8_t *>(c->dev + o_conv);
        int32_t *dits = reinterpret_cast<int32_t *>(c->dev + o_its);
        int32_t *dbad = reinterpret_cast<int32_t *>(c->dev + o_bad);
        const float *dllr = reinterpret_cast<const float *>(c->dev);
        const int64_t *doff = reinterpret_cast<const int64_t *>(c->dev + o_off);
        const int32_t *dcode = reinterpret_cast<const int32_t *>(c->dev + o_code);
        if (team > 1) {
            team_turn = std::unique_lock<std::mutex>(g_team_mu[device]);
            if (atsc3_minsum_resident_team_dev(dllr, nblk, doff, dcode, rc->desc, 1, max_checks,
                                               e.grows, e.n, e.max_dc, e.max_dv, iters, alpha,
                                               0, nullptr, nullptr, nullptr, 0, grid,
                                               team | ATSC3_TEAM_DIRECT, c->dev + o_ws, wsb,
                                               dbits, dconv, dits, dbad, st) != ATSC3_OK)
                break;
        } else if (atsc3_minsum_resident_dev(dllr, nblk, doff, dcode, rc->desc, 1, max_checks,
                                             e.grows, e.n, e.max_dc, e.max_dv, iters, alpha, 0,
                                             nullptr, grid, c->dev + o_ws, wsb, dbits, dconv,
                                             dits, dbad, st) != ATSC3_OK) {
            break;
        }
        uint8_t *hb = reinterpret_cast<uint8_t *>(c->host + h_out);
        uint8_t *hc = hb + nb * nn;
        int32_t *hi = reinterpret_cast<int32_t *>(hc + nb);
        int32_t *hd = hi + nb;
        if (hipMemcpyAsync(hb, dbits, nb * nn, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hc, dconv, nb, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hi, dits, nb * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(hd, dbad, nb * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            break;
        if (team > 1) {
            // a bounded team wait that ran out leaves its = -1: report it
            // as a device error (nothing written), as any failed decode
            const int32_t *hits = reinterpret_cast<const int32_t *>(hc + nb);
            bool aborted = false;
            for (size_t b = 0; b < nb; ++b) aborted = aborted || hits[b] < 0;
            if (aborted) break;
        }
        b0 = 0;
        for (const DropCall &k : calls) {
            const size_t kb = (size_t)k.nblk;
            std::memcpy(k.bits, hb + b0 * nn, kb * nn);
            std::memcpy(k.conv, hc + b0, kb);
            std::memcpy(k.its, hi + b0, kb * 4);
            std::memcpy(k.bad, hd + b0, kb * 4);
            b0 += kb;
        }
        rc_out = ATSC3_OK;
    } while (0);
    (void)hipStreamSynchronize(c->st);
    release_ctx(c);
    return rc_out;
}

// Calls that arrive together are decoded together.  The reference drives
// its kernel from a thread pool, each call a few blocks (lab/m9_fast.py:
// 893-925): on the GPU every call would pay a launch's whole iteration
// latency and its own transfers.  A call that finds an open batch of the
// same code, iteration limit, alpha and device joins it and waits.  One
// that finds none leads a new batch; when other threads finished a call of
// the same code within kRecentUs and are idle, it keeps the batch open until
// as many calls as those threads have joined (or kBatchCap blocks, or
// kJoinWindowUs pass), then runs it and wakes the members.  A lone caller
// never waits.  Every block is decoded as if alone (the
// resident engine's codewords are independent), so results are those of
// separate calls; a failed decode returns its status to every member,
// writing nothing.
constexpr int32_t kBatchCap = 256;
constexpr int kJoinWindowUs = 500;
constexpr int kRecentUs = 10000;
constexpr int kRecentCalls = 32;

struct DropKey {
    const ResCode *rc;
    int device, n;
    int32_t iters;
    uint32_t alpha_bits;
    bool operator==(const DropKey &o) const {
        return rc == o.rc && device == o.device && n == o.n && iters == o.iters &&
               alpha_bits == o.alpha_bits;
    }
};

struct DropBatch {
    DropKey key;
    std::vector<DropCall> calls;
    int32_t total = 0;
    bool closed = false, done = false;
    int32_t status = ATSC3_E_HIP;
    std::condition_variable cv;
};

struct RecentCall {
    std::thread::id thread;
    DropKey key;
    std::chrono::steady_clock::time_point end;
};

std::mutex g_batch_mu;
// batches still accepting calls; each batch is owned by its leader and its
// members together, so a member can read the status after the leader left
std::vector<std::shared_ptr<DropBatch>> g_open_batches;
std::vector<std::pair<std::thread::id, DropKey>> g_in_flight;
RecentCall g_recent[kRecentCalls];
int g_recent_next = 0;

// Other threads that may join a batch of `key` soon: those that finished a
// call of `key` within kRecentUs and are not inside a call now (a thread in
// flight with the same key leads or belongs to another batch already)
// (g_batch_mu held).
int other_threads(const DropKey &key, std::thread::id self) {
    const auto now = std::chrono::steady_clock::now();
    std::vector<std::thread::id> seen;
    for (const RecentCall &r : g_recent) {
        if (!(r.key == key) || r.thread == self ||
            now - r.end >= std::chrono::microseconds(kRecentUs))
            continue;
        bool busy = false;
        for (const auto &f : g_in_flight) busy = busy || f.first == r.thread;
        if (!busy && std::find(seen.begin(), seen.end(), r.thread) == seen.end())
            seen.push_back(r.thread);
    }
    return (int)seen.size();
}

void finish_call(const DropKey &key, std::thread::id self) {
    g_in_flight.erase(std::find(g_in_flight.begin(), g_in_flight.end(),
                                std::make_pair(self, key)));
    g_recent[g_recent_next] = RecentCall{self, key, std::chrono::steady_clock::now()};
    g_recent_next = (g_recent_next + 1) % kRecentCalls;
}

int32_t resident_dropin(const ResCode *rc, int device, const float *llr, int32_t nblk, int32_t n,
                        int32_t iters, float alpha, uint8_t *bits, uint8_t *conv, int32_t *its,
                        int32_t *bad) {
    const DropCall mine{llr, nblk, bits, conv, its, bad};
    DropKey key{rc, device, n, iters, 0u};
    std::memcpy(&key.alpha_bits, &alpha, sizeof(float));
    const std::thread::id self = std::this_thread::get_id();
    std::unique_lock<std::mutex> lk(g_batch_mu);
    g_in_flight.emplace_back(self, key);
    for (const std::shared_ptr<DropBatch> &open : g_open_batches) {
        DropBatch &b = *open;
        if (b.closed || !(b.key == key) || b.total + nblk > kBatchCap) continue;
        const std::shared_ptr<DropBatch> keep = open;
        b.calls.push_back(mine);
        b.total += nblk;
        b.cv.notify_all();
        b.cv.wait(lk, [&b] { return b.done; });
        finish_call(key, self);
        return b.status;
    }
    const std::shared_ptr<DropBatch> batch = std::make_shared<DropBatch>();
    DropBatch &b = *batch;
    b.key = key;
    b.calls.push_back(mine);
    b.total = nblk;
    const int others = nblk < kBatchCap ? other_threads(key, self) : 0;
    if (others > 0) {
        g_open_batches.push_back(batch);
        b.cv.wait_for(lk, std::chrono::microseconds(kJoinWindowUs), [&b, others] {
            return (int)b.calls.size() > others || b.total >= kBatchCap;
        });
        g_open_batches.erase(std::find(g_open_batches.begin(), g_open_batches.end(), batch));
    }
    b.closed = true;
    lk.unlock();
    // no call joins a closed batch: its call list is fixed from here on
    const int32_t st = resident_decode_calls(rc, device, n, iters, alpha, b.calls);
    lk.lock();
    b.status = st;
    b.done = true;
    finish_call(key, self);
    b.cv.notify_all();
    return st;
}

uint64_t csr_hash(int32_t n, int32_t nc, int32_t nedges, const int32_t *cptr,
                  const int32_t *cidx, const int32_t *vptr, const int32_t *vedge) {
    CsrHash ch;
    ch.add(cptr, (size_t)nc + 1);
    ch.add(cidx, (size_t)nedges);
    ch.add(vptr, (size_t)n + 1);
    ch.add(vedge, (size_t)nedges);
    return ch.h + (uint64_t)n * 0x100000001B3ull + (uint64_t)nc * 0x1000193ull +
           (uint64_t)nedges;
}

}  // namespace

extern "C" int32_t atsc3_dropin_engine(int32_t n, int32_t nc, int32_t nedges,
                                       const int32_t *cptr, const int32_t *cidx,
                                       const int32_t *vptr, const int32_t *vedge) {
    if (n <= 0 || nc < 0 || nedges < 0 || !cptr || !cidx || !vptr || !vedge) return ATSC3_E_ARG;
    const uint64_t h = csr_hash(n, nc, nedges, cptr, cidx, vptr, vedge);
    std::lock_guard<std::mutex> lock(g_drop_mu);
    load_table_file();
    for (const ResEntry &x : g_res_file)
        if (x.hash == h && x.n == n && x.nc == nc && x.ne == nedges) return 1;
    return 0;
}

extern "C" int32_t ldpc_minsum_f32(const float *llr, int32_t nblk, int32_t n, int32_t nc,
                                   int32_t nedges, const int32_t *cptr,
                                   const int32_t *cidx, const int32_t *vptr,
                                   const int32_t *vedge, int32_t iters, float alpha,
                                   uint8_t *bits, uint8_t *conv, int32_t *its,
                                   int32_t *bad, float *work) {
    (void)work;
    if (nblk < 0 || n <= 0 || nc < 0 || nedges < 0 || iters < 0) return ATSC3_E_ARG;
    if (!cptr) return ATSC3_E_ARG;
    if (atsc3::check_degrees(cptr, nc) != ATSC3_OK) return ATSC3_E_DEGREE;
    if (nblk == 0) return ATSC3_OK;
    if (!llr || !cidx || !vptr || !vedge || !bits || !conv || !its || !bad) return ATSC3_E_ARG;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return ATSC3_E_HIP;

    const uint64_t h = csr_hash(n, nc, nedges, cptr, cidx, vptr, vedge);
    // an A/322 code whose resident tables the build wrote: the resident
    // engine's FP32 mode (bit-exact with the reference kernel)
    if (alpha >= 0.0f && iters < (1 << 24)) {
        bool rfailed = false;
        const ResCode *rc = resident_code(device, h, n, nc, nedges, &rfailed);
        if (rfailed) return ATSC3_E_HIP;
        if (rc)
            return resident_dropin(rc, device, llr, nblk, n, iters, alpha, bits, conv, its, bad);
    }
    bool failed = false;
    const int32_t *tabs =
        cached_tables(device, n, nc, nedges, h, cptr, cidx, vptr, vedge, &failed);
    if (failed) return ATSC3_E_HIP;
    const size_t tab_words = (size_t)nc + 1 + nedges + (size_t)n + 1 + nedges;

    const size_t nb = (size_t)nblk, nn = (size_t)n;
    const int chunk = 65535;                      // grid.y limit of the phase kernels
    const int first_chunk = nblk < chunk ? nblk : chunk;
    // device: llr | workspace | bits | conv | its | bad (one chunk), [tables]
    const size_t d_llr = 0;
    const size_t d_ws = up256(d_llr + (size_t)first_chunk * nn * 4);
    const size_t d_bits = d_ws + up256(atsc3::csr_phase_ws_bytes(first_chunk, n, nedges));
    const size_t d_conv = d_bits + up256((size_t)first_chunk * nn);
    const size_t d_its = d_conv + up256((size_t)first_chunk);
    const size_t d_bad = d_its + up256((size_t)first_chunk * 4);
    const size_t d_tab = d_bad + up256((size_t)first_chunk * 4);
    const size_t d_need = d_tab + (tabs ? 0 : tab_words * 4);
    // pinned host: llr in (one chunk) | all outputs
    const size_t h_out = up256((size_t)first_chunk * nn * 4);
    const size_t h_need = h_out + nb * nn + nb + 8 * nb;

    DropCtx *c = acquire_ctx(device);
    if (!c) return ATSC3_E_HIP;
    int32_t rc = ATSC3_E_HIP;
    do {
        if (!grow(c, d_need, h_need)) break;
        hipStream_t st = c->st;
        if (!tabs) {
            int32_t *d = reinterpret_cast<int32_t *>(c->dev + d_tab);
            size_t o = 0;
            if (hipMemcpyAsync(d + o, cptr, ((size_t)nc + 1) * 4, hipMemcpyHostToDevice, st) ||
                hipMemcpyAsync(d + (o += (size_t)nc + 1), cidx, (size_t)nedges * 4,
                               hipMemcpyHostToDevice, st) ||
                hipMemcpyAsync(d + (o += nedges), vptr, (nn + 1) * 4, hipMemcpyHostToDevice,
                               st) ||
                hipMemcpyAsync(d + (o += nn + 1), vedge, (size_t)nedges * 4,
                               hipMemcpyHostToDevice, st))
                break;
            tabs = d;
        }
        const int32_t *t_cptr = tabs, *t_cidx = tabs + nc + 1;
        const int32_t *t_vptr = t_cidx + nedges, *t_vedge = t_vptr + n + 1;
        uint8_t *hb = reinterpret_cast<uint8_t *>(c->host + h_out);
        uint8_t *hc = hb + nb * nn;
        int32_t *hi = reinterpret_cast<int32_t *>(hc + nb);
        int32_t *hd = hi + nb;
        bool ok = true;
        for (int b0 = 0; b0 < nblk && ok; b0 += chunk) {
            const int k = nblk - b0 < chunk ? nblk - b0 : chunk;
            std::memcpy(c->host, llr + (size_t)b0 * nn, (size_t)k * nn * 4);
            ok = hipMemcpyAsync(c->dev + d_llr, c->host, (size_t)k * nn * 4,
                                hipMemcpyHostToDevice, st) == hipSuccess;
            ok = ok && atsc3::csr_phase_kc_f32(
                           reinterpret_cast<const float *>(c->dev + d_llr), k, n, nc, nedges,
                           t_cptr, t_cidx, t_vptr, t_vedge, iters, alpha, c->dev + d_ws,
                           reinterpret_cast<uint8_t *>(c->dev + d_bits),
                           reinterpret_cast<uint8_t *>(c->dev + d_conv),
                           reinterpret_cast<int32_t *>(c->dev + d_its),
                           reinterpret_cast<int32_t *>(c->dev + d_bad), st) == ATSC3_OK;
            ok = ok &&
                 hipMemcpyAsync(hb + (size_t)b0 * nn, c->dev + d_bits, (size_t)k * nn,
                                hipMemcpyDeviceToHost, st) == hipSuccess &&
                 hipMemcpyAsync(hc + b0, c->dev + d_conv, k, hipMemcpyDeviceToHost, st) ==
                     hipSuccess &&
                 hipMemcpyAsync(hi + b0, c->dev + d_its, (size_t)k * 4, hipMemcpyDeviceToHost,
                                st) == hipSuccess &&
                 hipMemcpyAsync(hd + b0, c->dev + d_bad, (size_t)k * 4, hipMemcpyDeviceToHost,
                                st) == hipSuccess;
            // the staging buffer of the next chunk's LLRs is reused
            ok = ok && hipStreamSynchronize(st) == hipSuccess;
        }
        if (!ok) break;
        std::memcpy(bits, hb, nb * nn);
        std::memcpy(conv, hc, nb);
        std::memcpy(its, hi, nb * 4);
        std::memcpy(bad, hd, nb * 4);
        rc = ATSC3_OK;
    } while (0);
    (void)hipStreamSynchronize(c->st);
    release_ctx(c);
    return rc;
}
