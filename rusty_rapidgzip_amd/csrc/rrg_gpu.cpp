// rrg_gpu.cpp -- host driver and C ABI of the MI355X gzip/DEFLATE decoder.
//
// Orchestration that the reference runs on CPU threads (pipeline.rs:468-1041
// for single/multi-member gzip, pipeline.rs:1357-1636 for BGZF) is restated
// here around the HIP kernels of rrg_kernels.hip:
//
//   1. block finder over chunk targets (pipeline.rs:491-551) and, between
//      them, finer segment targets so that thousands of lanes decode at once;
//   2. persistent lane-per-segment inflate (u16 cells, markers);
//   3. anchored verification chain with re-decode of false starts
//      (pipeline.rs:742-856), on the host over the per-segment results;
//   4. tail chain + marker resolution + narrowing on the device;
//   5. CRC pieces on the device, folded per member with GF(2) shifts on the
//      host, checked against every trailer (pipeline.rs:858-952);
//   6. in-order output chunks with the reference's chunk boundaries.
#include "rrg_gpu.h"

#include <errno.h>
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rrg_engine.h"

namespace rrg {
namespace host {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
thread_local hipError_t t_last_hip = hipSuccess;

// allocation-failure hook: > 0 = device allocations that still succeed
// before the next one fails (once); 0 = off
static std::atomic<long> g_fail_alloc{0};

hipError_t dev_malloc(void** p, size_t n) {
    long v = g_fail_alloc.load();
    while (v > 0) {
        if (g_fail_alloc.compare_exchange_weak(v, v - 1)) {
            if (v == 1) {
                *p = nullptr;
                return hipErrorOutOfMemory;
            }
            break;
        }
    }
    return hipMalloc(p, n);
}

void set_err(rrg_error* e, int kind, const char* fmt, ...) {
    if (!e) return;
    std::memset(e, 0, sizeof *e);
    e->kind = kind;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(e->message, sizeof e->message, fmt, ap);
    va_end(ap);
}

const char* detail_msg(uint32_t d) {
    switch (d) {
        case DET_HLIT_HDIST: return "dynamic block: HLIT/HDIST out of range";
        case DET_CODE16_FIRST: return "code 16 with no previous";
        case DET_CODE16_OVERRUN: return "code 16 overruns lengths";
        case DET_CODE17_OVERRUN: return "code 17 overruns lengths";
        case DET_CODE18_OVERRUN: return "code 18 overruns lengths";
        case DET_BAD_CL_SYMBOL: return "bad code-length symbol";
        case DET_EOB_ZERO: return "EOB symbol has zero length";
        case DET_OVERSUBSCRIBED: return "oversubscribed huffman tree";
        case DET_NO_CODE: return "no matching huffman code";
        case DET_LITLEN_RANGE: return "literal/length symbol out of range";
        case DET_DIST_RANGE: return "distance symbol out of range";
        case DET_DIST_FAR: return "back-reference distance out of bounds";
        case DET_STORED_NLEN: return "stored block: LEN/NLEN mismatch";
        case DET_RESERVED_BTYPE: return "reserved block type";
        case DET_MARKER_RANGE: return "marker references bytes outside the available prefix window";
        default: return "invalid deflate stream";
    }
}

void status_err(rrg_error* e, uint32_t st, uint32_t det) {
    switch (st) {
        case ST_EOF: set_err(e, RRG_E_EOF, "unexpected end of input"); break;
        case ST_INVALID: set_err(e, RRG_E_INVALID, "invalid deflate stream: %s", detail_msg(det)); break;
        case ST_BADMAGIC: set_err(e, RRG_E_BADMAGIC, "not a gzip file (bad magic)"); break;
        case ST_METHOD: set_err(e, RRG_E_METHOD, "unsupported compression method"); break;
        case ST_RESERVED: set_err(e, RRG_E_RESERVED, "reserved gzip flag bits set"); break;
        case ST_TRUNCATED: set_err(e, RRG_E_TRUNCATED, "truncated gzip stream"); break;
        default: set_err(e, RRG_E_ARG, "output exceeds the decoder's capacity"); break;
    }
}

// gzip.rs:252-300; >= 0: header length, < 0: -kind
long parse_header(const uint8_t* p, size_t n, uint32_t* method) {
    if (n < 10) return -RRG_E_TRUNCATED;
    if (p[0] != 0x1f || p[1] != 0x8b) return -RRG_E_BADMAGIC;
    if (p[2] != 8) {
        if (method) *method = p[2];
        return -RRG_E_METHOD;
    }
    const uint8_t flg = p[3];
    if (flg & 0xe0) {
        if (method) *method = flg;
        return -RRG_E_RESERVED;
    }
    size_t pos = 10;
    if (flg & 4) {
        if (pos + 2 > n) return -RRG_E_TRUNCATED;
        pos += 2 + (p[pos] | ((size_t)p[pos + 1] << 8));
        if (pos > n) return -RRG_E_TRUNCATED;
    }
    for (int f = 8; f <= 16; f <<= 1) {
        if (flg & f) {
            if (pos > n) return -RRG_E_TRUNCATED;
            const void* z = std::memchr(p + pos, 0, n - pos);
            if (!z) return -RRG_E_TRUNCATED;
            pos = (size_t)((const uint8_t*)z - p) + 1;
        }
    }
    if (flg & 2) {
        if (pos + 2 > n) return -RRG_E_TRUNCATED;
        pos += 2;
    }
    return (long)pos;
}

// gzip.rs:311-351
uint32_t bgzf_block_size(const uint8_t* p, size_t n) {
    if (n < 12 || p[0] != 0x1f || p[1] != 0x8b || p[2] != 8 || !(p[3] & 4)) return 0;
    const size_t xlen = p[10] | ((size_t)p[11] << 8);
    if (12 + xlen > n) return 0;
    size_t q = 12;
    const size_t end = 12 + xlen;
    while (q + 4 <= end) {
        const size_t slen = p[q + 2] | ((size_t)p[q + 3] << 8);
        if (q + 4 + slen > end) return 0;
        if (p[q] == 'B' && p[q + 1] == 'C' && slen == 2) return (uint32_t)(p[q + 4] | (p[q + 5] << 8)) + 1;
        q += 4 + slen;
    }
    return 0;
}

// GF(2) CRC-32 arithmetic (reflected, P = 0xEDB88320).
uint32_t gf2_mul(uint32_t a, uint32_t b) {
    uint32_t prod = 0;
    for (int i = 0; i < 32; ++i) {
        if (a & (0x80000000u >> i)) prod ^= b;
        b = (b & 1u) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
    }
    return prod;
}
// Multiplication by a fixed a: gf2_mul is linear in its second operand, so
// four tables of 256 products (one per byte of b) replace the 32 steps.
struct GfMulBy {
    uint32_t t[4][256];
    explicit GfMulBy(uint32_t a) {
        for (int k = 0; k < 4; ++k)
            for (uint32_t b = 0; b < 256; ++b) t[k][b] = gf2_mul(a, b << (8 * k));
    }
    uint32_t operator()(uint32_t b) const {
        return t[0][b & 255u] ^ t[1][(b >> 8) & 255u] ^ t[2][(b >> 16) & 255u] ^ t[3][b >> 24];
    }
};
struct X2N {
    uint32_t t[64];
    X2N() {
        t[0] = 0x40000000u;  // x^1
        for (int k = 1; k < 64; ++k) t[k] = gf2_mul(t[k - 1], t[k - 1]);
    }
};
const X2N& x2n() {
    static X2N x;
    return x;
}
// x^(8 d 16^k) mod P for every hexadecimal digit d of n at position k
struct X8Nib {
    uint32_t t[16][16];
    X8Nib() {
        for (int k = 0; k < 16; ++k) {
            const uint32_t base = x2n().t[3 + 4 * k];  // x^(8 * 16^k) = x^(2^(3 + 4k))
            t[k][0] = 0x80000000u;                     // 1
            for (int d = 1; d < 16; ++d) t[k][d] = gf2_mul(t[k][d - 1], base);
        }
    }
};
const X8Nib& x8nib() {
    static X8Nib x;
    return x;
}
// x^(8n) mod P: one product per nonzero hexadecimal digit of n.  Lengths
// repeat (BGZF members, full CRC pieces), so the last result is kept.
uint32_t x8n(uint64_t n) {
    thread_local uint64_t last_n = ~0ull;
    thread_local uint32_t last_r = 0;
    if (n == last_n) return last_r;
    const X8Nib& tab = x8nib();
    uint32_t r = 0x80000000u;
    const uint64_t n0 = n;
    for (int k = 0; n; ++k, n >>= 4)
        if (n & 15) r = gf2_mul(tab.t[k][n & 15], r);
    last_n = n0;
    last_r = r;
    return r;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------

// Result of one file's decode.

Geometry pick_geometry(Ctx& c, const rrg_cfg& cfg) {
    Geometry G;
    if (cfg.lanes_per_block <= 0 || cfg.lanes_per_block >= 64) {
        // one wavefront per segment (the default): LUT of 9 to 11 bits
        G.g.wave = true;
        G.g.lut_bits = cfg.lut_bits >= 9 && cfg.lut_bits <= 11 ? cfg.lut_bits : cfg.lut_bits == 8 ? 9 : 10;
        G.g.lanes_per_block = 64;
        int per_cu = inflate_wave_blocks_per_cu(G.g.lut_bits);
        if (per_cu < 1) per_cu = 1;
        G.g.blocks = per_cu * c.n_cu;
        // a wave's scratch is about that of 16 lanes of the lane-per-segment kernel
        if (c.max_lanes) G.g.blocks = std::max<int>(1, std::min<int>(G.g.blocks, (int)(c.max_lanes / 16)));
        // a wave takes segments one after the other: about four each
        G.lanes_in_flight = (uint64_t)G.g.blocks * 4;
        G.scratch_bytes = (uint64_t)G.g.blocks * inflate_wave_scratch_bytes();
        return G;
    }
    G.g.lut_bits = (cfg.lut_bits == 8 || cfg.lut_bits == 10) ? cfg.lut_bits : 9;
    int lanes = cfg.lanes_per_block;
    if (lanes > 64) lanes = 64;
    while (lanes > 1 && (size_t)lanes * inflate_lds_per_lane(G.g.lut_bits) > 65536) lanes /= 2;
    G.g.lanes_per_block = lanes;
    int per_cu = inflate_blocks_per_cu(G.g.lut_bits, lanes);
    if (per_cu < 1) per_cu = 1;
    G.g.blocks = per_cu * c.n_cu;
    if (c.max_lanes && (uint64_t)G.g.blocks * (uint64_t)lanes > c.max_lanes)
        G.g.blocks = std::max<int>(1, (int)(c.max_lanes / (uint64_t)lanes));
    G.lanes_in_flight = (uint64_t)G.g.blocks * (uint64_t)lanes;
    G.scratch_bytes = G.lanes_in_flight * inflate_scratch_u16_per_lane() * sizeof(uint16_t);
    return G;
}

// Token slots of a speculative segment spanning `span` input bytes: 1.5
// tokens per input byte (gzip -6 text and FASTQ need 0.6-0.8), plus room for
// the rest of a block of up to 32 KiB of input.  A segment decodes on to the
// first block boundary at or after its hint, so when the next segment's start
// is a false find (no block starts there) its decode runs into the next
// segment's range; that room keeps it from overflowing, and the false start
// then costs one re-decode of the next segment only.  Beyond it the kernel
// counts the tokens needed and the segment is regrown.  A multiple of 4.
uint64_t tok_slots(uint64_t span) { return ((span * 3) / 2 + 16384 + 3) & ~3ull; }

// Device bytes a decode of `len` compressed bytes holds at its peak: the
// input, the token regions of its segments (tok_slots over segments of at
// least 24 KiB), the u16 cells of the output and `out` bytes of output.
uint64_t engine_footprint(uint64_t len, uint64_t out) {
    const uint64_t n_seg = len / (24 * 1024) + 1;
    const uint64_t toks = (len * 3) / 2 + n_seg * (16384 + 4);
    // the tail chain keeps an entry and an exit window per block of segments
    // (about 4 blocks per CU: at most ~1024 blocks)
    const uint64_t tail = std::min<uint64_t>(n_seg, 1024) * 2 * 65536 + 65536 * 64;
    return len + 64 + 4 * toks + 2 * (out + 64 * n_seg) + out + tail;
}

// Member-record slots of a segment spanning `span` input bytes: a few member
// ends, or, when block ends are recorded, about one block per 4 KiB of input
// (more are found by regrowing).
uint32_t mb_slots(uint64_t span, bool block_ends) {
    return block_ends ? (uint32_t)std::min<uint64_t>(8 + span / 4096, 1u << 20) : 4u;
}

// Give a segment a larger token region (or member-record region) after it
// overflowed; the old region is abandoned.  A token overflow reports the
// tokens the whole segment needs (the kernel counts on without storing).
void grow_segment(SegDesc& d, uint32_t status, uint64_t& tok_used, uint64_t& mb_used, uint64_t needed) {
    if (status == ST_MB_OVERFLOW) {
        d.mb_base = (uint32_t)mb_used;
        d.mb_cap = d.mb_cap * 16 + 16;
        mb_used += d.mb_cap;
    } else {
        const uint64_t cap = (std::max<uint64_t>(needed + 64, 2 * d.tok_cap) + 3) & ~3ull;
        d.tok_base = (tok_used + 3) & ~3ull;
        d.tok_cap = std::min<uint64_t>(cap, 0xFFFFFFF0ull);
        tok_used = d.tok_base + d.tok_cap;
    }
}

inline bool overflowed(uint32_t st) { return st == ST_OVERFLOW || st == ST_MB_OVERFLOW; }

// Copy n bytes of host memory to device address d on c.s through the pinned
// staging buffer: the call returns at once, also behind running kernels.  The
// staged bytes stay put until the copy has run (regions are handed out in
// order and reused only after the stream has drained).
int stage_upload(Ctx& c, void* d, const void* h, size_t n, rrg_error* err) {
    if (n == 0) return RRG_OK;
    if (c.no_alloc && n > c.stg_cap) {
        // no staging buffer may be made here: a plain copy
        HIPCHK(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, c.s));
        return RRG_OK;
    }
    if (c.stg_used + n > c.stg_cap) {
        // full: wait for the copies in flight, then start over (larger if need be)
        HIPCHK(hipStreamSynchronize(c.s));
        c.stg_used = 0;
        if (n > c.stg_cap) {
            if (c.stg) (void)hipHostFree(c.stg);
            c.stg = nullptr;
            c.stg_cap = 0;
            const size_t cap = std::max<size_t>(2 * n, 8u << 20);
            HIPCHK(hipHostMalloc((void**)&c.stg, cap, hipHostMallocDefault));
            c.stg_cap = cap;
        }
    }
    uint8_t* const p = c.stg + c.stg_used;
    std::memcpy(p, h, n);
    c.stg_used = (c.stg_used + n + 255) & ~(size_t)255;
    HIPCHK(hipMemcpyAsync(d, p, n, hipMemcpyHostToDevice, c.s));
    return RRG_OK;
}

// Moves every queued download to its destination; call after a wait that
// covers the copies.
void finish_downloads(Ctx& c) {
    for (const Ctx::Download& d : c.dl_pend) std::memcpy(d.dst, c.dl + d.off, d.n);
    c.dl_pend.clear();
    c.dl_used = 0;
}

// Copy n bytes from device address s to host memory dst on c.s, landing in
// pinned memory (a pageable destination makes the runtime stage the copy
// itself, synchronously): the bytes reach dst at the next finish_downloads(),
// which the caller runs after its wait on the stream.
int download(Ctx& c, void* dst, const void* s, size_t n, rrg_error* err) {
    if (n == 0) return RRG_OK;
    if (c.dl_used + n > c.dl_cap) {
        if (c.no_alloc && n > c.dl_cap) {
            HIPCHK(hipMemcpyAsync(dst, s, n, hipMemcpyDeviceToHost, c.s));
            return RRG_OK;
        }
        // full: land what is queued, then start over (larger if need be)
        HIPCHK(hipStreamSynchronize(c.s));
        finish_downloads(c);
        if (n > c.dl_cap) {
            if (c.dl) (void)hipHostFree(c.dl);
            c.dl = nullptr;
            c.dl_cap = 0;
            const size_t cap = std::max<size_t>(2 * n, 4u << 20);
            HIPCHK(hipHostMalloc((void**)&c.dl, cap, hipHostMallocDefault));
            c.dl_cap = cap;
        }
    }
    HIPCHK(hipMemcpyAsync(c.dl + c.dl_used, s, n, hipMemcpyDeviceToHost, c.s));
    c.dl_pend.push_back(Ctx::Download{dst, c.dl_used, n});
    c.dl_used = (c.dl_used + n + 255) & ~(size_t)255;
    return RRG_OK;
}

// The expansion's time, once the stream has passed its end event.
void collect_expand_time(Ctx& c) {
    if (!c.x_timed) return;
    float ms = 0;
    if (hipEventElapsedTime(&ms, c.evx[0], c.evx[1]) == hipSuccess) c.prof.ms_expand += ms;
    c.x_timed = false;
}

// One launch of inflate pass 1 over the segment indices in `work`; results
// for every segment are copied back into `res`.
int run_inflate(Ctx& c, const Geometry& G, std::vector<SegDesc>& desc, std::vector<SegResult>& res,
                uint64_t tok_used, uint64_t mb_used, const std::vector<uint32_t>& work, rrg_error* err,
                std::vector<MemberRec>* recs) {
    const size_t S = desc.size();
    if (work.empty()) return RRG_OK;
    HIPCHK(c.sorted.ensure(G.scratch_bytes));
    // tokens of segments decoded by earlier launches stay where they are
    HIPCHK(c.toks.grow_keep((tok_used + 64) * sizeof(uint32_t), c.toks.cap, c.s));
    HIPCHK(c.mrec.grow_keep((mb_used + 1) * sizeof(MemberRec), c.mrec.cap, c.s));
    HIPCHK(c.segs.ensure(S * sizeof(SegDesc)));
    // results of segments decoded by earlier launches of this decode stay in
    // the buffer (every launch copies all S back), so growing keeps them
    HIPCHK(c.res.grow_keep(S * sizeof(SegResult), c.res.cap, c.s));
    HIPCHK(c.work.ensure(work.size() * sizeof(uint32_t)));
    HIPCHK(c.counter.ensure(16));
    if (int rc = stage_upload(c, c.segs.p, desc.data(), S * sizeof(SegDesc), err)) return rc;
    if (int rc = stage_upload(c, c.work.p, work.data(), work.size() * sizeof(uint32_t), err)) return rc;
    HIPCHK(hipMemsetAsync(c.counter.p, 0, 16, c.s));
    const double before = c.prof.ms_inflate;
    {
        StageTimer t(c, &c.prof.ms_inflate);
        if (G.g.wave)
            HIPCHK(launch_tokens_wave(c.in.as<uint64_t>(), c.segs.as<SegDesc>(), c.res.as<SegResult>(),
                                      c.mrec.as<MemberRec>(), c.work.as<uint32_t>(), (uint32_t)work.size(),
                                      c.counter.as<uint32_t>(), c.toks.as<uint32_t>(), c.sorted.as<uint8_t>(), G.g,
                                      c.s));
        else
            HIPCHK(launch_tokens(c.in.as<uint64_t>(), c.segs.as<SegDesc>(), c.res.as<SegResult>(),
                                 c.mrec.as<MemberRec>(), c.work.as<uint32_t>(), (uint32_t)work.size(),
                                 c.counter.as<uint32_t>(), c.toks.as<uint32_t>(), c.sorted.as<uint16_t>(), G.g, c.s));
        t.stop();
    }
    c.prof.n_inflate_launches++;
    res.resize(S);
    if (int rc = download(c, res.data(), c.res.p, S * sizeof(SegResult), err)) return rc;
    // the member records of every segment so far come back with the same wait
    if (recs) {
        recs->resize(mb_used);
        if (mb_used)
            if (int rc = download(c, recs->data(), c.mrec.p, mb_used * sizeof(MemberRec), err)) return rc;
    }
    HIPCHK(hipStreamSynchronize(c.s));
    finish_downloads(c);
    c.stg_used = 0;  // the stream has drained: no staged upload is pending
    if (c.verbose) {
        // the launch's largest outputs and its status counts
        std::vector<std::pair<uint64_t, uint32_t>> big;
        size_t n_over = 0, n_err = 0;
        uint64_t toks = 0, bytes = 0;
        for (uint32_t j : work) {
            big.push_back({res[j].n_cells, j});
            n_over += overflowed(res[j].status);
            n_err += res[j].status != ST_OK && !overflowed(res[j].status);
            toks += res[j].n_tok;
            bytes += (res[j].end_bit - desc[j].start_bit) / 8;
        }
        std::sort(big.rbegin(), big.rend());
        std::fprintf(stderr, "[rrg-gpu] inflate launch %llu: %zu segments in %.3f ms (%zu out of room, %zu failed, "
                     "%.3f tokens per input byte); largest outputs:", (unsigned long long)c.prof.n_inflate_launches,
                     work.size(), c.prof.ms_inflate - before, n_over, n_err, bytes ? (double)toks / (double)bytes : 0.0);
        for (size_t k = 0; k < big.size() && k < 4; ++k)
            std::fprintf(stderr, " %llu cells from %llu B", (unsigned long long)big[k].first,
                         (unsigned long long)((res[big[k].second].end_bit - desc[big[k].second].start_bit) / 8));
        // the input spans (the lane's work), largest and by quantile
        std::vector<uint64_t> span;
        for (uint32_t j : work) span.push_back((res[j].end_bit - desc[j].start_bit) / 8);
        std::sort(span.begin(), span.end());
        if (!span.empty())
            std::fprintf(stderr, "; input spans: median %llu B, p99 %llu B, max %llu B",
                         (unsigned long long)span[span.size() / 2], (unsigned long long)span[span.size() * 99 / 100],
                         (unsigned long long)span.back());
        std::fprintf(stderr, "\n");
    }
    return RRG_OK;
}

// Inflate pass 2 over `items` (tokens to cells, or to bytes of `out`), the
// largest first so that no long expansion starts last.  Returns without
// waiting: the caller prepares the next stage meanwhile, and the time is
// collected after a later wait (collect_expand_time).
int run_expand(Ctx& c, std::vector<ExpandItem>& items, DevBuf& out, rrg_error* err) {
    if (items.empty()) return RRG_OK;
    {
        // largest first, by 4 Ki-cell buckets (a counting sort: the items are
        // many and their sizes close)
        uint64_t top = 0;
        for (const ExpandItem& it : items) top = std::max<uint64_t>(top, it.n_cells >> 12);
        top = std::min<uint64_t>(top, 1u << 16);
        auto key = [&](const ExpandItem& it) { return (size_t)(top - std::min<uint64_t>(it.n_cells >> 12, top)); };
        std::vector<uint32_t> cnt(top + 2, 0);
        for (const ExpandItem& it : items) ++cnt[key(it) + 1];
        for (size_t b = 1; b < cnt.size(); ++b) cnt[b] += cnt[b - 1];
        std::vector<ExpandItem> by(items.size());
        for (const ExpandItem& it : items) by[cnt[key(it)]++] = it;
        items.swap(by);
    }
    HIPCHK(c.xitems.ensure(items.size() * sizeof(ExpandItem)));
    HIPCHK(c.counter.ensure(16));
    if (int rc = stage_upload(c, c.xitems.p, items.data(), items.size() * sizeof(ExpandItem), err)) return rc;
    HIPCHK(hipMemsetAsync(c.counter.p, 0, 16, c.s));
    HIPCHK(hipEventRecord(c.evx[0], c.s));
    HIPCHK(launch_expand(c.toks.as<uint32_t>(), c.xitems.as<ExpandItem>(), (uint32_t)items.size(),
                         c.counter.as<uint32_t>(), c.in.as<uint8_t>(), c.cells.as<uint16_t>(), out.as<uint8_t>(),
                         c.n_cu, c.s));
    HIPCHK(hipEventRecord(c.evx[1], c.s));
    c.x_timed = true;
    for (const ExpandItem& it : items) c.prof.n_tokens += it.n_tok;
    return RRG_OK;
}

// Cells of an expansion piece.  An item of at least 2 * kPiece cells (a
// segment of runs or short-period repeats expands to millions of cells per
// compressed KiB) would keep one wavefront busy long after the others end.
constexpr uint64_t kPiece = 1u << 19;

// Expansion items of chain segments with at least 2 * kPiece cells are cut at
// token boundaries into pieces of about kPiece cells (k_expand_cuts finds the
// tokens).  Each piece expands on a wavefront of its own with its cells
// numbered from its own start, so a source before the piece becomes a marker
// 0x8000 | offset before the piece, and each piece becomes a chain segment of
// its own that may hold markers: the tail chain and the resolve pass fill
// them in from the pieces before, as for any segment.  `item_cs[i]` is the
// chain segment of items[i] (SIZE_MAX: not a chain segment).  On return
// `of` maps each new chain segment to the one it was cut from (empty: nothing
// was cut), `cs_orig` holds the chain before the cut and `first_len[k]` the
// cells of its segment k's first piece (whose markers alone may name bytes
// before the segment).
int split_long_items(Ctx& c, std::vector<ExpandItem>& items, const std::vector<size_t>& item_cs,
                     std::vector<ChainSeg>& cs, std::vector<uint64_t>& cs_out, std::vector<ChainSeg>& cs_orig,
                     std::vector<uint32_t>& of, std::vector<uint64_t>& first_len, rrg_error* err) {
    cs_orig.clear();
    of.clear();
    first_len.clear();
    std::vector<size_t> big;
    for (size_t i = 0; i < items.size(); ++i)
        if (item_cs[i] != SIZE_MAX && !(items[i].flags & EXP_U8) && items[i].n_tok && items[i].n_cells >= 2 * kPiece)
            big.push_back(i);
    if (big.empty()) return RRG_OK;
    const uint32_t nb = (uint32_t)big.size();
    std::vector<ExpandItem> bi(nb);
    std::vector<uint32_t> firsts(nb);
    uint64_t R = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        bi[b] = items[big[b]];
        firsts[b] = (uint32_t)R;
        R += (bi[b].n_cells + kPiece - 1) / kPiece - 1;
    }
    const size_t items_bytes = ((size_t)nb * sizeof(ExpandItem) + 255) & ~(size_t)255;
    HIPCHK(c.xcut.ensure(items_bytes + (size_t)nb * 4));
    HIPCHK(c.cuts.ensure((size_t)R * sizeof(uint2)));
    if (int rc = stage_upload(c, c.xcut.p, bi.data(), (size_t)nb * sizeof(ExpandItem), err)) return rc;
    uint8_t* const d_first = static_cast<uint8_t*>(c.xcut.p) + items_bytes;
    if (int rc = stage_upload(c, d_first, firsts.data(), (size_t)nb * 4, err)) return rc;
    // every record is written once; a record left at ~0 would leave its item whole
    HIPCHK(hipMemsetAsync(c.cuts.p, 0xff, (size_t)R * sizeof(uint2), c.s));
    HIPCHK(launch_expand_cuts(c.toks.as<uint32_t>(), c.xcut.as<ExpandItem>(), reinterpret_cast<const uint32_t*>(d_first),
                              nb, (uint32_t)kPiece, c.cuts.as<uint2>(), c.s));
    std::vector<uint2> cuts(R);
    HIPCHK(hipMemcpyAsync(cuts.data(), c.cuts.p, (size_t)R * sizeof(uint2), hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    // the pieces of each cut item: (first token, first cell), in order
    std::vector<std::vector<std::pair<uint32_t, uint64_t>>> pieces(cs.size());
    std::vector<ExpandItem> out_items;
    out_items.reserve(items.size() + R);
    size_t bpos = 0;
    for (size_t i = 0; i < items.size(); ++i) {
        if (bpos < nb && big[bpos] == i) {
            const ExpandItem& it = items[i];
            const uint32_t b = (uint32_t)bpos++;
            const uint64_t nr = (it.n_cells + kPiece - 1) / kPiece - 1;
            std::vector<std::pair<uint32_t, uint64_t>> pc{{0u, 0ull}};
            bool ok = true;
            for (uint64_t r = 0; r < nr && ok; ++r) {
                const uint2 q = cuts[firsts[b] + r];
                if (q.x == pc.back().first && q.y == pc.back().second) continue;  // a long stored token holds several
                ok = q.x > pc.back().first && q.y > pc.back().second && q.x < it.n_tok && q.y < it.n_cells;
                pc.push_back({q.x, (uint64_t)q.y});
            }
            if (ok && pc.size() > 1) {
                for (size_t p = 0; p < pc.size(); ++p) {
                    const uint64_t t1 = p + 1 < pc.size() ? pc[p + 1].first : it.n_tok;
                    const uint64_t c1 = p + 1 < pc.size() ? pc[p + 1].second : it.n_cells;
                    out_items.push_back(ExpandItem{it.tok_base + pc[p].first, it.dst + pc[p].second, it.in_byte,
                                                   c1 - pc[p].second, (uint32_t)(t1 - pc[p].first), it.flags});
                }
                pieces[item_cs[i]] = std::move(pc);
                continue;
            }
        }
        out_items.push_back(items[i]);
    }
    items.swap(out_items);
    // the chain with each cut segment replaced by its pieces
    std::vector<ChainSeg> cs2;
    std::vector<uint64_t> out2;
    cs2.reserve(cs.size() + R);
    out2.reserve(cs.size() + R);
    first_len.resize(cs.size());
    for (size_t k = 0; k < cs.size(); ++k) {
        const std::vector<std::pair<uint32_t, uint64_t>>& pc = pieces[k];
        first_len[k] = pc.size() > 1 ? pc[1].second : cs[k].len;
        if (pc.size() <= 1) {
            cs2.push_back(cs[k]);
            out2.push_back(cs_out[k]);
            of.push_back((uint32_t)k);
            continue;
        }
        for (size_t p = 0; p < pc.size(); ++p) {
            const uint64_t c1 = p + 1 < pc.size() ? pc[p + 1].second : cs[k].len;
            ChainSeg s = cs[k];
            s.cells_base = cs[k].cells_base + pc[p].second;
            s.len = c1 - pc[p].second;
            if (p) {
                s.spec = 1;   // cells before the piece are markers
                s.first = 0;  // and the pieces before it are reachable
            }
            cs2.push_back(s);
            out2.push_back(cs_out[k] + pc[p].second);
            of.push_back((uint32_t)k);
        }
    }
    if (c.verbose)
        std::fprintf(stderr, "[rrg-gpu] expansion: %u long items cut into %zu pieces\n", nb,
                     cs2.size() - cs.size() + nb);
    cs.swap(cs2);
    cs_out.swap(out2);
    cs_orig.swap(cs2);
    return RRG_OK;
}

// The kernels after the expansion, for the chain segments `cs` into `out`
// (out_off[k]: chain segment k's first output byte) and the CRC ranges
// [lo, hi) of `out`: tail chain, resolve + narrow, CRC pieces.  Every
// descriptor goes up first and the error flags are cleared before the
// kernels run; the kernels then go as one launch of a graph built once per
// context (c.graph_post) or as separate launches, and one copy back and one
// wait end the stage.
// flags[k] != 0 when a marker of segment k reaches before the start of its
// file's output; raw[i] = the raw CRC of range i.
int resolve_and_crc(Ctx& c, std::vector<ChainSeg>& cs, const std::vector<uint64_t>& out_off, DevBuf& out,
                    const std::vector<std::pair<uint64_t, uint64_t>>& ranges, std::vector<uint32_t>& flags,
                    std::vector<uint32_t>& raw, rrg_error* err) {
    const uint32_t n = (uint32_t)cs.size();
    flags.assign(n, 0);
    // tail chain: one zone of up to 32 Ki cells per segment
    uint64_t tz = 0;
    bool any_spec = false;
    for (uint32_t j = 0; j < n; ++j) {
        cs[j].tz_base = tz;
        tz += std::min<uint64_t>(cs[j].len, 32768);
        any_spec |= cs[j].spec != 0;
    }
    const bool tail = n && any_spec && tz;
    // ~2 blocks per CU keeps phase 1 parallel while phase 2 stays short
    // (config 2: 1, 2, 4, 8 and 16 per CU took 1.16, 1.11, 1.16, 1.36 and 1.77 ms)
    const uint32_t G = std::max<uint32_t>(1, (uint32_t)(n / std::max(1, 2 * c.n_cu)));
    // resolve tiles: up to 512 KiB of output each, aligned in output
    // coordinates (a tile loads its segment's 32 KiB marker window once: a
    // larger tile loads it fewer times; 256 KiB tiles took 2.4% longer)
    const uint64_t TILE = 512 * 1024;
    std::vector<ResolveTile> tiles;
    for (uint32_t j = 0; j < n; ++j) {
        const uint64_t ob = out_off[j], oe = ob + cs[j].len;
        uint64_t p = ob;
        while (p < oe) {
            const uint64_t q = std::min<uint64_t>(oe, (p / TILE + 1) * TILE);
            ResolveTile t;
            t.cell_addr = cs[j].cells_base + (p - ob);
            t.out_off = p;
            t.seg = j;
            t.n = (uint32_t)(q - p);
            tiles.push_back(t);
            p = q;
        }
    }
    // CRC pieces of at most 256 KiB
    const uint64_t PIECE = 256 * 1024;
    std::vector<uint64_t> po;
    std::vector<uint32_t> pl;
    std::vector<size_t> first(ranges.size() + 1);
    for (size_t i = 0; i < ranges.size(); ++i) {
        first[i] = po.size();
        for (uint64_t p = ranges[i].first; p < ranges[i].second; p += PIECE) {
            po.push_back(p);
            pl.push_back((uint32_t)std::min<uint64_t>(PIECE, ranges[i].second - p));
        }
    }
    first[ranges.size()] = po.size();
    std::vector<uint32_t> pc(po.size());
    // descriptors up, flags cleared
    if (n) {
        HIPCHK(c.chain.ensure(n * sizeof(ChainSeg)));
        if (int rc = stage_upload(c, c.chain.p, cs.data(), n * sizeof(ChainSeg), err)) return rc;
        HIPCHK(c.err.ensure((size_t)n * 4));
        HIPCHK(hipMemsetAsync(c.err.p, 0, (size_t)n * 4, c.s));
    }
    if (tail) HIPCHK(c.tz.ensure(tail_scratch_bytes(n, G)));
    if (!tiles.empty()) {
        HIPCHK(c.tiles.ensure(tiles.size() * sizeof(ResolveTile)));
        if (int rc = stage_upload(c, c.tiles.p, tiles.data(), tiles.size() * sizeof(ResolveTile), err)) return rc;
    }
    if (!po.empty()) {
        HIPCHK(c.poff.ensure(po.size() * 8));
        HIPCHK(c.plen.ensure(pl.size() * 4));
        HIPCHK(c.pcrc.ensure(pc.size() * 4));
        if (int rc = stage_upload(c, c.poff.p, po.data(), po.size() * 8, err)) return rc;
        if (int rc = stage_upload(c, c.plen.p, pl.data(), pl.size() * 4, err)) return rc;
    }
    PostLaunch pl_;
    pl_.tail = tail;
    pl_.cs = c.chain.as<ChainSeg>();
    pl_.n_seg = n;
    pl_.G = G;
    pl_.cells = c.cells.as<uint16_t>();
    pl_.scratch = tail ? c.tz.as<uint16_t>() : nullptr;
    pl_.err = c.err.as<uint32_t>();
    pl_.tiles = c.tiles.as<ResolveTile>();
    pl_.n_tiles = (uint32_t)tiles.size();
    pl_.out = out.as<uint8_t>();
    pl_.poff = c.poff.as<uint64_t>();
    pl_.plen = c.plen.as<uint32_t>();
    pl_.n_pieces = (uint32_t)po.size();
    pl_.pcrc = c.pcrc.as<uint32_t>();
    if (c.graph_post) {
        // one launch of the graph for these kernels (built once per context)
        if (tail || pl_.n_tiles || pl_.n_pieces) {
            StageTimer t(c, &c.prof.ms_resolve);
            HIPCHK(launch_post(pl_, &c.post_graphs, c.s));
            t.stop();
            c.prof.n_graph_launches++;
        }
    } else {
        PostLaunch only = pl_;
        if (tail) {
            only.n_tiles = only.n_pieces = 0;
            StageTimer t(c, &c.prof.ms_tail);
            HIPCHK(launch_post(only, nullptr, c.s));
            t.stop();
        }
        if (pl_.n_tiles) {
            only = pl_;
            only.tail = false;
            only.n_pieces = 0;
            StageTimer t(c, &c.prof.ms_resolve);
            HIPCHK(launch_post(only, nullptr, c.s));
            t.stop();
        }
        if (pl_.n_pieces) {
            only = pl_;
            only.tail = false;
            only.n_tiles = 0;
            StageTimer t(c, &c.prof.ms_crc);
            HIPCHK(launch_post(only, nullptr, c.s));
            t.stop();
        }
    }
    if (tail) c.prof.n_tail_rounds += (n + G - 1) / G;
    if (n)
        if (int rc = download(c, flags.data(), c.err.p, (size_t)n * 4, err)) return rc;
    if (!po.empty())
        if (int rc = download(c, pc.data(), c.pcrc.p, pc.size() * 4, err)) return rc;
    HIPCHK(hipStreamSynchronize(c.s));
    finish_downloads(c);
    collect_expand_time(c);
    raw.assign(ranges.size(), 0);
    static const GfMulBy full_shift(x8n(PIECE));  // most pieces are whole
    for (size_t i = 0; i < ranges.size(); ++i) {
        uint32_t r = 0;
        for (size_t k = first[i]; k < first[i + 1]; ++k)
            r = (pl[k] == PIECE ? full_shift(r) : gf2_mul(x8n(pl[k]), r)) ^ pc[k];
        raw[i] = r;
    }
    return RRG_OK;
}

inline uint32_t std_crc(uint32_t raw, uint64_t len) { return raw ^ gf2_mul(x8n(len), 0xffffffffu) ^ 0xffffffffu; }

// ---------------------------------------------------------------------------
// batch engine
//
// A batch is a set of gzip files resident in c.in.  Every stage runs once for
// the whole batch: one finder launch over the chunk/segment targets of every
// file, one persistent inflate launch over all segments (BGZF members and
// speculative segments of uneven size share the work queue), host chain walks
// whose re-decodes are again gathered into one launch per round, one tail
// chain + resolve over the concatenated chains, one CRC launch over every
// member of every file.  A single-file decode is a batch of one.
// ---------------------------------------------------------------------------

// Host planning: BGZF framing walk (pipeline.rs:1357-1420) or the first header
// (lib.rs:463-472, gzip.rs:252-300).
void plan_job(Job& J) {
    const uint8_t* h = J.h;
    const uint64_t len = J.len;
    uint64_t rest = 0;
    if (!J.stream && bgzf_block_size(h, len)) {
        J.bgzf = true;
        uint64_t pos = 0;
        while (pos < len) {
            const uint32_t sz = bgzf_block_size(h + pos, len - pos);
            if (!sz) break;
            if (pos + sz > len) {
                J.walk_trunc = true;
                break;
            }
            J.mem.push_back({pos, sz});
            pos += sz;
        }
        if (J.walk_trunc || pos >= len) return;
        rest = pos;  // remainder: the first non-BGZF member onwards (pipeline.rs:1622-1633)
    }
    uint32_t meth = 0;
    const long hl = parse_header(h + rest, len - rest, &meth);
    if (hl < 0) {
        J.pre_failed = true;
        status_err(&J.pre_err, (uint32_t)-hl, 0);
        J.pre_err.method = meth;
        return;
    }
    J.body = rest + (uint64_t)hl;
}

// Finder limit flag: any verified block start will do (segment targets).
// Chunk targets keep the reference's candidate order (pipeline.rs:491-551,
// block_finder.rs:109-141), since the chunk boundaries must be its own.
constexpr uint64_t kAnyStart = 1ull << 63;

// Targets of one member-path body: chunk targets (pipeline.rs:497-503) and
// segment targets between them.
void make_targets(const Job& J, const Geometry& G, std::vector<uint64_t>& targets, std::vector<uint64_t>& limits,
                  std::vector<uint64_t>& scan_ends, std::vector<uint8_t>& is_chunk) {
    const uint64_t base = J.chain_start();
    const uint64_t total = J.last_end() - base;
    const uint64_t chunk_bits = (uint64_t)std::max<size_t>(J.chunk_size, 65536) * 8;
    uint64_t seg_bytes = J.segment_bytes;
    if (!seg_bytes) {
        // about one segment per lane in flight (a lane decodes one segment in
        // roughly the time the launch takes; four per wavefront of the
        // wave-per-segment geometry), between 24 and 64 KiB (96 KiB for that
        // geometry, whose wavefronts take their segments from one queue):
        // shorter segments cost finder work (a segment cannot be shorter than
        // a deflate block, so targets closer than the blocks are wasted
        // scans), longer ones idle lanes
        seg_bytes = (total / 8) / std::max<uint64_t>(1, G.lanes_in_flight);
        seg_bytes = std::min<uint64_t>(std::max<uint64_t>(seg_bytes, 24 * 1024), (G.g.wave ? 96 : 64) * 1024);
    }
    const uint64_t D = std::max<uint64_t>(1, chunk_bits / (seg_bytes * 8));
    const uint64_t seg_bits = J.serial ? seg_bytes * 8 : chunk_bits / D;
    if (J.stream || J.serial || J.slab) {
        // chunk starts already fixed: segment targets inside each chunk
        for (size_t k = 0; k < J.stream_B.size(); ++k) {
            const uint64_t lo = J.stream_B[k];
            const uint64_t hi = k + 1 < J.stream_B.size() ? J.stream_B[k + 1] : J.last_end();
            for (uint64_t t = lo + seg_bits; t < hi; t += seg_bits) {
                targets.push_back(t);
                limits.push_back(J.end_bits() | kAnyStart);
                scan_ends.push_back(std::min(hi, t + 2 * seg_bits));
                is_chunk.push_back(0);
            }
        }
        return;
    }
    for (uint64_t k = 0; k * chunk_bits < total; ++k) {
        if (k) {
            targets.push_back(base + k * chunk_bits);
            limits.push_back(J.end_bits());
            scan_ends.push_back(J.end_bits());
            is_chunk.push_back(1);
        }
        for (uint64_t j = 1; j < D; ++j) {
            const uint64_t t = k * chunk_bits + j * seg_bits;
            if (t >= total) break;
            targets.push_back(base + t);
            // a segment target is only a hint: search two segments' worth of
            // input (long stretches without block starts are not scanned through)
            limits.push_back(J.end_bits() | kAnyStart);
            scan_ends.push_back(std::min(J.end_bits(), base + t + 2 * seg_bits));
            is_chunk.push_back(0);
        }
    }
}

// A member-path segment that has to be decoded again: the one at chain
// position `pos` of job `job`, from bit `start` (its predecessor's end).
struct Redo {
    size_t job, pos;
    uint64_t start;
    uint32_t status;  // ST_OK: false start; ST_OVERFLOW / ST_MB_OVERFLOW: region too small
    uint64_t needed;  // tokens the segment needs (ST_OVERFLOW)
    uint64_t end_bit = UINT64_MAX;  // where its decode ended (ST_OVERFLOW; the start may be unverified)
};

// Anchored verification chain of one job (pipeline.rs:774-809): accept a
// segment iff it decoded OK from where its predecessor ended; otherwise
// record it for re-decode from there (optimistically continuing with the next
// segment's own start, verified next round).
void chain_round(Job& J, size_t job, const std::vector<SegDesc>& desc, const std::vector<SegResult>& res,
                 std::vector<Redo>& redo) {
    const size_t before = redo.size();
    uint64_t expected = J.chain_start();
    J.n_chain = 0;
    J.fatal = -1;
    J.saw_final = false;
    const size_t n = J.order.size();
    for (size_t k = 0; k < n; ++k) {
        const uint32_t j = J.order[k];
        const SegResult& r = res[j];
        if (desc[j].start_bit == expected && r.status == ST_OK) {
            ++J.n_chain;
            expected = r.end_bit;
            if (r.final_block) {
                J.saw_final = true;
                break;
            }
            continue;
        }
        if (desc[j].start_bit == expected && !overflowed(r.status)) {  // genuine error at a true boundary
            if (redo.size() == before) J.fatal = (long)k;
            break;
        }
        const bool verified = redo.size() == before;  // the chain before it holds no re-decode
        redo.push_back(Redo{job, k, expected, desc[j].start_bit == expected ? r.status : (uint32_t)ST_OK, r.n_tok});
        if (desc[j].start_bit == expected && r.status == ST_OVERFLOW) redo.back().end_bit = r.end_bit;
        if (verified && desc[j].start_bit == expected && r.status == ST_OVERFLOW) {
            // out of cells at a verified start, the lane decoded on without
            // storing: where it ends is known, so the next segment is checked
            // in this same round
            if (r.final_block) break;
            expected = r.end_bit;
        } else if (k + 1 < n) {
            expected = desc[J.order[k + 1]].start_bit;
        }
    }
    J.chain_end = expected;
}

// Errors, CRC/ISIZE checks and chunk emission of the member-path part of a
// job (pipeline.rs:783-809, 858-952).  `base` is the job-relative output
// offset of the part; `off[k]` the part-relative offset of chain segment k.
void finish_member_path(Job& J, const std::vector<SegDesc>& desc, const std::vector<SegResult>& res,
                        const std::vector<MemberRec>& recs, const std::vector<uint32_t>& flags,
                        const std::vector<uint32_t>& raw, size_t r_lo, uint64_t base, const std::vector<uint64_t>& off,
                        const std::vector<uint64_t>& bad) {
    const size_t n_chain = J.n_chain;
    const size_t S = J.order.size();
    Decoded& dec = J.dec;
    std::vector<uint32_t> chunk_of(S);
    {
        uint32_t k = 0;
        for (size_t j = 0; j < S; ++j) {
            if (j && J.seg_is_chunk[j]) ++k;
            chunk_of[j] = k;
        }
    }
    // speculation failure = a chunk whose guessed start was not where the chain
    // arrived, or whose speculative decode failed
    for (size_t j = 1; j < n_chain; ++j)
        if (J.seg_is_chunk[j] && (J.orig_start[j] != desc[J.order[j]].start_bit || !J.first_ok[j]))
            ++dec.spec_failures;
    const uint64_t body_abs = J.resume ? J.body_ref : J.in_off + J.body;
    // earliest failure, as the chunk index it stops the stream at
    uint32_t fail_chunk = UINT32_MAX;
    rrg_error ferr{};
    if (J.fatal >= 0) {
        fail_chunk = chunk_of[J.fatal];
        status_err(&ferr, res[J.order[J.fatal]].status, res[J.order[J.fatal]].detail);
    }
    // a back-reference before its member's start fails its chunk's decode,
    // ahead of any later error in that chunk
    for (size_t k = 0; k < n_chain; ++k)
        if (bad[k] != UINT64_MAX) {
            if (chunk_of[k] <= fail_chunk) {
                fail_chunk = chunk_of[k];
                set_err(&ferr, RRG_E_INVALID, "invalid deflate stream: %s", detail_msg(DET_DIST_FAR));
            }
            break;
        }
    for (size_t k = 0; k < n_chain; ++k)
        if (flags[J.c_lo + (J.bg_hi - J.bg_lo) + k]) {
            if (chunk_of[k] < fail_chunk) {
                fail_chunk = chunk_of[k];
                set_err(&ferr, RRG_E_INVALID, "invalid deflate stream: %s", detail_msg(DET_MARKER_RANGE));
            }
            break;
        }
    // member records in chain order, with their output ranges
    size_t mi = 0;
    uint64_t mstart = 0;
    bool stop = false;
    for (size_t j = 0; j < n_chain && !stop; ++j) {
        for (uint32_t k = 0; k < res[J.order[j]].n_mb; ++k, ++mi) {
            const MemberRec& m = recs[desc[J.order[j]].mb_base + k];
            const uint64_t e = off[j] + m.out_off;
            uint64_t len = e - mstart;
            mstart = e;
            const uint32_t ck = chunk_of[j];
            if (ck >= fail_chunk) {
                stop = true;
                break;
            }
            uint32_t r = raw[r_lo + mi];
            if (mi == 0 && J.resume) {
                // the member began in an earlier slab
                r = gf2_mul(x8n(len), J.carry_raw) ^ r;
                len += J.carry_len;
            }
            const uint32_t got = std_crc(r, len);
            const uint64_t tb = m.trailer_byte - body_abs;
            const uint64_t mnum = J.carry_member + mi;
            if (got != m.crc) {
                fail_chunk = ck;
                set_err(&ferr, RRG_E_CRC,
                        "CRC32 mismatch in member #%llu (uncompressed bytes: %llu, trailer at byte %llu): expected "
                        "%#010x, got %#010x",
                        (unsigned long long)mnum, (unsigned long long)len, (unsigned long long)tb, m.crc, got);
                ferr.expected = m.crc;
                ferr.got = got;
                ferr.got64 = len;
                ferr.member = (uint32_t)mnum;
                ferr.trailer_byte = tb;
                stop = true;
                break;
            }
            if ((uint32_t)len != m.isize) {
                fail_chunk = ck;
                set_err(&ferr, RRG_E_ISIZE, "ISIZE mismatch in member #%llu (trailer at byte %llu): expected %u, got %llu",
                        (unsigned long long)mnum, (unsigned long long)tb, m.isize, (unsigned long long)len);
                ferr.expected = m.isize;
                ferr.got64 = len;
                ferr.member = (uint32_t)mnum;
                ferr.trailer_byte = tb;
                stop = true;
                break;
            }
        }
    }
    if (J.slab && !stop) {
        // the member still open where the slab's chain ends
        J.members_done = J.carry_member + mi;
        const uint64_t tail = off[n_chain] - mstart;
        J.open_raw = J.open_r == SIZE_MAX ? 0u : raw[J.open_r];
        J.open_len = tail;
        if (mi == 0 && J.resume) {
            J.open_raw = gf2_mul(x8n(tail), J.carry_raw) ^ J.open_raw;
            J.open_len += J.carry_len;
        }
    }
    if (fail_chunk == UINT32_MAX && !J.saw_final && J.stop_bit == UINT64_MAX) {
        // the chain ran past every segment without reaching the end of the file
        fail_chunk = n_chain ? chunk_of[n_chain - 1] + 1 : 0;
        set_err(&ferr, RRG_E_IO, "parallel decode produced no final block");
    }
    // emit chunks: [start of chunk k, start of chunk k+1) over the chain
    const uint32_t n_chunks_chain = n_chain ? chunk_of[n_chain - 1] + 1 : 0;
    const uint32_t emit = std::min(n_chunks_chain, fail_chunk);
    uint32_t emitted = 0;
    for (size_t j = 1; j <= n_chain && emitted < emit; ++j)
        if (j == n_chain || chunk_of[j] != chunk_of[j - 1]) {
            dec.chunk_ends.push_back(base + off[j]);
            ++emitted;
        }
    if (fail_chunk != UINT32_MAX) {
        dec.failed = true;
        dec.error = ferr;
    }
}

// Errors, CRC/ISIZE checks and chunk emission of a single-worker job: the
// serial path (pipeline.rs:1211-1345) or, on non-seekable input, the streaming
// serial path (lib.rs:244-361).  Chunks end at block ends: after a non-final
// block all but the member's last 32 KiB are sent once more than that is
// pending (the serial path waits for chunk size + 32 KiB), and a member's
// final block sends the rest before its trailer is checked.  An error stops
// the stream after the chunks sent before the failing block.  `bad[k]` is the
// first cell of chain segment k produced by a back-reference before its
// member's start (UINT64_MAX: none).
void finish_serial(Job& J, const std::vector<SegDesc>& desc, const std::vector<SegResult>& res,
                   const std::vector<MemberRec>& recs, const std::vector<uint32_t>& raw, size_t r_lo, uint64_t base,
                   const std::vector<uint64_t>& off, const std::vector<uint64_t>& bad) {
    constexpr uint64_t W = 32768;
    const bool streaming = J.serial == 2;
    const uint64_t T = std::max<uint64_t>(J.chunk_size, 65536);
    Decoded& dec = J.dec;
    dec.spec_failures = 0;
    uint64_t sent = 0, mstart = 0;  // part-relative: end of the output sent, start of the current member
    size_t mi = 0, ri = r_lo;
    rrg_error ferr{};
    bool failed = false;
    auto emit = [&](uint64_t e) {
        dec.chunk_ends.push_back(base + e);
        sent = e;
    };
    for (size_t k = 0; k < J.n_out && !failed; ++k) {
        const uint32_t j = J.order[k];
        const SegResult& r = res[j];
        const uint64_t bad_at = bad[k] == UINT64_MAX ? UINT64_MAX : off[k] + bad[k];
        for (uint32_t q = 0; q < r.n_mb && !failed; ++q) {
            const MemberRec& m = recs[desc[j].mb_base + q];
            const uint64_t e = off[k] + m.out_off;
            if (e > bad_at) break;  // this block holds the failing back-reference
            if (m.trailer_byte == MB_BLOCK_END) {
                if (streaming ? e > sent + W : e >= sent + T + W) emit(e - W);
                continue;
            }
            if (e > sent) emit(e);
            const uint64_t len = e - mstart;
            const uint32_t got = std_crc(raw[ri++], len);
            // file offset of the trailer (pipeline.rs:1318); the streaming path reports 0 (lib.rs:329,338)
            const uint64_t tb = streaming ? 0 : m.trailer_byte - J.in_off;
            if (got != m.crc) {
                set_err(&ferr, RRG_E_CRC,
                        "CRC32 mismatch in member #%zu (uncompressed bytes: %llu, trailer at byte %llu): expected "
                        "%#010x, got %#010x",
                        mi, (unsigned long long)len, (unsigned long long)tb, m.crc, got);
                ferr.expected = m.crc;
                ferr.got = got;
                ferr.got64 = len;
                ferr.member = (uint32_t)mi;
                ferr.trailer_byte = tb;
                failed = true;
            } else if ((uint32_t)len != m.isize) {
                set_err(&ferr, RRG_E_ISIZE, "ISIZE mismatch in member #%zu (trailer at byte %llu): expected %u, got %llu",
                        mi, (unsigned long long)tb, m.isize, (unsigned long long)len);
                ferr.expected = m.isize;
                ferr.got64 = len;
                ferr.member = (uint32_t)mi;
                ferr.trailer_byte = tb;
                failed = true;
            }
            mstart = e;
            ++mi;
        }
        if (failed) break;
        if (bad_at != UINT64_MAX) {
            set_err(&ferr, RRG_E_INVALID, "invalid deflate stream: %s", detail_msg(DET_DIST_FAR));
            failed = true;
        } else if ((long)k == J.fatal) {
            const uint32_t st = r.status;
            if (st == ST_TRUNCATED) {
                // the member's final block was sent before its trailer was found cut off
                const uint64_t e = off[k] + r.n_cells;
                if (e > sent) emit(e);
            }
            // the streaming path reads a block again after more input, and at
            // the end of the input reports Truncated (lib.rs:200-205)
            if (streaming && (st == ST_EOF || st == ST_TRUNCATED)) set_err(&ferr, RRG_E_TRUNCATED, "truncated gzip stream");
            else status_err(&ferr, st, r.detail);
            failed = true;
        }
    }
    if (!failed && !J.saw_final) {
        set_err(&ferr, RRG_E_IO, "decode produced no final block");
        failed = true;
    }
    if (failed) {
        dec.failed = true;
        dec.error = ferr;
    }
}

// The first cell of each laid-out chain segment that a back-reference before
// the start of its member produced, where that is an error: on the serial
// paths always (fast_inflate.rs:1704-1715, inflate.rs:339-346); on the
// speculative path when the member began inside the current chunk, since the
// reference decodes such a member on its own (pipeline.rs:153-254) and only a
// chunk's first member may reach before the chunk.  A segment that starts
// inside a member decodes such a back-reference into markers; only segments
// that start less than 32 KiB after their member's start can hold one.
int member_reach(Ctx& c, const std::vector<Job>& jobs, const std::vector<SegDesc>& desc,
                 const std::vector<SegResult>& res, const std::vector<MemberRec>& recs, const std::vector<ChainSeg>& cs,
                 const std::vector<uint64_t>& first_len, std::vector<std::vector<uint64_t>>& bad, rrg_error* err) {
    constexpr uint64_t W = 32768;
    std::vector<ReachScan> items;
    std::vector<std::pair<size_t, size_t>> where;  // (job, chain position)
    for (size_t ji = 0; ji < jobs.size(); ++ji) {
        const Job& J = jobs[ji];
        if (J.body == UINT64_MAX) continue;
        bad[ji].assign(J.n_out, UINT64_MAX);
        const size_t nb = J.bg_hi - J.bg_lo;
        uint64_t mstart = J.off[nb];  // job-relative start of the current member
        uint64_t cstart = J.off[nb];  // job-relative start of the current chunk
        for (size_t k = 0; k < J.n_out; ++k) {
            const uint32_t j = J.order[k];
            const uint64_t o = J.off[nb + k];
            bool check = J.serial != 0;
            if (!J.serial) {
                if (J.seg_is_chunk[k]) cstart = o;
                check = !J.seg_is_chunk[k] && mstart > cstart;  // not the chunk's first member
            }
            if (check && (desc[j].flags & SEG_SPEC) && o - mstart < W && res[j].n_cells) {
                // a segment expanded in pieces: the markers of the later
                // pieces name cells of the segment itself
                const uint64_t len = first_len.empty() ? res[j].n_cells
                                                       : std::min<uint64_t>(res[j].n_cells, first_len[J.c_lo + nb + k]);
                items.push_back(ReachScan{cs[J.c_lo + nb + k].cells_base, (uint32_t)len, (uint32_t)(o - mstart)});
                where.push_back({ji, k});
            }
            for (uint32_t q = 0; q < res[j].n_mb; ++q) {
                const MemberRec& m = recs[desc[j].mb_base + q];
                if (m.trailer_byte != MB_BLOCK_END) mstart = o + m.out_off;
            }
        }
    }
    if (items.empty()) return RRG_OK;
    const size_t n = items.size();
    HIPCHK(c.reach.ensure(n * sizeof(ReachScan)));
    HIPCHK(c.first.ensure(n * sizeof(uint32_t)));
    std::vector<uint32_t> first(n, ~0u);
    HIPCHK(hipMemcpyAsync(c.reach.p, items.data(), n * sizeof(ReachScan), hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.first.p, first.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c.s));
    HIPCHK(launch_marker_reach(c.reach.as<ReachScan>(), (uint32_t)n, c.cells.as<uint16_t>(), c.first.as<uint32_t>(),
                               c.s));
    HIPCHK(hipMemcpyAsync(first.data(), c.first.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    for (size_t i = 0; i < n; ++i)
        if (first[i] != ~0u) bad[where[i].first][where[i].second] = first[i];
    return RRG_OK;
}

// One finder launch; every find must lie in [target, scan end) of its file.
// Exact targets may carry `stops` (their walk ends at the first window
// starting there) and report `exits` (the chain position where it ended) and,
// with want_recs, their first 256 chain positions (left on the device).
int run_finder(Ctx& c, const std::vector<uint64_t>& targets, const std::vector<uint64_t>& limits,
               const std::vector<uint64_t>& scan_ends, std::vector<uint64_t>& found, rrg_error* err,
               const std::vector<uint64_t>* stops, std::vector<uint64_t>* exits, bool want_recs) {
    found.assign(targets.size(), UINT64_MAX);
    if (targets.empty()) return RRG_OK;
    const size_t n = targets.size();
    HIPCHK(c.targets.ensure(n * 8));
    HIPCHK(c.limits.ensure(n * 8));
    HIPCHK(c.scan_ends.ensure(n * 8));
    HIPCHK(c.found.ensure(n * 8));
    if (int rc = stage_upload(c, c.targets.p, targets.data(), n * 8, err)) return rc;
    if (int rc = stage_upload(c, c.limits.p, limits.data(), n * 8, err)) return rc;
    if (int rc = stage_upload(c, c.scan_ends.p, scan_ends.data(), n * 8, err)) return rc;
    if (stops) {
        HIPCHK(c.stops.ensure(n * 8));
        if (int rc = stage_upload(c, c.stops.p, stops->data(), n * 8, err)) return rc;
    }
    if (exits) HIPCHK(c.exits.ensure(n * 8));
    if (want_recs) HIPCHK(c.recs.ensure(n * 32));
    std::vector<uint32_t> exact;  // targets that keep the reference's candidate order
    for (size_t i = 0; i < n; ++i)
        if (!(limits[i] & kAnyStart)) exact.push_back((uint32_t)i);
    if (!exact.empty()) {
        HIPCHK(c.exact_idx.ensure(exact.size() * 4));
        if (int rc = stage_upload(c, c.exact_idx.p, exact.data(), exact.size() * 4, err)) return rc;
    }
    StageTimer t(c, &c.prof.ms_find);
    HIPCHK(launch_find_blocks(c.in.as<uint64_t>(), (uint64_t)c.in_used * 8, c.targets.as<uint64_t>(),
                              c.limits.as<uint64_t>(), c.scan_ends.as<uint64_t>(),
                              stops ? c.stops.as<uint64_t>() : nullptr, c.found.as<uint64_t>(),
                              exits ? c.exits.as<uint64_t>() : nullptr, want_recs ? c.recs.as<uint64_t>() : nullptr,
                              (uint32_t)n, c.exact_idx.as<uint32_t>(), (uint32_t)exact.size(), c.s, c.s2, c.evf[0],
                              c.evf[1]));
    if (int rc = download(c, found.data(), c.found.p, n * 8, err)) return rc;
    if (exits) {
        exits->assign(n, 0);
        if (int rc = download(c, exits->data(), c.exits.p, n * 8, err)) return rc;
    }
    t.stop();
    HIPCHK(hipStreamSynchronize(c.s));
    finish_downloads(c);
    c.stg_used = 0;  // the stream has drained: no staged upload is pending
    if (c.verbose) {
        uint64_t longest = 0, none = 0;
        for (size_t i = 0; i < n; ++i) {
            const uint64_t e = found[i] != UINT64_MAX ? found[i] : exits ? (*exits)[i] : scan_ends[i];
            longest = std::max(longest, e - targets[i]);
            none += found[i] == UINT64_MAX;
        }
        std::fprintf(stderr, "[rrg-gpu] finder: %zu targets, %llu without a find, longest scan %llu KiB\n", n,
                     (unsigned long long)none, (unsigned long long)(longest / 8192));
    }
    // bit offsets scanned: from each target to its find, or to where its scan ended
    for (size_t i = 0; i < n; ++i) {
        const uint64_t e = found[i] != UINT64_MAX ? found[i] : exits ? std::min((*exits)[i], scan_ends[i]) : scan_ends[i];
        if (e > targets[i]) c.prof.find_bits += e - targets[i];
    }
    // a find is a bit offset inside the target's scan range; anything else is a device fault
    for (size_t i = 0; i < n; ++i)
        if (found[i] != UINT64_MAX && (found[i] < targets[i] || found[i] >= scan_ends[i])) {
            set_err(err, RRG_E_DEVICE, "block finder returned %llu for target %llu", (unsigned long long)found[i],
                    (unsigned long long)targets[i]);
            return RRG_E_DEVICE;
        }
    return RRG_OK;
}

// An exact target whose bounded first walk ended at chain position `p`
// without a find, before the end of its file (`limit`, in bits).
struct OpenTarget {
    size_t idx;  // index into the caller's found[]
    uint64_t p, limit;
    uint64_t span = 0;  // input (bits) the next round of legs covers
};

// Exact finds over long stretches without a verified candidate (stored or
// incompressible data can run for MiBs): the next part of each open target's
// range is cut into legs of kLegBits, all legs of all targets walked at once
// (one wavefront each, recording their first chain positions), and
// k_leg_fixup stitches each target's legs along its true chain.  The part
// doubles per round (1 MiB, 2 MiB, ... 64 MiB) until every target has its
// find or reached its file end: legs past a find are walked for nothing, and
// doubling bounds them by the distance walked before.
int resolve_open(Ctx& c, std::vector<OpenTarget> open, std::vector<uint64_t>& found, rrg_error* err) {
    constexpr uint64_t kLegBits = 8ull * 16384;
    constexpr uint64_t kFirstSpan = 8ull * (1ull << 20), kMaxSpan = 8ull * (64ull << 20);
    for (OpenTarget& o : open) o.span = kFirstSpan;
    while (!open.empty()) {
        std::vector<uint64_t> targets, limits, scan_ends, stops;
        std::vector<LegTarget> tg;
        for (const OpenTarget& o : open) {
            LegTarget T{};
            T.p = o.p;
            T.limit = o.limit;
            T.end = std::min(o.limit, o.p + o.span);
            T.first = (uint32_t)targets.size();
            for (uint64_t a = o.p; a < T.end; a += kLegBits) {
                targets.push_back(a);
                limits.push_back(o.limit);
                scan_ends.push_back(o.limit);
                stops.push_back(std::min(T.end, a + kLegBits));
            }
            T.n_legs = (uint32_t)(targets.size() - T.first);
            tg.push_back(T);
        }
        std::vector<uint64_t> lf, lx;
        int rc = run_finder(c, targets, limits, scan_ends, lf, err, &stops, &lx, true);
        if (rc) return rc;
        const size_t n = tg.size();
        HIPCHK(c.legtg.ensure(n * sizeof(LegTarget)));
        HIPCHK(c.lfound.ensure(n * 8));
        HIPCHK(c.lnext.ensure(n * 8));
        HIPCHK(hipMemcpyAsync(c.legtg.p, tg.data(), n * sizeof(LegTarget), hipMemcpyHostToDevice, c.s));
        std::vector<uint64_t> of(n), on(n);
        {
            StageTimer t(c, &c.prof.ms_find);
            HIPCHK(launch_leg_fixup(c.in.as<uint64_t>(), (uint64_t)c.in_used * 8, c.legtg.as<LegTarget>(),
                                    c.targets.as<uint64_t>(), c.found.as<uint64_t>(), c.exits.as<uint64_t>(),
                                    c.recs.as<uint64_t>(), c.lfound.as<uint64_t>(), c.lnext.as<uint64_t>(),
                                    (uint32_t)n, c.s));
            HIPCHK(hipMemcpyAsync(of.data(), c.lfound.p, n * 8, hipMemcpyDeviceToHost, c.s));
            HIPCHK(hipMemcpyAsync(on.data(), c.lnext.p, n * 8, hipMemcpyDeviceToHost, c.s));
            t.stop();
        }
        HIPCHK(hipStreamSynchronize(c.s));
        std::vector<OpenTarget> still;
        for (size_t i = 0; i < n; ++i) {
            const OpenTarget& o = open[i];
            if (of[i] != UINT64_MAX) {
                if (of[i] < o.p || of[i] >= o.limit) {
                    set_err(err, RRG_E_DEVICE, "leg fix-up returned %llu for %llu", (unsigned long long)of[i],
                            (unsigned long long)o.p);
                    return RRG_E_DEVICE;
                }
                found[o.idx] = of[i];
            } else if (on[i] + 13 <= o.limit && on[i] > o.p) {
                still.push_back(OpenTarget{o.idx, on[i], o.limit, std::min(kMaxSpan, 2 * o.span)});
            }
        }
        if (c.verbose)
            std::fprintf(stderr, "[rrg-gpu] finder legs: %zu targets, %zu legs, %zu still open\n", n, targets.size(),
                         still.size());
        open.swap(still);
    }
    return RRG_OK;
}

// Runs of a finder launch whose exact targets (limits without kAnyStart) had
// `stops`: continue those that ended open.
int finish_open(Ctx& c, const std::vector<uint64_t>& targets, const std::vector<uint64_t>& limits,
                const std::vector<uint64_t>& exits, std::vector<uint64_t>& found, rrg_error* err) {
    std::vector<OpenTarget> open;
    for (size_t i = 0; i < targets.size(); ++i)
        if (!(limits[i] & kAnyStart) && found[i] == UINT64_MAX && exits[i] > targets[i] && exits[i] + 13 <= limits[i])
            open.push_back(OpenTarget{i, exits[i], limits[i]});
    return resolve_open(c, open, found, err);
}

// Turn the re-decode requests of one settle round into segments to decode:
// each requested segment restarts at its verified start and, when it spans
// enough input, is cut at the blocks the finder reports inside it (one finder
// launch for all requests); the pieces join the job's chain after it.
int split_redo(Ctx& c, std::vector<Job>& jobs, std::vector<Redo>& redo, std::vector<SegDesc>& desc,
               const std::vector<SegResult>& res, uint64_t& tok_used, uint64_t& mb_used,
               std::vector<uint32_t>& again, rrg_error* err) {
    if (redo.empty()) return RRG_OK;
    constexpr uint64_t kPieceBits = 8 * 2048;  // at least 2 KiB of input per piece
    constexpr uint64_t kMaxPieces = 32;
    std::vector<uint64_t> targets, limits, scan_ends;
    std::vector<size_t> t_first(redo.size() + 1);
    for (size_t i = 0; i < redo.size(); ++i) {
        t_first[i] = targets.size();
        const Redo& R = redo[i];
        const Job& J = jobs[R.job];
        const SegDesc& d = desc[J.order[R.pos]];
        const uint64_t end = d.hint_bit == UINT64_MAX ? J.end_bits() : d.hint_bit;
        const uint64_t span = end > R.start ? end - R.start : 0;
        const uint64_t pieces = std::min<uint64_t>(kMaxPieces, span / kPieceBits);
        for (uint64_t p = 1; p < pieces; ++p) {
            targets.push_back(R.start + span * p / pieces);
            limits.push_back(J.end_bits() | kAnyStart);
            scan_ends.push_back(end);  // finds past the segment are not used
        }
    }
    t_first[redo.size()] = targets.size();
    std::vector<uint64_t> found;
    const int frc = run_finder(c, targets, limits, scan_ends, found, err);
    if (frc) return frc;
    // token slots granted beyond the default sizes in this round (a start
    // that is not verified yet may be a false one, whose counts mean nothing)
    uint64_t guessed = 0;
    constexpr uint64_t kGuessBudget = 1ull << 30;
    // later positions first, so earlier insertions do not move them
    std::vector<size_t> idx(redo.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        return redo[a].job != redo[b].job ? redo[a].job < redo[b].job : redo[a].pos > redo[b].pos;
    });
    for (size_t i : idx) {
        const Redo& R = redo[i];
        Job& J = jobs[R.job];
        const uint32_t j = J.order[R.pos];
        SegDesc& d = desc[j];
        const uint64_t hint = d.hint_bit;
        const uint64_t end = hint == UINT64_MAX ? J.end_bits() : hint;
        // a decode from a true start that ran past its hint passed no block
        // start in [hint, end_bit): cuts stay below the hint, and the last
        // piece's region covers the decode up to end_bit
        const uint64_t last_end = R.end_bit != UINT64_MAX && R.end_bit > end ? R.end_bit : end;
        // expected tokens per input byte in this range: what the segment
        // itself counted when it ran out of room, else what its chain
        // predecessor produced (pieces of a dense stretch would otherwise
        // overflow again)
        double ratio = 0;
        if (R.status == ST_OVERFLOW && R.end_bit != UINT64_MAX && R.end_bit > R.start + 8 * 1024) {
            ratio = (double)R.needed / (double)((R.end_bit - R.start) / 8);
        } else if (R.pos > 0) {
            const uint32_t q = J.order[R.pos - 1];
            if (res[q].status == ST_OK && res[q].end_bit == R.start && res[q].end_bit > desc[q].start_bit + 8 * 1024)
                ratio = (double)res[q].n_tok / (double)((res[q].end_bit - desc[q].start_bit) / 8);
        }
        // at most 4 tokens per input byte (a token takes at least 2 bits, a
        // literal triple at least 3); one piece's region at most kMaxGuess
        // tokens (beyond it the kernel counts and the piece is regrown)
        constexpr uint64_t kMaxGuess = 1ull << 26;
        ratio = std::min(ratio, 4.0);
        auto guess = [&](uint64_t span) -> uint64_t {
            const uint64_t g = std::min<uint64_t>(kMaxGuess, ((uint64_t)((double)span * ratio * 1.25) + 16384 + 3) & ~3ull);
            if (guessed + g > kGuessBudget) return 0;
            guessed += g;
            return g;
        };
        if (overflowed(R.status)) {
            grow_segment(d, R.status, tok_used, mb_used, R.needed);
            ++c.prof.n_grown;
        }
        d.start_bit = R.start;
        std::vector<uint64_t> cuts;
        for (size_t t = t_first[i]; t < t_first[i + 1]; ++t)
            if (found[t] != UINT64_MAX && found[t] > (cuts.empty() ? R.start : cuts.back()) && found[t] < end)
                cuts.push_back(found[t]);
        {
            // the segment's own region, for its (new) range
            const uint64_t span0 = ((cuts.empty() ? last_end : cuts[0]) - R.start) / 8;
            const uint64_t want = std::max<uint64_t>(guess(span0), ratio > 0 ? 0 : tok_slots(span0));
            if (!overflowed(R.status) && d.tok_cap < want) {
                d.tok_base = (tok_used + 3) & ~3ull;
                d.tok_cap = want;
                tok_used = d.tok_base + want;
            }
        }
        again.push_back(j);
        if (cuts.empty()) continue;
        d.hint_bit = cuts[0];
        std::vector<uint32_t> ins;
        for (size_t q = 0; q < cuts.size(); ++q) {
            SegDesc p{};
            p.start_bit = cuts[q];
            p.hint_bit = q + 1 < cuts.size() ? cuts[q + 1] : hint;
            p.end_byte = d.end_byte;
            const uint64_t span = ((q + 1 < cuts.size() ? cuts[q + 1] : last_end) - cuts[q]) / 8;
            p.tok_base = (tok_used + 3) & ~3ull;
            p.tok_cap = std::max<uint64_t>(tok_slots(span), ratio > 0 ? guess(span) : 0);
            tok_used = p.tok_base + p.tok_cap;
            p.flags = SEG_SPEC | (d.flags & SEG_BLOCK_ENDS);
            p.mb_base = (uint32_t)mb_used;
            p.mb_cap = mb_slots(span, p.flags & SEG_BLOCK_ENDS);
            mb_used += p.mb_cap;
            ins.push_back((uint32_t)desc.size());
            again.push_back((uint32_t)desc.size());
            desc.push_back(p);
        }
        const size_t at = R.pos + 1;
        J.order.insert(J.order.begin() + at, ins.begin(), ins.end());
        J.seg_is_chunk.insert(J.seg_is_chunk.begin() + at, ins.size(), 0);
        J.orig_start.insert(J.orig_start.begin() + at, cuts.begin(), cuts.end());
        J.first_ok.insert(J.first_ok.begin() + at, ins.size(), 1);
    }
    return RRG_OK;
}

// Decode every job of the batch (inputs already in c.in) into `out`.
int decode_batch(Ctx& c, const Geometry& G, std::vector<Job>& jobs, DevBuf& out, rrg_error* err) {
    // downloads queued by a call that failed before its wait are dropped
    c.dl_pend.clear();
    c.dl_used = 0;
    const uint64_t total_bits = (uint64_t)c.in_used * 8;
    // 1. host plan
    for (Job& J : jobs)
        if (!J.resume) plan_job(J);
    // 2. block finder over every member-path body
    // streaming chunking (streaming.rs:278-346): target = previous boundary +
    // chunk_bits, first verified candidate at or after it, until none is left
    for (Job& J : jobs) {
        if (J.body == UINT64_MAX || J.slab) continue;
        if (J.serial) {
            // one chunk: the single-worker paths cut chunks at block ends instead
            J.stream_B.assign(1, J.abs_bits(J.body));
            continue;
        }
        if (!J.stream) continue;
        const uint64_t chunk_bits = (uint64_t)std::max<size_t>(J.chunk_size, 65536) * 8;
        J.stream_B.assign(1, J.abs_bits(J.body));
        for (uint64_t t = J.stream_B.back() + chunk_bits; t < J.end_bits(); t = J.stream_B.back() + chunk_bits) {
            std::vector<uint64_t> f, x;
            const std::vector<uint64_t> tv{t}, lv{J.end_bits()}, sv{t + kExactCapBits};
            int frc = run_finder(c, tv, lv, lv, f, err, &sv, &x);
            if (!frc) frc = finish_open(c, tv, lv, x, f, err);
            if (frc) return frc;
            if (f[0] == UINT64_MAX) break;
            J.stream_B.push_back(f[0]);
        }
    }
    std::vector<uint64_t> targets, limits, scan_ends;
    std::vector<uint8_t> is_chunk;
    for (Job& J : jobs) {
        J.t_lo = targets.size();
        if (J.body != UINT64_MAX) make_targets(J, G, targets, limits, scan_ends, is_chunk);
        J.t_hi = targets.size();
    }
    // chunk targets walk at most kExactCapBits here and continue in legs
    std::vector<uint64_t> stops(targets.size()), exits;
    for (size_t i = 0; i < targets.size(); ++i) stops[i] = targets[i] + kExactCapBits;
    std::vector<uint64_t> found;
    int rc = run_finder(c, targets, limits, scan_ends, found, err, &stops, &exits);
    if (!rc) rc = finish_open(c, targets, limits, exits, found, err);
    if (rc) return rc;
    // 3. segment starts: chunk starts (the reference's, strictly increasing)
    // and every distinct segment find.  A segment cannot be shorter than a
    // deflate block, and one that spans two blocks (its target and the next
    // found the same block start) doubles the longest lane's work: with one
    // lane per segment, gaps well above the median get one more target, a
    // quarter of the way in.  The wavefront-per-segment geometry takes its
    // work longest first and does without that second finder pass.
    {
        std::vector<uint64_t> gt, gl, ge;
        std::vector<std::pair<size_t, size_t>> gat;  // (job, gap index)
        for (size_t ji = 0; ji < jobs.size(); ++ji) {
            Job& J = jobs[ji];
            if (J.body == UINT64_MAX) continue;
            std::vector<uint64_t> B{J.chain_start()};
            if (J.stream || J.serial || J.slab) B = J.stream_B;
            for (size_t i = J.t_lo; i < J.t_hi; ++i)
                if (is_chunk[i] && found[i] != UINT64_MAX && found[i] > B.back()) B.push_back(found[i]);
            // segment finds come in target order, which is almost always their
            // order too: merge them with the chunk starts (sort only when not)
            std::vector<uint64_t> F;
            F.reserve(J.t_hi - J.t_lo);
            for (size_t i = J.t_lo; i < J.t_hi; ++i)
                if (!is_chunk[i] && found[i] != UINT64_MAX) F.push_back(found[i]);
            if (!std::is_sorted(F.begin(), F.end())) std::sort(F.begin(), F.end());
            std::vector<uint64_t> M(B.size() + F.size());
            std::merge(B.begin(), B.end(), F.begin(), F.end(), M.begin());
            M.erase(std::unique(M.begin(), M.end()), M.end());
            J.stream_B = B;
            J.orig_start = M;
            if (M.size() < 16 || G.g.wave) continue;
            std::vector<uint64_t> span(M.size() - 1);
            for (size_t j = 0; j + 1 < M.size(); ++j) span[j] = M[j + 1] - M[j];
            std::nth_element(span.begin(), span.begin() + (long)(span.size() / 2), span.end());
            const uint64_t med = span[span.size() / 2];
            for (size_t j = 0; j < M.size(); ++j) {
                // the last segment runs to the end of the file
                const uint64_t nx = j + 1 < M.size() ? M[j + 1] : J.last_end();
                const uint64_t g = nx - M[j];
                if (g > med + med / 2 && g > 8 * 16384) {
                    // a block start hidden in the gap lies about halfway: search from a quarter on
                    gt.push_back(M[j] + g / 4);
                    gl.push_back(J.end_bits() | kAnyStart);
                    ge.push_back(nx);
                    gat.push_back({ji, j});
                }
            }
        }
        if (!gt.empty()) {
            std::vector<uint64_t> gf;
            rc = run_finder(c, gt, gl, ge, gf, err);
            if (rc) return rc;
            std::vector<std::vector<uint64_t>> extra(jobs.size());
            for (size_t i = 0; i < gt.size(); ++i)
                if (gf[i] != UINT64_MAX) extra[gat[i].first].push_back(gf[i]);
            for (size_t ji = 0; ji < jobs.size(); ++ji) {
                if (extra[ji].empty()) continue;
                std::vector<uint64_t>& M = jobs[ji].orig_start;
                std::vector<uint64_t> N(M.size() + extra[ji].size());
                std::sort(extra[ji].begin(), extra[ji].end());
                std::merge(M.begin(), M.end(), extra[ji].begin(), extra[ji].end(), N.begin());
                N.erase(std::unique(N.begin(), N.end()), N.end());
                M.swap(N);
            }
        }
    }
    // 4. segment descriptors
    std::vector<SegDesc> desc;
    uint64_t tok_used = 0, mb_used = 0;
    for (Job& J : jobs) {
        // BGZF members: one single-member segment each, output sized from ISIZE
        J.bg_lo = desc.size();
        J.hdr_err.assign(J.mem.size(), 0);
        for (size_t j = 0; j < J.mem.size(); ++j) {
            const uint8_t* p = J.h + J.mem[j].off;
            const uint64_t sz = J.mem[j].size;
            const long hl = parse_header(p, sz, nullptr);
            uint64_t isz = 0;
            if (sz >= 18)
                isz = (uint64_t)p[sz - 4] | ((uint64_t)p[sz - 3] << 8) | ((uint64_t)p[sz - 2] << 16) |
                      ((uint64_t)p[sz - 1] << 24);
            SegDesc d{};
            if (hl < 0 || sz < 18) {
                J.hdr_err[j] = 1;
                d.start_bit = d.hint_bit = 0;  // empty decode
                d.end_byte = 0;
            } else {
                d.start_bit = J.abs_bits(J.mem[j].off + (uint64_t)hl);
                d.hint_bit = UINT64_MAX;
                d.end_byte = J.in_off + J.len;
            }
            // a member's tokens stand for at least one byte each and take at
            // least two bits each: at most ISIZE and 4 per input byte
            d.tok_base = tok_used;
            d.tok_cap = (std::min<uint64_t>(isz, 4 * sz) + 64 + 3) & ~3ull;
            tok_used += d.tok_cap;
            d.flags = SEG_SINGLE_MEMBER;
            d.mb_base = (uint32_t)mb_used;
            d.mb_cap = 1;
            mb_used += 1;
            desc.push_back(d);
        }
        J.bg_hi = desc.size();
        // member path: chunk boundaries exactly as the reference (strictly
        // increasing finds), segments at every distinct find
        J.mp_lo = J.mp_hi = desc.size();
        if (J.body == UINT64_MAX) continue;
        std::vector<uint64_t>& M = J.orig_start;  // the segment starts (chunk starts marked below)
        std::vector<uint64_t>& B = J.stream_B;    // the chunk starts
        const size_t S = M.size();
        J.seg_is_chunk.assign(S, 0);
        for (size_t j = 0, bi = 0; j < S && bi < B.size(); ++j)
            if (M[j] == B[bi]) {
                J.seg_is_chunk[j] = 1;
                ++bi;
            }
        for (size_t j = 0; j < S; ++j) {
            SegDesc d{};
            d.start_bit = M[j];
            d.hint_bit = j + 1 < S ? M[j + 1] : J.stop_bit;
            d.end_byte = J.in_off + J.len;
            const uint64_t span = ((j + 1 < S ? M[j + 1] : J.last_end()) - M[j]) / 8;
            d.tok_base = tok_used;
            d.tok_cap = tok_slots(span);
            tok_used += d.tok_cap;
            d.flags = (j || J.resume ? SEG_SPEC : 0u) | (J.serial ? SEG_BLOCK_ENDS : 0u);
            d.mb_base = (uint32_t)mb_used;
            d.mb_cap = mb_slots(span, J.serial);
            mb_used += d.mb_cap;
            desc.push_back(d);
        }
        J.mp_hi = desc.size();
        J.order.resize(S);
        for (size_t j = 0; j < S; ++j) J.order[j] = (uint32_t)(J.mp_lo + j);
    }
    const size_t S = desc.size();
    // 4. one inflate launch over every segment of the batch
    std::vector<SegResult> res;
    std::vector<uint32_t> all(S);
    for (size_t j = 0; j < S; ++j) all[j] = (uint32_t)j;
    if (G.g.wave) {
        // a wave takes segments one after another: the longest first, so the
        // launch does not end waiting on a long one started last (spans in
        // 2 KiB buckets, largest bucket first, stable within a bucket)
        auto span_of = [&](uint32_t j) {
            const uint64_t end = desc[j].hint_bit != UINT64_MAX ? desc[j].hint_bit : desc[j].end_byte * 8;
            return end > desc[j].start_bit ? (end - desc[j].start_bit) >> 14 : 0;
        };
        uint64_t top = 0;
        for (uint32_t j : all) top = std::max<uint64_t>(top, span_of(j));
        top = std::min<uint64_t>(top, 4095);
        std::vector<uint32_t> cnt(top + 2, 0), sorted(S);
        for (uint32_t j : all) ++cnt[top - std::min<uint64_t>(span_of(j), top)];
        for (size_t b = 1; b < cnt.size(); ++b) cnt[b] += cnt[b - 1];
        for (size_t k = S; k-- > 0;) {
            const uint32_t j = all[k];
            sorted[--cnt[top - std::min<uint64_t>(span_of(j), top)]] = j;
        }
        all.swap(sorted);
    }
    std::vector<MemberRec> recs;  // member records, from the last inflate launch
    rc = run_inflate(c, G, desc, res, tok_used, mb_used, all, err, &recs);
    if (rc) return rc;
    for (Job& J : jobs) {
        // a capacity overflow is not a decode failure (the region is grown below)
        J.first_ok.resize(J.order.size());
        for (size_t k = 0; k < J.order.size(); ++k)
            J.first_ok[k] = res[J.order[k]].status == ST_OK || overflowed(res[J.order[k]].status);
    }
    // 5. settle: chain walks and capacity retries, re-decodes batched per round.
    // A member-path segment decoded again is split at block boundaries found
    // inside it, so the extra round costs a fraction of a segment's decode.
    for (int round = 0;; ++round) {
        std::vector<uint32_t> again;
        std::vector<Redo> redo;
        for (size_t ji = 0; ji < jobs.size(); ++ji) {
            Job& J = jobs[ji];
            if (round < 6)
                for (size_t j = J.bg_lo; j < J.bg_hi; ++j)
                    if (overflowed(res[j].status)) {
                        grow_segment(desc[j], res[j].status, tok_used, mb_used, res[j].n_tok);
                        ++c.prof.n_grown;
                        again.push_back((uint32_t)j);
                    }
            if (J.body != UINT64_MAX) chain_round(J, ji, desc, res, redo);
        }
        if (again.empty() && redo.empty()) break;
        if (round > 4 * (int)desc.size() + 64) {
            set_err(err, RRG_E_DEVICE, "verification chain did not settle");
            return RRG_E_DEVICE;
        }
        if (c.verbose) {
            size_t n_false = 0, n_over = 0;
            for (const Redo& R : redo) (R.status == ST_OK ? n_false : n_over) += 1;
            std::fprintf(stderr, "[rrg-gpu] settle round %d: %zu false starts, %zu out of room, %zu regrown alone\n",
                         round, n_false, n_over, again.size());
        }
        rc = split_redo(c, jobs, redo, desc, res, tok_used, mb_used, again, err);
        if (rc) return rc;
        c.prof.n_redecoded += again.size();
        rc = run_inflate(c, G, desc, res, tok_used, mb_used, again, err, &recs);
        if (rc) return rc;
    }
    c.prof.n_segments += desc.size();
    // the serial paths also send output of the failing segment (the blocks
    // before the failure, or a member whose trailer is cut off)
    for (Job& J : jobs) J.n_out = J.n_chain + (J.serial && J.fatal >= 0 ? 1 : 0);
    // 6. output layout (each file at a 256-byte aligned offset), the chain and
    // the expansion.  Chain segments get exact cell regions in chain order
    // (64-cell aligned: the expansion stores whole lines); BGZF members hold
    // no markers and expand straight into the output bytes.
    std::vector<ChainSeg> cs;
    std::vector<uint64_t> cs_out;
    std::vector<ExpandItem> items;
    std::vector<size_t> item_cs;  // chain segment of each item (SIZE_MAX: a BGZF member)
    uint64_t out_total = 0, cells_total = 0;
    // the tokens of a segment that overflowed its region are not all stored
    auto expandable = [&](size_t j) -> uint32_t {
        return overflowed(res[j].status) ? 0u : (uint32_t)std::min<uint64_t>(res[j].n_tok, desc[j].tok_cap);
    };
    struct Widen {
        const uint8_t* src;
        uint64_t cell, n;
    };
    std::vector<Widen> windows;
    for (Job& J : jobs) {
        const bool win = J.resume && J.win_len && J.body != UINT64_MAX;
        if (win) {
            // the output before a resumed slab: a marker-free segment just
            // before the job's output, which the tail chain carries into the
            // slab's first segment
            out_total = ((out_total + J.win_len + 255) & ~255ull) - J.win_len;
            cs.push_back(ChainSeg{cells_total, J.win_len, 0, 0u, 1u});
            cs_out.push_back(out_total);
            windows.push_back(Widen{J.win_src, cells_total, J.win_len});
            cells_total = (cells_total + J.win_len + 63) & ~63ull;
            out_total += J.win_len;
        }
        J.dec.out_base = out_total;
        J.c_lo = cs.size();
        J.off.clear();
        uint64_t o = 0;
        for (size_t j = J.bg_lo; j < J.bg_hi; ++j) {
            J.off.push_back(o);
            cs.push_back(ChainSeg{0, 0, 0, 0u, 1u});
            cs_out.push_back(out_total + o);
            if (res[j].n_cells) {
                items.push_back(ExpandItem{desc[j].tok_base, out_total + o, desc[j].start_bit >> 3, res[j].n_cells,
                                           expandable(j), EXP_U8});
                item_cs.push_back(SIZE_MAX);
            }
            o += res[j].n_cells;
        }
        for (size_t k = 0; k < J.n_out; ++k) {
            const size_t j = J.order[k];
            J.off.push_back(o);
            const uint64_t base = cells_total;
            cells_total = (cells_total + res[j].n_cells + 63) & ~63ull;
            cs.push_back(ChainSeg{base, res[j].n_cells, 0, (desc[j].flags & SEG_SPEC) ? 1u : 0u,
                                  k == 0 && !win ? 1u : 0u});
            cs_out.push_back(out_total + o);
            if (res[j].n_cells) {
                items.push_back(ExpandItem{desc[j].tok_base, base, desc[j].start_bit >> 3, res[j].n_cells,
                                           expandable(j), 0u});
                item_cs.push_back(cs.size() - 1);
            }
            o += res[j].n_cells;
        }
        J.off.push_back(o);
        J.U = o;
        out_total = (out_total + o + 255) & ~255ull;
    }
    HIPCHK(out.ensure(out_total + 64));
    HIPCHK(c.cells.ensure((cells_total + 64) * sizeof(uint16_t)));
    // long items expand in pieces, each a chain segment of its own (the
    // member checks below still see the original chain)
    std::vector<ChainSeg> cs_orig;  // the chain before the cut (empty: nothing was cut)
    std::vector<uint32_t> cs_of;
    std::vector<uint64_t> first_len;
    rc = split_long_items(c, items, item_cs, cs, cs_out, cs_orig, cs_of, first_len, err);
    if (rc) return rc;
    // the host lays out the tail chain, resolve tiles and CRC pieces while
    // the expansion runs
    rc = run_expand(c, items, out, err);
    if (rc) return rc;
    for (const Widen& w : windows)
        HIPCHK(launch_widen(w.src, c.cells.as<uint16_t>() + w.cell, w.n, c.s));
    // back-references before their member's start (read before the tail pass
    // rewrites markers)
    std::vector<std::vector<uint64_t>> bad(jobs.size());
    rc = member_reach(c, jobs, desc, res, recs, cs_orig.empty() ? cs : cs_orig, first_len, bad, err);
    if (rc) return rc;
    // 7. tail chain, resolve and the CRC of every member of every file (BGZF
    // members, then member-path members)
    std::vector<std::pair<uint64_t, uint64_t>> ranges;
    for (Job& J : jobs) {
        J.r_lo = ranges.size();
        const uint64_t ob = J.dec.out_base;
        const size_t nb = J.bg_hi - J.bg_lo;
        for (size_t k = 0; k < nb; ++k) ranges.push_back({ob + J.off[k], ob + J.off[k + 1]});
        uint64_t mstart = J.off[nb];
        for (size_t k = 0; k < J.n_out; ++k) {
            const size_t j = J.order[k];
            for (uint32_t q = 0; q < res[j].n_mb; ++q) {
                if (recs[desc[j].mb_base + q].trailer_byte == MB_BLOCK_END) continue;
                const uint64_t e = J.off[nb + k] + recs[desc[j].mb_base + q].out_off;
                ranges.push_back({ob + mstart, ob + e});
                mstart = e;
            }
        }
        J.open_r = SIZE_MAX;
        if (J.slab && J.body != UINT64_MAX) {
            // the output of the member still open at the end of the slab
            J.open_r = ranges.size();
            ranges.push_back({ob + mstart, ob + J.off[nb + J.n_out]});
        }
    }
    std::vector<uint32_t> flags, raw;
    rc = resolve_and_crc(c, cs, cs_out, out, ranges, flags, raw, err);
    if (rc) return rc;
    if (!cs_of.empty()) {
        // per original chain segment
        std::vector<uint32_t> f(cs_orig.size(), 0);
        for (size_t k = 0; k < cs_of.size(); ++k) f[cs_of[k]] |= flags[k];
        flags.swap(f);
    }
    // 8. per-file verdicts and chunk boundaries
    for (Job& J : jobs) {
        Decoded& dec = J.dec;
        const size_t nb = J.bg_hi - J.bg_lo;
        if (J.bgzf) {
            // first failing member (libdeflate_gzip_decompress semantics: header, deflate, CRC, ISIZE)
            size_t bad = nb;
            for (size_t k = 0; k < nb; ++k) {
                const size_t j = J.bg_lo + k;
                const SegResult& r = res[j];
                bool ok = !J.hdr_err[k] && r.status == ST_OK && r.n_mb == 1;
                if (ok) {
                    const MemberRec& m = recs[desc[j].mb_base];
                    ok = std_crc(raw[J.r_lo + k], r.n_cells) == m.crc && (uint32_t)r.n_cells == m.isize &&
                         m.trailer_byte + 8 <= J.in_off + J.mem[k].off + J.mem[k].size;
                }
                if (!ok) {
                    bad = k;
                    break;
                }
            }
            // batches ramp 64 KiB -> chunk size (pipeline.rs:1488-1551)
            const uint64_t tmax = std::max<size_t>(J.chunk_size, 65536);
            uint64_t target = J.bgzf_target0, cur = 0;
            size_t first = 0;
            for (size_t k = 0; k < nb; ++k) {
                cur += J.mem[k].size;
                if (cur >= target || k + 1 == nb) {
                    if (bad >= first && bad <= k) break;  // this batch holds the failing member
                    dec.chunk_ends.push_back(J.off[k + 1]);
                    if (cur >= target) target = std::min(target * 2, tmax);
                    cur = 0;
                    first = k + 1;
                }
            }
            J.bgzf_target_end = target;
            if (bad < nb) {
                dec.failed = true;
                set_err(&dec.error, RRG_E_INVALID, "invalid deflate stream: libdeflate gzip decode failed");
                continue;
            }
            if (J.walk_trunc) {
                dec.failed = true;
                set_err(&dec.error, RRG_E_TRUNCATED, "truncated gzip stream");
                continue;
            }
        }
        if (J.pre_failed) {
            dec.failed = true;
            dec.error = J.pre_err;
            continue;
        }
        if (J.body == UINT64_MAX) continue;
        std::vector<uint64_t> off_mp(J.off.begin() + nb, J.off.end());
        for (uint64_t& v : off_mp) v -= J.off[nb];
        if (J.serial) {
            finish_serial(J, desc, res, recs, raw, J.r_lo + nb, J.off[nb], off_mp, bad[&J - jobs.data()]);
            continue;
        }
        finish_member_path(J, desc, res, recs, flags, raw, J.r_lo + nb, J.off[nb], off_mp, bad[&J - jobs.data()]);
        if (J.stream) dec.spec_failures = 0;  // the streaming path counts none (streaming.rs:389-395)
    }
    return RRG_OK;
}

// Decode the single file already in c.in (and at host pointer h).
int decode_loaded(Ctx& c, const uint8_t* h, uint64_t len, const rrg_cfg& cfg, Decoded& dec, rrg_error* err,
                  bool stream = false) {
    const Geometry G = pick_geometry(c, cfg);
    std::vector<Job> jobs(1);
    jobs[0].in_off = 0;
    jobs[0].len = len;
    jobs[0].h = h;
    jobs[0].stream = stream;
    jobs[0].serial = cfg.num_threads == 1 ? (stream ? 2 : 1) : 0;
    jobs[0].chunk_size = cfg.chunk_size_bytes;
    jobs[0].segment_bytes = cfg.segment_bytes;
    const int rc = decode_batch(c, G, jobs, c.out, err);
    if (rc) return rc;
    dec = jobs[0].dec;
    return RRG_OK;
}

int load_host(Ctx& c, const uint8_t* file, size_t len, rrg_error* err) {
    HIPCHK(c.in.ensure(len + 64));
    HIPCHK(hipMemcpyAsync(c.in.p, file, len, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemsetAsync(c.in.as<uint8_t>() + len, 0, 64, c.s));
    c.in_used = len + 64;
    return RRG_OK;
}

void begin_profile(Ctx& c, uint64_t len) {
    std::memset(&c.prof, 0, sizeof c.prof);
    c.x_timed = false;
    c.verbose = false;
    c.prof.compressed_bytes = len;
}

void print_profile(const Ctx& c) {
    const rrg_profile& p = c.prof;
    std::fprintf(stderr,
                 "[rrg-gpu] %llu -> %llu bytes in %.3f ms | find %.3f inflate %.3f (x%llu, %llu re-decoded, %llu "
                 "grown) expand %.3f (%llu tokens) tail %.3f (%llu blocks) resolve %.3f crc %.3f | %llu segments, "
                 "%llu slabs, first chunk at %.3f ms, %.1f MiB of device memory\n",
                 (unsigned long long)p.compressed_bytes, (unsigned long long)p.uncompressed_bytes, p.ms_total,
                 p.ms_find, p.ms_inflate, (unsigned long long)p.n_inflate_launches, (unsigned long long)p.n_redecoded,
                 (unsigned long long)p.n_grown, p.ms_expand, (unsigned long long)p.n_tokens, p.ms_tail,
                 (unsigned long long)p.n_tail_rounds, p.ms_resolve, p.ms_crc, (unsigned long long)p.n_segments,
                 (unsigned long long)p.n_slabs, p.ms_first_chunk, (double)p.device_bytes_peak / 1048576.0);
}

int decode_device_impl(Ctx& c, const void* d_file, size_t len, const uint8_t* h_file, const rrg_cfg* cfg_in,
                       uint64_t* out_len, uint64_t* chunk_ends, size_t chunk_cap, size_t* n_chunks, rrg_stats* stats,
                       rrg_error* err) {
    HIPCHK(hipSetDevice(c.device));
    rrg_cfg cfg;
    if (cfg_in) cfg = *cfg_in;
    else rrg_cfg_default(&cfg);
    const double t0 = now_ms();
    begin_profile(c, len);
    c.verbose = cfg.verbose != 0;
    if (len == 0) {
        set_err(err, RRG_E_IO, "empty input");
        return RRG_E_IO;
    }
    HIPCHK(c.in.ensure(len + 64));
    HIPCHK(hipMemcpyAsync(c.in.p, d_file, len, hipMemcpyDeviceToDevice, c.s));
    HIPCHK(hipMemsetAsync(c.in.as<uint8_t>() + len, 0, 64, c.s));
    c.in_used = len + 64;
    std::vector<uint8_t> hcopy;
    const uint8_t* h = h_file;
    if (!h) {
        hcopy.resize(len);
        HIPCHK(hipMemcpyAsync(hcopy.data(), d_file, len, hipMemcpyDeviceToHost, c.s));
        HIPCHK(hipStreamSynchronize(c.s));
        h = hcopy.data();
    }
    Decoded dec;
    int rc = decode_loaded(c, h, len, cfg, dec, err);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c.s));
    const uint64_t total = dec.chunk_ends.empty() ? 0 : dec.chunk_ends.back();
    c.prof.uncompressed_bytes = total;
    c.prof.ms_total = now_ms() - t0;
    if (cfg.verbose) print_profile(c);
    if (out_len) *out_len = total;
    if (n_chunks) *n_chunks = dec.chunk_ends.size();
    if (chunk_ends)
        for (size_t i = 0; i < dec.chunk_ends.size() && i < chunk_cap; ++i) chunk_ends[i] = dec.chunk_ends[i];
    if (stats) {
        stats->compressed_bytes = len;
        stats->uncompressed_bytes = total;
        stats->chunks_decoded = dec.chunk_ends.size();
        stats->speculation_failures = dec.spec_failures;
    }
    if (dec.failed) {
        if (err) *err = dec.error;
        return dec.error.kind ? dec.error.kind : RRG_E_INVALID;
    }
    return RRG_OK;
}

bool init_ctx(Ctx& c, int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return false;
    if (hipSetDevice(device) != hipSuccess) return false;
    c.device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || hipStreamCreate(&c.s) != hipSuccess ||
        hipEventCreate(&c.ev0) != hipSuccess || hipEventCreate(&c.ev1) != hipSuccess ||
        hipEventCreate(&c.evx[0]) != hipSuccess || hipEventCreate(&c.evx[1]) != hipSuccess ||
        hipEventCreateWithFlags(&c.evc[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.evc[1], hipEventDisableTiming) != hipSuccess || hipStreamCreate(&c.s2) != hipSuccess ||
        hipEventCreateWithFlags(&c.evf[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.evf[1], hipEventDisableTiming) != hipSuccess)
        return false;
    c.n_cu = prop.multiProcessorCount;
    return true;
}

}  // namespace host
}  // namespace rrg

using namespace rrg;
using namespace rrg::host;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int rrg_gpu_abi_version(void) { return RRG_GPU_ABI_VERSION; }

void rrg_gpu_debug_fail_alloc(int n) { g_fail_alloc.store(n > 0 ? n : 0); }

void rrg_cfg_default(rrg_cfg* cfg) {
    std::memset(cfg, 0, sizeof *cfg);
    cfg->chunk_size_bytes = 4u << 20;
    cfg->lut_bits = 0;
    cfg->lanes_per_block = 64;
    cfg->launch_graph = 0;
}

void* rrg_gpu_create(int device) {
    Ctx* c = new Ctx();
    if (!init_ctx(*c, device)) {
        delete c;
        return nullptr;
    }
    return c;
}

void rrg_gpu_destroy(void* ctx) {
    if (!ctx) return;
    Ctx* c = static_cast<Ctx*>(ctx);
    (void)hipSetDevice(c->device);
    delete c;
}

void rrg_gpu_last_profile(void* ctx, rrg_profile* out) {
    if (ctx && out) *out = static_cast<Ctx*>(ctx)->prof;
}

int rrg_gpu_decode_device(void* ctx, const void* d_file, size_t len, const uint8_t* h_file, const rrg_cfg* cfg_in,
                          const uint8_t** d_out, uint64_t* out_len, uint64_t* chunk_ends, size_t chunk_cap,
                          size_t* n_chunks, rrg_stats* stats, rrg_error* err) {
    if (err) std::memset(err, 0, sizeof *err);
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (!ctx) {
        set_err(err, RRG_E_ARG, "null context");
        return RRG_E_ARG;
    }
    Ctx& c = *static_cast<Ctx*>(ctx);
    const int rc = decode_device_impl(c, d_file, len, h_file, cfg_in, out_len, chunk_ends, chunk_cap, n_chunks,
                                      stats, err);
    if (d_out) *d_out = c.out.as<uint8_t>();
    return rc;
}

int rrg_gpu_autotune(void* ctx, const void* d_file, size_t len, const uint8_t* h_file, const rrg_cfg* base, int reps,
                     rrg_cfg* best, double* best_ms, rrg_error* err) {
    if (err) std::memset(err, 0, sizeof *err);
    if (!ctx || !best) {
        set_err(err, RRG_E_ARG, "null context or result");
        return RRG_E_ARG;
    }
    rrg_cfg b;
    if (base) b = *base;
    else rrg_cfg_default(&b);
    if (reps < 1) reps = 1;
    // decomposition (lanes per workgroup, literal LUT width) x segment size (0 = automatic)
    static const int kGeom[][2] = {{16, 8}, {16, 9}, {64, 9}, {64, 10}, {64, 11}};
    static const size_t kSeg[] = {0, 40960, 57344};
    std::vector<uint64_t> ref_ends, ends(1u << 20);
    double best_t = 1e300;
    bool first = true;
    for (const auto& geo : kGeom)
        for (size_t seg : kSeg) {
            const int lanes = geo[0], lut = geo[1];
            rrg_cfg cfg = b;
            cfg.lanes_per_block = lanes;
            cfg.lut_bits = lut;
            cfg.segment_bytes = seg;
            cfg.verbose = 0;
            double t = 1e300;
            for (int r = 0; r <= reps; ++r) {
                const uint8_t* d_out = nullptr;
                uint64_t out_len = 0;
                size_t n = 0;
                const double t0 = now_ms();
                const int rc = rrg_gpu_decode_device(ctx, d_file, len, h_file, &cfg, &d_out, &out_len, ends.data(),
                                                     ends.size(), &n, nullptr, err);
                const double dt = now_ms() - t0;
                if (rc) return rc;  // the sample must decode
                if (n > ends.size()) {
                    set_err(err, RRG_E_ARG, "autotune: more than %zu chunks", ends.size());
                    return RRG_E_ARG;
                }
                const std::vector<uint64_t> got(ends.begin(), ends.begin() + (long)n);
                if (first) {
                    ref_ends = got;
                    first = false;
                } else if (got != ref_ends) {
                    set_err(err, RRG_E_DEVICE, "autotune: chunk boundaries differ between configurations");
                    return RRG_E_DEVICE;
                }
                if (r) t = std::min(t, dt);  // the first run of each configuration warms up
            }
            if (t < best_t) {
                best_t = t;
                *best = cfg;
            }
        }
    best->verbose = b.verbose;
    if (best_ms) *best_ms = best_t;
    return RRG_OK;
}

static int decode_host(void* ctx, const uint8_t* file, size_t len, const rrg_cfg* cfg_in, rrg_chunk_cb cb, void* user,
                       rrg_stats* stats, rrg_error* err, bool stream) {
    if (err) std::memset(err, 0, sizeof *err);
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (!ctx) {
        set_err(err, RRG_E_ARG, "null context");
        return RRG_E_ARG;
    }
    Ctx& c = *static_cast<Ctx*>(ctx);
    HIPCHK(hipSetDevice(c.device));
    rrg_cfg cfg;
    if (cfg_in) cfg = *cfg_in;
    else rrg_cfg_default(&cfg);
    const double t0 = now_ms();
    begin_profile(c, len);
    c.verbose = cfg.verbose != 0;
    if (len == 0) {
        // the streaming path reads no gzip header (streaming.rs:174-196)
        if (stream) set_err(err, RRG_E_TRUNCATED, "truncated gzip stream");
        else set_err(err, RRG_E_IO, "empty input");
        return err ? err->kind : (stream ? RRG_E_TRUNCATED : RRG_E_IO);
    }
    if (!stream && cfg.num_threads != 1) {
        // the speculative chunking: in slabs, chunks handed out as they are ready
        const int src = decode_slabbed(c, file, len, cfg, cb, user, stats, err);
        if (src != kNotSlabbed) {
            c.prof.ms_total = now_ms() - t0;
            if (cfg.verbose) print_profile(c);
            return src;
        }
    }
    int rc = load_host(c, file, len, err);
    if (rc) return rc;
    Decoded dec;
    rc = decode_loaded(c, file, len, cfg, dec, err, stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c.s));
    // in-order emission (pipeline.rs:921-937)
    uint64_t prev = 0, sent = 0, total = 0;
    uint64_t biggest = 0;
    for (uint64_t e : dec.chunk_ends) {
        biggest = std::max(biggest, e - prev);
        prev = e;
    }
    prev = 0;
    // two pinned halves: the copy of chunk k+1 runs while the callback takes chunk k
    const size_t half = (biggest + 64 + 255) & ~(size_t)255;
    if (2 * half > c.pin_cap) {
        if (c.pin) (void)hipHostFree(c.pin);
        c.pin = nullptr;
        c.pin_cap = 0;
        HIPCHK(hipHostMalloc((void**)&c.pin, 2 * half, hipHostMallocDefault));
        c.pin_cap = 2 * half;
    }
    uint8_t* const buf[2] = {c.pin, c.pin + half};
    // a closed sink: the speculative path returns Ok counting the refused
    // chunk's bytes (pipeline.rs:1028-1030), the serial path without them
    // (pipeline.rs:1264-1266); the streaming serial path ignores it and
    // decodes on, counting every chunk (lib.rs:296-298)
    const bool serial = cfg.num_threads == 1;
    const bool stream_serial = stream && serial;
    const std::vector<uint64_t>& E = dec.chunk_ends;
    auto start_of = [&](size_t k) -> uint64_t { return k ? E[k - 1] : 0; };
    auto issue = [&](size_t k) -> hipError_t {
        const uint64_t a = start_of(k), n = E[k] - a;
        hipError_t e = hipSuccess;
        if (n) e = hipMemcpyAsync(buf[k & 1], c.out.as<uint8_t>() + a, n, hipMemcpyDeviceToHost, c.s);
        if (e == hipSuccess) e = hipEventRecord(c.evc[k & 1], c.s);
        return e;
    };
    bool closed = false;
    if (!E.empty()) HIPCHK(issue(0));
    for (size_t k = 0; k < E.size(); ++k) {
        const uint64_t n = E[k] - start_of(k);
        if (!closed) {
            if (k + 1 < E.size()) HIPCHK(issue(k + 1));
            HIPCHK(hipEventSynchronize(c.evc[k & 1]));
            if (cb && cb(user, buf[k & 1], n)) {
                closed = true;
                if (!stream_serial) {
                    if (!serial) total += n;
                    break;
                }
            }
        }
        total += n;
        ++sent;
    }
    HIPCHK(hipStreamSynchronize(c.s));  // no copy into the staging buffer outlives the call
    c.prof.uncompressed_bytes = total;
    c.prof.ms_total = now_ms() - t0;
    if (cfg.verbose) print_profile(c);
    if (stats) {
        stats->compressed_bytes = stream_serial ? 0 : len;  // lib.rs:402-407
        stats->uncompressed_bytes = total;
        stats->chunks_decoded = sent;
        stats->speculation_failures = dec.spec_failures;
    }
    if (closed && !stream_serial) {
        // speculative path: Ok with partial stats; BGZF: BrokenPipe (pipeline.rs:1028-1030, 1575-1578)
        if (!stream && bgzf_block_size(file, len)) {
            set_err(err, RRG_E_IO, "sink closed");
            return RRG_E_IO;
        }
        return RRG_OK;
    }
    if (dec.failed) {
        if (err) *err = dec.error;
        return dec.error.kind ? dec.error.kind : RRG_E_INVALID;
    }
    return RRG_OK;
}


int rrg_gpu_decode(void* ctx, const uint8_t* file, size_t len, const rrg_cfg* cfg_in, rrg_chunk_cb cb, void* user,
                   rrg_stats* stats, rrg_error* err) {
    return decode_host(ctx, file, len, cfg_in, cb, user, stats, err, false);
}

// Everything readable from fd, after the bytes already in buf.
static int read_all(int fd, std::vector<uint8_t>& buf, rrg_error* err) {
    size_t n = buf.size();
    for (;;) {
        if (buf.size() < n + (1u << 20)) buf.resize(std::max<size_t>(2 * buf.size(), n + (1u << 20)));
        const ssize_t got = read(fd, buf.data() + n, buf.size() - n);
        if (got < 0) {
            if (errno == EINTR) continue;
            set_err(err, RRG_E_IO, "read failed: %s", std::strerror(errno));
            return RRG_E_IO;
        }
        if (got == 0) break;
        n += (size_t)got;
    }
    buf.resize(n);
    return RRG_OK;
}

int rrg_gpu_decode_fd(void* ctx, int fd, const rrg_cfg* cfg, rrg_chunk_cb cb, void* user, rrg_stats* stats,
                      rrg_error* err) {
    if (err) std::memset(err, 0, sizeof *err);
    if (stats) std::memset(stats, 0, sizeof *stats);
    std::vector<uint8_t> buf;
    if (ctx && (!cfg || cfg->num_threads != 1)) {
        // the streaming chunking, decoded in slabs while the input is read
        Ctx& c = *static_cast<Ctx*>(ctx);
        HIPCHK(hipSetDevice(c.device));
        rrg_cfg cf;
        if (cfg) cf = *cfg;
        else rrg_cfg_default(&cf);
        const double t0 = now_ms();
        begin_profile(c, 0);
        c.verbose = cf.verbose != 0;
        const int src = decode_stream_slabbed(c, fd, buf, cf, cb, user, stats, err);
        if (src != kNotSlabbed) {
            c.prof.compressed_bytes = stats ? stats->compressed_bytes : 0;
            c.prof.ms_total = now_ms() - t0;
            if (cf.verbose) print_profile(c);
            return src;
        }
    }
    // the single-worker streaming path, or no deflate body to cut: the whole
    // input, then one decode
    const int rc = read_all(fd, buf, err);
    if (rc) return rc;
    return decode_host(ctx, buf.data(), buf.size(), cfg, cb, user, stats, err, true);
}

int rrg_gpu_read_gz(void* ctx, const char* path, const rrg_cfg* cfg, rrg_chunk_cb cb, void* user, rrg_stats* stats,
                    rrg_error* err) {
    if (err) std::memset(err, 0, sizeof *err);
    if (stats) std::memset(stats, 0, sizeof *stats);
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        set_err(err, RRG_E_IO, "cannot open %s", path);
        return RRG_E_IO;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        set_err(err, RRG_E_IO, "cannot stat %s", path);
        return RRG_E_IO;
    }
    if (!S_ISREG(st.st_mode) || st.st_size == 0) {
        // non-seekable or empty: the streaming path (lib.rs:378-431)
        const int rc = rrg_gpu_decode_fd(ctx, fd, cfg, cb, user, stats, err);
        close(fd);
        return rc;
    }
    const size_t len = (size_t)st.st_size;
    void* m = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (m == MAP_FAILED) {
        set_err(err, RRG_E_IO, "mmap failed for %s", path);
        return RRG_E_IO;
    }
    (void)madvise(m, len, MADV_SEQUENTIAL);
    const int rc = rrg_gpu_decode(ctx, static_cast<const uint8_t*>(m), len, cfg, cb, user, stats, err);
    munmap(m, len);
    return rc;
}

}  // extern "C"
