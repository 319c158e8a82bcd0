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

#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rrg_launch.h"

namespace {
using namespace rrg;

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
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
        case DET_DIST_FAR: return "back-reference before the start of the member";
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
// x^(8n) mod P
uint32_t x8n(uint64_t n) {
    uint32_t r = 0x80000000u;
    int k = 3;
    while (n) {
        if (n & 1) r = gf2_mul(x2n().t[k], r);
        n >>= 1;
        ++k;
    }
    return r;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n) {
        if (n <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t c = n + n / 8 + 4096;
        hipError_t e = hipMalloc(&p, c);
        if (e == hipSuccess) cap = c;
        return e;
    }
    // grow keeping the first `used` bytes
    hipError_t grow_keep(size_t n, size_t used, hipStream_t s) {
        if (n <= cap) return hipSuccess;
        void* q = nullptr;
        const size_t c = n + n / 4 + 4096;
        hipError_t e = hipMalloc(&q, c);
        if (e != hipSuccess) return e;
        if (p && used) {
            e = hipMemcpyAsync(q, p, used, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                (void)hipFree(q);
                return e;
            }
        }
        if (p) (void)hipFree(p);
        p = q;
        cap = c;
        return hipSuccess;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct Ctx {
    int device = 0;
    hipStream_t s = nullptr;
    int n_cu = 0;
    DevBuf in, targets, found, segs, res, mrec, work, counter, cells, sorted, chain, tz, lself, lptr, lcell, counters,
        tiles, out, poff, plen, pcrc, err;
    uint8_t* pin = nullptr;
    size_t pin_cap = 0;
    rrg_profile prof{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~Ctx() {
        DevBuf* all[] = {&in,      &targets, &found, &segs, &res,  &mrec,  &work,     &counter,
                         &cells,   &sorted,  &chain, &tz,   &lself, &lptr, &lcell,    &counters,
                         &tiles,   &out,     &poff,  &plen, &pcrc, &err};
        for (DevBuf* b : all) b->release();
        if (pin) (void)hipHostFree(pin);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (s) (void)hipStreamDestroy(s);
    }
};

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) {                                                              \
            set_err(err, RRG_E_DEVICE, "HIP error %s at %s:%d", hipGetErrorString(_e), __FILE__, \
                    __LINE__);                                                               \
            return RRG_E_DEVICE;                                                             \
        }                                                                                    \
    } while (0)

// Times one group of kernels on the context stream.
struct StageTimer {
    Ctx& c;
    double* acc;
    StageTimer(Ctx& cc, double* a) : c(cc), acc(a) { (void)hipEventRecord(c.ev0, c.s); }
    void stop() {
        (void)hipEventRecord(c.ev1, c.s);
        (void)hipEventSynchronize(c.ev1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c.ev0, c.ev1);
        *acc += ms;
    }
};

// Result of one decode, in device output coordinates.
struct Decoded {
    std::vector<uint64_t> chunk_ends;  // emitted chunk end offsets (absolute in c.out)
    uint64_t spec_failures = 0;
    bool failed = false;
    rrg_error error{};
};

struct Geometry {
    InflateGeom g;
    uint64_t lanes_in_flight;
};

Geometry pick_geometry(Ctx& c, const rrg_cfg& cfg) {
    Geometry G;
    G.g.lut_bits = (cfg.lut_bits == 8 || cfg.lut_bits == 10) ? cfg.lut_bits : 9;
    int lanes = cfg.lanes_per_block > 0 ? cfg.lanes_per_block : 32;
    if (lanes > 64) lanes = 64;
    while (lanes > 1 && (size_t)lanes * inflate_lds_per_lane(G.g.lut_bits) > 65536) lanes /= 2;
    G.g.lanes_per_block = lanes;
    int per_cu = inflate_blocks_per_cu(G.g.lut_bits, lanes);
    if (per_cu < 1) per_cu = 1;
    G.g.blocks = per_cu * c.n_cu;
    G.lanes_in_flight = (uint64_t)G.g.blocks * (uint64_t)lanes;
    return G;
}

// Give segment j a larger output region (cells or member records) after it
// overflowed; the old region is abandoned.
void grow_segment(SegDesc& d, uint32_t status, uint64_t& cells_used, uint64_t& mb_used) {
    if (status == ST_MB_OVERFLOW) {
        d.mb_base = (uint32_t)mb_used;
        d.mb_cap = d.mb_cap * 16 + 16;
        mb_used += d.mb_cap;
    } else {
        const uint64_t cap = d.cells_cap * 8 + 65536;
        d.cells_base = (cells_used + 7) & ~7ull;
        d.cells_cap = cap;
        cells_used = d.cells_base + cap;
    }
}

// One launch of the inflate kernel over the segment indices in `work`; results
// for every segment are copied back into `res`.
int run_inflate(Ctx& c, const Geometry& G, uint64_t nbits, uint64_t in_len, std::vector<SegDesc>& desc,
                std::vector<SegResult>& res, uint64_t cells_used, uint64_t mb_used, const std::vector<uint32_t>& work,
                rrg_error* err) {
    const size_t S = desc.size();
    if (work.empty()) return RRG_OK;
    HIPCHK(c.sorted.ensure((size_t)G.g.blocks * G.g.lanes_per_block * 480 * sizeof(uint16_t)));
    HIPCHK(c.cells.grow_keep((cells_used + 64) * sizeof(uint16_t), c.cells.cap, c.s));
    HIPCHK(c.mrec.grow_keep((mb_used + 1) * sizeof(MemberRec), c.mrec.cap, c.s));
    HIPCHK(c.segs.ensure(S * sizeof(SegDesc)));
    HIPCHK(c.res.ensure(S * sizeof(SegResult)));
    HIPCHK(c.work.ensure(work.size() * sizeof(uint32_t)));
    HIPCHK(c.counter.ensure(16));
    HIPCHK(hipMemcpyAsync(c.segs.p, desc.data(), S * sizeof(SegDesc), hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.work.p, work.data(), work.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemsetAsync(c.counter.p, 0, 16, c.s));
    {
        StageTimer t(c, &c.prof.ms_inflate);
        HIPCHK(launch_inflate(c.in.as<uint64_t>(), nbits, in_len, c.segs.as<SegDesc>(), c.res.as<SegResult>(),
                              c.mrec.as<MemberRec>(), c.work.as<uint32_t>(), (uint32_t)work.size(),
                              c.counter.as<uint32_t>(), c.cells.as<uint16_t>(), c.sorted.as<uint16_t>(), G.g, c.s));
        t.stop();
    }
    c.prof.n_inflate_launches++;
    res.resize(S);
    HIPCHK(hipMemcpyAsync(res.data(), c.res.p, S * sizeof(SegResult), hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    return RRG_OK;
}

inline bool overflowed(uint32_t st) { return st == ST_OVERFLOW || st == ST_MB_OVERFLOW; }

// Geometry of the accepted chain, shared by the resolve / CRC stages.
struct ChainOut {
    std::vector<uint64_t> off;  // output offset (relative to the path's base) per chain segment, plus total
};

// Tail chain + resolve + narrow for chain segments [0, n) into c.out at out_base.
int resolve_chain(Ctx& c, const std::vector<SegDesc>& desc, const std::vector<SegResult>& res, uint32_t n,
                  uint64_t out_base, const std::vector<uint64_t>& off, uint32_t* marker_err_seg, rrg_error* err) {
    *marker_err_seg = UINT32_MAX;
    if (n == 0) return RRG_OK;
    std::vector<ChainSeg> cs(n);
    uint64_t tz = 0;
    for (uint32_t j = 0; j < n; ++j) {
        cs[j].cells_base = desc[j].cells_base;
        cs[j].len = res[j].n_cells;
        cs[j].tz_base = tz;
        cs[j].spec = (desc[j].flags & SEG_SPEC) ? 1u : 0u;
        cs[j].pad = 0;
        tz += std::min<uint64_t>(res[j].n_cells, 32768);
    }
    HIPCHK(c.chain.ensure(n * sizeof(ChainSeg)));
    HIPCHK(hipMemcpyAsync(c.chain.p, cs.data(), n * sizeof(ChainSeg), hipMemcpyHostToDevice, c.s));
    HIPCHK(c.counters.ensure(64));
    bool any_spec = false;
    for (uint32_t j = 0; j < n; ++j) any_spec |= cs[j].spec != 0;
    if (any_spec && tz) {
        StageTimer t(c, &c.prof.ms_tail);
        // ~4 blocks per CU keeps phase 1 parallel while phase 2 stays short
        const uint32_t G = std::max<uint32_t>(1, (uint32_t)(n / std::max(1, 4 * c.n_cu)));
        const uint32_t nb = (n + G - 1) / G;
        HIPCHK(c.tz.ensure((size_t)nb * 32768 * 2 * sizeof(uint16_t)));
        HIPCHK(c.counters.ensure(16));
        const uint32_t none = UINT32_MAX;
        HIPCHK(hipMemcpyAsync(c.counters.p, &none, 4, hipMemcpyHostToDevice, c.s));
        uint16_t* exitb = c.tz.as<uint16_t>();
        uint16_t* entry = exitb + (size_t)nb * 32768;
        HIPCHK(launch_tail(c.chain.as<ChainSeg>(), n, G, c.cells.as<uint16_t>(), exitb, entry,
                           c.counters.as<uint32_t>(), c.s));
        uint32_t e1 = UINT32_MAX;
        HIPCHK(hipMemcpyAsync(&e1, c.counters.p, 4, hipMemcpyDeviceToHost, c.s));
        t.stop();
        HIPCHK(hipStreamSynchronize(c.s));
        *marker_err_seg = e1;
        c.prof.n_tail_rounds = nb;
    }
    // resolve tiles: 64 KiB of output each, aligned in output coordinates
    const uint64_t TILE = 65536;
    std::vector<ResolveTile> tiles;
    for (uint32_t j = 0; j < n; ++j) {
        const uint64_t ob = out_base + off[j], oe = ob + res[j].n_cells;
        uint64_t p = ob;
        while (p < oe) {
            const uint64_t q = std::min<uint64_t>(oe, (p / TILE + 1) * TILE);
            ResolveTile t;
            t.cell_addr = desc[j].cells_base + (p - ob);
            t.out_off = p;
            t.seg = j;
            t.n = (uint32_t)(q - p);
            tiles.push_back(t);
            p = q;
        }
    }
    if (!tiles.empty()) {
        StageTimer t(c, &c.prof.ms_resolve);
        HIPCHK(c.tiles.ensure(tiles.size() * sizeof(ResolveTile)));
        HIPCHK(c.err.ensure(16));
        HIPCHK(hipMemcpyAsync(c.tiles.p, tiles.data(), tiles.size() * sizeof(ResolveTile), hipMemcpyHostToDevice,
                              c.s));
        const uint32_t none = UINT32_MAX;
        HIPCHK(hipMemcpyAsync(c.err.p, &none, 4, hipMemcpyHostToDevice, c.s));
        HIPCHK(launch_resolve(c.tiles.as<ResolveTile>(), (uint32_t)tiles.size(), c.chain.as<ChainSeg>(),
                              c.cells.as<uint16_t>(), c.out.as<uint8_t>(), c.err.as<uint32_t>(), c.s));
        uint32_t e2 = UINT32_MAX;
        HIPCHK(hipMemcpyAsync(&e2, c.err.p, 4, hipMemcpyDeviceToHost, c.s));
        t.stop();
        HIPCHK(hipStreamSynchronize(c.s));
        *marker_err_seg = std::min(*marker_err_seg, e2);
    }
    return RRG_OK;
}

// CRC pieces for byte ranges [lo, hi) of c.out; returns the raw CRC of each range.
int crc_ranges(Ctx& c, const std::vector<std::pair<uint64_t, uint64_t>>& ranges, std::vector<uint32_t>& raw,
               rrg_error* err) {
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
    if (!po.empty()) {
        StageTimer t(c, &c.prof.ms_crc);
        HIPCHK(c.poff.ensure(po.size() * 8));
        HIPCHK(c.plen.ensure(pl.size() * 4));
        HIPCHK(c.pcrc.ensure(pc.size() * 4));
        HIPCHK(hipMemcpyAsync(c.poff.p, po.data(), po.size() * 8, hipMemcpyHostToDevice, c.s));
        HIPCHK(hipMemcpyAsync(c.plen.p, pl.data(), pl.size() * 4, hipMemcpyHostToDevice, c.s));
        HIPCHK(launch_crc_pieces(c.out.as<uint8_t>(), c.poff.as<uint64_t>(), c.plen.as<uint32_t>(),
                                 (uint32_t)po.size(), c.pcrc.as<uint32_t>(), c.s));
        HIPCHK(hipMemcpyAsync(pc.data(), c.pcrc.p, pc.size() * 4, hipMemcpyDeviceToHost, c.s));
        t.stop();
        HIPCHK(hipStreamSynchronize(c.s));
    }
    raw.assign(ranges.size(), 0);
    const uint32_t full_shift = x8n(PIECE);
    for (size_t i = 0; i < ranges.size(); ++i) {
        uint32_t r = 0;
        for (size_t k = first[i]; k < first[i + 1]; ++k) r = gf2_mul(pl[k] == PIECE ? full_shift : x8n(pl[k]), r) ^ pc[k];
        raw[i] = r;
    }
    return RRG_OK;
}

inline uint32_t std_crc(uint32_t raw, uint64_t len) { return raw ^ gf2_mul(x8n(len), 0xffffffffu) ^ 0xffffffffu; }

// parallel_decode_member (pipeline.rs:468-1041) over the body at file offset
// body_off of the file in c.in; output appended to c.out at out_base.
int member_path(Ctx& c, const Geometry& G, uint64_t body_off, uint64_t file_len, const rrg_cfg& cfg,
                uint64_t out_base, Decoded& dec, rrg_error* err) {
    const uint64_t nbits = file_len * 8;
    const uint64_t base = body_off * 8;
    const uint64_t total = nbits - base;
    const uint64_t chunk_bits = (uint64_t)std::max<size_t>(cfg.chunk_size_bytes, 65536) * 8;
    // segment size: enough segments to keep every lane busy twice over
    uint64_t seg_bytes = cfg.segment_bytes;
    if (!seg_bytes) {
        seg_bytes = (total / 8) / std::max<uint64_t>(1, 2 * G.lanes_in_flight);
        seg_bytes = std::max<uint64_t>(seg_bytes, 32 * 1024);
    }
    uint64_t D = std::max<uint64_t>(1, chunk_bits / (seg_bytes * 8));
    const uint64_t seg_bits = chunk_bits / D;

    // targets: chunk targets (pipeline.rs:497-503) and segment targets between them
    std::vector<uint64_t> targets;
    std::vector<uint8_t> is_chunk;
    for (uint64_t k = 0; k * chunk_bits < total; ++k) {
        if (k) {
            targets.push_back(base + k * chunk_bits);
            is_chunk.push_back(1);
        }
        for (uint64_t j = 1; j < D; ++j) {
            const uint64_t t = k * chunk_bits + j * seg_bits;
            if (t >= total) break;
            targets.push_back(base + t);
            is_chunk.push_back(0);
        }
    }
    std::vector<uint64_t> found(targets.size());
    if (!targets.empty()) {
        HIPCHK(c.targets.ensure(targets.size() * 8));
        HIPCHK(c.found.ensure(targets.size() * 8));
        HIPCHK(hipMemcpyAsync(c.targets.p, targets.data(), targets.size() * 8, hipMemcpyHostToDevice, c.s));
        StageTimer t(c, &c.prof.ms_find);
        HIPCHK(launch_find_blocks(c.in.as<uint64_t>(), nbits, c.targets.as<uint64_t>(), c.found.as<uint64_t>(),
                                  (uint32_t)targets.size(), c.s));
        HIPCHK(hipMemcpyAsync(found.data(), c.found.p, found.size() * 8, hipMemcpyDeviceToHost, c.s));
        t.stop();
        HIPCHK(hipStreamSynchronize(c.s));
    }
    // a find is a bit offset inside the body; anything else is a device fault
    for (size_t i = 0; i < found.size(); ++i)
        if (found[i] != UINT64_MAX && (found[i] < targets[i] || found[i] >= nbits)) {
            set_err(err, RRG_E_DEVICE, "block finder returned %llu for target %llu",
                    (unsigned long long)found[i], (unsigned long long)targets[i]);
            return RRG_E_DEVICE;
        }
    // chunk boundaries exactly as the reference (strictly increasing finds)
    std::vector<uint64_t> B{base};
    for (size_t i = 0; i < targets.size(); ++i)
        if (is_chunk[i] && found[i] != UINT64_MAX && found[i] > B.back()) B.push_back(found[i]);
    std::vector<uint64_t> M(B);
    for (size_t i = 0; i < targets.size(); ++i)
        if (!is_chunk[i] && found[i] != UINT64_MAX) M.push_back(found[i]);
    std::sort(M.begin(), M.end());
    M.erase(std::unique(M.begin(), M.end()), M.end());
    const size_t S = M.size();
    c.prof.n_segments += S;
    std::vector<uint8_t> seg_is_chunk(S, 0);
    {
        size_t bi = 0;
        for (size_t j = 0; j < S && bi < B.size(); ++j)
            if (M[j] == B[bi]) {
                seg_is_chunk[j] = 1;
                ++bi;
            }
    }
    // segment descriptors
    std::vector<SegDesc> desc(S);
    uint64_t cells_used = 0, mb_used = 0;
    for (size_t j = 0; j < S; ++j) {
        const uint64_t hint = j + 1 < S ? M[j + 1] : UINT64_MAX;
        const uint64_t span = ((j + 1 < S ? M[j + 1] : nbits) - M[j]) / 8;
        SegDesc& d = desc[j];
        d.start_bit = M[j];
        d.hint_bit = hint;
        d.cells_base = cells_used;
        d.cells_cap = ((span * 4 + 16384) + 7) & ~7ull;
        cells_used += d.cells_cap;
        d.flags = j ? SEG_SPEC : 0u;
        d.mb_base = (uint32_t)mb_used;
        d.mb_cap = 4;
        mb_used += 4;
        d.pad = 0;
    }
    std::vector<uint64_t> orig_start(M);
    std::vector<SegResult> res;
    std::vector<uint32_t> all(S);
    for (size_t j = 0; j < S; ++j) all[j] = (uint32_t)j;
    int rc = run_inflate(c, G, nbits, file_len, desc, res, cells_used, mb_used, all, err);
    if (rc) return rc;
    // a capacity overflow is not a decode failure (the region is grown below)
    std::vector<uint8_t> first_ok(S);
    for (size_t j = 0; j < S; ++j) first_ok[j] = res[j].status == ST_OK || overflowed(res[j].status);

    // anchored chain (pipeline.rs:774-809) with batched re-decode rounds
    size_t n_chain = 0;
    long fatal = -1;
    bool saw_final = false;
    for (int round = 0;; ++round) {
        uint64_t expected = base;
        std::vector<uint32_t> redo;
        n_chain = 0;
        fatal = -1;
        saw_final = false;
        for (size_t j = 0; j < S; ++j) {
            const SegResult& r = res[j];
            if (desc[j].start_bit == expected && r.status == ST_OK) {
                ++n_chain;
                expected = r.end_bit;
                if (r.final_block) {
                    saw_final = true;
                    break;
                }
                continue;
            }
            if (desc[j].start_bit == expected && overflowed(r.status)) {
                // verified start, output region too small: grow and decode again
                grow_segment(desc[j], r.status, cells_used, mb_used);
                redo.push_back((uint32_t)j);
                if (j + 1 < S) expected = desc[j + 1].start_bit;
                continue;
            }
            if (desc[j].start_bit == expected) {  // genuine error at a true boundary
                if (redo.empty()) fatal = (long)j;
                break;
            }
            desc[j].start_bit = expected;
            redo.push_back((uint32_t)j);
            if (j + 1 < S) expected = desc[j + 1].start_bit;  // optimistic; verified next round
        }
        if (redo.empty()) break;
        if (round > (int)S + 64) {
            set_err(err, RRG_E_DEVICE, "verification chain did not settle");
            return RRG_E_DEVICE;
        }
        rc = run_inflate(c, G, nbits, file_len, desc, res, cells_used, mb_used, redo, err);
        if (rc) return rc;
    }
    // chunk index of each segment (reference chunk k starts at B[k])
    std::vector<uint32_t> chunk_of(S);
    {
        uint32_t k = 0;
        for (size_t j = 0; j < S; ++j) {
            if (j && seg_is_chunk[j]) ++k;
            chunk_of[j] = k;
        }
    }
    // speculation failure = a chunk whose guessed start was not where the chain
    // arrived, or whose speculative decode failed (pipeline.rs:783-809)
    uint64_t spec = 0;
    for (size_t j = 1; j < n_chain; ++j)
        if (seg_is_chunk[j] && (orig_start[j] != desc[j].start_bit || !first_ok[j])) ++spec;
    dec.spec_failures += spec;

    // output offsets of the accepted chain
    std::vector<uint64_t> off(n_chain + 1, 0);
    for (size_t j = 0; j < n_chain; ++j) off[j + 1] = off[j] + res[j].n_cells;
    const uint64_t U = off[n_chain];
    HIPCHK(c.out.grow_keep(out_base + U + 64, out_base, c.s));
    uint32_t merr = UINT32_MAX;
    rc = resolve_chain(c, desc, res, (uint32_t)n_chain, out_base, off, &merr, err);
    if (rc) return rc;

    // member CRC / ISIZE (pipeline.rs:867-916)
    std::vector<MemberRec> recs(mb_used);
    if (mb_used) {
        HIPCHK(hipMemcpyAsync(recs.data(), c.mrec.p, mb_used * sizeof(MemberRec), hipMemcpyDeviceToHost, c.s));
        HIPCHK(hipStreamSynchronize(c.s));
    }
    struct Mem {
        uint64_t lo, hi;
        const MemberRec* rec;
        uint32_t seg;
    };
    std::vector<Mem> mems;
    uint64_t mstart = 0;
    for (size_t j = 0; j < n_chain; ++j)
        for (uint32_t k = 0; k < res[j].n_mb; ++k) {
            const MemberRec& r = recs[desc[j].mb_base + k];
            const uint64_t e = off[j] + r.out_off;
            mems.push_back({mstart, e, &r, (uint32_t)j});
            mstart = e;
        }
    std::vector<std::pair<uint64_t, uint64_t>> ranges;
    for (const Mem& m : mems) ranges.push_back({out_base + m.lo, out_base + m.hi});
    std::vector<uint32_t> raw;
    rc = crc_ranges(c, ranges, raw, err);
    if (rc) return rc;

    // earliest failure, as the chunk index it stops the stream at
    uint32_t fail_chunk = UINT32_MAX;
    rrg_error ferr{};
    if (fatal >= 0) {
        fail_chunk = chunk_of[fatal];
        status_err(&ferr, res[fatal].status, res[fatal].detail);
    }
    if (merr != UINT32_MAX && chunk_of[merr] < fail_chunk) {
        fail_chunk = chunk_of[merr];
        set_err(&ferr, RRG_E_INVALID, "invalid deflate stream: %s", detail_msg(DET_MARKER_RANGE));
    }
    for (size_t i = 0; i < mems.size(); ++i) {
        const Mem& m = mems[i];
        const uint64_t len = m.hi - m.lo;
        const uint32_t got = std_crc(raw[i], len);
        const uint32_t ck = chunk_of[m.seg];
        if (ck >= fail_chunk) break;
        if (got != m.rec->crc) {
            fail_chunk = ck;
            set_err(&ferr, RRG_E_CRC,
                    "CRC32 mismatch in member #%zu (uncompressed bytes: %llu, trailer at byte %llu): expected "
                    "%#010x, got %#010x",
                    i, (unsigned long long)len, (unsigned long long)(m.rec->trailer_byte - body_off),
                    m.rec->crc, got);
            ferr.expected = m.rec->crc;
            ferr.got = got;
            ferr.got64 = len;
            ferr.member = (uint32_t)i;
            ferr.trailer_byte = m.rec->trailer_byte - body_off;
            break;
        }
        if ((uint32_t)len != m.rec->isize) {
            fail_chunk = ck;
            set_err(&ferr, RRG_E_ISIZE, "ISIZE mismatch in member #%zu (trailer at byte %llu): expected %u, got %llu",
                    i, (unsigned long long)(m.rec->trailer_byte - body_off), m.rec->isize, (unsigned long long)len);
            ferr.expected = m.rec->isize;
            ferr.got64 = len;
            ferr.member = (uint32_t)i;
            ferr.trailer_byte = m.rec->trailer_byte - body_off;
            break;
        }
    }
    if (fail_chunk == UINT32_MAX && !saw_final) {
        // the chain ran past every segment without reaching the end of the file
        fail_chunk = n_chain ? chunk_of[n_chain - 1] + 1 : 0;
        set_err(&ferr, RRG_E_IO, "parallel decode produced no final block");
    }
    // emit chunks: [start of chunk k, start of chunk k+1) over the chain
    const uint32_t n_chunks_chain = n_chain ? chunk_of[n_chain - 1] + 1 : 0;
    const uint32_t emit = std::min(n_chunks_chain, fail_chunk);
    for (size_t j = 1; j <= n_chain && dec.chunk_ends.size() < emit; ++j)
        if (j == n_chain || chunk_of[j] != chunk_of[j - 1]) dec.chunk_ends.push_back(out_base + off[j]);
    if (fail_chunk != UINT32_MAX) {
        dec.failed = true;
        dec.error = ferr;
    }
    return RRG_OK;
}

// parallel_decode_bgzf (pipeline.rs:1357-1636)
int bgzf_path(Ctx& c, const Geometry& G, const uint8_t* h, uint64_t len, const rrg_cfg& cfg, Decoded& dec,
              rrg_error* err) {
    struct Mb {
        uint64_t off, size;
    };
    std::vector<Mb> mem;
    uint64_t pos = 0;
    bool walk_trunc = false;
    while (pos < len) {
        const uint32_t sz = bgzf_block_size(h + pos, len - pos);
        if (!sz) break;
        if (pos + sz > len) {
            walk_trunc = true;
            break;
        }
        mem.push_back({pos, sz});
        pos += sz;
    }
    const size_t S = mem.size();
    std::vector<SegDesc> desc(S);
    std::vector<int> hdr_err(S, 0);
    uint64_t cells_used = 0, mb_used = 0;
    for (size_t j = 0; j < S; ++j) {
        const uint8_t* p = h + mem[j].off;
        const long hl = parse_header(p, mem[j].size, nullptr);
        uint64_t isz = 0;
        if (mem[j].size >= 18)
            isz = (uint64_t)p[mem[j].size - 4] | ((uint64_t)p[mem[j].size - 3] << 8) |
                  ((uint64_t)p[mem[j].size - 2] << 16) | ((uint64_t)p[mem[j].size - 1] << 24);
        SegDesc& d = desc[j];
        if (hl < 0 || mem[j].size < 18) {
            hdr_err[j] = 1;
            d.start_bit = d.hint_bit = 0;  // empty decode
        } else {
            d.start_bit = (mem[j].off + (uint64_t)hl) * 8;
            d.hint_bit = UINT64_MAX;
        }
        d.cells_base = cells_used;
        d.cells_cap = ((std::min<uint64_t>(isz, 1u << 22) + 64) + 7) & ~7ull;
        cells_used += d.cells_cap;
        d.flags = SEG_SINGLE_MEMBER;
        d.mb_base = (uint32_t)mb_used;
        d.mb_cap = 1;
        mb_used += 1;
        d.pad = 0;
    }
    c.prof.n_segments += S;
    std::vector<SegResult> res;
    std::vector<uint32_t> all(S);
    for (size_t j = 0; j < S; ++j) all[j] = (uint32_t)j;
    int rc = run_inflate(c, G, len * 8, len, desc, res, cells_used, mb_used, all, err);
    if (rc) return rc;
    for (int round = 0; round < 6; ++round) {
        std::vector<uint32_t> again;
        for (size_t j = 0; j < S; ++j)
            if (overflowed(res[j].status)) {
                grow_segment(desc[j], res[j].status, cells_used, mb_used);
                again.push_back((uint32_t)j);
            }
        if (again.empty()) break;
        rc = run_inflate(c, G, len * 8, len, desc, res, cells_used, mb_used, again, err);
        if (rc) return rc;
    }
    std::vector<uint64_t> off(S + 1, 0);
    for (size_t j = 0; j < S; ++j) off[j + 1] = off[j] + res[j].n_cells;
    const uint64_t U = off[S];
    HIPCHK(c.out.ensure(U + 64));
    uint32_t merr = UINT32_MAX;
    rc = resolve_chain(c, desc, res, (uint32_t)S, 0, off, &merr, err);
    if (rc) return rc;
    std::vector<MemberRec> recs(mb_used);
    if (mb_used) {
        HIPCHK(hipMemcpyAsync(recs.data(), c.mrec.p, mb_used * sizeof(MemberRec), hipMemcpyDeviceToHost, c.s));
        HIPCHK(hipStreamSynchronize(c.s));
    }
    std::vector<std::pair<uint64_t, uint64_t>> ranges(S);
    for (size_t j = 0; j < S; ++j) ranges[j] = {off[j], off[j + 1]};
    std::vector<uint32_t> raw;
    rc = crc_ranges(c, ranges, raw, err);
    if (rc) return rc;
    // first failing member (libdeflate_gzip_decompress semantics: header, deflate, CRC, ISIZE)
    size_t bad = S;
    for (size_t j = 0; j < S; ++j) {
        const SegResult& r = res[j];
        bool ok = !hdr_err[j] && r.status == ST_OK && r.n_mb == 1;
        if (ok) {
            const MemberRec& m = recs[desc[j].mb_base];
            ok = std_crc(raw[j], r.n_cells) == m.crc && (uint32_t)r.n_cells == m.isize &&
                 m.trailer_byte + 8 <= mem[j].off + mem[j].size;
        }
        if (!ok) {
            bad = j;
            break;
        }
    }
    // batches ramp 64 KiB -> chunk size (pipeline.rs:1488-1551)
    const uint64_t tmax = std::max<size_t>(cfg.chunk_size_bytes, 65536);
    uint64_t target = 65536, cur = 0;
    size_t first = 0;
    for (size_t j = 0; j < S; ++j) {
        cur += mem[j].size;
        if (cur >= target || j + 1 == S) {
            if (bad >= first && bad <= j) break;  // this batch holds the failing member
            dec.chunk_ends.push_back(off[j + 1]);
            if (cur >= target) target = std::min(target * 2, tmax);
            cur = 0;
            first = j + 1;
        }
    }
    if (bad < S) {
        dec.failed = true;
        set_err(&dec.error, RRG_E_INVALID, "invalid deflate stream: libdeflate gzip decode failed");
        return RRG_OK;
    }
    if (walk_trunc) {
        dec.failed = true;
        set_err(&dec.error, RRG_E_TRUNCATED, "truncated gzip stream");
        return RRG_OK;
    }
    if (pos < len) {
        // remainder: the first non-BGZF member onwards (pipeline.rs:1622-1633)
        uint32_t meth = 0;
        const long hl = parse_header(h + pos, len - pos, &meth);
        if (hl < 0) {
            dec.failed = true;
            status_err(&dec.error, (uint32_t)-hl, 0);
            dec.error.method = meth;
            return RRG_OK;
        }
        Decoded rest;
        rc = member_path(c, G, pos + (uint64_t)hl, len, cfg, U, rest, err);
        if (rc) return rc;
        for (uint64_t e : rest.chunk_ends) dec.chunk_ends.push_back(e);
        if (rest.failed) {
            dec.failed = true;
            dec.error = rest.error;
        }
    }
    (void)G;
    return RRG_OK;
}

// Decode the file already in c.in (and, for framing walks, at host pointer h).
int decode_loaded(Ctx& c, const uint8_t* h, uint64_t len, const rrg_cfg& cfg, Decoded& dec, rrg_error* err) {
    const Geometry G = pick_geometry(c, cfg);
    if (bgzf_block_size(h, len)) return bgzf_path(c, G, h, len, cfg, dec, err);
    uint32_t meth = 0;
    const long hl = parse_header(h, len, &meth);
    if (hl < 0) {
        dec.failed = true;
        status_err(&dec.error, (uint32_t)-hl, 0);
        dec.error.method = meth;
        return RRG_OK;
    }
    return member_path(c, G, (uint64_t)hl, len, cfg, 0, dec, err);
}

int load_host(Ctx& c, const uint8_t* file, size_t len, rrg_error* err) {
    HIPCHK(c.in.ensure(len + 64));
    HIPCHK(hipMemcpyAsync(c.in.p, file, len, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemsetAsync(c.in.as<uint8_t>() + len, 0, 64, c.s));
    return RRG_OK;
}

void begin_profile(Ctx& c, uint64_t len) {
    std::memset(&c.prof, 0, sizeof c.prof);
    c.prof.compressed_bytes = len;
}

void print_profile(const Ctx& c) {
    const rrg_profile& p = c.prof;
    std::fprintf(stderr,
                 "[rrg-gpu] %llu -> %llu bytes in %.3f ms | find %.3f inflate %.3f (x%llu) tail %.3f (%llu markers, "
                 "%llu rounds) resolve %.3f crc %.3f | %llu segments\n",
                 (unsigned long long)p.compressed_bytes, (unsigned long long)p.uncompressed_bytes, p.ms_total,
                 p.ms_find, p.ms_inflate, (unsigned long long)p.n_inflate_launches, p.ms_tail,
                 (unsigned long long)p.n_markers_tail, (unsigned long long)p.n_tail_rounds, p.ms_resolve, p.ms_crc,
                 (unsigned long long)p.n_segments);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int rrg_gpu_abi_version(void) { return RRG_GPU_ABI_VERSION; }

void rrg_cfg_default(rrg_cfg* cfg) {
    std::memset(cfg, 0, sizeof *cfg);
    cfg->chunk_size_bytes = 4u << 20;
    cfg->lut_bits = 9;
    cfg->lanes_per_block = 32;
}

void* rrg_gpu_create(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    Ctx* c = new Ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || hipStreamCreate(&c->s) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return nullptr;
    }
    c->n_cu = prop.multiProcessorCount;
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
    HIPCHK(hipSetDevice(c.device));
    rrg_cfg cfg;
    if (cfg_in) cfg = *cfg_in;
    else rrg_cfg_default(&cfg);
    const double t0 = now_ms();
    begin_profile(c, len);
    if (len == 0) {
        set_err(err, RRG_E_IO, "empty input");
        return RRG_E_IO;
    }
    HIPCHK(c.in.ensure(len + 64));
    HIPCHK(hipMemcpyAsync(c.in.p, d_file, len, hipMemcpyDeviceToDevice, c.s));
    HIPCHK(hipMemsetAsync(c.in.as<uint8_t>() + len, 0, 64, c.s));
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
    if (d_out) *d_out = c.out.as<uint8_t>();
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

int rrg_gpu_decode(void* ctx, const uint8_t* file, size_t len, const rrg_cfg* cfg_in, rrg_chunk_cb cb, void* user,
                   rrg_stats* stats, rrg_error* err) {
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
    if (len == 0) {
        set_err(err, RRG_E_IO, "empty input");
        return RRG_E_IO;
    }
    int rc = load_host(c, file, len, err);
    if (rc) return rc;
    Decoded dec;
    rc = decode_loaded(c, file, len, cfg, dec, err);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c.s));
    // in-order emission (pipeline.rs:921-937)
    uint64_t prev = 0, sent = 0, total = 0;
    uint64_t biggest = 0;
    for (uint64_t e : dec.chunk_ends) {
        biggest = std::max(biggest, e - prev);
        prev = e;
    }
    if (biggest > c.pin_cap) {
        if (c.pin) (void)hipHostFree(c.pin);
        c.pin = nullptr;
        c.pin_cap = 0;
        HIPCHK(hipHostMalloc((void**)&c.pin, biggest + 64, hipHostMallocDefault));
        c.pin_cap = biggest + 64;
    }
    prev = 0;
    bool closed = false;
    for (uint64_t e : dec.chunk_ends) {
        const uint64_t n = e - prev;
        if (n) HIPCHK(hipMemcpyAsync(c.pin, c.out.as<uint8_t>() + prev, n, hipMemcpyDeviceToHost, c.s));
        HIPCHK(hipStreamSynchronize(c.s));
        total += n;
        prev = e;
        if (cb && cb(user, c.pin, n)) {
            closed = true;
            break;
        }
        ++sent;
    }
    c.prof.uncompressed_bytes = total;
    c.prof.ms_total = now_ms() - t0;
    if (cfg.verbose) print_profile(c);
    if (stats) {
        stats->compressed_bytes = len;
        stats->uncompressed_bytes = total;
        stats->chunks_decoded = sent;
        stats->speculation_failures = dec.spec_failures;
    }
    if (closed) {
        // speculative path: Ok with partial stats; BGZF: BrokenPipe (pipeline.rs:1028-1030, 1575-1578)
        if (bgzf_block_size(file, len)) {
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
    if (fstat(fd, &st) != 0 || !S_ISREG(st.st_mode) || st.st_size == 0) {
        close(fd);
        set_err(err, RRG_E_IO, "%s: not a non-empty regular file (pipes and empty files are not supported)", path);
        return RRG_E_IO;
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
