// rrg_engine.h -- internal interface of the host driver (not part of the C ABI).
//
// The batch engine (rrg_gpu.cpp) decodes a set of gzip files resident in one
// device input buffer; the single-file entry points, the batched entry point
// and the serving loop (rrg_serve.cpp) are built on it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "rrg_gpu.h"
#include "rrg_launch.h"

namespace rrg {
namespace host {

// A fixed region of device memory that DevBufs are carved from instead of
// being allocated (bump allocation; a grown buffer abandons its old block).
// A request that does not fit fails the call with `exhausted` set.
struct Arena {
    uint8_t* base = nullptr;
    size_t cap = 0, used = 0, high = 0;
    bool exhausted = false;
    size_t fixed_short = 0;  // bytes a caller-owned (fixed) buffer was asked to hold beyond its capacity
    void* take(size_t n) {
        const size_t a = (used + 255) & ~(size_t)255;
        if (a > cap || n > cap - a) {
            exhausted = true;
            return nullptr;
        }
        used = a + n;
        if (used > high) high = used;
        return base + a;
    }
};

// hipMalloc, or hipErrorOutOfMemory when the allocation-failure hook
// (rrg_gpu_debug_fail_alloc) says so.
hipError_t dev_malloc(void** p, size_t n);

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    Arena* arena = nullptr;  // carve from this arena instead of allocating
    bool fixed = false;      // caller-owned memory of capacity cap: never reallocated
    hipError_t take(void** q, size_t n) {
        if (fixed) {
            if (arena) arena->fixed_short = n;
            return hipErrorOutOfMemory;
        }
        if (arena) {
            *q = arena->take(n);
            return *q ? hipSuccess : hipErrorOutOfMemory;
        }
        return dev_malloc(q, n);
    }
    void drop() {
        if (p && !arena && !fixed) (void)hipFree(p);
    }
    hipError_t ensure(size_t n) {
        if (n <= cap) return hipSuccess;
        drop();
        p = nullptr;
        cap = 0;
        const size_t c = arena ? n : n + n / 8 + 4096;
        hipError_t e = take(&p, c);
        if (e == hipSuccess) cap = c;
        return e;
    }
    // grow keeping the first `used` bytes
    hipError_t grow_keep(size_t n, size_t used, hipStream_t s) {
        if (n <= cap) return hipSuccess;
        void* q = nullptr;
        const size_t c = arena ? n + n / 4 : n + n / 4 + 4096;
        hipError_t e = take(&q, c);
        if (e != hipSuccess) return e;
        if (p && used) {
            e = hipMemcpyAsync(q, p, used, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                if (!arena) (void)hipFree(q);
                return e;
            }
        }
        drop();
        p = q;
        cap = c;
        return hipSuccess;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
    void release() {
        drop();
        p = nullptr;
        cap = 0;
    }
};

struct Ctx {
    int device = 0;
    hipStream_t s = nullptr;
    hipStream_t s2 = nullptr;  // a second stream (default priority) for work that overlaps s
    hipStream_t su = nullptr;  // uploads of the next slab's input (slab decode; made on first use)
    hipStream_t sc = nullptr;  // copies of decoded chunks to the host (slab decode; made on first use)
    hipEvent_t evf[2] = {nullptr, nullptr};  // fork / join between s and s2
    int n_cu = 0;
    uint64_t max_lanes = 0;  // cap on inflate lanes in flight (0 = as many as fit)
    DevBuf in, targets, limits, scan_ends, found, segs, res, mrec, work, counter, toks, xitems, cells, sorted, chain, tz,
        tiles, out, poff, plen, pcrc, err, reach, first, stops, exits, recs, legtg, lfound, lnext, exact_idx,
        xcut, cuts;
    DevBuf in2, out2;  // the other input and output buffer of the slab decode
    DevBuf win;        // the last 32 KiB of output before the next slab
    size_t in_used = 0;  // bytes of c.in holding the current batch (files + padding)
    std::vector<std::vector<uint64_t>> batch_ends;  // chunk ends of the last batched call, per file
    uint8_t* pin = nullptr;
    size_t pin_cap = 0;
    rrg_profile prof{};
    bool verbose = false;  // Config.verbose of the current call: per-round lines on stderr
    bool graph_post = false;  // tail, resolve and CRC as one graph launch (batched forms)
    PostGraphs post_graphs;   // those graphs, built once, their nodes set per batch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evc[2] = {nullptr, nullptr};  // staging copies of the chunk emission
    // the expansion's start and end: read after a later wait, so that the
    // host prepares the next stage while the expansion runs
    hipEvent_t evx[2] = {nullptr, nullptr};
    bool x_timed = false;
    // pinned staging of uploads queued behind running kernels (handed out in
    // order; reset only where the stream has drained)
    uint8_t* stg = nullptr;
    size_t stg_cap = 0, stg_used = 0;
    bool no_alloc = false;  // a call that must not allocate: no staging buffer is made or grown
    // pinned landing area of device-to-host copies, moved to their
    // destinations after the wait that ends them (finish_downloads)
    struct Download {
        void* dst;
        size_t off, n;
    };
    uint8_t* dl = nullptr;
    size_t dl_cap = 0, dl_used = 0;
    std::vector<Download> dl_pend;
    // every device buffer of the context
    std::vector<DevBuf*> bufs() {
        return {&in,    &targets, &limits, &scan_ends, &found, &segs,  &res,   &mrec, &work, &counter,
                &toks,  &xitems,  &cells,  &sorted,    &chain, &tz,    &tiles, &out,  &poff, &plen,
                &pcrc,  &err,     &reach,  &first,     &stops, &exits, &recs,  &legtg, &lfound, &lnext, &exact_idx,
                &xcut,  &cuts,    &in2,    &out2,      &win};
    }
    ~Ctx() {
        for (DevBuf* b : bufs()) b->release();
        if (pin) (void)hipHostFree(pin);
        if (stg) (void)hipHostFree(stg);
        if (dl) (void)hipHostFree(dl);
        for (hipEvent_t e : evx)
            if (e) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : evc)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : evf)
            if (e) (void)hipEventDestroy(e);
        if (s2) (void)hipStreamDestroy(s2);
        post_graphs.release();
        if (su) (void)hipStreamDestroy(su);
        if (sc) (void)hipStreamDestroy(sc);
        if (s) (void)hipStreamDestroy(s);
    }
};

// The HIP error behind the last RRG_E_DEVICE returned on this thread (the
// server tells an allocation failure from a device fault by it).
extern thread_local hipError_t t_last_hip;

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) {                                                              \
            t_last_hip = _e;                                                                 \
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

struct Decoded {
    std::vector<uint64_t> chunk_ends;  // output chunk end offsets, relative to out_base
    uint64_t out_base = 0;             // the file's first output byte in the batch output buffer
    uint64_t spec_failures = 0;
    bool failed = false;
    rrg_error error{};
};

struct Geometry {
    InflateGeom g;
    uint64_t lanes_in_flight;  // segments in flight (the automatic segment size aims at one each)
    uint64_t scratch_bytes;    // inflate scratch (per lane or per wave)
};

struct Job {
    // input
    uint64_t in_off = 0;         // the file's first byte in c.in (multiple of 8)
    uint64_t len = 0;            // file length
    const uint8_t* h = nullptr;  // host copy of the file: headers and BGZF framing are read here
    size_t chunk_size = 4u << 20;
    size_t segment_bytes = 0;
    // non-seekable input (streaming.rs): chunk boundaries are a chain of finds,
    // each one chunk past the previous boundary; no BGZF special case
    bool stream = false;
    std::vector<uint64_t> stream_B;  // chunk starts (absolute bits)
    // single worker (Config.num_threads == 1): 1 = the serial path of
    // pipeline.rs:1211-1345, 2 = the streaming serial path of lib.rs:244-361
    // (non-seekable input).  Chunks end at block ends; no speculation counted.
    int serial = 0;
    // One slab of a bounded-memory decode (rrg_slab.cpp).  `slab`: the member
    // path's chunk starts are given in stream_B (stream_B[0] is where the
    // chain starts) and its last segment stops at the first block start at or
    // past stop_bit (UINT64_MAX: the final block).  `resume`: the chain starts
    // inside the stream at stream_B[0], after output that earlier slabs
    // produced; no header is parsed.
    bool slab = false, resume = false;
    uint64_t stop_bit = UINT64_MAX;
    uint64_t win_len = 0;              // output bytes just before the slab that markers may reach (<= 32 KiB)
    const uint8_t* win_src = nullptr;  // device: those bytes
    uint64_t body_ref = 0;             // c.in byte that reported trailer offsets count from (wrapping)
    uint32_t carry_raw = 0;            // raw CRC and length of the open member's output before the slab
    uint64_t carry_len = 0;
    uint64_t carry_member = 0;         // members whose trailers came before the slab
    uint64_t bgzf_target0 = 65536;     // BGZF batch target the slab's first batch starts with
    // results of a slab: the open member at the end of the chain
    uint32_t open_raw = 0;
    uint64_t open_len = 0, members_done = 0;
    size_t open_r = SIZE_MAX;  // its CRC range
    uint64_t chain_end = 0;    // end bit of the member-path chain
    uint64_t bgzf_target_end = 65536;
    // host plan
    struct Mb {
        uint64_t off, size;  // file offsets
    };
    std::vector<Mb> mem;           // BGZF members
    std::vector<uint8_t> hdr_err;  // per BGZF member
    bool bgzf = false, walk_trunc = false;
    bool pre_failed = false;  // header error found while planning (reported after any BGZF output)
    rrg_error pre_err{};
    uint64_t body = UINT64_MAX;  // file offset of the speculative (member-path) deflate body; none if MAX
    // targets and segments (index ranges in the batch arrays)
    size_t t_lo = 0, t_hi = 0;
    size_t bg_lo = 0, bg_hi = 0, mp_lo = 0, mp_hi = 0;
    std::vector<uint32_t> order;        // member-path segments (desc indices) in stream order
    std::vector<uint8_t> seg_is_chunk;  // per position in `order`: the segment starts a reference chunk
    std::vector<uint64_t> orig_start;
    std::vector<uint8_t> first_ok;
    size_t n_chain = 0;
    size_t n_out = 0;  // chain segments laid out: the chain, and on the serial paths its failing segment
    long fatal = -1;
    bool saw_final = false;
    // output
    std::vector<uint64_t> off;  // job-relative output offset: BGZF members, then chain segments, then the end
    uint64_t U = 0;
    size_t c_lo = 0;  // first chain segment of this job in the batch chain
    size_t r_lo = 0;  // first CRC range
    Decoded dec;

    uint64_t abs_bits(uint64_t file_byte) const { return (in_off + file_byte) * 8; }
    uint64_t end_bits() const { return (in_off + len) * 8; }
    // where the member-path chain starts, and where its last segment may end
    uint64_t chain_start() const { return resume ? stream_B[0] : abs_bits(body); }
    uint64_t last_end() const { return stop_bit != UINT64_MAX ? stop_bit : end_bits(); }
};

void set_err(rrg_error* e, int kind, const char* fmt, ...);
void status_err(rrg_error* e, uint32_t st, uint32_t det);
long parse_header(const uint8_t* p, size_t n, uint32_t* method);
uint32_t bgzf_block_size(const uint8_t* p, size_t n);
double now_ms();
Geometry pick_geometry(Ctx& c, const rrg_cfg& cfg);
// Device bytes held at the peak of a decode of len compressed bytes into out bytes.
uint64_t engine_footprint(uint64_t len, uint64_t out);
void begin_profile(Ctx& c, uint64_t len);
void print_profile(const Ctx& c);
// Decode every job (inputs already in c.in[0, c.in_used)) into `out`.
int decode_batch(Ctx& c, const Geometry& G, std::vector<Job>& jobs, DevBuf& out, rrg_error* err);
// Exact walks of the block finder (block_finder.rs:109-141) in c.in: found[i]
// = the first verified block start on the skip chain from targets[i] below
// scan_ends[i] (UINT64_MAX if none); limits[i] = the end of the target's
// file.  Walks are cut at stops[i] (reported in exits[i]) and finish_open
// continues the cut ones to their end.
int run_finder(Ctx& c, const std::vector<uint64_t>& targets, const std::vector<uint64_t>& limits,
               const std::vector<uint64_t>& scan_ends, std::vector<uint64_t>& found, rrg_error* err,
               const std::vector<uint64_t>* stops = nullptr, std::vector<uint64_t>* exits = nullptr,
               bool want_recs = false);
int finish_open(Ctx& c, const std::vector<uint64_t>& targets, const std::vector<uint64_t>& limits,
                const std::vector<uint64_t>& exits, std::vector<uint64_t>& found, rrg_error* err);
constexpr uint64_t kExactCapBits = 8ull * 131072;  // bits an exact walk covers before it continues in legs
// rrg_gpu_decode on a file in host memory, in slabs under a device-byte
// budget, with the chunks of decoded slabs handed to cb while later slabs
// decode (rrg_slab.cpp).  Not for the single-worker or streaming chunking.
// Returns kNotSlabbed, having done nothing, for a file without a first
// deflate body or BGZF member to cut (the whole-file decode reports it).
constexpr int kNotSlabbed = -1;
int decode_slabbed(Ctx& c, const uint8_t* file, size_t len, const rrg_cfg& cfg, rrg_chunk_cb cb, void* user,
                   rrg_stats* stats, rrg_error* err);
// The same over a non-seekable input read as the decode goes (the streaming
// chunking, streaming.rs:278-346).  kNotSlabbed: no deflate body after the
// first header; `head` then holds the bytes read so far.
int decode_stream_slabbed(Ctx& c, int fd, std::vector<uint8_t>& head, const rrg_cfg& cfg, rrg_chunk_cb cb,
                          void* user, rrg_stats* stats, rrg_error* err);
// The device-resident decode of rrg_gpu_decode_device into c.out (the file at
// d_file in device memory); the out-parameters as there, minus d_out.
int decode_device_impl(Ctx& c, const void* d_file, size_t len, const uint8_t* h_file, const rrg_cfg* cfg_in,
                       uint64_t* out_len, uint64_t* chunk_ends, size_t chunk_cap, size_t* n_chunks, rrg_stats* stats,
                       rrg_error* err);
// Create the context's stream and events on `device`; false on failure.
bool init_ctx(Ctx& c, int device);

}  // namespace host
}  // namespace rrg
