// rrg_slab.cpp -- rrg_gpu_decode / rrg_gpu_read_gz on a file in slabs:
// bounded device memory, with chunks handed out while later slabs decode.
//
// The reference hands each chunk to the consumer as soon as it and every
// chunk before it are ready, and bounds the decoded chunks waiting for the
// consumer (pipeline.rs:272-375, 858-952).  Here the compressed stream is cut
// into slabs of whole chunks, each decoded by one run of the batch engine
// into one of two device output buffers, with three threads at work:
//
//   worker thread  : finds slab s's chunk starts and decodes slab s
//   upload thread  : copies slab s+1's compressed bytes to the device
//   calling thread : copies slab s-1's chunks to pinned host memory and hands
//                    them to the callback, in order
//
// Chunk starts are the reference's (pipeline.rs:489-551): targets every
// chunk_size bytes from the first deflate body, each found by the exact block
// finder, kept when strictly past the previous start.  Slab s takes the
// targets [a_s, b_s); its chain ends at the first block start at or past the
// kept find of target b_s (its `stop`), which is where slab s+1's chain
// starts.  Carried from one slab to the next: that bit position, the last
// 32 KiB of output (slab s+1's markers resolve against it), the raw CRC and
// length of the member still open and the number of members before it.  A
// chunk starts at the first block start the chain reaches at or past its
// find, so cutting the chain at chunk starts changes no chunk boundary,
// DecodeStats count or error.
//
// BGZF files are cut at batch boundaries (pipeline.rs:1488-1551), each slab
// starting its batches at the target the ramp had reached.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <thread>
#include <functional>
#include <cerrno>
#include <cstring>
#include <unistd.h>

#include "rrg_engine.h"

namespace rrg {
namespace host {
namespace {

constexpr uint64_t kFirstSlab = 32ull << 20;   // compressed bytes of the first slab (the first chunk comes sooner)
constexpr uint64_t kGrowth = 4;                // each slab takes up to this many times its predecessor's input
constexpr uint64_t kSafeBits = 8ull * 65536;   // finds closer than this to the end of a slab's input are redone
constexpr uint64_t kWindow = 32768;            // deflate window
constexpr double kRatio0 = 8.0;                // output bytes per input byte assumed before the first slab

// Device bytes of a slab of `len` compressed bytes with `ratio` output bytes
// per input byte: the engine's working set plus the second input and output
// buffer.
double slab_bytes(uint64_t len, double ratio) {
    const uint64_t out = (uint64_t)((double)len * ratio);
    return (double)engine_footprint(len, out) + (double)len + (double)out + (double)kWindow;
}

// Largest slab input (at least 64 KiB) whose estimate `est` stays within `budget`.
template <class F>
uint64_t max_slab_input(uint64_t budget, F est) {
    uint64_t lo = 65536, hi = 1ull << 40;
    if (est(lo) > (double)budget) return lo;
    while (hi - lo > 65536) {
        const uint64_t mid = lo + (hi - lo) / 2;
        (est(mid) <= (double)budget ? lo : hi) = mid;
    }
    return lo;
}

uint64_t device_bytes(Ctx& c) {
    uint64_t n = 0;
    for (DevBuf* b : c.bufs()) n += b->cap;
    return n;
}

// A decoded slab handed to the calling thread.
struct SlabOut {
    int buf = 0;                  // 0: c.out, 1: c.out2
    uint64_t base = 0;            // first output byte of the slab's chunks in that buffer
    std::vector<uint64_t> ends;   // chunk ends relative to base
    bool failed = false;          // the stream stops with `error` after these chunks
    rrg_error error{};
    bool last = false;
};

struct Pipe {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<SlabOut> q;
    uint64_t consumed = 0;  // slabs whose chunks the calling thread has copied out
    bool closed = false;    // the calling thread stopped taking chunks
    bool done = false;      // the worker has pushed its last slab (or failed)
    int rc = RRG_OK;        // the worker's own failure (device, arguments)
    rrg_error err{};
    uint64_t spec = 0;      // speculation failures of the slabs decoded so far
    uint64_t slabs = 0;

    void push(SlabOut&& s) {
        std::lock_guard<std::mutex> lk(mu);
        q.push_back(std::move(s));
        cv.notify_all();
    }
    // true when slab `s` may write the output buffer slab s-2 used
    bool wait_free(uint64_t s) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return closed || s < 2 || consumed >= s - 1; });
        return !closed;
    }
    bool is_closed() {
        std::lock_guard<std::mutex> lk(mu);
        return closed;
    }
};

// The compressed bytes [lo, hi) of the file in a device input buffer, copied
// on the context's upload stream by a thread of its own.  The buffer is sized
// on the calling thread, so only that thread touches the context's buffers.
struct Upload {
    std::thread th;
    int rc = RRG_OK;
    rrg_error e{};

    void start(Ctx& c, DevBuf& dst, const uint8_t* file, uint64_t lo, uint64_t hi) {
        (void)join(nullptr);
        rrg_error* err = &e;
        rc = [&]() -> int {
            HIPCHK(dst.ensure(hi - lo + 64));
            return RRG_OK;
        }();
        if (rc) return;
        uint8_t* d = dst.as<uint8_t>();
        th = std::thread([this, &c, d, file, lo, hi] {
            rrg_error* err = &e;
            rc = [&]() -> int {
                HIPCHK(hipSetDevice(c.device));
                HIPCHK(hipMemcpyAsync(d, file + lo, hi - lo, hipMemcpyHostToDevice, c.su));
                HIPCHK(hipMemsetAsync(d + (hi - lo), 0, 64, c.su));
                HIPCHK(hipStreamSynchronize(c.su));
                return RRG_OK;
            }();
        });
    }
    int join(rrg_error* err) {
        if (th.joinable()) th.join();
        if (rc && err) *err = e;
        const int r = rc;
        rc = RRG_OK;
        return r;
    }
    ~Upload() {
        if (th.joinable()) th.join();
    }
};

// Synchronous upload into c.in (a slab whose input had to grow).
int load_range(Ctx& c, const uint8_t* file, uint64_t lo, uint64_t hi, rrg_error* err) {
    HIPCHK(c.in.ensure(hi - lo + 64));
    HIPCHK(hipMemcpyAsync(c.in.p, file + lo, hi - lo, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemsetAsync(c.in.as<uint8_t>() + (hi - lo), 0, 64, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    c.in_used = hi - lo + 64;
    return RRG_OK;
}

// The member-path chain as it crosses from one slab to the next.
struct Carry {
    uint64_t chain = 0;    // where the next slab's chain starts (file bits): a block start
    uint64_t guess = 0;    // the next slab's first chunk start as found
    uint32_t raw = 0;      // raw CRC and length of the output of the member still open
    uint64_t len = 0;
    uint64_t member = 0;   // members whose trailers came before
    uint64_t win_len = 0;  // output bytes before the chain held in c.win (<= 32 KiB)
};

// Compressed bytes of a non-seekable input (pipe, socket) read as the decode
// needs them; bytes before the current slab are dropped.
struct PipeSource {
    int fd = -1;
    std::vector<uint8_t> buf;  // bytes [base, base + n) of the stream
    uint64_t base = 0, n = 0;
    bool eof = false;

    uint64_t end() const { return base + n; }
    const uint8_t* at(uint64_t pos) const { return buf.data() + (pos - base); }
    // read until bytes up to `want` are held or the stream ends
    int fill(uint64_t want, rrg_error* err) {
        while (!eof && end() < want) {
            if (buf.size() - n < (1u << 20)) buf.resize(std::max<size_t>(n + (1u << 20), buf.size() + buf.size() / 2));
            const size_t ask = (size_t)std::min<uint64_t>(buf.size() - n, want - end());
            const ssize_t got = read(fd, buf.data() + n, ask);
            if (got < 0) {
                if (errno == EINTR) continue;
                set_err(err, RRG_E_IO, "read failed: %s", std::strerror(errno));
                return RRG_E_IO;
            }
            if (got == 0) eof = true;
            n += (uint64_t)got;
        }
        return RRG_OK;
    }
    void drop_before(uint64_t pos) {
        if (pos <= base) return;
        const uint64_t k = std::min(pos - base, n);
        std::memmove(buf.data(), buf.data() + k, n - k);
        n -= k;
        base += k;
    }
};

struct Worker {
    Ctx& c;
    const uint8_t* file;  // the file in host memory (nullptr for a pipe)
    uint64_t len;         // its length
    const rrg_cfg& cfg;
    Pipe& P;
    Geometry G;
    uint64_t budget = 0;
    uint64_t fixed = 0;                  // device bytes held whatever the slab size
    double slack = 1.25;                 // device bytes held over the estimate (buffers grow in steps)
    uint64_t in_done = 0, out_done = 0;  // compressed and output bytes of the slabs decoded
    uint64_t biggest = 0;                // largest slab input decoded so far
    struct Mb {
        uint64_t off, size;
    };
    std::vector<Mb> mem;  // BGZF members (file offsets)

    double ratio() const { return in_done > (1u << 20) ? (double)out_done / (double)in_done : kRatio0; }
    double estimate(uint64_t in, double r) const { return slack * slab_bytes(in, r) + (double)fixed; }
    uint64_t next_input(uint64_t prev) const {
        const double r = std::max(ratio(), 1.0);
        const uint64_t cap = max_slab_input(budget, [&](uint64_t in) { return estimate(in, r); });
        return std::min(cap, prev ? prev * kGrowth : kFirstSlab);
    }
    // after a slab of `in` input bytes and `out` output bytes: the peak so
    // far, and, under an explicit budget (the context then holds no more than
    // the budget from earlier calls), the slack the estimate needs to cover
    // what the buffers hold
    void measured(uint64_t in, uint64_t out) {
        const uint64_t held = device_bytes(c);
        c.prof.device_bytes_peak = std::max<uint64_t>(c.prof.device_bytes_peak, held);
        if (cfg.device_budget_bytes && in >= biggest && in) {
            biggest = in;
            const double e = slab_bytes(in, (double)out / (double)in);
            if (held > fixed) slack = std::max(slack, (double)(held - fixed) / e);
        }
    }

    int decode_one(std::vector<Job>& jobs, int buf, rrg_error* err) {
        DevBuf& out = buf ? c.out2 : c.out;
        const int rc = decode_batch(c, G, jobs, out, err);
        c.prof.device_bytes_peak = std::max<uint64_t>(c.prof.device_bytes_peak, device_bytes(c));
        return rc;
    }

    // Input of a slab whose chain may have run into the end of its input:
    // c.in made to hold [lo, end of the stream); sets hi and h_lo.
    using Rest = std::function<int(uint64_t lo, uint64_t& hi, const uint8_t*& h_lo, rrg_error* err)>;

    int member_slab(uint64_t s, uint64_t body, bool stream, const uint8_t* h_lo, uint64_t lo, uint64_t hi,
                    bool at_end, const std::vector<uint64_t>& Gs, uint64_t stop, Carry& K, const Rest& rest,
                    bool& done, rrg_error* err);
    int member_path(uint64_t body, rrg_error* err);
    int stream_path(PipeSource& src, uint64_t body, rrg_error* err);
    int bgzf(rrg_error* err);
};

// Decode member-path slab s, whose compressed bytes [lo, hi) are in c.in (and
// at h_lo on the host): chunk starts as found Gs (Gs[0] = K.guess, strictly
// increasing), its chain from K.chain to the first block start at or past
// `stop` (UINT64_MAX: the final block).  Hands the chunks to the calling
// thread and moves K on; `done` when the stream ends here.
int Worker::member_slab(uint64_t s, uint64_t body, bool stream, const uint8_t* h_lo, uint64_t lo, uint64_t hi,
                        bool at_end, const std::vector<uint64_t>& Gs, uint64_t stop, Carry& K, const Rest& rest,
                        bool& done, rrg_error* err) {
    done = false;
    // chunks whose start as found the chain has already passed are empty
    // (the first block start at or past it is the chain start)
    size_t m = 0;
    while (m < Gs.size() && Gs[m] <= K.chain) ++m;
    uint64_t spec_pre = 0;
    for (size_t i = 0; i + 1 < m; ++i) spec_pre += Gs[i] != K.chain;
    const uint64_t spec_first = Gs[m - 1] != K.chain ? 1 : 0;  // m >= 1: Gs[0] = guess <= chain
    SlabOut so;
    so.buf = (int)(s & 1);
    so.ends.assign(m - 1, 0);
    if (!P.wait_free(s)) {
        done = true;
        return RRG_OK;
    }
    if (stop <= K.chain) {
        // every chunk of the slab is empty: nothing to decode
        so.ends.push_back(0);
        if (!stream) P.spec += spec_pre + spec_first;
        P.slabs++;
        P.push(std::move(so));
        K.guess = stop;
        return RRG_OK;
    }
    std::vector<Job> jobs(1);
    Job& J = jobs[0];
    auto setup = [&]() {
        J = Job{};
        J.h = h_lo;
        J.len = hi - lo;
        J.chunk_size = cfg.chunk_size_bytes;
        J.segment_bytes = cfg.segment_bytes;
        J.stream = stream;  // no BGZF framing on a non-seekable input
        J.slab = true;
        J.resume = s > 0;
        J.stop_bit = stop == UINT64_MAX ? UINT64_MAX : stop - lo * 8;
        J.stream_B.assign(1, K.chain - lo * 8);
        for (size_t i = m; i < Gs.size(); ++i) J.stream_B.push_back(Gs[i] - lo * 8);
        if (J.resume) {
            J.body = 0;
            J.win_len = K.win_len;
            J.win_src = c.win.as<uint8_t>();
            J.body_ref = body - lo;  // wraps: trailer offsets count from the first deflate body
            J.carry_raw = K.raw;
            J.carry_len = K.len;
            J.carry_member = K.member;
        }
    };
    setup();
    int rc = decode_one(jobs, so.buf, err);
    if (rc) return rc;
    bool last = stop == UINT64_MAX || J.saw_final;
    if (stop != UINT64_MAX && (J.saw_final || J.dec.failed) && !at_end) {
        // the chain may have run into the end of the slab's input: decode
        // the slab again with the rest of the stream behind it
        if ((rc = rest(lo, hi, h_lo, err))) return rc;
        setup();
        if ((rc = decode_one(jobs, so.buf, err))) return rc;
        last = J.saw_final;
    }
    const uint64_t chain_end = J.chain_end + lo * 8;
    if (c.verbose)
        std::fprintf(stderr,
                     "[rrg-gpu] slab %llu: input bytes %llu..%llu, %zu chunks, %llu output bytes, %.1f MiB of device "
                     "memory\n",
                     (unsigned long long)s, (unsigned long long)(K.chain / 8), (unsigned long long)(chain_end / 8),
                     so.ends.size() + J.dec.chunk_ends.size(), (unsigned long long)J.U,
                     (double)device_bytes(c) / 1048576.0);
    so.base = J.dec.out_base;
    for (uint64_t e : J.dec.chunk_ends) so.ends.push_back(e);
    so.failed = J.dec.failed;
    so.error = J.dec.error;
    so.last = last || so.failed;
    // the streaming path counts no speculation failures (streaming.rs:389-395)
    if (!stream) P.spec += spec_pre + J.dec.spec_failures + (J.n_chain ? spec_first : 0);
    P.slabs++;
    in_done += chain_end / 8 - K.chain / 8;
    out_done += J.U;
    measured(chain_end / 8 - K.chain / 8, J.U);
    if (!so.last) {
        // the window for the next slab: the last 32 KiB of output, which may
        // reach back into this slab's own window
        const uint64_t W = std::min<uint64_t>(kWindow, K.win_len + J.U);
        const uint8_t* ob = (so.buf ? c.out2 : c.out).as<uint8_t>();
        HIPCHK(c.win.ensure(kWindow));
        if (W) HIPCHK(hipMemcpyAsync(c.win.p, ob + J.dec.out_base + J.U - W, W, hipMemcpyDeviceToDevice, c.s));
        HIPCHK(hipStreamSynchronize(c.s));
        K.win_len = W;
        K.raw = J.open_raw;
        K.len = J.open_len;
        K.member = J.members_done;
        K.chain = chain_end;
        K.guess = stop;
    }
    done = so.last;
    P.push(std::move(so));
    return RRG_OK;
}

// One finder launch for exact targets `tv` (bits in c.in) over the input
// [0, data_end); finds at or past scan_end come back as UINT64_MAX.
int exact_finds(Ctx& c, const std::vector<uint64_t>& tv, uint64_t data_end, uint64_t scan_end,
                std::vector<uint64_t>& fv, bool& any_late, rrg_error* err) {
    std::vector<uint64_t> lv(tv.size(), data_end), sv(tv.size(), scan_end), stv, xv;
    for (uint64_t t : tv) stv.push_back(t + kExactCapBits);
    int rc = run_finder(c, tv, lv, sv, fv, err, &stv, &xv);
    if (!rc) rc = finish_open(c, tv, lv, xv, fv, err);
    any_late = false;
    for (uint64_t& f : fv)
        if (f != UINT64_MAX && f >= scan_end) {
            f = UINT64_MAX;
            any_late = true;
        }
    return rc;
}

// ---------------------------------------------------------------------------
// The member path of a file: slabs of chunk targets.
// ---------------------------------------------------------------------------
int Worker::member_path(uint64_t body, rrg_error* err) {
    const uint64_t body_bit = body * 8, end_bit = len * 8;
    const uint64_t chunk_bytes = std::max<size_t>(cfg.chunk_size_bytes, 65536);
    const uint64_t chunk_bits = chunk_bytes * 8;
    // targets k = 1..nT at body + k * chunk_bits, below the end of the file
    // (pipeline.rs:497-503)
    const uint64_t nT = (end_bit - body_bit - 1) / chunk_bits;
    auto target = [&](uint64_t k) { return body_bit + k * chunk_bits; };
    // targets per slab: a slab with n targets holds n + 1 chunks
    auto targets_for = [&](uint64_t input) { return input / chunk_bytes > 1 ? input / chunk_bytes - 1 : 0; };
    // input past a slab's stop target: room for that target's walk and for
    // the decode past the stop to the next block start (more is uploaded
    // when either runs out)
    const uint64_t margin = std::min<uint64_t>(4ull << 20, std::max<uint64_t>(256ull << 10, chunk_bytes / 2));
    auto input_end = [&](uint64_t b) {
        return b > nT ? (uint64_t)len : std::min<uint64_t>(len, target(b) / 8 + margin);
    };
    std::vector<uint64_t> f(nT + 2, UINT64_MAX);  // finds (file bits)
    std::vector<uint8_t> known(nT + 2, 0);

    uint64_t a = 1;             // first target of the slab; b: its stop target
    uint64_t lastB = body_bit;  // the last chunk start kept (file bits)
    Carry K;
    K.chain = K.guess = body_bit;
    uint64_t b = a + targets_for(next_input(0));
    uint64_t lo = 0, hi = input_end(b);
    Upload up;
    up.start(c, c.in2, file, lo, hi);
    const Rest rest = [&](uint64_t lo_, uint64_t& hi_, const uint8_t*& h_lo, rrg_error* e) -> int {
        int r = up.join(e);
        if (!r) r = load_range(c, file, lo_, len, e);
        hi_ = len;
        h_lo = file + lo_;
        return r;
    };
    for (uint64_t s = 0;; ++s) {
        int rc = up.join(err);
        if (rc) return rc;
        if (P.is_closed()) return RRG_OK;
        std::swap(c.in, c.in2);
        c.in_used = hi - lo + 64;
        // 1. the finds of targets a..b, the input growing until each is
        // known; b moves on while its find is not kept
        for (;;) {
            // every target this input can settle: a..b, and the later ones
            // inside it (kept for the slabs after, or for b moving on)
            const uint64_t base = lo * 8;
            const uint64_t data_end = (hi - lo) * 8;
            const bool to_end = hi == len;
            // a candidate needs real input behind it for its verification
            const uint64_t scan_end = to_end ? data_end : data_end - std::min(data_end, kSafeBits);
            std::vector<uint64_t> ks, tv, fv;
            for (uint64_t k = a; k <= nT; ++k) {
                if (k > b && target(k) >= base + scan_end) break;
                if (!known[k]) {
                    ks.push_back(k);
                    tv.push_back(target(k) - base);
                }
            }
            if (!ks.empty()) {
                bool late = false;
                if ((rc = exact_finds(c, tv, data_end, scan_end, fv, late, err))) return rc;
                bool grow = false;
                for (size_t i = 0; i < ks.size(); ++i) {
                    if (fv[i] != UINT64_MAX) {
                        f[ks[i]] = fv[i] + base;
                        known[ks[i]] = 1;
                    } else if (to_end) {
                        known[ks[i]] = 1;  // no block start up to the end of the file
                    } else if (ks[i] <= b) {
                        grow = true;
                    }
                }
                if (grow) {
                    hi = std::min<uint64_t>(len, lo + 4 * (hi - lo) + margin);
                    if ((rc = load_range(c, file, lo, hi, err))) return rc;
                    continue;
                }
            }
            if (c.verbose)
                std::fprintf(stderr, "[rrg-gpu] slab %llu plan: targets %llu..%llu of %llu, input %llu..%llu\n",
                             (unsigned long long)s, (unsigned long long)a, (unsigned long long)b,
                             (unsigned long long)nT, (unsigned long long)lo, (unsigned long long)hi);
            if (b > nT) break;  // no target left: the slab runs to the final block
            uint64_t kept = lastB;
            for (uint64_t k = a; k < b; ++k)
                if (f[k] != UINT64_MAX) kept = std::max(kept, f[k]);
            if (f[b] != UINT64_MAX && f[b] > kept) break;
            ++b;  // target b keeps no chunk start: the slab takes it
            const uint64_t need = input_end(b);
            if (need > hi) {
                hi = need;
                if ((rc = load_range(c, file, lo, hi, err))) return rc;
            }
        }
        const bool last_planned = b > nT;
        // 2. the slab's chunk starts as found (strictly increasing)
        std::vector<uint64_t> Gs{K.guess};
        for (uint64_t k = a; k < b && k <= nT; ++k)
            if (f[k] != UINT64_MAX && f[k] > lastB) {
                Gs.push_back(f[k]);
                lastB = f[k];
            }
        const uint64_t stop = last_planned ? UINT64_MAX : f[b];
        if (!last_planned) lastB = stop;
        // 3. the next slab's input goes up while this one decodes; the next
        // chain starts at or past `stop`, its finds at or past target b + 1
        const uint64_t a_next = b + 1;
        uint64_t b_next = 0, lo_next = 0, hi_next = 0;
        if (!last_planned) {
            const uint64_t slab_in = std::max<uint64_t>(stop / 8 - std::min(stop / 8, K.chain / 8), chunk_bytes);
            b_next = a_next + targets_for(next_input(slab_in));
            lo_next = (std::min(stop, target(a_next)) / 8) & ~7ull;
            hi_next = input_end(b_next);
            up.start(c, c.in2, file, lo_next, hi_next);
        }
        // 4. decode
        bool done = false;
        rc = member_slab(s, body, false, file + lo, lo, hi, hi == len, Gs, stop, K, rest, done, err);
        if (rc || done) {
            (void)up.join(nullptr);
            return rc;
        }
        a = a_next;
        b = b_next;
        lo = lo_next;
        hi = hi_next;
    }
}

// ---------------------------------------------------------------------------
// The member path of a non-seekable input (streaming.rs:278-346): chunk k
// starts at the first verified block at or after chunk k-1's start plus the
// chunk size, so the starts are found one after the other as bytes arrive.
// ---------------------------------------------------------------------------
int Worker::stream_path(PipeSource& src, uint64_t body, rrg_error* err) {
    const uint64_t chunk_bytes = std::max<size_t>(cfg.chunk_size_bytes, 65536);
    const uint64_t chunk_bits = chunk_bytes * 8;
    const uint64_t margin = std::min<uint64_t>(4ull << 20, std::max<uint64_t>(256ull << 10, chunk_bytes / 2));
    Carry K;
    K.chain = K.guess = body * 8;
    uint64_t lo = 0, prev_in = 0;
    // c.in := [lo, hi) of the stream, reading up to `want` first
    auto load = [&](uint64_t want, uint64_t& hi, rrg_error* e) -> int {
        int r = src.fill(want, e);
        if (r) return r;
        hi = std::min(src.end(), want);
        return load_range(c, src.at(lo), 0, hi - lo, e);
    };
    const Rest rest = [&](uint64_t lo_, uint64_t& hi_, const uint8_t*& h_lo, rrg_error* e) -> int {
        int r = src.fill(UINT64_MAX, e);
        if (!r) r = load_range(c, src.at(lo_), 0, src.end() - lo_, e);
        hi_ = src.end();
        h_lo = src.at(lo_);
        return r;
    };
    for (uint64_t s = 0;; ++s) {
        if (P.is_closed()) return RRG_OK;
        const uint64_t want = next_input(prev_in);
        uint64_t hi = 0;
        int rc = load(K.chain / 8 + want + margin, hi, err);
        if (rc) return rc;
        // the chunk starts of the slab, one find after the other, until the
        // slab holds `want` input bytes (the find that ends it is its stop)
        std::vector<uint64_t> Gs{K.guess};
        uint64_t stop = UINT64_MAX;
        for (;;) {
            const bool at_end = src.eof && hi == src.end();
            const uint64_t t = Gs.back() + chunk_bits;
            if (at_end && t >= hi * 8) break;  // no target left: to the final block
            const uint64_t data_end = (hi - lo) * 8;
            const uint64_t scan_end = at_end ? data_end : data_end - std::min(data_end, kSafeBits);
            std::vector<uint64_t> fv;
            bool late = false;
            if (t < lo * 8 + scan_end) {
                if ((rc = exact_finds(c, {t - lo * 8}, data_end, scan_end, fv, late, err))) return rc;
            } else {
                fv.assign(1, UINT64_MAX);
            }
            if (fv[0] == UINT64_MAX) {
                if (at_end) break;  // no block start up to the end of the stream
                // not known in this input: read more
                if ((rc = load(std::max(hi + 4 * (hi - lo), t / 8 + margin), hi, err))) return rc;
                continue;
            }
            const uint64_t fb = fv[0] + lo * 8;
            if (fb / 8 - K.chain / 8 >= want) {
                stop = fb;
                break;
            }
            Gs.push_back(fb);
        }
        const bool at_end = src.eof && hi == src.end();
        bool done = false;
        rc = member_slab(s, body, true, src.at(lo), lo, hi, at_end, Gs, stop, K, rest, done, err);
        if (rc || done) return rc;
        prev_in = std::max<uint64_t>(stop / 8 - std::min(stop / 8, lo), chunk_bytes);
        // the next slab reads from its first chunk's start on
        lo = (stop / 8) & ~7ull;
        src.drop_before(lo);
    }
}

// ---------------------------------------------------------------------------
// BGZF: slabs of whole batches.
// ---------------------------------------------------------------------------
int Worker::bgzf(rrg_error* err) {
    // batch ends (in members) and the ramp target each batch starts with
    const uint64_t tmax = std::max<size_t>(cfg.chunk_size_bytes, 65536);
    std::vector<size_t> batch_end;
    std::vector<uint64_t> batch_target;
    {
        uint64_t tg = 65536, cur = 0, t_first = 65536;
        for (size_t k = 0; k < mem.size(); ++k) {
            cur += mem[k].size;
            if (cur >= tg || k + 1 == mem.size()) {
                batch_end.push_back(k + 1);
                batch_target.push_back(t_first);
                if (cur >= tg) tg = std::min(tg * 2, tmax);
                cur = 0;
                t_first = tg;
            }
        }
    }
    // whole batches from `first` up to about `want` input bytes; the last
    // slab takes everything after its members too (a cut member, or the
    // gzip members that follow the BGZF ones)
    auto plan = [&](size_t first, uint64_t want, size_t& end, uint64_t& lo_, uint64_t& hi_) {
        const size_t m0 = first ? batch_end[first - 1] : 0;
        size_t e = first;
        uint64_t got = 0;
        while (e < batch_end.size() && (e == first || got < want)) {
            const size_t m1 = batch_end[e++];
            got = mem[m1 - 1].off + mem[m1 - 1].size - mem[m0].off;
        }
        end = e;
        lo_ = mem[m0].off;
        hi_ = e == batch_end.size() ? len : mem[batch_end[e - 1] - 1].off + mem[batch_end[e - 1] - 1].size;
    };
    size_t bi = 0, be = 0;  // the slab's batches [bi, be)
    uint64_t lo = 0, hi = 0;
    plan(0, next_input(0), be, lo, hi);
    Upload up;
    up.start(c, c.in2, file, lo, hi);
    for (uint64_t s = 0;; ++s) {
        int rc = up.join(err);
        if (rc) return rc;
        if (P.is_closed()) return RRG_OK;
        std::swap(c.in, c.in2);
        c.in_used = hi - lo + 64;
        const bool last = be == batch_end.size();
        size_t be_next = be;
        uint64_t lo_next = 0, hi_next = 0;
        if (!last) {
            plan(be, next_input(hi - lo), be_next, lo_next, hi_next);
            up.start(c, c.in2, file, lo_next, hi_next);
        }
        SlabOut so;
        so.buf = (int)(s & 1);
        if (!P.wait_free(s)) {
            (void)up.join(nullptr);
            return RRG_OK;
        }
        std::vector<Job> jobs(1);
        Job& J = jobs[0];
        J.h = file + lo;
        J.len = hi - lo;
        J.chunk_size = cfg.chunk_size_bytes;
        J.segment_bytes = cfg.segment_bytes;
        J.bgzf_target0 = batch_target[bi];
        if ((rc = decode_one(jobs, so.buf, err))) {
            (void)up.join(nullptr);
            return rc;
        }
        if (c.verbose)
            std::fprintf(stderr, "[rrg-gpu] slab %llu: BGZF batches %zu..%zu, input bytes %llu..%llu\n",
                         (unsigned long long)s, bi, be, (unsigned long long)lo, (unsigned long long)hi);
        so.base = J.dec.out_base;
        so.ends = J.dec.chunk_ends;
        so.failed = J.dec.failed;
        so.error = J.dec.error;
        so.last = last || so.failed;
        P.spec += J.dec.spec_failures;
        P.slabs++;
        in_done += hi - lo;
        out_done += J.U;
        measured(hi - lo, J.U);
        const bool stop_now = so.last;
        P.push(std::move(so));
        if (stop_now) return up.join(err);
        bi = be;
        be = be_next;
        lo = lo_next;
        hi = hi_next;
    }
}

// Budget, lane cap and worker of one slab decode; the context's lane cap is
// restored when it goes.
struct Setup {
    Ctx& c;
    uint64_t lanes_before;
    uint64_t budget = 0;
    explicit Setup(Ctx& cc) : c(cc), lanes_before(cc.max_lanes) {}
    ~Setup() { c.max_lanes = lanes_before; }
    int init(const rrg_cfg& cfg, rrg_error* err) {
        // the upload and chunk-copy streams, made on the first slab decode
        if (!c.su) HIPCHK(hipStreamCreate(&c.su));
        if (!c.sc) HIPCHK(hipStreamCreate(&c.sc));
        budget = cfg.device_budget_bytes;
        if (!budget) {
            size_t free_b = 0, total_b = 0;
            HIPCHK(hipMemGetInfo(&free_b, &total_b));
            budget = std::max<uint64_t>(free_b / 2, 256ull << 20);
        }
        // a context holding more than an explicit budget from earlier calls
        // gives its buffers back first
        if (cfg.device_budget_bytes && device_bytes(c) > budget) {
            HIPCHK(hipDeviceSynchronize());
            for (DevBuf* d : c.bufs()) d->release();
        }
        // lanes (and their scratch) for the largest slab the budget allows:
        // about one per 16 KiB of its input
        const uint64_t lanes =
            std::max<uint64_t>(1024, max_slab_input(budget, [](uint64_t in) { return slab_bytes(in, 1.0); }) / 16384);
        if (!c.max_lanes || lanes < c.max_lanes) c.max_lanes = lanes;
        return RRG_OK;
    }
    void worker(Worker& W) const {
        W.budget = budget;
        // per-lane scratch, both input buffers' margins and the small buffers
        W.fixed = W.G.scratch_bytes + 2 * (4ull << 20) + (2ull << 20);
    }
};

// The calling thread's side of a slab decode: run `work` on a worker thread
// and hand every chunk of every decoded slab to cb in order (pipeline.rs:
// 921-937), through two pinned staging halves, the copy of chunk k+1 running
// while the callback takes chunk k.  A closed sink ends the decode: Ok
// counting the refused chunk's bytes (pipeline.rs:1028-1030), BrokenPipe for
// BGZF (pipeline.rs:1575-1578).
int run_pipeline(Ctx& c, Pipe& P, const std::function<int(rrg_error*)>& work, rrg_chunk_cb cb, void* user,
                 bool is_bgzf, std::chrono::steady_clock::time_point t0, rrg_stats* stats, uint64_t* compressed,
                 rrg_error* err) {
    std::thread worker([&] {
        rrg_error e{};
        int rc = RRG_OK;
        if (hipSetDevice(c.device) != hipSuccess) {
            rc = RRG_E_DEVICE;
            set_err(&e, RRG_E_DEVICE, "cannot select device %d", c.device);
        } else {
            rc = work(&e);
        }
        std::lock_guard<std::mutex> lk(P.mu);
        P.rc = rc;
        P.err = e;
        P.done = true;
        P.cv.notify_all();
    });
    uint64_t total = 0, sent = 0;
    bool closed = false, first = true;
    int rc = RRG_OK;
    rrg_error ferr{};
    bool failed = false;
    for (;;) {
        SlabOut so;
        {
            std::unique_lock<std::mutex> lk(P.mu);
            P.cv.wait(lk, [&] { return !P.q.empty() || P.done; });
            if (P.q.empty()) break;
            so = std::move(P.q.front());
            P.q.pop_front();
        }
        const uint8_t* ob = (so.buf ? c.out2 : c.out).as<uint8_t>() + so.base;
        uint64_t biggest = 0, prev = 0;
        for (uint64_t e : so.ends) {
            biggest = std::max(biggest, e - prev);
            prev = e;
        }
        const size_t half = (biggest + 64 + 255) & ~(size_t)255;
        hipError_t he = hipSuccess;
        if (2 * half > c.pin_cap) {
            if (c.pin) (void)hipHostFree(c.pin);
            c.pin = nullptr;
            c.pin_cap = 0;
            he = hipHostMalloc((void**)&c.pin, 2 * half, hipHostMallocDefault);
            if (he == hipSuccess) c.pin_cap = 2 * half;
        }
        uint8_t* const buf[2] = {c.pin, c.pin + c.pin_cap / 2};
        auto start_of = [&](size_t k) -> uint64_t { return k ? so.ends[k - 1] : 0; };
        auto issue = [&](size_t k) -> hipError_t {
            const uint64_t a = start_of(k), n = so.ends[k] - a;
            hipError_t e = hipSuccess;
            if (n) e = hipMemcpyAsync(buf[k & 1], ob + a, n, hipMemcpyDeviceToHost, c.sc);
            if (e == hipSuccess) e = hipEventRecord(c.evc[k & 1], c.sc);
            return e;
        };
        if (he == hipSuccess && !so.ends.empty()) he = issue(0);
        for (size_t k = 0; k < so.ends.size() && he == hipSuccess; ++k) {
            const uint64_t n = so.ends[k] - start_of(k);
            if (k + 1 < so.ends.size() && (he = issue(k + 1)) != hipSuccess) break;
            if ((he = hipEventSynchronize(c.evc[k & 1])) != hipSuccess) break;
            if (first) {
                c.prof.ms_first_chunk =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                first = false;
            }
            if (cb && cb(user, buf[k & 1], n)) {
                closed = true;
                total += n;
                break;
            }
            total += n;
            ++sent;
        }
        (void)hipStreamSynchronize(c.sc);  // no copy into the staging buffer outlives the slab
        if (he != hipSuccess) {
            rc = RRG_E_DEVICE;
            set_err(&ferr, RRG_E_DEVICE, "HIP error %s in the chunk copies", hipGetErrorString(he));
            break;
        }
        {
            std::lock_guard<std::mutex> lk(P.mu);
            P.consumed++;
            if (closed) P.closed = true;
            P.cv.notify_all();
        }
        if (closed) break;
        if (so.failed) {
            failed = true;
            ferr = so.error;
            break;
        }
        if (so.last) break;
    }
    {
        // stop the worker if it is still decoding, and wait for it
        std::lock_guard<std::mutex> lk(P.mu);
        P.closed = true;
        P.cv.notify_all();
    }
    worker.join();
    if (!rc && !failed && !closed && P.rc) {
        rc = P.rc;
        ferr = P.err;
    }
    c.prof.uncompressed_bytes = total;
    c.prof.n_slabs = P.slabs;
    if (stats) {
        stats->compressed_bytes = *compressed;
        stats->uncompressed_bytes = total;
        stats->chunks_decoded = sent;
        stats->speculation_failures = P.spec;
    }
    if (rc) {
        if (err) *err = ferr;
        return rc;
    }
    if (closed) {
        if (is_bgzf) {
            set_err(err, RRG_E_IO, "sink closed");
            return RRG_E_IO;
        }
        return RRG_OK;
    }
    if (failed) {
        if (err) *err = ferr;
        return ferr.kind ? ferr.kind : RRG_E_INVALID;
    }
    return RRG_OK;
}

}  // namespace

int decode_slabbed(Ctx& c, const uint8_t* file, size_t len, const rrg_cfg& cfg, rrg_chunk_cb cb, void* user,
                   rrg_stats* stats, rrg_error* err) {
    const auto t0 = std::chrono::steady_clock::now();
    // which form: BGZF batches, or the member path from the first header;
    // a file with neither is left to the whole-file decode
    const bool is_bgzf = bgzf_block_size(file, len) != 0;
    std::vector<Worker::Mb> mem;
    long hl = 0;
    if (is_bgzf) {
        uint64_t pos = 0;
        while (pos < len) {
            const uint32_t sz = bgzf_block_size(file + pos, len - pos);
            if (!sz || pos + sz > len) break;
            mem.push_back({pos, sz});
            pos += sz;
        }
        if (mem.empty()) return kNotSlabbed;
    } else {
        hl = parse_header(file, len, nullptr);
        if (hl < 0 || (uint64_t)hl >= len) return kNotSlabbed;
    }
    Setup S(c);
    int rc = S.init(cfg, err);
    if (rc) return rc;
    Pipe P;
    Worker W{c, file, len, cfg, P, pick_geometry(c, cfg)};
    S.worker(W);
    W.mem = std::move(mem);
    uint64_t compressed = len;
    return run_pipeline(
        c, P, [&](rrg_error* e) { return is_bgzf ? W.bgzf(e) : W.member_path((uint64_t)hl, e); }, cb, user, is_bgzf,
        t0, stats, &compressed, err);
}

int decode_stream_slabbed(Ctx& c, int fd, std::vector<uint8_t>& head, const rrg_cfg& cfg, rrg_chunk_cb cb,
                          void* user, rrg_stats* stats, rrg_error* err) {
    const auto t0 = std::chrono::steady_clock::now();
    PipeSource src;
    src.fd = fd;
    // the first header, read as far as it reaches
    long hl = -RRG_E_TRUNCATED;
    for (uint64_t want = 65536;; want *= 4) {
        int rc = src.fill(want, err);
        if (rc) return rc;
        hl = parse_header(src.at(0), src.end(), nullptr);
        if (hl != -RRG_E_TRUNCATED || src.eof) break;
    }
    if (hl < 0 || (uint64_t)hl >= src.end()) {
        // no deflate body to cut: the caller decodes what was read and the rest
        head.assign(src.at(0), src.at(0) + src.end());
        return kNotSlabbed;
    }
    Setup S(c);
    int rc = S.init(cfg, err);
    if (rc) return rc;
    Pipe P;
    Worker W{c, nullptr, 0, cfg, P, pick_geometry(c, cfg)};
    S.worker(W);
    uint64_t compressed = 0;
    rc = run_pipeline(
        c, P,
        [&](rrg_error* e) {
            const int r = W.stream_path(src, (uint64_t)hl, e);
            compressed = src.end();
            return r;
        },
        cb, user, false, t0, stats, &compressed, err);
    return rc;
}

}  // namespace host
}  // namespace rrg
