// rrg_serve.cpp -- batched decode and the serving loop (C ABI).
//
// rrg_gpu_decode_batch lays a set of gzip files out in one device input
// buffer (each file 256-byte aligned and followed by at least 64 zero bytes,
// the padding every kernel's bounded over-read relies on) and runs the batch
// engine once for all of them.
//
// The server turns a stream of requests into such batches:
//
//   submit ──> queue ──[admission: device bytes under the budget]──> stager
//   stager: H2D copy of a batch into a free input buffer (two in rotation)
//   decoder: engine pass over the staged batch while the stager fills the
//            other buffer (continuous batching: whatever arrived meanwhile
//            forms the next batch)
//   completions: per-request output slice of the batch's output buffer, held
//            (and counted against the budget) until the caller releases it.
//
// The admission budget plays the role of the reference's InflightCap
// (pipeline.rs:272-375): it bounds the memory held by work that has been
// started but not yet consumed.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rrg_engine.h"

using namespace rrg;
using namespace rrg::host;

namespace {

constexpr uint64_t kAlign = 256;
constexpr uint64_t kPad = 64;

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// One input of a batch as the loader sees it.
struct Src {
    const void* d;     // device pointer or null
    const uint8_t* h;  // host pointer or null
    uint64_t len;
};

// Lay the files out in `in` and make host copies available for planning.
// Returns the input offsets.
int load_files(Ctx& c, DevBuf& in, hipStream_t s, const std::vector<Src>& src, std::vector<uint64_t>& in_off,
               std::vector<std::vector<uint8_t>>& hcopy, size_t& used, rrg_error* err) {
    const size_t n = src.size();
    in_off.resize(n);
    uint64_t off = 0;
    for (size_t k = 0; k < n; ++k) {
        in_off[k] = off;
        off = align_up(off + src[k].len + kPad);
    }
    used = std::max<uint64_t>(off, kPad);
    HIPCHK(in.ensure(used));
    hcopy.assign(n, {});
    for (size_t k = 0; k < n; ++k) {
        uint8_t* dst = in.as<uint8_t>() + in_off[k];
        const uint64_t len = src[k].len;
        if (len) {
            if (src[k].d) HIPCHK(hipMemcpyAsync(dst, src[k].d, len, hipMemcpyDeviceToDevice, s));
            else HIPCHK(hipMemcpyAsync(dst, src[k].h, len, hipMemcpyHostToDevice, s));
        }
        const uint64_t end = (k + 1 < n) ? in_off[k + 1] : used;
        HIPCHK(hipMemsetAsync(dst + len, 0, end - (in_off[k] + len), s));
        if (!src[k].h && len) {
            hcopy[k].resize(len);
            HIPCHK(hipMemcpyAsync(hcopy[k].data(), src[k].d, len, hipMemcpyDeviceToHost, s));
        }
    }
    HIPCHK(hipStreamSynchronize(s));
    (void)c;
    return RRG_OK;
}

// Engine pass over files already in c.in; fills per-file results.
struct FileResult {
    uint64_t out_off = 0, out_len = 0;
    std::vector<uint64_t> ends;  // relative to out_off
    rrg_stats stats{};
    rrg_error error{};
};

int run_batch(Ctx& c, const rrg_cfg& cfg, const std::vector<Src>& src, const std::vector<uint64_t>& in_off,
              const std::vector<std::vector<uint8_t>>& hcopy, DevBuf& out, std::vector<FileResult>& results,
              rrg_error* err) {
    const size_t n = src.size();
    const Geometry G = pick_geometry(c, cfg);
    std::vector<Job> jobs;
    std::vector<size_t> job_of(n, SIZE_MAX);
    results.assign(n, FileResult{});
    for (size_t k = 0; k < n; ++k) {
        if (src[k].len == 0) {
            set_err(&results[k].error, RRG_E_IO, "empty input");
            continue;
        }
        Job J;
        J.in_off = in_off[k];
        J.len = src[k].len;
        J.h = src[k].h ? src[k].h : hcopy[k].data();
        J.chunk_size = cfg.chunk_size_bytes;
        J.segment_bytes = cfg.segment_bytes;
        J.serial = cfg.num_threads == 1 ? 1 : 0;
        job_of[k] = jobs.size();
        jobs.push_back(std::move(J));
    }
    if (!jobs.empty()) {
        // the batched and serving forms may run tail, resolve and CRC as one graph
        const bool graph_before = c.graph_post;
        c.graph_post = cfg.launch_graph != 0;
        const int rc = decode_batch(c, G, jobs, out, err);
        c.graph_post = graph_before;
        if (rc) return rc;
    }
    for (size_t k = 0; k < n; ++k) {
        FileResult& R = results[k];
        R.stats.compressed_bytes = src[k].len;
        if (job_of[k] == SIZE_MAX) continue;
        const Decoded& d = jobs[job_of[k]].dec;
        R.out_off = d.out_base;
        R.ends = d.chunk_ends;
        R.out_len = R.ends.empty() ? 0 : R.ends.back();
        R.stats.uncompressed_bytes = R.out_len;
        R.stats.chunks_decoded = R.ends.size();
        R.stats.speculation_failures = d.spec_failures;
        if (d.failed) {
            R.error = d.error;
            if (!R.error.kind) R.error.kind = RRG_E_INVALID;
        }
    }
    return RRG_OK;
}

// ---------------------------------------------------------------------------
// server
// ---------------------------------------------------------------------------
double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct OutBlock {  // one batch's output buffer, shared by its requests
    DevBuf buf;
    int live = 0;  // requests not yet released
};

struct Request {
    uint64_t id = 0, tag = 0;
    std::vector<uint8_t> data;
    uint64_t est = 0;  // admission charge while in flight
    std::chrono::steady_clock::time_point t_submit;
    double ms_queue = 0, ms_total = 0;
    // result
    OutBlock* block = nullptr;
    FileResult res;
    uint64_t batch = 0;
    uint32_t batch_items = 0;
    bool done = false, released = false;
};

struct Staged {
    std::vector<Request*> reqs;
    DevBuf in;
    size_t used = 0;
    std::vector<uint64_t> in_off;
    int rc = RRG_OK;  // staging failed (an allocation failure: only this batch fails)
    rrg_error err{};
};

// Decode engines of a server: each a context (stream, buffers) with a thread
// of its own, so that one batch's kernels run while another is being planned,
// verified or laid out on the host.
constexpr int kEngines = 2;

struct Server {
    Ctx c;                     // engine 0 (also the device the server was made on)
    Ctx c1;                    // engine 1
    hipStream_t up = nullptr;  // staging stream
    rrg_cfg cfg{};
    uint64_t budget = 0, max_batch_bytes = 0;
    int max_items = 0;

    std::mutex m;
    std::condition_variable cv;  // any state change
    std::deque<Request*> queue;
    std::deque<std::unique_ptr<Staged>> staged;    // ready for the decoder
    std::vector<std::unique_ptr<Staged>> free_in;  // input buffers not in use
    std::deque<uint64_t> completed;
    std::unordered_map<uint64_t, Request*> reqs;
    std::vector<OutBlock*> free_out, all_out;
    uint64_t next_id = 1, held = 0, pending = 0, batches = 0;
    bool stopping = false, failed = false;
    rrg_error fail_err{};
    std::thread stager, decoder, decoder1;
    int engines = kEngines;

    Ctx& engine(int e) { return e ? c1 : c; }

    // admission charge of a request: the device bytes the engine holds for
    // it while it decodes (input, cell regions and output; the output size
    // is estimated from the last member's ISIZE, at least 4x the input)
    static uint64_t estimate(const std::vector<uint8_t>& d) {
        uint64_t isz = 0;
        const size_t n = d.size();
        if (n >= 4)
            isz = (uint64_t)d[n - 4] | ((uint64_t)d[n - 3] << 8) | ((uint64_t)d[n - 2] << 16) |
                  ((uint64_t)d[n - 1] << 24);
        return engine_footprint(n, std::max<uint64_t>(isz, 4 * (uint64_t)n)) + kAlign + kPad;
    }

    // After an allocation failure: give back every cached device buffer of
    // engine `ec` so that its next batch starts from free memory.
    static void drop_cached(Ctx& ec) {
        for (DevBuf* b : ec.bufs()) b->release();
    }

    void stage_loop() {
        (void)hipSetDevice(c.device);
        for (;;) {
            std::unique_ptr<Staged> st;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] {
                    if (stopping && queue.empty()) return true;
                    if (queue.empty() || free_in.empty()) return false;
                    // FIFO admission: the head request must fit (or nothing is held)
                    return held == 0 || held + queue.front()->est <= budget;
                });
                if (queue.empty()) return;  // stopping
                st = std::move(free_in.back());
                free_in.pop_back();
                uint64_t bytes = 0;
                while (!queue.empty() && (int)st->reqs.size() < max_items) {
                    Request* r = queue.front();
                    if (!st->reqs.empty() && (bytes + r->data.size() > max_batch_bytes || held + r->est > budget))
                        break;
                    if (st->reqs.empty() && held != 0 && held + r->est > budget) break;
                    queue.pop_front();
                    held += r->est;
                    bytes += r->data.size();
                    st->reqs.push_back(r);
                }
            }
            std::vector<Src> src;
            for (Request* r : st->reqs) src.push_back(Src{nullptr, r->data.data(), r->data.size()});
            std::vector<std::vector<uint8_t>> hcopy;
            rrg_error e{};
            t_last_hip = hipSuccess;
            const int rc = load_files(c, st->in, up, src, st->in_off, hcopy, st->used, &e);
            std::lock_guard<std::mutex> lk(m);
            if (rc) {
                st->rc = rc;
                st->err = e;
                if (t_last_hip == hipErrorOutOfMemory) st->in.release();
                else {  // a device fault: nothing decodes after it
                    failed = true;
                    fail_err = e;
                }
            }
            staged.push_back(std::move(st));
            cv.notify_all();
        }
    }

    void decode_loop(int eng) {
        Ctx& c = engine(eng);
        (void)hipSetDevice(c.device);
        for (;;) {
            std::unique_ptr<Staged> st;
            OutBlock* ob = nullptr;
            bool was_failed = false;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return !staged.empty() || (stopping && queue.empty() && pending == 0); });
                if (staged.empty()) return;
                st = std::move(staged.front());
                staged.pop_front();
                if (!free_out.empty()) {
                    ob = free_out.back();
                    free_out.pop_back();
                } else {
                    ob = new OutBlock();
                    all_out.push_back(ob);
                }
                was_failed = failed;
            }
            const auto t_start = std::chrono::steady_clock::now();
            for (Request* r : st->reqs)
                r->ms_queue = std::chrono::duration<double, std::milli>(t_start - r->t_submit).count();
            std::vector<Src> src;
            for (Request* r : st->reqs) src.push_back(Src{nullptr, r->data.data(), r->data.size()});
            std::vector<FileResult> results;
            rrg_error e = st->err;
            int rc = was_failed ? RRG_E_DEVICE : st->rc;
            bool oom = false;
            if (!rc) {
                // the staged buffer becomes the context's input for this pass
                std::swap(c.in, st->in);
                c.in_used = st->used;
                begin_profile(c, 0);
                t_last_hip = hipSuccess;
                rc = run_batch(c, cfg, src, st->in_off, {}, ob->buf, results, &e);
                std::swap(c.in, st->in);
                if (cfg.verbose) {
                    uint64_t in_b = 0, out_b = 0;
                    for (const Request* r : st->reqs) in_b += r->data.size();
                    for (const FileResult& R : results) out_b += R.out_len;
                    c.prof.compressed_bytes = in_b;
                    c.prof.uncompressed_bytes = out_b;
                    c.prof.ms_total = ms_since(t_start);
                    std::fprintf(stderr, "[rrg-gpu] engine %d, batch of %zu requests:\n", eng, st->reqs.size());
                    print_profile(c);
                }
                if (rc && t_last_hip == hipErrorOutOfMemory) {
                    // this batch did not fit: it fails alone, later batches go on
                    oom = true;
                    drop_cached(c);
                    ob->buf.release();
                }
            } else if (!was_failed) {
                oom = true;  // staging ran out of memory
            }
            std::lock_guard<std::mutex> lk(m);
            const uint64_t seq = ++batches;
            for (size_t k = 0; k < st->reqs.size(); ++k) {
                Request* r = st->reqs[k];
                held -= r->est;
                if (rc) {
                    r->res = FileResult{};
                    r->res.error = (failed && !oom) ? fail_err : e;
                    if (!r->res.error.kind) r->res.error.kind = rc;
                } else {
                    r->res = std::move(results[k]);
                    held += r->res.out_len;
                }
                r->block = ob;
                ++ob->live;
                r->batch = seq;
                r->batch_items = (uint32_t)st->reqs.size();
                r->ms_total = ms_since(r->t_submit);
                r->done = true;
                std::vector<uint8_t>().swap(r->data);
                completed.push_back(r->id);
                --pending;
            }
            if (ob->live == 0) free_out.push_back(ob);
            if (rc && !failed && !oom) {
                failed = true;
                fail_err = e;
            }
            st->reqs.clear();
            st->rc = RRG_OK;
            st->err = rrg_error{};
            free_in.push_back(std::move(st));
            cv.notify_all();
        }
    }
};

}  // namespace

extern "C" {

int rrg_gpu_decode_batch(void* ctx, rrg_batch_item* items, size_t n, const rrg_cfg* cfg_in, const uint8_t** d_out,
                         rrg_error* err) {
    if (err) std::memset(err, 0, sizeof *err);
    if (!ctx || (n && !items)) {
        set_err(err, RRG_E_ARG, "null context or items");
        return RRG_E_ARG;
    }
    Ctx& c = *static_cast<Ctx*>(ctx);
    HIPCHK(hipSetDevice(c.device));
    rrg_cfg cfg;
    if (cfg_in) cfg = *cfg_in;
    else rrg_cfg_default(&cfg);
    const double t0 = now_ms();
    std::vector<Src> src(n);
    uint64_t total_in = 0;
    for (size_t k = 0; k < n; ++k) {
        if (items[k].len && !items[k].d_file && !items[k].h_file) {
            set_err(err, RRG_E_ARG, "batch item %zu has no input", k);
            return RRG_E_ARG;
        }
        src[k] = Src{items[k].d_file, items[k].h_file, items[k].len};
        total_in += items[k].len;
    }
    begin_profile(c, total_in);
    std::vector<uint64_t> in_off;
    std::vector<std::vector<uint8_t>> hcopy;
    int rc = load_files(c, c.in, c.s, src, in_off, hcopy, c.in_used, err);
    if (rc) return rc;
    std::vector<FileResult> results;
    rc = run_batch(c, cfg, src, in_off, hcopy, c.out, results, err);
    if (rc) return rc;
    c.batch_ends.assign(n, {});
    uint64_t total_out = 0;
    for (size_t k = 0; k < n; ++k) {
        FileResult& R = results[k];
        c.batch_ends[k] = std::move(R.ends);
        items[k].out_off = R.out_off;
        items[k].out_len = R.out_len;
        items[k].chunk_ends = c.batch_ends[k].data();
        items[k].n_chunks = c.batch_ends[k].size();
        items[k].stats = R.stats;
        items[k].error = R.error;
        total_out += R.out_len;
    }
    c.prof.uncompressed_bytes = total_out;
    c.prof.ms_total = now_ms() - t0;
    if (cfg.verbose) print_profile(c);
    if (d_out) *d_out = c.out.as<uint8_t>();
    return RRG_OK;
}

void* rrg_server_create(int device, const rrg_server_opts* opts) {
    Server* S = new Server();
    rrg_server_opts o{};
    if (opts) o = *opts;
    else rrg_cfg_default(&o.cfg);
    S->engines = o.engines == 1 ? 1 : kEngines;
    if (!init_ctx(S->c, device) || (S->engines > 1 && !init_ctx(S->c1, device)) ||
        hipStreamCreate(&S->up) != hipSuccess) {
        delete S;
        return nullptr;
    }
    if (!o.cfg.chunk_size_bytes) o.cfg.chunk_size_bytes = 4u << 20;
    S->cfg = o.cfg;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    S->budget = o.hbm_budget_bytes ? o.hbm_budget_bytes : free_b / 2;
    S->max_batch_bytes = o.max_batch_bytes ? o.max_batch_bytes : (512ull << 20);
    S->max_items = o.max_batch_items > 0 ? o.max_batch_items : 4096;
    // an input buffer per engine being decoded, and one being staged
    for (int i = 0; i < S->engines + 1; ++i) S->free_in.push_back(std::make_unique<Staged>());
    S->stager = std::thread([S] { S->stage_loop(); });
    S->decoder = std::thread([S] { S->decode_loop(0); });
    if (S->engines > 1) S->decoder1 = std::thread([S] { S->decode_loop(1); });
    return S;
}

uint64_t rrg_server_submit(void* srv, const uint8_t* file, size_t len, uint64_t tag) {
    if (!srv || (len && !file)) return 0;
    Server& S = *static_cast<Server*>(srv);
    Request* r = new Request();
    r->tag = tag;
    r->data.assign(file, file + len);
    r->est = Server::estimate(r->data);
    r->t_submit = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(S.m);
    if (S.stopping) {
        delete r;
        return 0;
    }
    r->id = S.next_id++;
    S.reqs[r->id] = r;
    S.queue.push_back(r);
    ++S.pending;
    S.cv.notify_all();
    return r->id;
}

int rrg_server_poll(void* srv, rrg_completion* out, int max, int timeout_ms) {
    if (!srv || !out || max <= 0) return 0;
    Server& S = *static_cast<Server*>(srv);
    std::unique_lock<std::mutex> lk(S.m);
    auto ready = [&] { return !S.completed.empty(); };
    if (timeout_ms < 0) S.cv.wait(lk, ready);
    else S.cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    int k = 0;
    while (k < max && !S.completed.empty()) {
        const uint64_t id = S.completed.front();
        S.completed.pop_front();
        auto it = S.reqs.find(id);
        if (it == S.reqs.end()) continue;  // released before it was polled
        const Request* r = it->second;
        rrg_completion& o = out[k++];
        std::memset(&o, 0, sizeof o);
        o.id = r->id;
        o.tag = r->tag;
        o.d_out = r->block ? r->block->buf.as<uint8_t>() + r->res.out_off : nullptr;
        o.out_len = r->res.out_len;
        o.n_chunks = r->res.ends.size();
        o.batch = r->batch;
        o.batch_items = r->batch_items;
        o.ms_queue = r->ms_queue;
        o.ms_total = r->ms_total;
        o.stats = r->res.stats;
        o.error = r->res.error;
    }
    return k;
}

size_t rrg_server_chunk_ends(void* srv, uint64_t id, uint64_t* ends, size_t cap) {
    if (!srv) return 0;
    Server& S = *static_cast<Server*>(srv);
    std::lock_guard<std::mutex> lk(S.m);
    auto it = S.reqs.find(id);
    if (it == S.reqs.end() || !it->second->done) return 0;
    const std::vector<uint64_t>& e = it->second->res.ends;
    for (size_t i = 0; i < e.size() && i < cap; ++i) ends[i] = e[i];
    return e.size();
}

int rrg_server_copy_out(void* srv, uint64_t id, uint8_t* host, size_t cap) {
    if (!srv) return RRG_E_ARG;
    Server& S = *static_cast<Server*>(srv);
    const uint8_t* p = nullptr;
    uint64_t n = 0;
    {
        std::lock_guard<std::mutex> lk(S.m);
        auto it = S.reqs.find(id);
        if (it == S.reqs.end() || !it->second->done || it->second->released) return RRG_E_ARG;
        const Request* r = it->second;
        p = r->block->buf.as<uint8_t>() + r->res.out_off;
        n = std::min<uint64_t>(r->res.out_len, cap);
    }
    if (n && hipMemcpy(host, p, n, hipMemcpyDeviceToHost) != hipSuccess) return RRG_E_DEVICE;
    return RRG_OK;
}

int rrg_server_release(void* srv, uint64_t id) {
    if (!srv) return RRG_E_ARG;
    Server& S = *static_cast<Server*>(srv);
    std::lock_guard<std::mutex> lk(S.m);
    auto it = S.reqs.find(id);
    if (it == S.reqs.end() || !it->second->done) return RRG_E_ARG;
    Request* r = it->second;
    S.held -= std::min<uint64_t>(S.held, r->res.out_len);
    if (r->block && --r->block->live == 0) S.free_out.push_back(r->block);
    S.reqs.erase(it);
    // a completion not yet polled is dropped with it
    auto cit = std::find(S.completed.begin(), S.completed.end(), id);
    if (cit != S.completed.end()) S.completed.erase(cit);
    delete r;
    S.cv.notify_all();
    return RRG_OK;
}

uint64_t rrg_server_pending(void* srv) {
    if (!srv) return 0;
    Server& S = *static_cast<Server*>(srv);
    std::lock_guard<std::mutex> lk(S.m);
    return S.pending;
}

void rrg_server_destroy(void* srv) {
    if (!srv) return;
    Server* S = static_cast<Server*>(srv);
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stopping = true;
        // outputs nobody will release no longer hold the budget
        S->held = 0;
        S->cv.notify_all();
    }
    S->stager.join();
    S->decoder.join();
    if (S->decoder1.joinable()) S->decoder1.join();
    (void)hipSetDevice(S->c.device);
    for (auto& kv : S->reqs) delete kv.second;
    for (OutBlock* b : S->all_out) {
        b->buf.release();
        delete b;
    }
    for (auto& st : S->free_in) st->in.release();
    if (S->up) (void)hipStreamDestroy(S->up);
    delete S;
}

}  // extern "C"
