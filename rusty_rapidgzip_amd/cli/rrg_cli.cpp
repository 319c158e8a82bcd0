// rrg_cli.cpp -- `rapidgzip-gpu`: the command-line drop-in of the reference's
// `rapidgzip-rs` (crates/rusty-rapidgzip-bin/src/main.rs:12-105) over the
// C ABI of librrg_gpu.so.
//
//   rapidgzip-gpu [-P N] [--chunk-size BYTES] [-v] INPUT
//
// INPUT is a .gz file or `-` for stdin; the decompressed stream goes to
// stdout.  As in the reference, `-` becomes /dev/stdin, so a redirected
// regular file is still mapped and a pipe takes the streaming path.  The
// reference recycles drained output buffers back to its workers
// (main.rs:92-97); here every chunk arrives in one of the decoder's reused
// pinned staging buffers and is written to stdout straight from it, so no
// buffer is allocated per chunk either.  With -v a rate line goes to stderr
// about once a second (main.rs:70-91) and the decoder prints its stage
// timings at the end.
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <chrono>
#include <string>

#include "rrg_gpu.h"

namespace {

struct Sink {
    bool verbose = false;
    std::chrono::steady_clock::time_point start, last;
    uint64_t since_last = 0, total = 0;
    int write_errno = 0;
};

double secs(std::chrono::steady_clock::duration d) { return std::chrono::duration<double>(d).count(); }

int on_chunk(void* user, const uint8_t* data, size_t len) {
    Sink& s = *static_cast<Sink*>(user);
    size_t off = 0;
    while (off < len) {
        const ssize_t w = write(STDOUT_FILENO, data + off, len - off);
        if (w < 0) {
            if (errno == EINTR) continue;
            s.write_errno = errno;
            return 1;  // a closed stdout ends the stream, like a dropped receiver
        }
        off += (size_t)w;
    }
    s.since_last += len;
    s.total += len;
    if (s.verbose) {
        const auto now = std::chrono::steady_clock::now();
        const double el = secs(now - s.last);
        if (el >= 1.0) {
            const double mib = 1024.0 * 1024.0;
            fprintf(stderr, "[rapidgzip +%.2fs] %.1f MB/s (avg %.1f MB/s, %.1f MB total)\n", secs(now - s.start),
                    s.since_last / el / mib, s.total / secs(now - s.start) / mib, s.total / mib);
            s.last = now;
            s.since_last = 0;
        }
    }
    return 0;
}

void usage(FILE* f) {
    fprintf(f,
            "usage: rapidgzip-gpu [-P N] [--chunk-size BYTES] [-v] INPUT\n"
            "  INPUT             .gz file, or - for stdin\n"
            "  -P, --threads N   worker threads (0 = auto; 1 = the single-worker chunking)\n"
            "  --chunk-size N    approximate chunk size in bytes (default 4194304)\n"
            "  -v, --verbose     rate lines and stage timings on stderr\n");
}

bool parse_size(const char* s, size_t* out) {
    char* end = nullptr;
    errno = 0;
    const unsigned long long v = strtoull(s, &end, 10);
    if (errno || !end || *end || end == s) return false;
    *out = (size_t)v;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    rrg_cfg cfg;
    rrg_cfg_default(&cfg);
    const char* input = nullptr;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        size_t v = 0;
        if (a == "-h" || a == "--help") {
            usage(stdout);
            return 0;
        } else if (a == "-V" || a == "--version") {
            printf("rapidgzip-gpu (C ABI %d)\n", rrg_gpu_abi_version());
            return 0;
        } else if (a == "-v" || a == "--verbose") {
            cfg.verbose = 1;
        } else if ((a == "-P" || a == "--threads") && i + 1 < argc && parse_size(argv[i + 1], &v)) {
            cfg.num_threads = (int)v;
            ++i;
        } else if (a.rfind("-P", 0) == 0 && a.size() > 2 && parse_size(a.c_str() + 2, &v)) {
            cfg.num_threads = (int)v;
        } else if (a.rfind("--threads=", 0) == 0 && parse_size(a.c_str() + 10, &v)) {
            cfg.num_threads = (int)v;
        } else if (a == "--chunk-size" && i + 1 < argc && parse_size(argv[i + 1], &v)) {
            cfg.chunk_size_bytes = v;
            ++i;
        } else if (a.rfind("--chunk-size=", 0) == 0 && parse_size(a.c_str() + 13, &v)) {
            cfg.chunk_size_bytes = v;
        } else if (!input && (a == "-" || a[0] != '-')) {
            input = argv[i];
        } else {
            fprintf(stderr, "rapidgzip-gpu: bad argument '%s'\n", argv[i]);
            usage(stderr);
            return 2;
        }
    }
    if (!input) {
        usage(stderr);
        return 2;
    }
    const char* path = strcmp(input, "-") == 0 ? "/dev/stdin" : input;
    void* ctx = rrg_gpu_create(0);
    if (!ctx) {
        fprintf(stderr, "rapidgzip-gpu: no usable HIP device\n");
        return 1;
    }
    Sink sink;
    sink.verbose = cfg.verbose != 0;
    sink.start = sink.last = std::chrono::steady_clock::now();
    rrg_stats st;
    rrg_error err;
    const int rc = rrg_gpu_read_gz(ctx, path, &cfg, on_chunk, &sink, &st, &err);
    rrg_gpu_destroy(ctx);
    if (sink.write_errno && sink.write_errno != EPIPE) {
        fprintf(stderr, "rapidgzip-gpu: write failed: %s\n", strerror(sink.write_errno));
        return 1;
    }
    if (rc != RRG_OK) {
        fprintf(stderr, "rapidgzip-gpu: %s\n", err.message);
        return 1;
    }
    if (sink.verbose)
        fprintf(stderr, "[rapidgzip] %llu -> %llu bytes in %llu chunks, %llu speculation failures\n",
                (unsigned long long)st.compressed_bytes, (unsigned long long)st.uncompressed_bytes,
                (unsigned long long)st.chunks_decoded, (unsigned long long)st.speculation_failures);
    return 0;
}
