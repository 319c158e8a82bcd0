/*
 * rrg_gpu.h -- C ABI of the MI355X gzip/DEFLATE decoder (librrg_gpu.so).
 *
 * Plain pointers and sizes only.  One context per host thread; a context owns
 * one HIP stream and its device buffers and is not thread-safe (the same
 * contract as the reference's per-worker native decoders,
 * crates/rusty-rapidgzip/src/libdeflate_ffi.rs:28-35).
 *
 * Reference interfaces replaced (paths relative to crates/rusty-rapidgzip/src/):
 *   rrg_gpu_read_gz        read_gz(path, Sender<Arc<Vec<u8>>>, Config)   lib.rs:366-496
 *   rrg_gpu_decode         read_gz over an in-memory (mmapped) file       lib.rs:435-484
 *   rrg_gpu_decode_fd      read_gz over a non-seekable source (pipe,      lib.rs:378-431,
 *                          stdin): the streaming-parallel chunking        streaming.rs:278-346,389-600
 *   rrg_cfg                Config{num_threads, chunk_size_bytes, verbose} lib.rs:49-100
 *   rrg_stats              DecodeStats                                    lib.rs:102-108
 *   rrg_error              Error / GzipError / DeflateError               lib.rs:110-118, gzip.rs:25-61,
 *                                                                          deflate/mod.rs:22-28
 *   rrg_chunk_cb           Sender::send of one in-order Arc<Vec<u8>>      pipeline.rs:921-937
 *   rrg_gpu_create/destroy the opaque per-worker handle pattern of
 *                          isal_shim.c:191-261 (rrg_isal_alloc/free)
 *   rrg_gpu_autotune       no reference counterpart: picks the GPU decomposition
 *   rrg_gpu_decode_device  no reference counterpart: the device-resident
 *                          form (input and output in HBM) used to measure
 *                          the kernels without PCIe transfers
 *   rrg_gpu_decode_batch   many independent read_gz calls (lib.rs:366-496)
 *                          served by one pass of every kernel; BGZF members
 *                          (pipeline.rs:1357-1636) and speculative segments
 *                          of all files share one persistent work queue
 *   rrg_server_*           serving form: a request loop whose admission
 *                          holds device bytes under a budget, the analogue
 *                          of InflightCap (pipeline.rs:272-375), with
 *                          continuous batching over rrg_gpu_decode_batch
 */
#ifndef RRG_GPU_H
#define RRG_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RRG_GPU_ABI_VERSION 6

/* Error kinds; values match the variants of the reference error enums. */
enum rrg_error_kind {
    RRG_OK = 0,
    RRG_E_EOF = 1,          /* DeflateError::UnexpectedEof */
    RRG_E_INVALID = 2,      /* DeflateError::Invalid(message) */
    RRG_E_BADMAGIC = 3,     /* GzipError::BadMagic */
    RRG_E_METHOD = 4,       /* GzipError::UnsupportedMethod */
    RRG_E_RESERVED = 5,     /* GzipError::ReservedFlags */
    RRG_E_TRUNCATED = 6,    /* GzipError::Truncated */
    RRG_E_CRC = 7,          /* GzipError::CrcMismatch */
    RRG_E_ISIZE = 8,        /* GzipError::IsizeMismatch */
    RRG_E_IO = 9,           /* Error::Io (open/mmap failure, empty file, no final block) */
    RRG_E_DEVICE = 10,      /* HIP runtime failure */
    RRG_E_ARG = 11,         /* bad argument / capacity too small */
};

typedef struct {
    int kind;               /* rrg_error_kind */
    uint32_t expected;      /* CrcMismatch.expected / IsizeMismatch.expected */
    uint32_t got;           /* CrcMismatch.got */
    uint64_t got64;         /* IsizeMismatch.got / CrcMismatch.uncompressed */
    uint32_t member;        /* 0-based gzip member index */
    uint64_t trailer_byte;  /* trailer offset in the deflate body (as the reference reports it) */
    uint32_t method;        /* UnsupportedMethod(u8) / ReservedFlags(u8) */
    char message[192];
} rrg_error;

typedef struct {
    uint64_t compressed_bytes;
    uint64_t uncompressed_bytes;
    uint64_t chunks_decoded;
    uint64_t speculation_failures;
} rrg_stats;

typedef struct {
    size_t chunk_size_bytes;  /* Config.chunk_size_bytes, default 4 MiB */
    int num_threads;          /* Config.num_threads.  1 = the reference's single-worker chunking
                                 (pipeline.rs:481-488, lib.rs:391-408): chunks end at block ends,
                                 no speculation is counted; any other value (0 = auto) = the
                                 speculative path's chunking.  The GPU decode is the same. */
    int verbose;              /* Config.verbose: 0 = Off, 1 = On (stage timings on stderr) */
    /* GPU decomposition (no reference counterpart; 0 = automatic) */
    size_t segment_bytes;     /* compressed bytes per decoded segment */
    int lut_bits;             /* literal/length LUT bits: 0 = automatic (10 with one wavefront per
                                 segment, 9 with one lane per segment); 9-11 and 8-10 respectively */
    int lanes_per_block;      /* 64 (default): one wavefront per segment, its lanes decoding each
                                 Huffman block together; 8, 16 or 32: one lane per segment, that
                                 many segments per workgroup */
    size_t device_budget_bytes; /* rrg_gpu_decode / rrg_gpu_read_gz on a file: device memory the
                                 decode aims to stay within (0 = half the device's free memory).
                                 The file is decoded in slabs of whole chunks sized to it, and the
                                 chunks of each slab go to the callback while the next slab
                                 decodes.  A slab holds at least one chunk, so a chunk whose
                                 output alone exceeds the budget still decodes. */
    int launch_graph;         /* rrg_gpu_decode_batch and the server: the tail, resolve and CRC
                                 kernels of a batch as one launch of a HIP graph built once per
                                 context (1) or as separate launches (0, the default: on a batch
                                 of 64 x 256 KiB files the graph launch measured 0.5 ms slower
                                 than the seven kernel launches it replaces) */
} rrg_cfg;

/* Return nonzero to close the stream (the reference's dropped Receiver). */
typedef int (*rrg_chunk_cb)(void* user, const uint8_t* data, size_t len);

/* Per-call stage timings of the last decode (HIP events on the context's stream). */
typedef struct {
    double ms_total;      /* host wall clock of the whole call */
    double ms_find;       /* block finder kernel */
    double ms_inflate;    /* inflate pass 1 (Huffman decode to tokens), retries included */
    double ms_tail;       /* tail-chain kernels */
    double ms_resolve;    /* resolve + narrow kernel */
    double ms_crc;        /* CRC kernel */
    uint64_t n_segments;
    uint64_t n_redecoded;     /* segments decoded again (false starts, grown output regions) */
    uint64_t n_inflate_launches;
    uint64_t n_tail_rounds;
    uint64_t compressed_bytes;
    uint64_t uncompressed_bytes;
    uint64_t n_grown;         /* token regions grown after an overflow */
    double ms_expand;         /* inflate pass 2 (tokens to cells / bytes) */
    uint64_t n_tokens;        /* tokens of the segments expanded */
    uint64_t find_bits;       /* bit offsets the block finder scanned (targets to finds or scan ends) */
    double ms_first_chunk;    /* host wall clock from the call to the first chunk handed to the callback */
    uint64_t n_slabs;         /* slabs the file was decoded in (host-output decode of a file) */
    uint64_t device_bytes_peak; /* device memory the context held at the peak of the call */
    uint64_t n_graph_launches;  /* graph launches of tail + resolve + CRC (their time is in ms_resolve) */
} rrg_profile;

int rrg_gpu_abi_version(void);
void rrg_cfg_default(rrg_cfg* cfg);

/* Context on HIP device `device`; NULL on failure. */
void* rrg_gpu_create(int device);
void rrg_gpu_destroy(void* ctx);

/* read_gz: decode the gzip file at `path`, handing each in-order output chunk
 * to cb.  Returns RRG_OK or an error kind (details in *err). */
int rrg_gpu_read_gz(void* ctx, const char* path, const rrg_cfg* cfg, rrg_chunk_cb cb, void* user,
                    rrg_stats* stats, rrg_error* err);

/* Same over a file already in host memory.  The file is decoded in slabs of
 * whole chunks under cfg->device_budget_bytes, and each slab's chunks reach
 * cb while the next slab decodes (cb runs on the calling thread, one chunk
 * at a time, in order). */
int rrg_gpu_decode(void* ctx, const uint8_t* file, size_t len, const rrg_cfg* cfg, rrg_chunk_cb cb, void* user,
                   rrg_stats* stats, rrg_error* err);

/* read_gz over a non-seekable source (pipe, stdin, socket): chunk k starts
 * at the first verified block at or after the start of chunk k-1 plus
 * chunk_size (the reference's streaming dispatcher), BGZF is not
 * special-cased and speculation_failures is 0, as on the reference's
 * streaming path.  The input is read as the decode goes, a slab of chunks at
 * a time, and each slab's chunks reach cb before the rest of the input is
 * read; host memory holds about one slab of input.  With num_threads == 1
 * (the single-worker streaming path) the input is read to its end first.
 * rrg_gpu_read_gz takes this path for non-regular and empty files.  The GPU
 * still verifies every chunk start against its predecessor's end. */
int rrg_gpu_decode_fd(void* ctx, int fd, const rrg_cfg* cfg, rrg_chunk_cb cb, void* user, rrg_stats* stats,
                      rrg_error* err);

/* Autotuner (no reference counterpart; the tuning the reference leaves to
 * Config.num_threads).  Decodes the device-resident sample d_file (a complete
 * gzip file) under each combination of decomposition (one lane per segment,
 * 16 per workgroup, lut_bits {8, 9}; one wavefront per segment, lut_bits
 * {9, 10, 11}) and segment_bytes {automatic, 40 KiB, 56 KiB}, on top of `base`
 * (chunk_size_bytes and num_threads are kept), reps timed runs after one
 * warm-up each, and returns the fastest configuration in *best and its time
 * in *best_ms.  Every configuration must produce the same chunk boundaries
 * (the output does not depend on them); a difference is RRG_E_DEVICE. */
int rrg_gpu_autotune(void* ctx, const void* d_file, size_t len, const uint8_t* h_file, const rrg_cfg* base, int reps,
                     rrg_cfg* best, double* best_ms, rrg_error* err);

/* Device-resident form.  d_file: the gzip file in device memory.  h_file:
 * optional host copy (used only to walk BGZF member headers; NULL = copy the
 * headers down on demand).  On return *d_out points at the decompressed bytes
 * in device memory (owned by ctx, valid until the next call), chunk_ends[i] is
 * the end offset of output chunk i (the reference's Vec boundaries), *n_chunks
 * their count (at most chunk_cap are written). */
int rrg_gpu_decode_device(void* ctx, const void* d_file, size_t len, const uint8_t* h_file, const rrg_cfg* cfg,
                          const uint8_t** d_out, uint64_t* out_len, uint64_t* chunk_ends, size_t chunk_cap,
                          size_t* n_chunks, rrg_stats* stats, rrg_error* err);

void rrg_gpu_last_profile(void* ctx, rrg_profile* out);

/* No reference counterpart: rrg_gpu_decode_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The call allocates nothing: every device
 * buffer of the decode comes from one 256 MiB device array of the library,
 * held per device from its first use until the process ends, and the output
 * goes to the caller's d_out (256-byte aligned, out_cap >= decompressed size
 * + 64 bytes).  Inputs of at most 8 MiB (compressed); a larger input, an
 * output buffer too small or a decode that needs more workspace is RRG_E_ARG.
 * The array is shared by every call on a device and keeps the last call's
 * contents: calls that overlap on one device race.  Results are those of
 * rrg_gpu_decode_device. */
int rrg_gpu_decode_device_noalloc(void* ctx, const void* d_file, size_t len, const uint8_t* h_file,
                                  const rrg_cfg* cfg, uint8_t* d_out, size_t out_cap, uint64_t* out_len,
                                  uint64_t* chunk_ends, size_t chunk_cap, size_t* n_chunks, rrg_stats* stats,
                                  rrg_error* err);

/* ---- batched form ----------------------------------------------------- */

/* One file of a batch.  Inputs: d_file (device memory) or, when NULL, h_file
 * (host memory, uploaded); h_file is optional when d_file is given (it spares
 * a device-to-host copy of the headers).  Outputs are written by the call. */
typedef struct {
    const void* d_file;
    const uint8_t* h_file;
    size_t len;
    uint64_t out_off;            /* first decompressed byte of this file in *d_out */
    uint64_t out_len;            /* bytes decoded (the streamed prefix when error.kind != 0) */
    const uint64_t* chunk_ends;  /* chunk end offsets relative to out_off; owned by ctx until its next call */
    size_t n_chunks;
    rrg_stats stats;
    rrg_error error;             /* error.kind == RRG_OK: the file decoded and every trailer checked */
} rrg_batch_item;

/* Decode n independent gzip files with one launch of each kernel.  Each file
 * gets exactly the result, chunk boundaries and error of a separate
 * rrg_gpu_decode_device call.  Returns RRG_OK when the batch ran (per-file
 * errors are in items[i].error) or an error kind for the batch as a whole
 * (device failure, bad argument).  *d_out: the batch output, owned by ctx and
 * valid until its next call. */
int rrg_gpu_decode_batch(void* ctx, rrg_batch_item* items, size_t n, const rrg_cfg* cfg, const uint8_t** d_out,
                         rrg_error* err);

/* ---- serving form ----------------------------------------------------- */

typedef struct {
    size_t hbm_budget_bytes;  /* device bytes that admitted requests may hold (inputs in flight + outputs not yet
                                 released); 0 = half of the free device memory at creation */
    size_t max_batch_bytes;   /* compressed bytes per batch; 0 = 512 MiB */
    int max_batch_items;      /* files per batch; 0 = 4096 */
    rrg_cfg cfg;              /* decode configuration of every request */
    int engines;              /* batches decoded at once, each by a decode context of its own on its own
                                 stream (0 = 2; 1 = one batch at a time) */
} rrg_server_opts;

typedef struct {
    uint64_t id;              /* from rrg_server_submit */
    uint64_t tag;             /* caller's tag */
    const uint8_t* d_out;     /* decompressed bytes in device memory, valid until rrg_server_release(id) */
    uint64_t out_len;
    uint64_t n_chunks;
    uint64_t batch;           /* sequence number of the batch that decoded it */
    uint32_t batch_items;     /* files in that batch */
    double ms_queue;          /* submit -> start of its batch's decode */
    double ms_total;          /* submit -> completion */
    rrg_stats stats;
    rrg_error error;          /* error.kind == RRG_OK on success */
} rrg_completion;

/* Start a server (one device context, a staging thread and a decode thread). */
void* rrg_server_create(int device, const rrg_server_opts* opts);
/* Queue a request; the input is copied, so `file` may be reused on return.
 * Returns its id (>= 1), or 0 when the server is shutting down. */
uint64_t rrg_server_submit(void* srv, const uint8_t* file, size_t len, uint64_t tag);
/* Up to `max` completions in completion order; waits up to timeout_ms (< 0:
 * forever) for the first one.  Returns the number written. */
int rrg_server_poll(void* srv, rrg_completion* out, int max, int timeout_ms);
/* Copy a completed request's chunk ends (relative to its output) into ends[0..cap). */
size_t rrg_server_chunk_ends(void* srv, uint64_t id, uint64_t* ends, size_t cap);
/* Copy a completed request's output to host memory (at most cap bytes); RRG_OK or an error kind. */
int rrg_server_copy_out(void* srv, uint64_t id, uint8_t* host, size_t cap);
/* Return a completed request's output memory to the budget. */
int rrg_server_release(void* srv, uint64_t id);
/* Requests submitted and not yet completed. */
uint64_t rrg_server_pending(void* srv);
/* Finish the queued requests, stop the threads, free device memory. */
void rrg_server_destroy(void* srv);

/* ---- testing hook (no reference counterpart) ---------------------------- */

/* Make the n-th next device allocation of the library fail with HIP's
 * out-of-memory error (once; n <= 0 turns the hook off).  Used by the tests
 * of the server's recovery from an allocation failure. */
void rrg_gpu_debug_fail_alloc(int n);

#ifdef __cplusplus
}
#endif
#endif
