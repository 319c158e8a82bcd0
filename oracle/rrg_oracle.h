/*
 * rrg_oracle.h -- CPU restatement of rusty_rapidgzip's gzip/DEFLATE decode path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU path is
 * compared against (and the timed "port" CPU baseline).  Nothing in the product
 * (rusty_rapidgzip_amd/) links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference crate, crates/rusty-rapidgzip/src/).
 */
#ifndef RRG_ORACLE_H
#define RRG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error kinds (lib.rs:110-118, gzip.rs:25-61, deflate/mod.rs:22-28). */
enum {
    ORC_OK = 0,
    ORC_E_EOF = 1,          /* DeflateError::UnexpectedEof */
    ORC_E_INVALID = 2,      /* DeflateError::Invalid(msg) */
    ORC_E_BADMAGIC = 3,     /* GzipError::BadMagic */
    ORC_E_METHOD = 4,       /* GzipError::UnsupportedMethod */
    ORC_E_RESERVED = 5,     /* GzipError::ReservedFlags */
    ORC_E_TRUNCATED = 6,    /* GzipError::Truncated */
    ORC_E_CRC = 7,          /* GzipError::CrcMismatch */
    ORC_E_ISIZE = 8,        /* GzipError::IsizeMismatch */
    ORC_E_IO = 9,           /* Error::Io (empty input, no final block) */
    ORC_E_NOMEM = 10,
};

typedef struct {
    int kind;
    uint32_t expected, got32;
    uint64_t got64;          /* IsizeMismatch.got / CrcMismatch.uncompressed */
    uint32_t member;
    uint64_t trailer_byte;
    const char* msg;
} orc_error;

typedef struct {
    uint64_t compressed_bytes, uncompressed_bytes, chunks_decoded, speculation_failures;
} orc_stats;

/* ---- primitives ---------------------------------------------------------- */
uint32_t orc_crc32(uint32_t crc, const uint8_t* p, size_t n);
uint32_t orc_crc32_combine(uint32_t crc_a, uint32_t crc_b, uint64_t len_b);

/* gzip.rs:252-300; returns header length or -kind */
long orc_parse_header(const uint8_t* p, size_t n);
/* gzip.rs:311-351; returns BSIZE+1 or 0 when not BGZF */
uint32_t orc_parse_bgzf_block_size(const uint8_t* p, size_t n);

/* block_finder.rs:109-141.  Returns 1 and sets *out when found. */
int orc_find_next_dynamic_block(const uint8_t* data, size_t len, uint64_t from_bit,
                                uint64_t until_bit, uint64_t* out);
/* block_finder.rs:196-262: 1 when the bits at `pos` (a block's BFINAL bit)
 * hold a dynamic header that parses in the finder's strict mode. */
int orc_strict_header_ok(const uint8_t* data, size_t len, uint64_t pos);
/* block_finder.rs:57-101 prefilter table (13-bit window -> skip). */
uint8_t orc_next_lut(uint32_t window13);

/* safe_inflate.rs:617-644 (puff-style textbook inflate of raw DEFLATE).
 * Writes at most out_cap bytes; *out_len = bytes produced, *consumed = input
 * bytes consumed.  Returns ORC_OK or an error kind.  */
int orc_safe_inflate(const uint8_t* in, size_t in_len, uint8_t* out, size_t out_cap,
                     size_t* out_len, size_t* consumed);

/* Speculative chunk decode: pipeline.rs:1080-1185 (decode_one_chunk) with the
 * first member in u16 cells (fast_inflate.rs:1794-2234) and later members as
 * plain bytes.  Cells: bit15 set = marker, low 15 bits = prefix offset; else low
 * 8 bits = byte. */
typedef struct {
    uint64_t out_off;      /* byte offset in the chunk where the member ended */
    uint32_t crc, isize;   /* trailer fields */
    uint64_t trailer_byte; /* byte offset of the trailer in body */
} orc_member_boundary;

typedef struct {
    uint16_t* cells;
    size_t n, cap;
    uint64_t start_bit, end_hint, actual_end_bit;
    int final_block;
    orc_member_boundary* mb;
    size_t n_mb, cap_mb;
    int err;
    const char* msg;
} orc_chunk;

int orc_decode_chunk(const uint8_t* body, size_t len, uint64_t start_bit, uint64_t end_hint,
                     orc_chunk* c);
void orc_chunk_free(orc_chunk* c);

/* speculative.rs:175-213 restated over cells: out[i] = marker ? tail[tl-1-off] : byte.
 * Returns ORC_OK or ORC_E_INVALID (offset >= tail length). */
int orc_resolve_cells(const uint16_t* cells, size_t n, const uint8_t* tail, size_t tail_len,
                      uint8_t* out);

/* read_gz (lib.rs:366-496 + pipeline.rs:468-1041, 1357-1636) over an in-memory
 * file.  Each in-order output chunk is handed to cb (return nonzero = sink
 * closed).  nthreads workers are used for the finder, decode, resolve and
 * CRC stages. */
typedef int (*orc_sink_cb)(void* user, const uint8_t* p, size_t n);
int orc_read_gz(const uint8_t* file, size_t len, size_t chunk_size, int nthreads,
                orc_sink_cb cb, void* user, orc_stats* st, orc_error* err);
/* nthreads == 1 (0 = the machine's parallelism, pipeline.rs:1187-1196) takes
 * the serial path of pipeline.rs:1211-1345 for non-BGZF members. */

/* The streaming serial path for non-seekable inputs with one worker
 * (lib.rs:244-361, dispatched at lib.rs:383-408). */
int orc_read_gz_stream_serial(const uint8_t* file, size_t len, orc_sink_cb cb, void* user,
                              orc_stats* st, orc_error* err);

/* All boundaries the block finder yields for a body (pipeline.rs:491-551),
 * written to out (cap entries, first is 0).  Returns the count. */
size_t orc_chunk_boundaries(const uint8_t* body, size_t len, size_t chunk_size, int nthreads,
                            uint64_t* out, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
