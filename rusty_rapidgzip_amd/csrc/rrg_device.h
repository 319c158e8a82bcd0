// rrg_device.h -- data layout shared by the HIP kernels and the host driver.
//
// Everything here is plain-old-data that lives in HBM.  The host driver fills
// SegDesc, the kernels fill SegResult / MemberRec / cell buffers.
#pragma once
#include <stdint.h>

namespace rrg {

// A "segment" is the unit one lane decodes: from start_bit until the first
// block boundary at or after hint_bit (or until the final member ends).
// Reference counterpart: WorkItem{start_bit,end_bit_hint} (pipeline.rs:377-386).
struct SegDesc {
    uint64_t start_bit;
    uint64_t hint_bit;
    uint64_t tok_base;    // first token of this segment's token region (a multiple of 4)
    uint64_t tok_cap;     // capacity in tokens (a multiple of 4)
    uint64_t end_byte;    // end of the segment's file in the input buffer (bits past it are not read)
    uint32_t flags;       // SEG_SPEC | SEG_SINGLE_MEMBER | SEG_BLOCK_ENDS
    uint32_t mb_base;     // first MemberRec slot
    uint32_t mb_cap;      // MemberRec slots
    uint32_t pad;
};

enum : uint32_t {
    SEG_SPEC = 1u,           // first member may reference bytes before the segment (markers)
    SEG_SINGLE_MEMBER = 2u,  // stop after one member's trailer (BGZF members)
    SEG_BLOCK_ENDS = 4u,     // also record every non-final block end (single-worker chunking)
};

// MemberRec.trailer_byte of a record that marks a non-final block end
constexpr uint64_t MB_BLOCK_END = ~0ull;

struct SegResult {
    uint64_t n_cells;   // output cells the tokens stand for
    uint64_t end_bit;   // actual_end_bit (pipeline.rs:1176-1182)
    uint32_t status;    // ST_*
    uint32_t detail;    // DET_* for ST_INVALID
    uint32_t final_block;
    uint32_t n_mb;      // member boundaries recorded
    uint32_t n_tok;     // tokens produced (with ST_OVERFLOW: tokens needed)
    uint32_t pad;
};

// Tokens (u32) of the first decode pass; bits 31..30 give the type.
//   literals  00: count (1..3) in bits 25..24, the bytes in bits 7..0, 15..8, 23..16
//   match     01: length - 3 in bits 23..16, distance - 1 in bits 14..0
//   stored    10: byte count (1..65535) in bits 15..0; the next token is its
//   offset    11: input byte offset of the stored bytes from the segment's
//                 first input byte, in bits 29..0 (no output of its own)
enum : uint32_t { TOK_LIT = 0u, TOK_MATCH = 1u, TOK_STORED = 2u, TOK_OFFSET = 3u };

// One segment of the expansion pass: its tokens become n_cells u16 cells at
// cells[dst] (or, with EXP_U8, bytes at out[dst]: members that cannot hold
// markers).
struct ExpandItem {
    uint64_t tok_base;
    uint64_t dst;
    uint64_t in_byte;  // the segment's first input byte (stored-token offsets are relative to it)
    uint64_t n_cells;
    uint32_t n_tok;
    uint32_t flags;    // EXP_U8
};
enum : uint32_t { EXP_U8 = 1u };

// Member boundary inside a segment (pipeline.rs:417-427 MemberBoundary), or,
// with SEG_BLOCK_ENDS, a non-final block end (trailer_byte == MB_BLOCK_END).
struct MemberRec {
    uint64_t out_off;       // cell offset in the segment where the member ended
    uint64_t trailer_byte;  // byte offset of the trailer in the body
    uint32_t crc;
    uint32_t isize;
};

enum : uint32_t {
    ST_OK = 0,
    ST_EOF = 1,          // DeflateError::UnexpectedEof
    ST_INVALID = 2,      // DeflateError::Invalid
    ST_BADMAGIC = 3,     // GzipError::BadMagic
    ST_METHOD = 4,       // GzipError::UnsupportedMethod
    ST_RESERVED = 5,     // GzipError::ReservedFlags
    ST_TRUNCATED = 6,    // GzipError::Truncated
    ST_OVERFLOW = 11,    // token capacity exceeded (host retries bigger)
    ST_MB_OVERFLOW = 12, // member-record capacity exceeded (host retries bigger)
    ST_TOO_BIG = 13,     // a segment's output reaches 2^32 cells, or a stored block lies 1 GiB past its start
};

enum : uint32_t {
    DET_NONE = 0,
    DET_HLIT_HDIST,
    DET_CODE16_FIRST,
    DET_CODE16_OVERRUN,
    DET_CODE17_OVERRUN,
    DET_CODE18_OVERRUN,
    DET_BAD_CL_SYMBOL,
    DET_EOB_ZERO,
    DET_OVERSUBSCRIBED,
    DET_NO_CODE,
    DET_LITLEN_RANGE,
    DET_DIST_RANGE,
    DET_DIST_FAR,        // back-reference before the member start
    DET_STORED_NLEN,
    DET_RESERVED_BTYPE,
    DET_MARKER_RANGE,    // marker references bytes before the start of the output
    DET_COUNT
};

// Tile of the resolve kernel: a run of one segment's cells written to out.
struct ResolveTile {
    uint64_t cell_addr;  // absolute cell index of the tile's first cell
    uint64_t out_off;    // global output offset of the tile's first byte
    uint32_t seg;        // chain-order segment index
    uint32_t n;          // cells in the tile
};

// An exact block-finder target continued in legs (k_leg_fixup).
struct LegTarget {
    uint64_t p;      // true chain position where the legs begin
    uint64_t limit;  // the end of the target's file (candidates and verification)
    uint64_t end;    // where the last leg stops
    uint32_t first, n_legs;
};

// A segment whose markers may name bytes of the previous member: `reach`
// bytes of its own member precede it (single-worker paths).
struct ReachScan {
    uint64_t cells_base;  // absolute cell index of the segment's first cell (a multiple of 8)
    uint32_t len;         // cells
    uint32_t reach;       // markers with prefix offset >= reach name bytes before the member
};

// Chain-order segment geometry used by the tail / resolve kernels.
struct ChainSeg {
    uint64_t cells_base;  // absolute cell index of the segment's first cell
    uint64_t len;         // cells
    uint64_t tz_base;     // index of the segment's tail zone in the dense tail arrays
    uint32_t spec;        // first member may hold markers
    uint32_t first;       // first segment of a file's output: nothing before it is reachable
};

}  // namespace rrg
