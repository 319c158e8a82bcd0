// rrg_launch.h -- host-side launchers for the HIP kernels in rrg_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rrg_device.h"

namespace rrg {

// Block finder: found[i] = first verified dynamic-block offset in
// [targets[i], scan end) (UINT64_MAX when none).  limits[i] without bit 63 is
// the end of the target's file (verification reads up to it); the scan ends
// there or at scan_ends[i] when given.  Without bit 63 the candidates are
// visited in the reference's NEXT_LUT skip order (its exact choice); with bit
// 63 every candidate is tried in bit order and verified stored-block headers
// count as starts too (some block start).  One wavefront per target.
// total_bits: size of the whole (padded) input buffer.  Exact targets may
// take d_stops (windows start below it) and report d_exits (the chain position
// where the walk ended) and d_recs (4 words per target: the chain positions of
// the first 256 bits after the target); each may be null.  Exact targets
// are walked on stream s; with s_seg non-null the segment targets are
// scanned concurrently on s_seg (forked from and joined back into s with the
// two events), else on s after them.  d_exact_idx lists the n_exact exact
// targets (their indices, ascending); the exact walk launches one workgroup
// per entry, so its few workgroups do not wait behind the segment scans for
// dispatch.
hipError_t launch_find_blocks(const uint64_t* d_in64, uint64_t total_bits, const uint64_t* d_targets,
                              const uint64_t* d_limits, const uint64_t* d_scan_ends, const uint64_t* d_stops,
                              uint64_t* d_found, uint64_t* d_exits, uint64_t* d_recs, uint32_t n_targets,
                              const uint32_t* d_exact_idx, uint32_t n_exact, hipStream_t s, hipStream_t s_seg,
                              hipEvent_t fork, hipEvent_t join);
// Exact targets continued in legs: per target, the first verified candidate
// on its chain over its legs (~0 if none) and the chain position after them.
hipError_t launch_leg_fixup(const uint64_t* d_in64, uint64_t total_bits, const LegTarget* d_tg, const uint64_t* d_leg_a,
                            const uint64_t* d_leg_found, const uint64_t* d_leg_exit, const uint64_t* d_leg_rec,
                            uint64_t* d_out_found, uint64_t* d_out_next, uint32_t n, hipStream_t s);

// Pass 1 of the inflate: persistent lane-per-segment Huffman decode of the
// segment indices in d_work[0..n_work) into tokens (rrg_device.h).
struct InflateGeom {
    int lanes_per_block;  // threads per workgroup (one lane decodes one segment); 64 with `wave`
    int blocks;           // persistent workgroups
    int lut_bits;         // literal/length LUT bits (8..10; 9..11 with `wave`)
    bool wave = false;    // one wavefront per segment (k_tokens_wave)
};
int inflate_lds_per_lane(int lut_bits);
int inflate_scratch_u16_per_lane();  // per-lane global scratch (d_sorted), in u16
int inflate_blocks_per_cu(int lut_bits, int lanes);
hipError_t launch_tokens(const uint64_t* d_in64, const SegDesc* d_segs, SegResult* d_res, MemberRec* d_mrec,
                         const uint32_t* d_work, uint32_t n_work, uint32_t* d_counter, uint32_t* d_toks,
                         uint16_t* d_sorted, const InflateGeom& g, hipStream_t s);
// The wave-cooperative form of pass 1 (one wavefront and one set of tables per
// segment; the 64 lanes decode each Huffman block body together): the same
// results and token stream.  d_scratch holds inflate_wave_scratch_bytes() per
// workgroup.
size_t inflate_wave_scratch_bytes();
int inflate_wave_lds_bytes(int lut_bits);
int inflate_wave_blocks_per_cu(int lut_bits);
hipError_t launch_tokens_wave(const uint64_t* d_in64, const SegDesc* d_segs, SegResult* d_res, MemberRec* d_mrec,
                              const uint32_t* d_work, uint32_t n_work, uint32_t* d_counter, uint32_t* d_toks,
                              uint8_t* d_scratch, const InflateGeom& g, hipStream_t s);

// Pass 2: one wavefront per item expands its tokens into u16 cells at
// d_cells[dst] (or bytes at d_out[dst] with EXP_U8); persistent waves take
// the items in order from *d_counter (cleared by the caller).
hipError_t launch_expand(const uint32_t* d_toks, const ExpandItem* d_items, uint32_t n_items, uint32_t* d_counter,
                         const uint8_t* d_in8, uint16_t* d_cells, uint8_t* d_out, int n_cu, hipStream_t s);

// Cut points of the n long items d_items[b] into pieces of about `piece`
// cells: d_cuts[d_first[b] + m - 1] = (index of the token holding cell
// m * piece, that token's first cell) for every m >= 1 with m * piece below
// the item's cells.
hipError_t launch_expand_cuts(const uint32_t* d_toks, const ExpandItem* d_items, const uint32_t* d_first, uint32_t n,
                              uint32_t piece, uint2* d_cuts, hipStream_t s);

// Tail chain: blocked scan over blocks of G chain segments.  d_scratch holds
// tail_scratch_bytes(n_seg, G); d_err[s] is set to 1 for a chain segment s
// whose marker reaches before the start of its file's output (d_err holds
// n_seg words, cleared by the caller).
size_t tail_scratch_bytes(uint32_t n_seg, uint32_t G);
hipError_t launch_tail(const ChainSeg* d_cs, uint32_t n_seg, uint32_t G, uint16_t* d_cells, uint16_t* d_scratch,
                       uint32_t* d_err, hipStream_t s);

// Resolve markers and narrow u16 cells to u8 output (d_err as for launch_tail).
hipError_t launch_resolve(const ResolveTile* d_tiles, uint32_t n_tiles, const ChainSeg* d_cs,
                          const uint16_t* d_cells, uint8_t* d_out, uint32_t* d_err, hipStream_t s);

// first[i] = the first cell of items[i] holding a marker at or beyond its
// reach (unchanged when there is none; the caller fills first[] with ~0u).
hipError_t launch_marker_reach(const ReachScan* d_items, uint32_t n, const uint16_t* d_cells, uint32_t* d_first,
                               hipStream_t s);

// The kernels after the expansion of one batch: tail chain (when `tail`),
// resolve + narrow (n_tiles > 0), CRC pieces (n_pieces > 0), in that order on
// one stream.
struct PostLaunch {
    bool tail = false;
    const ChainSeg* cs = nullptr;
    uint32_t n_seg = 0, G = 1;
    uint16_t* cells = nullptr;
    uint16_t* scratch = nullptr;
    uint32_t* err = nullptr;
    const ResolveTile* tiles = nullptr;
    uint32_t n_tiles = 0;
    uint8_t* out = nullptr;
    const uint64_t* poff = nullptr;
    const uint32_t* plen = nullptr;
    uint32_t n_pieces = 0;
    uint32_t* pcrc = nullptr;
};

// Graphs of those kernels, one per combination of the three parts, each
// built and instantiated once; a batch sets the nodes' grids and arguments
// in the instantiated graph and launches it.
struct PostGraphs {
    hipGraph_t graph[8] = {};
    hipGraphExec_t exec[8] = {};
    hipGraphNode_t node[8][7] = {};
    int n_nodes[8] = {};
    void release();
};

// With g: one graph launch; without: separate launches.
hipError_t launch_post(const PostLaunch& p, PostGraphs* g, hipStream_t s);

// cells[i] = src[i] for i < n: output bytes as marker-free cells (the
// window a resumed decode starts from).
hipError_t launch_widen(const uint8_t* d_src, uint16_t* d_cells, uint64_t n, hipStream_t s);

// Raw (init 0, no final xor) CRC-32 of each piece [off, off+len), len <= 256 KiB.
hipError_t launch_crc_pieces(const uint8_t* d_out, const uint64_t* d_piece_off, const uint32_t* d_piece_len,
                             uint32_t n_pieces, uint32_t* d_crc, hipStream_t s);

}  // namespace rrg
