// rrg_kernels.hip -- CDNA4 (gfx950) kernels for the gzip/DEFLATE decode path.
//
// Stages (reference counterparts, paths relative to crates/rusty-rapidgzip/src/):
//   k_find_blocks   block finder          deflate/block_finder.rs:57-301
//   k_tokens        speculative decode    pipeline.rs:1080-1185, deflate/fast_inflate.rs:1794-2234
//   k_expand        LZ77 copies / markers deflate/fast_inflate.rs:1973-2004, 2209-2222
//   k_tail_*        tail-window chain     pipeline.rs:742-856, 1043-1078
//   k_resolve       marker resolution     deflate/speculative.rs:168-213
//   k_crc_pieces    CRC-32 pieces         pipeline.rs:706-733 (crc32fast update)
//
// All of it is integer work: bit extraction, LDS table lookups, byte copies and
// GF(2) arithmetic.  No MFMA.
#include <hip/hip_runtime.h>

#include <cstring>

#include "rrg_device.h"
#include "rrg_hdr.h"
#include "rrg_launch.h"

namespace rrg {

// ---------------------------------------------------------------------------
// Bit access.  The compressed body is stored as little-endian u64 words padded
// with >= 4 zero words, so a 64-bit window at any position up to nbits+128 is
// in bounds.  Bits at or beyond nbits read as zero; every consumer checks
// availability against nbits before acting on what it read.
// ---------------------------------------------------------------------------
// 64 bits of the 128-bit value hi:lo starting at bit s (0..63), branch-free:
// both words are always used, so neither load is left pending on any path
__device__ __forceinline__ uint64_t funnel_lo(uint64_t lo, uint64_t hi, uint32_t s) {
    return (lo >> s) | ((hi << 1) << (63 - s));
}

__device__ __forceinline__ uint64_t bits64_at(const uint64_t* __restrict__ in64, uint64_t pos) {
    const uint64_t q = pos >> 6;
    const uint32_t s = (uint32_t)(pos & 63);
    const uint64_t lo = in64[q];
    const uint64_t hi = in64[q + 1];
    return funnel_lo(lo, hi, s);
}

// Eight consecutive cells from any cell index (gfx950 global memory takes
// 16-byte accesses at 2-byte alignment).
__device__ __forceinline__ uint4 load_cells8(const uint16_t* p) {
    uint4 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}
__device__ __forceinline__ void store_cells8(uint16_t* p, uint4 v) { __builtin_memcpy(p, &v, 16); }

__device__ __forceinline__ uint32_t byte_at(const uint64_t* __restrict__ in64, uint64_t i) {
    return (uint32_t)((in64[i >> 3] >> ((i & 7) * 8)) & 0xffu);
}

// RFC 1951 code-length order (deflate/tables.rs:4-6)
__constant__ uint8_t c_cl_order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
// position of each precode symbol in the header's length order (inverse of c_cl_order)
__constant__ uint8_t c_cl_pos[19] = {3, 17, 15, 13, 11, 9, 7, 5, 4, 6, 8, 10, 12, 14, 16, 18, 0, 1, 2};

// Length / distance bases computed instead of looked up (same values as
// deflate/tables.rs:10-30).
__device__ __forceinline__ uint32_t len_extra(uint32_t i) { return (i < 8 || i == 28) ? 0u : (i - 4) >> 2; }
__device__ __forceinline__ uint32_t len_base(uint32_t i) {
    return i < 8 ? i + 3 : (i == 28 ? 258u : ((4u + (i & 3)) << len_extra(i)) + 3u);
}
__device__ __forceinline__ uint32_t dist_extra(uint32_t d) { return d < 4 ? 0u : (d >> 1) - 1; }
__device__ __forceinline__ uint32_t dist_base(uint32_t d) {
    return d < 4 ? d + 1 : ((2u + (d & 1)) << dist_extra(d)) + 1u;
}

// ===========================================================================
// Block finder (block_finder.rs:109-141).
//
// The reference walks a chain of positions: at p it peeks 13 bits w, jumps by
// NEXT_LUT[w] (the first shift at which the zero-padded window passes the
// BFINAL=0/BTYPE=2/HLIT<=29/HDIST<=29 predicate, 13 if none), and at a jump of
// 0 verifies the candidate, moving on by one bit if it fails.  Because padded
// shifts of 11 and 12 never pass, some offsets are never visited; the answer
// is "the first verified offset ON that chain".  One wavefront per target
// reproduces it exactly: the 64 lanes evaluate step and verification for 64
// consecutive offsets in parallel, then the wave follows the chain through
// that window with readlane hops.
// ===========================================================================
__device__ __forceinline__ bool is_cand13(uint32_t b) {
    return !(b & 1u) && ((b >> 1) & 3u) == 2u && ((b >> 3) & 31u) <= 29u && ((b >> 8) & 31u) <= 29u;
}

// ---------------------------------------------------------------------------
// Finder layout: one wavefront per target, FIND_WAVES wavefronts per workgroup
// sharing the NEXT_LUT.  Each wave keeps a 66-word (4224-bit) region of the
// stream in LDS.  Per 64-offset window every lane evaluates NEXT_LUT and, at
// a jump of 0, the cheap precode test (header fields + strict Kraft over the
// 19 precode lengths); the wave then follows the reference's chain and, at a
// visited offset that passed the cheap test, runs the full strict header parse
// and the 64-symbol trial decode as wave-uniform code.  The first offset that
// passes everything is the answer -- the same order as the reference.
// ---------------------------------------------------------------------------
constexpr int FIND_WAVES = 4;

// Values every lane holds identically, re-declared as wave-uniform so the
// compiler keeps the chain walk in scalar registers with scalar branches.
__device__ __forceinline__ uint32_t uni32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni32((uint32_t)(v >> 32)) << 32) | uni32((uint32_t)v);
}
constexpr int REGION_WORDS = 66;

struct FindWave {
    uint64_t reg[REGION_WORDS];
    uint8_t lens[320];
    uint8_t clut[128];     // precode LUT: (symbol << 3) | length
    uint16_t cnt[32];      // code-length counts: literal/length [0..16), distance [16..32)
    uint16_t sorted[320];  // length-sorted symbols: literal/length [0..288), distance [288..320)
};

// 64 bits at stream position `pos` for wave-uniform code: from the LDS region
// when it holds them, else from global memory.
__device__ __forceinline__ uint64_t ubits(const uint64_t* __restrict__ in64, const FindWave& fw, uint64_t rbase,
                                          uint64_t pos) {
    const uint64_t q = pos >> 6;
    uint64_t lo, hi;
    if (q >= rbase && q + 1 < rbase + REGION_WORDS) {
        lo = fw.reg[q - rbase];
        hi = fw.reg[q - rbase + 1];
    } else {
        lo = in64[q];
        hi = in64[q + 1];
    }
    const uint32_t sh = (uint32_t)(pos & 63);
    return uni64(sh ? (lo >> sh) | (hi << (64 - sh)) : lo);
}

// Canonical decode over a length-sorted symbol table (RFC 1951 3.2.2), with
// the reference's lookup rules (10 bits available for the LUT probe, 15 for
// codes longer than 10).  `cnt` holds the code-length counts in registers
// (wave-uniform).  Returns the symbol or -1.
__device__ __forceinline__ int ucanon(const uint64_t* __restrict__ in64, const FindWave& fw, uint64_t rbase,
                                      uint64_t nbits, uint64_t& pos, const uint32_t (&cnt)[16],
                                      const uint16_t* sorted) {
    if (pos + 10 > nbits) return -1;
    const uint64_t w = ubits(in64, fw, rbase, pos);
    int code = 0, first = 0, index = 0;
#pragma unroll
    for (int l = 1; l <= 15; ++l) {
        code |= (int)((w >> (l - 1)) & 1u);
        const int count = (int)cnt[l];
        if (code - count < first) {
            if (l > 10 && pos + 15 > nbits) return -1;
            pos += (uint64_t)l;
            return (int)uni32(sorted[index + (code - first)]);
        }
        index += count;
        first += count;
        first <<= 1;
        code <<= 1;
    }
    return -1;
}

// counts + length-sorted symbols of lens[0..n) into cnt[0..16) / sorted[0..n)
// (symbol order within a length), by the whole wave: the counts by ballot per
// length, each symbol's place by its rank among the symbols of its length.
// Lane l (1..15) keeps length l's count and running offset.
__device__ void sort_lengths(const uint8_t* lens, int n, uint16_t* cnt, uint16_t* sorted, int lane) {
    const bool owner = lane >= 1 && lane < 16;
    uint32_t my_cnt = 0;
    for (int b = 0; b < n; b += 64) {
        const int sy = b + lane;
        const uint32_t L = sy < n ? lens[sy] : 0u;
        for (int l = 1; l < 16; ++l) {
            const uint32_t c = (uint32_t)__popcll(__ballot(L == (uint32_t)l));
            if (lane == l) my_cnt += c;
        }
    }
    // length l's first place: the counts of lengths 1..l-1
    uint32_t incl = owner ? my_cnt : 0u;
#pragma unroll
    for (int dd = 1; dd < 16; dd <<= 1) {
        const uint32_t o = (uint32_t)__shfl_up((int)incl, dd, 64);
        if (lane >= dd) incl += o;
    }
    uint32_t off = owner ? incl - my_cnt : 0u;
    if (lane < 16) cnt[lane] = (uint16_t)(owner ? my_cnt : 0u);
    const uint64_t below = (1ull << lane) - 1ull;
    for (int b = 0; b < n; b += 64) {
        const int sy = b + lane;
        const uint32_t L = sy < n ? lens[sy] : 0u;
        uint32_t place = 0;
        for (int l = 1; l < 16; ++l) {
            const uint64_t m = __ballot(L == (uint32_t)l);
            if (!m) continue;
            const uint32_t base = uni32((uint32_t)__builtin_amdgcn_readlane((int)off, l));
            if (L == (uint32_t)l) place = base + (uint32_t)__popcll(m & below);
            if (lane == l) off += (uint32_t)__popcll(m);
        }
        if (L) sorted[place] = (uint16_t)sy;
    }
}

// Strict dynamic header (block_finder.rs:196-262) at `p`, the block's first
// bit (BFINAL/BTYPE are the caller's), as wave-uniform code; the lanes
// cooperate to build the precode LUT and to fill runs of code lengths.  On
// success the canonical tables of both codes are in fw and *after is the
// first bit of the block's data.
__device__ bool strict_header(const uint64_t* __restrict__ in64, FindWave& fw, uint64_t rbase, uint64_t nbits,
                              uint64_t p, int lane, uint64_t* after, uint32_t* hdist_out) {
    const uint64_t w0 = ubits(in64, fw, rbase, p);
    const uint32_t hlit = 257u + (uint32_t)((w0 >> 3) & 31u);
    const uint32_t hdist = 1u + (uint32_t)((w0 >> 8) & 31u);
    const uint32_t hclen = 4u + (uint32_t)((w0 >> 13) & 15u);
    // field ranges and the precode (the scan's prefilter has checked them for
    // its candidates; a header behind a block end has not been)
    if (hlit > 286u || hdist > 30u || !precode_ok(w0, ubits(in64, fw, rbase, p + 64), p, nbits)) return false;
    // precode lengths and canonical codes, one symbol per lane (lanes 0..18),
    // counts and ranks by ballot: everything stays in registers.  Every lane
    // then builds the two LUT entries lane and lane+64.
    const uint64_t w1 = ubits(in64, fw, rbase, p + 17);
    const uint32_t k = lane < 19 ? c_cl_pos[lane] : 19u;  // position of symbol `lane` in the header order
    const uint32_t my_len = (k < hclen) ? (uint32_t)((w1 >> (3 * k)) & 7u) : 0u;
    uint32_t my_code = 0;
    {
        uint32_t code = 0, cnt_prev = 0;
        for (uint32_t l = 1; l < 8; ++l) {
            const uint64_t same = __ballot(my_len == l);
            code = (code + cnt_prev) << 1;  // first code of length l
            if (my_len == l) my_code = code + (uint32_t)__popcll(same & ((1ull << lane) - 1));
            cnt_prev = (uint32_t)__popcll(same);
        }
    }
    const uint32_t my_rev = my_len ? (__brev(my_code) >> (32 - my_len)) : 0u;
    {
        uint8_t e0 = 0, e1 = 0;
        for (int sy = 0; sy < 19; ++sy) {
            const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)my_len, sy);
            if (!l) continue;
            const uint32_t rev = (uint32_t)__builtin_amdgcn_readlane((int)my_rev, sy);
            const uint32_t m = (1u << l) - 1;
            if (((uint32_t)lane & m) == rev) e0 = (uint8_t)((sy << 3) | l);
            if (((uint32_t)(lane + 64) & m) == rev) e1 = (uint8_t)((sy << 3) | l);
        }
        fw.clut[lane] = e0;
        fw.clut[lane + 64] = e1;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint64_t pos = p + 17 + 3ull * hclen;
    const uint32_t total = hlit + hdist;
    uint32_t i = 0, last = 0;
    uint32_t sumL = 0, nL = 0, sumD = 0, nD = 0;
    // a run of n lengths L: the lanes fill lens[] together, the Kraft sums
    // and counts follow by arithmetic
    auto put = [&](uint32_t L, uint32_t n) -> bool {
        for (uint32_t b = 0; b < n; b += 64)
            if ((uint32_t)lane < n - b) fw.lens[i + b + (uint32_t)lane] = (uint8_t)L;
        if (L) {
            const uint32_t nl = i < hlit ? min(n, hlit - i) : 0u, nd = n - nl;
            sumL += nl << (15 - L);
            nL += nl;
            sumD += nd << (15 - L);
            nD += nd;
        }
        const bool crossed = i < hlit && i + n >= hlit;
        i += n;
        // oversubscription only grows: reject as soon as it appears
        if ((nL >= 2 && sumL > 32768u) || (nD >= 2 && sumD > 32768u)) return false;
        // the literal/length code is settled once its lengths are all read
        if (crossed) {
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (uni32(fw.lens[256]) == 0 || (nL >= 2 && sumL != 32768u)) return false;
        }
        return true;
    };
    while (i < total) {
        if (pos + 10 > nbits) return false;
        const uint64_t w = ubits(in64, fw, rbase, pos);
        const uint32_t e = uni32(fw.clut[(uint32_t)(w & 127u)]);
        const uint32_t l = e & 7u;
        if (!l) return false;
        pos += l;
        const uint32_t sym = e >> 3;
        const uint64_t x = w >> l;
        if (sym <= 15) {
            last = sym;
            if (!put(sym, 1)) return false;
        } else if (sym == 16) {
            if (i == 0 || pos + 2 > nbits) return false;
            const uint32_t n = 3u + (uint32_t)(x & 3u);
            pos += 2;
            if (i + n > total || !put(last, n)) return false;
        } else if (sym == 17) {
            if (pos + 3 > nbits) return false;
            const uint32_t n = 3u + (uint32_t)(x & 7u);
            pos += 3;
            if (i + n > total || !put(0, n)) return false;
            last = 0;
        } else {
            if (pos + 7 > nbits) return false;
            const uint32_t n = 11u + (uint32_t)(x & 127u);
            pos += 7;
            if (i + n > total || !put(0, n)) return false;
            last = 0;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (uni32(fw.lens[256]) == 0) return false;
    if ((nL >= 2 && sumL != 32768u) || (nD >= 2 && sumD != 32768u)) return false;
    sort_lengths(fw.lens, (int)hlit, fw.cnt, fw.sorted, lane);
    sort_lengths(fw.lens + hlit, (int)hdist, fw.cnt + 16, fw.sorted + 288, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    *after = pos;
    *hdist_out = hdist;
    return true;
}

// Decode up to max_syms symbols of the block whose tables are in fw
// (block_finder.rs:268-301): 0 = an invalid symbol or the end of the input,
// 1 = end of block (pos just after it), 2 = max_syms decoded.
__device__ int trial_syms(const uint64_t* __restrict__ in64, FindWave& fw, uint64_t rbase, uint64_t nbits,
                          uint64_t& pos, int max_syms) {
    uint32_t cL[16], cD[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) {
        cL[l] = uni32(fw.cnt[l]);
        cD[l] = uni32(fw.cnt[16 + l]);
    }
    for (int k = 0; k < max_syms; ++k) {
        const int sym = ucanon(in64, fw, rbase, nbits, pos, cL, fw.sorted);
        if (sym < 0) return 0;
        if (sym <= 255) continue;
        if (sym == 256) return 1;
        if (sym > 285) return 0;
        const uint32_t lx = len_extra((uint32_t)sym - 257);
        if (pos + lx > nbits) return 0;
        pos += lx;
        const int ds = ucanon(in64, fw, rbase, nbits, pos, cD, fw.sorted + 288);
        if (ds < 0 || ds >= 30) return 0;
        const uint32_t dx = dist_extra((uint32_t)ds);
        if (pos + dx > nbits) return 0;
        pos += dx;
    }
    return 2;
}

// The reference's candidate verification (block_finder.rs:175-189): strict
// header and a 64-symbol trial; *hdist_out = the header's HDIST.
__device__ bool deep_verify(const uint64_t* __restrict__ in64, FindWave& fw, uint64_t rbase, uint64_t nbits,
                            uint64_t p, int lane, uint32_t* hdist_out) {
    uint64_t pos;
    if (!strict_header(in64, fw, rbase, nbits, p, lane, &pos, hdist_out)) return false;
    return trial_syms(in64, fw, rbase, nbits, pos, 64) != 0;
}

// deep_verify with the REGION_WORDS words from p's word staged in LDS first:
// the header and the trial decode are a chain of dependent reads, each an LDS
// access instead of a trip to L2/HBM.  Reads past the region go to memory.
__device__ __forceinline__ bool verify_in_region(const uint64_t* __restrict__ in64, FindWave& fw, uint64_t nvalid,
                                                 uint64_t nbits, uint64_t p, int lane, uint32_t* hdist_out) {
    const uint64_t rb = p >> 6;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int k = lane; k < REGION_WORDS; k += 64) fw.reg[k] = rb + k < nvalid ? in64[rb + k] : 0ull;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    return deep_verify(in64, fw, rb, nbits, p, lane, hdist_out);
}

// Segment targets take any true block start, so they may pass over true ones.
// A header with at most this many distance codes constrains random bits far
// less than the rest (a code of one to three symbols is complete almost by
// itself): every false block start seen on 1.3 GB of gzip -6 FASTQ had one.
// The scan for a segment target takes the first verified candidate with more
// distance codes, and the first weaker one only if its range holds none.
constexpr uint32_t kWeakHdist = 4;

// Lanes of one 64-position window that the reference's skip chain visits when
// it enters at the window's first position: binary lifting over the doubling
// tables J (J[k][l] = the position 2^k chain steps after l, 64 = left the
// window) from position 0 toward each lane.
__device__ __forceinline__ uint64_t chain_lanes(const uint32_t (&J)[6], int lane) {
    uint32_t p = 0;
#pragma unroll
    for (int k = 5; k >= 0; --k) {
        const uint32_t q = (uint32_t)__shfl((int)J[k], (int)p, 64);
        if (q <= (uint32_t)lane) p = q;
    }
    return __ballot(p == (uint32_t)lane);
}

// One exact walk (block_finder.rs:109-141): candidates p need p + 13 <= limit,
// verification reads up to nbits, and windows start below `stop`.  With
// `rec`, the chain positions in [rec_base, rec_base + 256) are OR-ed into
// rec0..rec3.  With `mrg`, the walk ends at the first chain position that is
// set in the bitmap m0..m3 based at mrg_base (its own candidates before that
// point are checked; that position and later ones are the other chain's).
struct ExactScan {
    uint64_t stop, limit, nbits;
    bool rec;
    uint64_t rec_base;
    bool mrg;
    uint64_t mrg_base, m0, m1, m2, m3;
};

// Returns the first verified candidate on the chain from `cur` (~0 if none
// before the walk ends); `cur` is left on the chain position after the last
// window walked (every candidate before it was checked), or on the merge
// position, reported in *merged (~0 if none).
__device__ __forceinline__ uint64_t exact_walk(const uint64_t* __restrict__ in64, uint64_t nvalid, FindWave& fw,
                               const uint8_t* s_next, int lane, uint64_t& rbase, uint64_t& cur, const ExactScan& sc,
                               uint64_t& rec0, uint64_t& rec1, uint64_t& rec2, uint64_t& rec3, uint64_t* merged) {
    *merged = ~0ull;
    while (true) {
        const uint64_t base = cur;
        if (base + 13 > sc.limit || base >= sc.stop) return ~0ull;  // the reference's `while pos + 13 <= limit`
        // the window needs bits [base, base + 63 + 128)
        if (rbase == ~0ull || (base >> 6) + 3 >= rbase + REGION_WORDS || (base >> 6) < rbase) {
            rbase = base >> 6;
            for (int k = lane; k < REGION_WORDS; k += 64) fw.reg[k] = (rbase + k < nvalid) ? in64[rbase + k] : 0ull;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
        const uint64_t p = base + (uint64_t)lane;
        const uint32_t wi = (uint32_t)((p >> 6) - rbase);
        const uint32_t sh = (uint32_t)(p & 63);
        const uint64_t a0 = fw.reg[wi], a1 = fw.reg[wi + 1], a2 = fw.reg[wi + 2];
        const uint64_t lo = sh ? (a0 >> sh) | (a1 << (64 - sh)) : a0;
        const uint64_t hi = sh ? (a1 >> sh) | (a2 << (64 - sh)) : a1;
        uint32_t step = 255;
        bool cand = false;
        if (p + 13 <= sc.limit) {
            step = s_next[(uint32_t)(lo & 0x1fffu)];
            if (step == 0) cand = precode_ok(lo, hi, p, sc.nbits);
        }
        const uint64_t cm = __ballot(cand);
        // Jump tables over the window: J[k][l] = the position 2^k chain steps
        // after l (64 = left the window).  Whether a candidate lies on the
        // chain, and where the chain leaves the window, then take six
        // readlanes each instead of a walk of up to 64 dependent steps.
        const uint32_t nxu = (step == 0) ? 1u : step;
        uint32_t J[6];
        J[0] = min((uint32_t)lane + nxu, 64u);
#pragma unroll
        for (int k = 1; k < 6; ++k) {
            const uint32_t prev = J[k - 1];
            const uint32_t nxt = (uint32_t)__shfl((int)prev, (int)min(prev, 63u), 64);
            J[k] = prev >= 64u ? 64u : nxt;
        }
        const uint64_t room = sc.limit - 13 - base;  // base + 13 <= limit here
        const bool rec_now = sc.rec && base < sc.rec_base + 256;
        const bool need_on = rec_now || sc.mrg;
        const uint64_t onm = need_on ? uni64(chain_lanes(J, lane)) : 0ull;
        uint64_t cands = cm;
        uint32_t ml = 64;
        if (sc.mrg) {
            const uint64_t idx = p - sc.mrg_base;  // base >= mrg_base on every walk that merges
            bool hit = false;
            if (p >= sc.mrg_base && idx < 256) {
                const uint32_t w = (uint32_t)(idx >> 6);
                const uint64_t word = w == 0 ? sc.m0 : w == 1 ? sc.m1 : w == 2 ? sc.m2 : sc.m3;
                hit = ((word >> (idx & 63)) & 1ull) != 0;
            }
            const uint64_t hm = __ballot(hit) & onm;
            if (hm) {
                ml = (uint32_t)__builtin_ctzll(hm);
                cands &= (1ull << ml) - 1;  // from the merge position on, the other walk checked them
            }
        }
        uint32_t pth = 0;  // last chain position visited
        while (cands) {
            const uint32_t c = (uint32_t)__builtin_ctzll(cands);
            cands &= cands - 1;
            if (need_on) {
                if (!((onm >> c) & 1ull)) continue;
            } else {
#pragma unroll
                for (int k = 5; k >= 0; --k) {
                    const uint32_t q = uni32((uint32_t)__builtin_amdgcn_readlane((int)J[k], (int)pth));
                    if (q <= c) pth = q;
                }
                if (pth != c) continue;  // the chain jumps over this candidate
            }
            uint32_t hd = 0;
            if (uni32(deep_verify(in64, fw, rbase, sc.nbits, base + c, lane, &hd) ? 1u : 0u)) return base + c;
        }
        if (rec_now) {
            const uint64_t d = base - sc.rec_base;
            const uint32_t w0 = (uint32_t)(d >> 6), s0 = (uint32_t)(d & 63);
            const uint64_t lo_bits = onm << s0, hi_bits = s0 ? onm >> (64 - s0) : 0ull;
            if (w0 == 0) { rec0 |= lo_bits; rec1 |= hi_bits; }
            else if (w0 == 1) { rec1 |= lo_bits; rec2 |= hi_bits; }
            else if (w0 == 2) { rec2 |= lo_bits; rec3 |= hi_bits; }
            else { rec3 |= lo_bits; }
        }
        if (ml < 64) {
            cur = base + ml;
            *merged = cur;
            return ~0ull;
        }
        uint32_t last;
        if (need_on) {
            last = 63u - (uint32_t)__builtin_clzll(onm);  // position 0 is always on the chain
        } else {
#pragma unroll
            for (int k = 5; k >= 0; --k) {
                const uint32_t q = uni32((uint32_t)__builtin_amdgcn_readlane((int)J[k], (int)pth));
                if (q < 64u) pth = q;
            }
            last = pth;
        }
        cur = base + last + uni32((uint32_t)__builtin_amdgcn_readlane((int)nxu, (int)last));
        if (room < 64) return ~0ull;  // every later chain position is past the limit
    }
}

__device__ void fill_next_lut(uint8_t* s_next) {
    for (int i = (int)threadIdx.x; i < 8192; i += (int)blockDim.x) {
        uint32_t w = (uint32_t)i, step = 0;
        while (step < 13 && !is_cand13(w)) {
            w >>= 1;
            ++step;
        }
        s_next[i] = (uint8_t)step;
    }
}

// One wavefront per exact target (limits[t] bit 63 clear; segment targets,
// bit 63 set, are k_find_any's).  limits[t] = the end of the target's file;
// scan_ends[t] bounds the candidates, stops[t] the windows (null = no
// bound).  Reports found[t], exits[t] (the chain position where the walk
// ended) and, if recs is not null, recs[4t..4t+4) = the chain positions of
// the first 256 bits after the target.
__global__ __launch_bounds__(64 * FIND_WAVES) void k_find_blocks(const uint64_t* __restrict__ in64, uint64_t total_bits,
                                                                const uint64_t* __restrict__ targets,
                                                                const uint64_t* __restrict__ limits,
                                                                const uint64_t* __restrict__ scan_ends,
                                                                const uint64_t* __restrict__ stops,
                                                                uint64_t* __restrict__ found, uint64_t* __restrict__ exits,
                                                                uint64_t* __restrict__ recs, uint32_t n_targets) {
    __shared__ uint8_t s_next[8192];
    __shared__ FindWave s_wave[FIND_WAVES];
    const uint32_t t = blockIdx.x * FIND_WAVES + (threadIdx.x >> 6);
    const bool mine = t < n_targets && (limits[t] >> 63) == 0;
    if (!__syncthreads_or(mine)) return;  // no exact target in this workgroup
    fill_next_lut(s_next);
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63);
    if (!mine) return;
    FindWave& fw = s_wave[threadIdx.x >> 6];
    uint64_t cur = uni64(targets[t]);
    // search limit: the end of the target's file (several files may share the
    // buffer); candidates lie below `limit` (where the scan stops) and
    // verification may read up to `nbits`, the end of the file
    const uint64_t nbits = uni64(limits[t]);
    const uint64_t limit = scan_ends ? min(uni64(scan_ends[t]), nbits) : nbits;
    const uint64_t nvalid = (total_bits >> 6) + 2;  // words that exist (the input is padded)
    uint64_t result = ~0ull;
    uint64_t rbase = ~0ull;
    {
        ExactScan sc{};
        sc.stop = stops ? uni64(stops[t]) : ~0ull;
        sc.limit = limit;
        sc.nbits = nbits;
        sc.rec = recs != nullptr;
        sc.rec_base = cur;
        sc.mrg = false;
        uint64_t r0 = 0, r1 = 0, r2 = 0, r3 = 0, merged;
        result = exact_walk(in64, nvalid, fw, s_next, lane, rbase, cur, sc, r0, r1, r2, r3, &merged);
        if (lane == 0) {
            found[t] = result;
            if (exits) exits[t] = cur;
            if (recs) {
                recs[4ull * t] = r0;
                recs[4ull * t + 1] = r1;
                recs[4ull * t + 2] = r2;
                recs[4ull * t + 3] = r3;
            }
        }
        return;
    }
}

// Segment targets (limits[t] bit 63 set): any verified block start will do,
// so the candidates are tried in bit order without the reference's NEXT_LUT
// skip chain, which is only needed where the exact candidate it picks
// matters (k_find_blocks).  One wavefront per workgroup, so a wave that
// finishes early frees its slot at once.
//
// The scan runs in rounds of 4096 positions: lane l takes the 64 offsets of
// word w0 + l.  The 13-bit test runs bit-parallel on the lane's own words:
// bit j of X_k is bit k of offset j, and is_cand13 = !b0 & !b1 & b2 &
// !(b4 & b5 & b6 & b7) & !(b9 & b10 & b11 & b12).  Each lane then runs the
// precode test on its own candidates; the survivors (about one in 2000
// positions) are listed in bit order and, kPendBatch at a time, get the
// strict header parse one per lane; only what passes that gets the wave's
// cooperative verification, in bit order.  The scan therefore runs ahead of
// its find by at most kPendBatch precode survivors.
struct AnyWave : FindWave {
    uint64_t pend[128];  // precode survivors, bit order (bit 63: verified stored header)
    uint32_t pre3[512];  // precode3_entry(k)
};

__device__ __forceinline__ uint64_t zero_bytes(uint64_t v) {  // high bit of each zero byte (and maybe of bytes above one)
    return (v - 0x0101010101010101ull) & ~v & 0x8080808080808080ull;
}

__global__ __launch_bounds__(64) void k_find_any(const uint64_t* __restrict__ in64, uint64_t total_bits,
                                                 const uint64_t* __restrict__ targets,
                                                 const uint64_t* __restrict__ limits,
                                                 const uint64_t* __restrict__ scan_ends, uint64_t* __restrict__ found,
                                                 uint64_t* __restrict__ exits, uint32_t n_targets) {
    __shared__ AnyWave fw;
    const uint32_t t = blockIdx.x;
    if (t >= n_targets) return;
    const uint64_t lim_word = uni64(limits[t]);
    if ((lim_word >> 63) == 0) return;  // an exact target
    const int lane = (int)threadIdx.x;
    const uint64_t nbits = lim_word & ~(1ull << 63);
    const uint64_t limit = scan_ends ? min(uni64(scan_ends[t]), nbits) : nbits;
    const uint64_t nvalid = (total_bits >> 6) + 2;  // words that exist (the input is padded)
    const uint64_t first = uni64(targets[t]);
    for (int k = lane; k < 512; k += 64) fw.pre3[k] = precode3_entry((uint32_t)k);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // A flush costs about its slowest header parse whatever the number of
    // survivors in it: 48 at a time (fewer flushes per target) against 32
    // (less scanning past the block start) measured best.
    constexpr uint32_t kPendBatch = 48, kPendCap = 128;
    uint64_t result = ~0ull, weak = ~0ull;
    uint32_t npend = 0;
    bool end = false;
    auto fetch = [&](uint64_t q) __attribute__((always_inline)) -> uint64_t { return q < nvalid ? in64[q] : 0ull; };
    auto flush2 = [&]() __attribute__((always_inline)) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (uint32_t c0 = 0; c0 < npend && !end; c0 += 64) {
            const uint32_t k = c0 + (uint32_t)lane;
            bool ok = false;
            uint64_t e = 0;
            if (k < npend) {
                e = fw.pend[k];
                ok = (e >> 63) != 0 || strict_header_lane(fetch, nbits, e);
            }
            uint64_t m = __ballot(ok);
            const uint64_t stored = __ballot(ok && (e >> 63));
            while (m) {
                const uint32_t rel = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t lo = uni32((uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)e, (int)rel));
                const uint32_t hi = uni32((uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(e >> 32), (int)rel));
                const uint64_t pos = ((uint64_t)(hi & 0x7fffffffu) << 32) | lo;
                if ((stored >> rel) & 1ull) {
                    result = pos;
                    end = true;
                    break;
                }
                uint32_t hd = 0;
                if (uni32(verify_in_region(in64, fw, nvalid, nbits, pos, lane, &hd) ? 1u : 0u)) {
                    if (uni32(hd) > kWeakHdist) {
                        result = pos;
                        end = true;
                        break;
                    }
                    if (weak == ~0ull) weak = pos;
                }
            }
        }
        npend = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    };
    // positions p of a round: first <= p and p + 13 <= limit
    uint64_t w0 = first >> 6;
    uint64_t N0 = fetch(w0 + lane), N1 = fetch(w0 + lane + 1), N2 = fetch(w0 + lane + 2);
    while (!end && w0 * 64 + 13 <= limit) {
        const uint64_t W0 = N0, W1 = N1, W2 = N2;
        // the next round's words, in flight during this one
        N0 = fetch(w0 + 64 + lane);
        N1 = fetch(w0 + 65 + lane);
        N2 = fetch(w0 + 66 + lane);
        const uint64_t wb = (w0 + (uint64_t)lane) * 64;  // this lane's first offset
        auto X = [&](uint32_t k) __attribute__((always_inline)) -> uint64_t { return (W0 >> k) | (W1 << (64 - k)); };
        uint64_t cm = X(2) & ~(W0 | X(1) | (X(4) & X(5) & X(6) & X(7)) | (X(9) & X(10) & X(11) & X(12)));
        uint64_t valid = ~0ull;
        if (wb < first) valid = first - wb < 64 ? ~0ull << (first - wb) : 0ull;
        if (wb + 13 > limit)
            valid = 0;
        else if (limit - 13 - wb < 63)
            valid &= (2ull << (limit - 13 - wb)) - 1ull;
        cm &= valid;
        // Stored blocks are block starts too (decoding from one is as good as
        // from a dynamic header), and long stored stretches hold no dynamic
        // block at all.  A byte-aligned non-final stored header (low 3 bits
        // 0, LEN == ~NLEN) whose block is followed by another stored header
        // is accepted without further checks (2^-32 per byte on random data).
        // Bytes whose LEN/NLEN agree are found bit-parallel first.
        uint64_t sm = 0;
        {
            const uint64_t a_lo = (W0 >> 8) | (W1 << 56), a_hi = W1 >> 8;    // bytes k + 1
            const uint64_t b_lo = (W0 >> 24) | (W1 << 40), b_hi = W1 >> 24;  // bytes k + 3
            const uint64_t t_lo = a_lo ^ b_lo, t_hi = a_hi ^ b_hi;
            const uint64_t u = t_lo & ((t_lo >> 8) | (t_hi << 56));  // byte k: (b[k+1]^b[k+3]) & (b[k+2]^b[k+4])
            uint64_t hits = zero_bytes(~u) & zero_bytes(W0 & 0x0707070707070707ull);
            while (hits) {
                const uint32_t j = (uint32_t)__builtin_ctzll(hits) - 7;  // the byte's bit offset
                hits &= hits - 1;
                const uint64_t p = wb + j;
                if (!((valid >> j) & 1ull) || p + 40 > limit) continue;
                const uint64_t v = (W0 >> j) | (j ? W1 << (64 - j) : 0ull);
                if ((v & 7u) != 0 || (((v >> 8) ^ (v >> 24)) & 0xffffu) != 0xffffu) continue;
                const uint64_t nb = (p >> 3) + 5 + ((v >> 8) & 0xffffu);  // next header byte
                if ((nb + 5) * 8 > nbits) continue;
                const uint32_t h = byte_at(in64, nb), l0 = byte_at(in64, nb + 1), l1 = byte_at(in64, nb + 2),
                               n0 = byte_at(in64, nb + 3), n1 = byte_at(in64, nb + 4);
                if ((h & 6u) == 0 && ((l0 | (l1 << 8)) ^ (n0 | (n1 << 8))) == 0xffffu) sm |= 1ull << j;
            }
        }
        // the precode test of this lane's candidates (bits [p, p + 128) are
        // in W0..W2); a verified stored header passes as it is
        uint64_t pm = sm;
        for (uint64_t m = cm & ~sm; m; m &= m - 1) {
            const uint32_t j = (uint32_t)__builtin_ctzll(m);
            const uint64_t lo = j ? (W0 >> j) | (W1 << (64 - j)) : W0;
            const uint64_t hi = j ? (W1 >> j) | (W2 << (64 - j)) : W1;
            if (precode_ok_lut(fw.pre3, lo, hi, wb + j, nbits)) pm |= 1ull << j;
        }
        // list the survivors in bit order (lane order, then offset order),
        // at most kPendCap at a time
        const uint32_t mine = (uint32_t)__popcll(pm);
        uint32_t before = mine;  // inclusive prefix sum over the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)before, d, 64);
            if (lane >= d) before += o;
        }
        const uint32_t total = uni32((uint32_t)__shfl((int)before, 63, 64));
        before -= mine;
        for (uint32_t done = 0; done < total && !end;) {
            const uint32_t take = min(total - done, kPendCap - npend);
            uint32_t r = before;
            for (uint64_t m = pm; m; m &= m - 1, ++r) {
                if (r < done || r >= done + take) continue;
                const uint32_t j = (uint32_t)__builtin_ctzll(m);
                fw.pend[npend + (r - done)] = (wb + j) | (((sm >> j) & 1ull) << 63);
            }
            npend += take;
            done += take;
            if (npend >= kPendBatch) flush2();
        }
        w0 += 64;
    }
    if (!end && npend) flush2();
    if (result == ~0ull) result = weak;
    if (lane == 0) {
        found[t] = result;
        if (exits) exits[t] = min(w0 * 64, limit);
    }
}

// Exact targets without chain records (first walks): the reference's skip
// chain (block_finder.rs:109-141) in rounds of 4096 positions, lane l
// walking the chain through word w0 + l by NEXT_LUT steps.  A lane's entry
// is the previous lane's exit; the lanes first walk from guessed entries (the
// start of their word) and walk again from the previous lane's exit until no
// entry changes -- chains from nearby entries merge within a few steps, so
// this takes two or three walks, and the fixed point is the true chain (lane
// 0 enters at the true position, and each lane's entry equals its
// predecessor's exit).  The chain positions with a jump of 0 are the
// candidates; each lane runs the precode test on its own, the survivors are
// listed in chain order and, kPendBatch at a time, get the strict header
// parse one per lane and then the wave's verification in order: the first
// that verifies is the find.  The walk ends where the reference's loop would
// (p + 13 > limit) or at the first chain position at or past stops[t];
// exits[t] is that position.
__global__ __launch_bounds__(64) void k_find_exact(const uint64_t* __restrict__ in64, uint64_t total_bits,
                                                   const uint64_t* __restrict__ targets,
                                                   const uint64_t* __restrict__ limits,
                                                   const uint64_t* __restrict__ scan_ends,
                                                   const uint64_t* __restrict__ stops, uint64_t* __restrict__ found,
                                                   uint64_t* __restrict__ exits, uint32_t n_targets,
                                                   const uint32_t* __restrict__ exact_idx, uint32_t n_exact) {
    __shared__ uint8_t s_next[8192];
    __shared__ AnyWave fw;
    if (blockIdx.x >= n_exact) return;
    // these few latency-bound walks share the SIMDs with the segment scans
    // (VALU bound): the wave priority lets them issue first when ready
    __builtin_amdgcn_s_setprio(3);
    const uint32_t t = exact_idx[blockIdx.x];
    if (t >= n_targets) return;
    const uint64_t nbits = uni64(limits[t]);
    if (nbits >> 63) return;  // a segment target
    fill_next_lut(s_next);
    __syncthreads();
    const int lane = (int)threadIdx.x;
    const uint64_t limit = scan_ends ? min(uni64(scan_ends[t]), nbits) : nbits;
    const uint64_t stop = stops ? uni64(stops[t]) : ~0ull;
    const uint64_t nvalid = (total_bits >> 6) + 2;  // words that exist (the input is padded)
    constexpr uint32_t kPendBatch = 32, kPendCap = 128;
    auto fetch = [&](uint64_t q) __attribute__((always_inline)) -> uint64_t { return q < nvalid ? in64[q] : 0ull; };
    uint64_t result = ~0ull;
    uint32_t npend = 0;
    bool end = false;
    auto flush = [&]() __attribute__((always_inline)) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (uint32_t c0 = 0; c0 < npend && !end; c0 += 64) {
            const uint32_t k = c0 + (uint32_t)lane;
            bool ok = false;
            uint64_t e = 0;
            if (k < npend) {
                e = fw.pend[k];
                ok = strict_header_lane(fetch, nbits, e);
            }
            uint64_t m = __ballot(ok);
            while (m) {
                const uint32_t rel = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t lo = uni32((uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)e, (int)rel));
                const uint32_t hi = uni32((uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(e >> 32), (int)rel));
                const uint64_t pos = ((uint64_t)hi << 32) | lo;
                uint32_t hd = 0;
                if (uni32(verify_in_region(in64, fw, nvalid, nbits, pos, lane, &hd) ? 1u : 0u)) {
                    result = pos;
                    end = true;
                    break;
                }
            }
        }
        npend = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    };
    uint64_t P = uni64(targets[t]);  // the chain position entering the round
    while (!end) {
        if (P + 13 > limit || P >= stop) break;  // the walk ends at P
        const uint64_t w0 = P >> 6;
        const uint64_t W0 = fetch(w0 + lane), W1 = fetch(w0 + lane + 1), W2 = fetch(w0 + lane + 2);
        const uint64_t ws = (w0 + (uint64_t)lane) * 64, we = ws + 64;
        // the chain through this lane's word from `entry`: visited offsets V,
        // candidates C (jump 0), and X = the first position past the word or
        // where the walk ends
        uint64_t V = 0, C = 0, X = 0;
        auto walk = [&](uint64_t entry) __attribute__((always_inline)) {
            V = 0;
            C = 0;
            uint64_t pos = entry;
            while (pos < we && pos + 13 <= limit && pos < stop) {
                const uint32_t o = (uint32_t)(pos - ws);  // entry >= ws on every walk that reaches here
                const uint32_t w13 = (uint32_t)((o ? (W0 >> o) | (W1 << (64 - o)) : W0) & 0x1fffu);
                const uint32_t step = s_next[w13];
                V |= 1ull << o;
                if (step == 0) C |= 1ull << o;
                pos += step ? step : 1;
            }
            X = pos;
        };
        uint64_t entry = lane == 0 ? P : ws;
        walk(entry);
        while (true) {
            const uint64_t px = ((uint64_t)(uint32_t)__shfl_up((int)(uint32_t)(X >> 32), 1, 64) << 32) |
                                (uint32_t)__shfl_up((int)(uint32_t)X, 1, 64);
            const uint64_t ne = lane == 0 ? P : px;
            const bool changed = ne != entry;
            if (!__ballot(changed)) break;
            if (changed) {
                entry = ne;
                walk(entry);
            }
        }
        // candidates in chain order: the precode test per lane, survivors listed
        uint64_t pm = 0;
        for (uint64_t m = C; m; m &= m - 1) {
            const uint32_t j = (uint32_t)__builtin_ctzll(m);
            const uint64_t lo = j ? (W0 >> j) | (W1 << (64 - j)) : W0;
            const uint64_t hi = j ? (W1 >> j) | (W2 << (64 - j)) : W1;
            if (precode_ok(lo, hi, ws + j, nbits)) pm |= 1ull << j;
        }
        const uint32_t mine = (uint32_t)__popcll(pm);
        uint32_t before = mine;  // inclusive prefix sum over the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)before, d, 64);
            if (lane >= d) before += o;
        }
        const uint32_t total = uni32((uint32_t)__shfl((int)before, 63, 64));
        before -= mine;
        for (uint32_t done = 0; done < total && !end;) {
            const uint32_t take = min(total - done, kPendCap - npend);
            uint32_t r = before;
            for (uint64_t m = pm; m; m &= m - 1, ++r) {
                if (r < done || r >= done + take) continue;
                fw.pend[npend + (r - done)] = ws + (uint64_t)__builtin_ctzll(m);
            }
            npend += take;
            done += take;
            if (npend >= kPendBatch) flush();
        }
        // the chain leaves the round at lane 63's exit (or ended earlier, in
        // which case every later lane's exit is that same position)
        P = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(X >> 32), 63, 64) << 32) |
            (uint32_t)__shfl((int)(uint32_t)X, 63, 64);
        P = uni64(P);
    }
    if (!end && npend) flush();
    if (lane == 0) {
        found[t] = result;
        if (exits) exits[t] = P;
    }
}

// Exact targets that a bounded first scan left open, continued in legs: leg
// j of a target starts at a_j (leg 0 at the target's true chain position P)
// and was walked by its own wavefront, recording its first verified
// candidate, its exit and its first 256 chain positions.  The skip chain is a
// function of the position, so two walks that share one position agree from
// there on.  One wavefront per target walks the true chain from P into each
// leg until it meets that leg's recorded positions (a few bits, usually),
// checking the candidates on the way, and then takes over the leg's result.

__global__ __launch_bounds__(64) void k_leg_fixup(const uint64_t* __restrict__ in64, uint64_t total_bits,
                                                  const LegTarget* __restrict__ tg, const uint64_t* __restrict__ leg_a,
                                                  const uint64_t* __restrict__ leg_found,
                                                  const uint64_t* __restrict__ leg_exit,
                                                  const uint64_t* __restrict__ leg_rec, uint64_t* __restrict__ out_found,
                                                  uint64_t* __restrict__ out_next, uint32_t n) {
    __shared__ uint8_t s_next[8192];
    __shared__ FindWave fw;
    fill_next_lut(s_next);
    __syncthreads();
    const uint32_t t = blockIdx.x;
    if (t >= n) return;
    const int lane = (int)threadIdx.x;
    const LegTarget T = tg[t];
    const uint64_t nvalid = (total_bits >> 6) + 2;
    uint64_t rbase = ~0ull, P = uni64(T.p), result = ~0ull;
    uint64_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    for (uint32_t j = 0; j < T.n_legs && P + 13 <= T.limit; ++j) {
        const uint32_t k = T.first + j;
        const uint64_t a = uni64(leg_a[k]);
        const uint64_t leg_end = j + 1 < T.n_legs ? uni64(leg_a[k + 1]) : T.end;
        const uint64_t lf = uni64(leg_found[k]);
        if (P == a) {  // this leg's walk was the true chain
            if (lf != ~0ull) {
                result = lf;
                break;
            }
            P = uni64(leg_exit[k]);
            continue;
        }
        ExactScan sc{};
        sc.stop = a + 256;
        sc.limit = T.limit;
        sc.nbits = T.limit;
        sc.rec = false;
        sc.mrg = true;
        sc.mrg_base = a;
        sc.m0 = uni64(leg_rec[4ull * k]);
        sc.m1 = uni64(leg_rec[4ull * k + 1]);
        sc.m2 = uni64(leg_rec[4ull * k + 2]);
        sc.m3 = uni64(leg_rec[4ull * k + 3]);
        uint64_t cur = P, merged;
        uint64_t f = exact_walk(in64, nvalid, fw, s_next, lane, rbase, cur, sc, r0, r1, r2, r3, &merged);
        if (f != ~0ull) {
            result = f;
            break;
        }
        if (merged != ~0ull) {
            // from here the true chain is the leg's: its first verified
            // candidate (at or after the merge position) or its exit
            if (lf != ~0ull) {
                result = lf;
                break;
            }
            P = uni64(leg_exit[k]);
            continue;
        }
        // no merge within the recorded positions: walk the rest of the leg
        ExactScan sc2{};
        sc2.stop = leg_end;
        sc2.limit = T.limit;
        sc2.nbits = T.limit;
        f = exact_walk(in64, nvalid, fw, s_next, lane, rbase, cur, sc2, r0, r1, r2, r3, &merged);
        if (f != ~0ull) {
            result = f;
            break;
        }
        P = cur;
    }
    if (lane == 0) {
        out_found[t] = result;
        out_next[t] = P;
    }
}

hipError_t launch_find_blocks(const uint64_t* d_in64, uint64_t total_bits, const uint64_t* d_targets,
                              const uint64_t* d_limits, const uint64_t* d_scan_ends, const uint64_t* d_stops,
                              uint64_t* d_found, uint64_t* d_exits, uint64_t* d_recs, uint32_t n_targets,
                              const uint32_t* d_exact_idx, uint32_t n_exact, hipStream_t s, hipStream_t s_seg,
                              hipEvent_t fork, hipEvent_t join) {
    if (n_targets == 0) return hipSuccess;
    hipError_t e;
    // the exact walks go first: they are few but long, and their workgroups
    // must not queue behind the many segment scans
    hipStream_t ss = s;
    if (s_seg) {
        if ((e = hipEventRecord(fork, s)) != hipSuccess) return e;
        ss = s_seg;
    }
    if (d_recs)  // legs: chain records and merges need the window walk
        hipLaunchKernelGGL(k_find_blocks, dim3((n_targets + FIND_WAVES - 1) / FIND_WAVES), dim3(64 * FIND_WAVES), 0, s,
                           d_in64, total_bits, d_targets, d_limits, d_scan_ends, d_stops, d_found, d_exits, d_recs,
                           n_targets);
    else if (n_exact)
        hipLaunchKernelGGL(k_find_exact, dim3(n_exact), dim3(64), 0, s, d_in64, total_bits, d_targets, d_limits,
                           d_scan_ends, d_stops, d_found, d_exits, n_targets, d_exact_idx, n_exact);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (s_seg && (e = hipStreamWaitEvent(s_seg, fork, 0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_find_any, dim3(n_targets), dim3(64), 0, ss, d_in64, total_bits, d_targets, d_limits,
                       d_scan_ends, d_found, d_exits, n_targets);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (s_seg) {
        if ((e = hipEventRecord(join, s_seg)) != hipSuccess || (e = hipStreamWaitEvent(s, join, 0)) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_leg_fixup(const uint64_t* d_in64, uint64_t total_bits, const LegTarget* d_tg, const uint64_t* d_leg_a,
                            const uint64_t* d_leg_found, const uint64_t* d_leg_exit, const uint64_t* d_leg_rec,
                            uint64_t* d_out_found, uint64_t* d_out_next, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_leg_fixup, dim3(n), dim3(64), 0, s, d_in64, total_bits, d_tg, d_leg_a, d_leg_found,
                       d_leg_exit, d_leg_rec, d_out_found, d_out_next, n);
    return hipGetLastError();
}

// ===========================================================================
// Speculative inflate, pass 1 (k_tokens): one lane decodes one segment's
// Huffman symbols into tokens.  Persistent grid; lanes pull segment indices
// from a device work queue as they finish.  No output bytes are touched here:
// a lane's only global loads are its own input words, which it requests about
// 200 bits ahead of use, so a lane waits on LDS table lookups, not on memory.
// The copies happen in pass 2 (k_expand), cooperatively and only for the
// segments that the host's verification chain accepts.
//
// Per-lane LDS: literal/length LUT (2^LB u16), distance LUT (128 u16, also the
// 7-bit precode LUT while a header is read), code lengths (320 B), counts and
// canonical-code scratch.  LUT entry = (symbol << 4) | code length; a zero
// entry means a code longer than the LUT (or an unassigned code) and goes to
// the canonical slow path, which reads the length-sorted symbols that the
// builder stored in this lane's global scratch.
//
// A back-reference that reaches before the segment's first byte (first member
// only) is kept as a match; pass 2 turns its cells into markers
// 0x8000 | prefix_offset (fast_inflate.rs:2209-2222).  Later members must stay
// inside themselves (libdeflate_ffi.rs:64-123).
// ===========================================================================
constexpr int DIST_LB = 7;

// Per-lane LDS: literal/length LUT (2^LB u16 entries (sym << 4) | len),
// distance LUT (128 u8 entries (sym << 3) | len, len <= 7; the precode LUT
// while a header is read), four 16-entry u16 arrays (code-length counts of
// both tables, canonical next code and sorted offset while a table is built),
// the symbols of the codes the LUTs do not hold, in canonical order (the
// first kLongLit literal/length ones, u16, and every distance one, u8), and
// per table where the canonical walk for those codes starts (first code of
// length LUT bits + 1 | symbols with shorter codes << 16).
constexpr int kLongLit = 48;
constexpr int kLongDist = 32;
template <int LB>
struct LaneTables {
    static constexpr int kLit = 1 << LB;
    static constexpr int kBytes = kLit * 2 + (1 << DIST_LB) + 4 * 32 + 2 * kLongLit + kLongDist + 16;
};
// per-lane global scratch, in u16: 320 length-sorted symbols, 320 B of code
// lengths, 32 B of precode lengths
constexpr int LANE_SCRATCH_U16 = 496;

int inflate_scratch_u16_per_lane() { return LANE_SCRATCH_U16; }

int inflate_lds_per_lane(int lut_bits) {
    switch (lut_bits) {
        case 8: return LaneTables<8>::kBytes;
        case 10: return LaneTables<10>::kBytes;
        default: return LaneTables<9>::kBytes;
    }
}

// Canonical table build (huffman.rs:215-325 semantics, incomplete accepted).
// Returns false when oversubscribed with >= 2 symbols.
template <typename E, typename L>
__device__ bool build_table(const uint8_t* __restrict__ lens, int n, E* __restrict__ lut, int lbits,
                            uint16_t* __restrict__ cnt, uint16_t* __restrict__ tnext, uint16_t* __restrict__ toffs,
                            uint16_t* __restrict__ sorted, L* __restrict__ longs, int long_cap,
                            uint32_t* __restrict__ head) {
    // entry: (symbol << 4) | length in u16 tables, (symbol << 3) | length in u8 ones
    constexpr int kSh = sizeof(E) == 1 ? 3 : 4;
    for (int l = 0; l < 16; ++l) cnt[l] = 0;
    for (int s = 0; s < n; ++s) cnt[lens[s]]++;
    cnt[0] = 0;
    uint32_t total = 0, nsym = 0, nshort = 0;
    for (int l = 1; l <= 15; ++l) {
        total += (uint32_t)cnt[l] << (15 - l);
        nsym += cnt[l];
        if (l <= lbits) nshort += cnt[l];
    }
    if (nsym >= 2 && total > 32768u) return false;
    uint32_t code = 0, off = 0;
    for (int l = 1; l <= 15; ++l) {
        code = (code + cnt[l - 1]) << 1;
        tnext[l] = (uint16_t)code;
        toffs[l] = (uint16_t)off;
        off += cnt[l];
    }
    if (head) *head = (lbits < 15 ? (uint32_t)tnext[lbits + 1] : 0u) | (nshort << 16);
    uint32_t* lut32 = reinterpret_cast<uint32_t*>(lut);
    for (int i = 0; i < (int)(((1u << lbits) * sizeof(E)) / 4); ++i) lut32[i] = 0;
    for (int s = 0; s < n; ++s) {
        const int l = lens[s];
        if (!l) continue;
        const uint32_t c = tnext[l];
        tnext[l] = (uint16_t)(c + 1);
        const uint32_t at = toffs[l];
        if (sorted) sorted[at] = (uint16_t)s;
        toffs[l] = (uint16_t)(at + 1);
        if (l <= lbits) {
            const uint32_t rev = __brev(c) >> (32 - l);
            const E e = (E)((s << kSh) | l);
            for (uint32_t idx = rev; idx < (1u << lbits); idx += 1u << l) lut[idx] = e;
        } else if (longs && at - nshort < (uint32_t)long_cap) {
            longs[at - nshort] = (L)s;
        }
    }
    return true;
}

// Slow path for a LUT miss: the canonical walk (RFC 1951 3.2.2) over the
// code lengths above the LUT's LBITS, from the first code of length LBITS + 1
// and the count of shorter codes (`head`, from build_table).  The symbol
// comes from the LDS list of long-code symbols, or, past its kcap entries,
// from the length-sorted table in the lane's global scratch.  With `seq`
// nonzero the long codes stand for consecutive symbols from seq (the fixed
// literal code's 9-bit codes: 144-255).
template <int LBITS, typename L>
__device__ __forceinline__ bool decode_long(uint64_t bits, const uint16_t* __restrict__ cnt,
                                            const uint16_t* __restrict__ sorted, const L* __restrict__ longs,
                                            uint32_t kcap, uint32_t head, uint32_t seq, uint32_t* sym,
                                            uint32_t* len) {
    const uint32_t nshort = head >> 16;
    int first = (int)(head & 0xffffu), index = (int)nshort;
    int code = (int)(__brev((uint32_t)bits) >> (31 - LBITS));  // the first LBITS + 1 bits, first bit highest
#pragma unroll
    for (int l = LBITS + 1; l <= 15; ++l) {
        const int count = cnt[l];
        if (code - count < first) {
            const uint32_t at = (uint32_t)(index + (code - first));
            uint32_t v;
            if (seq) {
                v = seq + (at - nshort);
            } else if (at - nshort < kcap) {
                v = longs[at - nshort];
            } else {
                v = sorted[at];
                // use the loaded value here: the wait for this load then stays on
                // the slow path instead of becoming a full vmcnt wait on every symbol
                asm volatile("" : "+v"(v));
            }
            *sym = v;
            *len = (uint32_t)l;
            return true;
        }
        index += count;
        first += count;
        first <<= 1;
        code = (code << 1) | (int)((bits >> l) & 1u);
    }
    return false;
}

__device__ int parse_gz_header(const uint64_t* __restrict__ in64, uint64_t len, uint64_t at, uint64_t* hl) {
    // gzip.rs:252-300
    const uint64_t n = len - at;
    if (n < 10) return ST_TRUNCATED;
    if (byte_at(in64, at) != 0x1f || byte_at(in64, at + 1) != 0x8b) return ST_BADMAGIC;
    if (byte_at(in64, at + 2) != 8) return ST_METHOD;
    const uint32_t flg = byte_at(in64, at + 3);
    if (flg & 0xe0u) return ST_RESERVED;
    uint64_t p = 10;
    if (flg & 4u) {
        if (p + 2 > n) return ST_TRUNCATED;
        const uint64_t xlen = byte_at(in64, at + p) | (byte_at(in64, at + p + 1) << 8);
        p += 2 + xlen;
        if (p > n) return ST_TRUNCATED;
    }
    for (uint32_t f = 8; f <= 16; f <<= 1) {
        if (flg & f) {
            while (p < n && byte_at(in64, at + p) != 0) ++p;
            if (p >= n) return ST_TRUNCATED;
            ++p;
        }
    }
    if (flg & 2u) {
        if (p + 2 > n) return ST_TRUNCATED;
        p += 2;
    }
    *hl = p;
    return ST_OK;
}

enum : int { S_FETCH = 0, S_BLOCK_START, S_SYM, S_BLOCK_END, S_EXIT, S_FINISH };

// Largest cell count a segment may reach (its cell positions are 32-bit).
constexpr uint32_t kMaxSegCells = 0xF0000000u;

template <int LB>
// 3 waves per SIMD (<= 168 VGPRs): the per-lane LDS tables allow about as many
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_tokens(
    const uint64_t* __restrict__ in64, const SegDesc* __restrict__ segs, SegResult* __restrict__ res,
    MemberRec* __restrict__ mrec, const uint32_t* __restrict__ work, uint32_t n_work, uint32_t* __restrict__ counter,
    uint32_t* __restrict__ toks, uint16_t* __restrict__ sorted_scratch) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int LMASK = (1 << LB) - 1;
    const int tid = (int)threadIdx.x;
    uint8_t* my = lds + tid * LaneTables<LB>::kBytes;
    uint16_t* litlut = reinterpret_cast<uint16_t*>(my);
    uint8_t* distlut = my + 2 * (1 << LB);
    uint16_t* cntL = reinterpret_cast<uint16_t*>(distlut + (1 << DIST_LB));
    uint16_t* cntD = cntL + 16;
    uint16_t* tnext = cntD + 16;
    uint16_t* toffs = tnext + 16;
    uint16_t* longL = toffs + 16;
    uint8_t* longD = reinterpret_cast<uint8_t*>(longL + kLongLit);
    uint32_t* lhead = reinterpret_cast<uint32_t*>(longD + kLongDist);  // [0] literal/length, [1] distance
    // code lengths and length-sorted symbols live in per-lane global scratch: they
    // are touched once per block header, not per symbol
    uint16_t* sortedL = sorted_scratch + ((size_t)blockIdx.x * blockDim.x + tid) * LANE_SCRATCH_U16;
    uint16_t* sortedD = sortedL + 288;
    uint8_t* lens = reinterpret_cast<uint8_t*>(sortedL + 320);
    uint8_t* cltmp = reinterpret_cast<uint8_t*>(sortedL + 480);  // precode lengths while a header is read

    int state = S_FETCH;
    uint32_t seg = 0;
    uint64_t pos = 0, hint = 0, bitbuf = 0, seg_byte = 0;
    uint64_t nbits = 0;  // end of the segment's file in the input buffer (bits)
    uint32_t bitcnt = 0;
    // the input window of the symbol loop: C0, C1, C2 = in64[ci], in64[ci + 1],
    // in64[ci + 2].  A refill reads C0 and C1; C2 is the word loaded ahead,
    // at the previous word crossing, so a refill waits for a load issued
    // about 64 input bits earlier, never for one it issued itself.
    uint64_t ci = 0, C0 = 0, C1 = 0, C2 = 0;
    uint32_t* tk = toks;
    uint32_t tok_cap = 0, n_tok = 0;
    uint32_t out_n = 0, member_base = 0;
    // up to three literals wait in lit_acc; up to three tokens wait in q1..q3
    // (q0 is the oldest) and are stored four at a time as one 16-byte write
    uint32_t lit_acc = 0, lit_n = 0, q0 = 0, q1 = 0, q2 = 0, q3 = 0, nq = 0;
    uint32_t spec = 0, mflags = 0, bfinal = 0, n_mb = 0, mb_base = 0, mb_cap = 0, fixed_seq = 0;
    // counting: the token region is full; decode on without storing so the
    // segment reports the size it needs (one regrow instead of repeated ones)
    uint32_t counting = 0;
    uint32_t status = ST_OK, detail = DET_NONE, final_block = 0;

// 64 bits at pos (the input buffer is padded, so words up to 8 past the end
// exist).  The asm use completes the loads inside the refill, so that code
// merging with a refill-free path does not inherit a pending-load wait.
#define RRG_REFILL()                      \
    do {                                  \
        bitbuf = bits64_at(in64, pos);    \
        asm volatile("" : "+v"(bitbuf));  \
        bitcnt = 64;                      \
    } while (0)
// the same from the symbol loop's window, moved on to pos first: one word per
// crossing (the common case), or three fresh words after a jump
#define RRG_REFILL_W()                                                              \
    do {                                                                            \
        const uint64_t _q = pos >> 6;                                               \
        if (_q != ci) {                                                             \
            uint32_t _st = 3;                                                       \
            if (_q > ci && _q - ci < 3) _st = (uint32_t)(_q - ci);                   \
            ci = _q;                                                                \
            for (uint32_t _k = _st; _k > 0; --_k) {                                 \
                C0 = C1;                                                            \
                C1 = C2;                                                            \
                C2 = in64[ci + 3 - _k];                                             \
            }                                                                       \
        }                                                                           \
        bitbuf = funnel_lo(C0, C1, (uint32_t)(pos & 63));                           \
        bitcnt = 64;                                                                \
    } while (0)
#define RRG_CONSUME(n)           \
    do {                         \
        const uint32_t _n = (n); \
        bitbuf >>= _n;           \
        bitcnt -= _n;            \
        pos += _n;               \
    } while (0)
#define RRG_FAIL(st, det)   \
    do {                    \
        status = (st);      \
        detail = (det);     \
        state = S_FETCH;    \
        goto finish;        \
    } while (0)
// read n (<= 32) bits with the reference's availability rule
#define RRG_GETBITS(dst, n)                                         \
    do {                                                            \
        const uint32_t _k = (n);                                    \
        if (pos + _k > nbits) RRG_FAIL(ST_EOF, DET_NONE);           \
        if (bitcnt < _k) RRG_REFILL();                              \
        (dst) = _k ? (uint32_t)(bitbuf & ((1ull << _k) - 1)) : 0u; \
        RRG_CONSUME(_k);                                            \
    } while (0)
// one token into the queue; every fourth completes an aligned 16-byte group
#define RRG_EMIT(t)                                                                   \
    do {                                                                              \
        if (n_tok >= tok_cap) counting = 1;                                           \
        if (!counting) {                                                              \
            q0 = q1;                                                                  \
            q1 = q2;                                                                  \
            q2 = q3;                                                                  \
            q3 = (t);                                                                 \
            if (++nq == 4) {                                                          \
                *reinterpret_cast<uint4*>(tk + n_tok - 3) = make_uint4(q0, q1, q2, q3); \
                nq = 0;                                                               \
            }                                                                         \
        }                                                                             \
        ++n_tok;                                                                      \
    } while (0)
#define RRG_FLUSH_LITS()                                             \
    do {                                                             \
        if (lit_n) {                                                 \
            RRG_EMIT((TOK_LIT << 30) | (lit_n << 24) | lit_acc);     \
            lit_n = 0;                                               \
            lit_acc = 0;                                             \
        }                                                            \
    } while (0)

// One Huffman symbol into the token stream; CHK = 1 applies the reference's
// end-of-input checks.  `break` leaves the symbol loop at end-of-block,
// `continue` goes on with the next symbol.
#define RRG_SYMBOL(CHK)                                                                          \
    {                                                                                            \
        uint32_t e = litlut[bitbuf & LMASK];                                                     \
        uint32_t l = e & 15u, sym = e >> 4;                                                      \
        if (l == 0) {                                                                            \
            if ((CHK) && pos + 15 > nbits) RRG_FAIL(ST_EOF, DET_NONE);                           \
            if (!decode_long<LB>(bitbuf, cntL, sortedL, longL, kLongLit, lhead[0], fixed_seq, &sym, &l)) \
                RRG_FAIL(ST_INVALID, DET_NO_CODE);                                               \
        }                                                                                        \
        if ((CHK) && pos + l > nbits) RRG_FAIL(ST_EOF, DET_NONE);                                \
        RRG_CONSUME(l);                                                                          \
        if (sym < 256) {                                                                         \
            lit_acc |= sym << (8 * lit_n);                                                       \
            if (++lit_n == 3) {                                                                  \
                RRG_EMIT((TOK_LIT << 30) | (3u << 24) | lit_acc);                                \
                lit_n = 0;                                                                       \
                lit_acc = 0;                                                                     \
            }                                                                                    \
            ++out_n;                                                                             \
            continue;                                                                            \
        }                                                                                        \
        if (sym == 256) {                                                                        \
            state = S_BLOCK_END;                                                                 \
            break;                                                                               \
        }                                                                                        \
        if (sym > 285) RRG_FAIL(ST_INVALID, DET_LITLEN_RANGE);                                   \
        const uint32_t li = sym - 257;                                                           \
        const uint32_t lx = len_extra(li);                                                       \
        if ((CHK) && pos + lx > nbits) RRG_FAIL(ST_EOF, DET_NONE);                               \
        const uint32_t length = len_base(li) + (uint32_t)(bitbuf & ((1ull << lx) - 1));          \
        RRG_CONSUME(lx);                                                                         \
        uint32_t de = distlut[bitbuf & ((1u << DIST_LB) - 1)];                                   \
        uint32_t dl = de & 7u, dsym = de >> 3;                                                   \
        if (dl == 0) {                                                                           \
            if ((CHK) && pos + 15 > nbits) RRG_FAIL(ST_EOF, DET_NONE);                           \
            if (!decode_long<DIST_LB>(bitbuf, cntD, sortedD, longD, kLongDist, lhead[1], 0u, &dsym, &dl)) \
                RRG_FAIL(ST_INVALID, DET_NO_CODE);                                               \
        }                                                                                        \
        if ((CHK) && pos + dl > nbits) RRG_FAIL(ST_EOF, DET_NONE);                               \
        RRG_CONSUME(dl);                                                                         \
        if (dsym >= 30) RRG_FAIL(ST_INVALID, DET_DIST_RANGE);                                    \
        const uint32_t dx = dist_extra(dsym);                                                    \
        if ((CHK) && pos + dx > nbits) RRG_FAIL(ST_EOF, DET_NONE);                               \
        const uint32_t distance = dist_base(dsym) + (uint32_t)(bitbuf & ((1ull << dx) - 1));     \
        RRG_CONSUME(dx);                                                                         \
        /* a back-reference before the segment's first byte becomes markers in */                \
        /* the first member of a speculative segment (fast_inflate.rs:2209-2222); */             \
        /* elsewhere it is an error */                                                           \
        if (distance > out_n - member_base && !spec) RRG_FAIL(ST_INVALID, DET_DIST_FAR);         \
        RRG_FLUSH_LITS();                                                                        \
        RRG_EMIT((TOK_MATCH << 30) | ((length - 3) << 16) | (distance - 1));                     \
        out_n += length;                                                                         \
        if (out_n >= kMaxSegCells) RRG_FAIL(ST_TOO_BIG, DET_NONE);                               \
    }

    while (true) {
        if (state == S_FETCH) {
            const uint32_t w = atomicAdd(counter, 1u);
            if (w >= n_work) {
                state = S_EXIT;
            } else {
                seg = work[w];
                const SegDesc d = segs[seg];
                pos = d.start_bit;
                seg_byte = d.start_bit >> 3;
                hint = d.hint_bit;
                counting = 0;
                nbits = d.end_byte * 8;
                tk = toks + d.tok_base;
                tok_cap = (uint32_t)d.tok_cap;
                ci = ~0ull;  // the window is reloaded at the first symbol refill
                n_tok = 0;
                nq = 0;
                lit_n = 0;
                lit_acc = 0;
                out_n = 0;
                member_base = 0;
                spec = d.flags & SEG_SPEC;
                mflags = d.flags & (SEG_SINGLE_MEMBER | SEG_BLOCK_ENDS);
                mb_base = d.mb_base;
                mb_cap = d.mb_cap;
                n_mb = 0;
                status = ST_OK;
                detail = DET_NONE;
                final_block = 0;
                bfinal = 0;
                if (pos > nbits) {
                    status = ST_EOF;
                    goto finish;
                }
                RRG_REFILL();
                state = S_BLOCK_START;
            }
        }
        if (__ballot(state != S_EXIT) == 0) break;

        if (state == S_SYM) {
            // ---- Huffman symbols (fast_inflate.rs:2065-2229), as tokens ----
            // A symbol reads at most 48 bits (litlen <= 15 + 5 extra, dist <=
            // 15 + 13 extra), so with 48 bits buffered and 64 left before the
            // end of the input none of its end-of-input checks can fail: that
            // case skips them.
            for (int it = 0; it < 64; ++it) {
                if (bitcnt < 48) RRG_REFILL_W();
                if (pos + 64 <= nbits) {
                    RRG_SYMBOL(0)
                } else {
                    RRG_SYMBOL(1)
                }
            }
        } else if (state == S_BLOCK_START) {
            // ---- block header (fast_inflate.rs:1914-1935, inflate.rs:74-145) ----
            if (pos >= hint) {
                state = S_FETCH;
                goto finish;
            }
            uint32_t v, bt;
            RRG_GETBITS(v, 1);
            bfinal = v;
            RRG_GETBITS(bt, 2);
            if (bt == 0) {
                // stored block (fast_inflate.rs:1937-1963): one token for its bytes
                pos = (pos + 7) & ~7ull;
                RRG_REFILL();
                uint32_t ln, nln;
                RRG_GETBITS(ln, 16);
                RRG_GETBITS(nln, 16);
                if ((ln ^ nln) != 0xffffu) RRG_FAIL(ST_INVALID, DET_STORED_NLEN);
                if (pos + 8ull * ln > nbits) RRG_FAIL(ST_EOF, DET_NONE);
                if (ln) {
                    const uint64_t off = (pos >> 3) - seg_byte;
                    if (off + ln >= (1ull << 30) || out_n + ln >= kMaxSegCells) RRG_FAIL(ST_TOO_BIG, DET_NONE);
                    RRG_FLUSH_LITS();
                    // pieces of at most 256 bytes: no token stands for more than 258 cells
                    for (uint32_t k = 0; k < ln; k += 256) {
                        RRG_EMIT((TOK_STORED << 30) | min(ln - k, 256u));
                        RRG_EMIT((TOK_OFFSET << 30) | (uint32_t)(off + k));
                    }
                    out_n += ln;
                    pos += 8ull * ln;
                }
                RRG_REFILL();
                state = S_BLOCK_END;
            } else if (bt == 1) {
                // codes above LB bits are all 9-bit ones (LB is 8 here; 9 and 10 hold every code)
                fixed_seq = LB < 9 ? 144u : 0u;
                for (int s = 0; s < 288; ++s) lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
                build_table(lens, 288, litlut, LB, cntL, tnext, toffs, sortedL, longL, kLongLit, lhead);
                for (int s = 0; s < 30; ++s) lens[s] = 5;
                build_table(lens, 30, distlut, DIST_LB, cntD, tnext, toffs, sortedD, longD, kLongDist, lhead + 1);
                state = S_SYM;
            } else if (bt == 2) {
                uint32_t a, b2, c2;
                RRG_GETBITS(a, 5);
                RRG_GETBITS(b2, 5);
                RRG_GETBITS(c2, 4);
                const int hlit = (int)a + 257, hdist = (int)b2 + 1, hclen = (int)c2 + 4;
                if (hlit > 286 || hdist > 30) RRG_FAIL(ST_INVALID, DET_HLIT_HDIST);
                uint8_t* cl = cltmp;
                for (int i = 0; i < 19; ++i) cl[i] = 0;
                for (int i = 0; i < hclen; ++i) {
                    uint32_t x;
                    RRG_GETBITS(x, 3);
                    cl[c_cl_order[i]] = (uint8_t)x;
                }
                // precode LUT (7 bits) lives in the distance LUT while lengths are read
                if (!build_table(cl, 19, distlut, DIST_LB, cntD, tnext, toffs, nullptr, (uint8_t*)nullptr, 0,
                                 (uint32_t*)nullptr))
                    RRG_FAIL(ST_INVALID, DET_OVERSUBSCRIBED);
                const int total = hlit + hdist;
                int i = 0;
                uint8_t last = 0;  // previous code length (for code 16)
                while (i < total) {
                    if (pos + 10 > nbits) RRG_FAIL(ST_EOF, DET_NONE);
                    if (bitcnt < 32) RRG_REFILL();
                    const uint32_t e = distlut[bitbuf & 127u];
                    const uint32_t l = e & 7u;
                    if (l == 0) {
                        if (pos + 15 > nbits) RRG_FAIL(ST_EOF, DET_NONE);
                        RRG_FAIL(ST_INVALID, DET_NO_CODE);
                    }
                    RRG_CONSUME(l);
                    const uint32_t sym = e >> 3;
                    if (sym <= 15) {
                        last = (uint8_t)sym;
                        lens[i++] = last;
                    } else if (sym == 16) {
                        if (i == 0) RRG_FAIL(ST_INVALID, DET_CODE16_FIRST);
                        uint32_t x;
                        RRG_GETBITS(x, 2);
                        const int n = (int)x + 3;
                        if (i + n > total) RRG_FAIL(ST_INVALID, DET_CODE16_OVERRUN);
                        for (int j = 0; j < n; ++j) lens[i + j] = last;
                        i += n;
                    } else if (sym == 17) {
                        uint32_t x;
                        RRG_GETBITS(x, 3);
                        const int n = (int)x + 3;
                        if (i + n > total) RRG_FAIL(ST_INVALID, DET_CODE17_OVERRUN);
                        for (int j = 0; j < n; ++j) lens[i + j] = 0;
                        last = 0;
                        i += n;
                    } else {
                        uint32_t x;
                        RRG_GETBITS(x, 7);
                        const int n = (int)x + 11;
                        if (i + n > total) RRG_FAIL(ST_INVALID, DET_CODE18_OVERRUN);
                        for (int j = 0; j < n; ++j) lens[i + j] = 0;
                        last = 0;
                        i += n;
                    }
                }
                if (lens[256] == 0) RRG_FAIL(ST_INVALID, DET_EOB_ZERO);
                fixed_seq = 0;
                if (!build_table(lens, hlit, litlut, LB, cntL, tnext, toffs, sortedL, longL, kLongLit, lhead))
                    RRG_FAIL(ST_INVALID, DET_OVERSUBSCRIBED);
                if (!build_table(lens + hlit, hdist, distlut, DIST_LB, cntD, tnext, toffs, sortedD, longD, kLongDist,
                                 lhead + 1))
                    RRG_FAIL(ST_INVALID, DET_OVERSUBSCRIBED);
                state = S_SYM;
            } else {
                RRG_FAIL(ST_INVALID, DET_RESERVED_BTYPE);
            }
        } else if (state == S_BLOCK_END) {
            if (!bfinal) {
                if (mflags & SEG_BLOCK_ENDS) {
                    // the single-worker paths cut chunks at block ends (pipeline.rs:1254-1272)
                    if (n_mb >= mb_cap) RRG_FAIL(ST_MB_OVERFLOW, DET_NONE);
                    MemberRec m;
                    m.out_off = out_n;
                    m.trailer_byte = MB_BLOCK_END;
                    m.crc = 0;
                    m.isize = 0;
                    mrec[mb_base + n_mb] = m;
                    ++n_mb;
                }
                state = S_BLOCK_START;
            } else {
                // member end: trailer, then the next member's header (pipeline.rs:1125-1170)
                pos = (pos + 7) & ~7ull;
                const uint64_t tb = pos >> 3;
                if (tb + 8 > (nbits >> 3)) RRG_FAIL(ST_TRUNCATED, DET_NONE);
                const uint64_t w = bits64_at(in64, pos);
                if (n_mb >= mb_cap) RRG_FAIL(ST_MB_OVERFLOW, DET_NONE);
                MemberRec m;
                m.out_off = out_n;
                m.trailer_byte = tb;
                m.crc = (uint32_t)w;
                m.isize = (uint32_t)(w >> 32);
                mrec[mb_base + n_mb] = m;
                ++n_mb;
                const uint64_t after = tb + 8;
                pos = after * 8;
                if (after >= (nbits >> 3)) {
                    final_block = 1;
                    state = S_FETCH;
                    goto finish;
                }
                if (mflags & SEG_SINGLE_MEMBER) {
                    state = S_FETCH;
                    goto finish;
                }
                uint64_t hl = 0;
                const int hs = parse_gz_header(in64, nbits >> 3, after, &hl);
                if (hs != ST_OK) RRG_FAIL((uint32_t)hs, DET_NONE);
                pos = (after + hl) * 8;
                RRG_REFILL();
                member_base = out_n;
                spec = 0;
                state = S_BLOCK_START;
            }
        }
        continue;
    finish : {
        // the literals and tokens still waiting belong to the output before the
        // segment's end (or its failure, whose output the single-worker paths send)
        RRG_FLUSH_LITS();
        if (!counting) {
            const uint32_t b = n_tok - nq;  // first waiting token
            if (nq > 0) tk[b] = nq == 1 ? q3 : nq == 2 ? q2 : q1;
            if (nq > 1) tk[b + 1] = nq == 2 ? q3 : q2;
            if (nq > 2) tk[b + 2] = q3;
        }
        // n_tok = the tokens it needs (to the failure, when block ends are
        // recorded: the output before a failure is part of that decode)
        if (counting && (status == ST_OK || (mflags & SEG_BLOCK_ENDS))) status = ST_OVERFLOW;
        SegResult r;
        r.n_cells = out_n;
        r.end_bit = pos;
        r.status = status;
        r.detail = detail;
        r.final_block = final_block;
        r.n_mb = n_mb;
        r.n_tok = n_tok;
        r.pad = 0;
        res[seg] = r;
    }
    }
#undef RRG_SYMBOL
#undef RRG_REFILL_W
#undef RRG_REFILL
#undef RRG_CONSUME
#undef RRG_FAIL
#undef RRG_GETBITS
#undef RRG_EMIT
#undef RRG_FLUSH_LITS
}

int inflate_blocks_per_cu(int lut_bits, int lanes) {
    const size_t shm = (size_t)lanes * inflate_lds_per_lane(lut_bits);
    int nb = 0;
    hipError_t e;
    switch (lut_bits) {
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tokens<8>, lanes, shm); break;
        case 10: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tokens<10>, lanes, shm); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tokens<9>, lanes, shm); break;
    }
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_tokens(const uint64_t* d_in64, const SegDesc* d_segs, SegResult* d_res, MemberRec* d_mrec,
                         const uint32_t* d_work, uint32_t n_work, uint32_t* d_counter, uint32_t* d_toks,
                         uint16_t* d_sorted, const InflateGeom& g, hipStream_t s) {
    if (n_work == 0) return hipSuccess;
    const size_t shm = (size_t)g.lanes_per_block * inflate_lds_per_lane(g.lut_bits);
    switch (g.lut_bits) {
        case 8:
            hipLaunchKernelGGL(k_tokens<8>, dim3(g.blocks), dim3(g.lanes_per_block), shm, s, d_in64, d_segs, d_res,
                               d_mrec, d_work, n_work, d_counter, d_toks, d_sorted);
            break;
        case 10:
            hipLaunchKernelGGL(k_tokens<10>, dim3(g.blocks), dim3(g.lanes_per_block), shm, s, d_in64, d_segs, d_res,
                               d_mrec, d_work, n_work, d_counter, d_toks, d_sorted);
            break;
        default:
            hipLaunchKernelGGL(k_tokens<9>, dim3(g.blocks), dim3(g.lanes_per_block), shm, s, d_in64, d_segs, d_res,
                               d_mrec, d_work, n_work, d_counter, d_toks, d_sorted);
            break;
    }
    return hipGetLastError();
}

// ===========================================================================
// Inflate pass 1, wave-cooperative form: one wavefront per segment, one set
// of Huffman tables per wavefront in LDS, and the 64 lanes decoding one block
// body together.
//
// Lane 0 runs everything that is serial by nature -- block headers, table
// builds, stored blocks, member trailers and headers, the decode near the end
// of the input -- with the same state machine as k_tokens.  A Huffman block
// body is decoded in rounds.  A round from a known symbol start P gives lane k
// the window [P + kW, P + (k+1)W) of W bits (a multiple of B = 128):
//
//  1. Every lane decodes from the start of its window as if a symbol began
//     there, to the first symbol start at or past the window's end, counting
//     tokens and cells.  Lane 0 starts at a true symbol start; the others
//     usually start inside a code, but Huffman codes resynchronise: after a
//     few symbols a walk from a wrong start lands on a true symbol start and
//     from there follows the true walk (on FASTQ at gzip -6, median 7 units,
//     99% within 50).  At each crossing of a multiple of B (from P) a lane
//     records the first symbol start at or past it (LDS) and its token and
//     cell counts there (global scratch).
//  2. A lane that reached its window end walks on into the next windows, and
//     at each crossing compares its symbol start with the one recorded by the
//     lane owning that window.  Two walks that share one symbol start agree
//     from there on, so the first equal crossing is where the later lane's walk
//     has become the true one: the earlier lane ends there and the later one
//     starts there (a later lane that never synchronises is skipped).
//  3. Following these links from lane 0 gives the lanes whose walks are true,
//     each from its start crossing to the next one's; a wave prefix sum over
//     their token and cell counts places their output.
//  4. Those lanes decode their ranges again and store their tokens at their
//     final positions.
// Literal tokens (up to three literals) are closed at every crossing, so a
// range's token count is a difference of the counts recorded in step 1.  The
// walks of step 1 skip the end-of-input checks: a round only runs where every
// read stays 64 bits inside the input.  A lane on the true walk that meets an
// invalid code, or a back-reference before its member, hands the rest of the
// block back to lane 0's serial decode from the start of that lane's range,
// which stops at the failing symbol exactly as the lane-per-segment decode
// does.  The token stream is the one k_tokens writes except for where
// literal tokens close, so the expansion pass and the host are unchanged.
// ===========================================================================
constexpr uint32_t kPB = 128;                // crossing spacing (bits)
constexpr uint32_t kPWMax = 4096;            // window per lane (bits), at most
constexpr uint32_t kPWMin = 2 * kPB;         // smaller windows: lane 0 decodes alone
constexpr uint32_t kPNM = kPWMax / kPB;      // crossings per window
constexpr uint32_t kPMS = 16;                // crossings per window where a later lane can take over
constexpr int WDLB = 9;                      // distance LUT bits
constexpr size_t kWaveScratchBytes = 256;  // unused (the rounds keep their records in LDS)

size_t inflate_wave_scratch_bytes() { return kWaveScratchBytes; }

// lane 0's decode state, kept in LDS while the wave runs (so it holds no
// registers there)
struct SerialState {
    uint64_t pos, hint, seg_byte, nbits, tk_base, body_start;
    uint32_t state, seg, tok_cap, n_tok, out_n, member_base, spec, mflags, bfinal, n_mb, mb_base, mb_cap;
    uint32_t fixed_seq, counting, force_serial, parW, status, detail, final_block, hlit, hdist, pad;
};

struct WaveHand {
    uint64_t pos;      // lane 0 -> wave: body start; wave -> lane 0: where lane 0 goes on
    uint64_t nbits;
    uint64_t tk_base;  // the segment's token region
    uint32_t cmd;      // 0: no segment left, 1: decode a block body, 2: build its tables first
    uint32_t next;     // wave -> lane 0: S_BLOCK_END (body done), S_SYM (decode on serially), S_EXIT (too big)
    uint32_t n_tok, out_n, tok_cap, member_base, spec, counting, fixed_seq, W;
    uint32_t hlit, hdist;  // with cmd 2: the code lengths in lens to build the tables from
};

// Codes longer than a LUT, decoded from the cumulative Kraft sums: with the
// next 15 input bits read first-bit-highest as `rev`, a code has length l
// when lj[l-1] <= rev < lj[l] (lj[l] = sum over k <= l of count[k] << (15-k),
// the left-justified end of the length-l codes), and its index among the
// symbols sorted by (length, symbol) is (rev >> (15-l)) + bias[l] (bias[l] =
// symbols shorter than l - first code of length l).
struct LongTab {
    uint16_t lj[16];
    int16_t bias[16];
};

// The LUT entry (lit_entry / dist_entry) of a code longer than LB bits, 0 when
// no code matches, from the entries sorted by (length, symbol): the length is
// the first l with rev < lj[l].
template <int LB, int N, typename E>
__device__ __forceinline__ uint32_t wave_long(uint64_t bits, const LongTab& t, const E (&sorted)[N]) {
    const uint32_t rev = __brev((uint32_t)bits) >> 17;
    const uint4 q = *reinterpret_cast<const uint4*>(&t.lj[8]);
    const uint32_t v[8] = {q.x & 0xffffu, q.x >> 16, q.y & 0xffffu, q.y >> 16,
                           q.z & 0xffffu, q.z >> 16, q.w & 0xffffu, q.w >> 16};
    uint32_t n = 0;
#pragma unroll
    for (int l = LB + 1; l <= 15; ++l) n += rev >= v[l - 8] ? 1u : 0u;
    const uint32_t l = LB + 1 + n;
    // past every code (l == 16): the lookup below stays in the list and is dropped
    const uint32_t lc = l > 15 ? 15u : l;
    int idx = (int)(rev >> (15 - lc)) + (int)t.bias[lc];
    idx = idx < 0 ? 0 : idx > N - 1 ? N - 1 : idx;
    const uint32_t e = (uint32_t)sorted[idx];
    return l > 15 ? 0u : e | l;
}

// LUT entries of the wave kernel.  Literal/length (u16): bits 3..0 code
// length (0: longer than the LUT), bit 4 clear: a literal, its byte in bits
// 15..8; bit 4 set: bits 7..5 = extra bits of a length (0..5; 6 = symbol 286
// or 287, invalid; 7 = end of block) and bits 15..8 = its base - 3.
// Distance (u32): bits 3..0 code length, 7..4 extra bits, 23..8 base.
// (Length and distance bases and extra bits: deflate/tables.rs:10-30.)
__device__ __forceinline__ uint32_t lit_entry(uint32_t s, uint32_t l) {
    if (s < 256) return (s << 8) | l;
    if (s == 256) return (7u << 5) | 16u | l;
    if (s > 285) return (6u << 5) | 16u | l;
    const uint32_t li = s - 257;
    return ((len_base(li) - 3u) << 8) | (len_extra(li) << 5) | 16u | l;
}
__device__ __forceinline__ uint32_t dist_entry(uint32_t s, uint32_t l) {
    return (dist_base(s) << 8) | (dist_extra(s) << 4) | l;  // s < 30: HDIST <= 30, the fixed code has 30
}

// Canonical Huffman table from n code lengths (huffman.rs:215-325 semantics:
// incomplete codes accepted), built by the 64 lanes of the wave: counts by
// LDS atomics, each symbol's place in (length, symbol) order by ballots, and
// every LUT entry decoded from the cumulative Kraft sums.  LUT entries are
// lit_entry / dist_entry for codes of at most TB bits, 0 otherwise.  Returns
// false (the same in every lane) when the code is oversubscribed with at least
// two symbols.  Called by the whole wave (it synchronises).
template <int TB, bool DIST, typename E>
__device__ bool wave_build(const uint8_t* __restrict__ lens, uint32_t n, void* __restrict__ lutv,
                           E* __restrict__ sorted, LongTab& lt, uint32_t* __restrict__ c32, uint32_t lane) {
    if (lane < 16) c32[lane] = 0;
    __syncthreads();
    for (uint32_t s = lane; s < n; s += 64) {
        const uint32_t l = lens[s];
        if (l) atomicAdd(&c32[l], 1u);
    }
    __syncthreads();
    uint32_t cnt[16], cur[16], ljv[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) cnt[l] = l ? uni32(c32[l]) : 0u;
    uint32_t total = 0, nsym = 0, code = 0, off = 0;
#pragma unroll
    for (int l = 1; l <= 15; ++l) {
        total += cnt[l] << (15 - l);
        nsym += cnt[l];
    }
    if (nsym >= 2 && total > 32768u) return false;
    ljv[0] = 0;
#pragma unroll
    for (int l = 1; l <= 15; ++l) {
        code = (code + cnt[l - 1]) << 1;
        cur[l] = off;
        ljv[l] = ljv[l - 1] + (cnt[l] << (15 - l));
        if (lane == (uint32_t)l) {
            lt.lj[l] = (uint16_t)ljv[l];
            lt.bias[l] = (int16_t)((int)off - (int)code);
        }
        off += cnt[l];
    }
    if (lane == 0) {
        lt.lj[0] = 0;
        lt.bias[0] = 0;
    }
    // symbols in (length, symbol) order
    for (uint32_t b = 0; b < n; b += 64) {
        const uint32_t s = b + lane;
        const uint32_t l = s < n ? lens[s] : 0u;
        uint32_t at = 0;
#pragma unroll
        for (int ll = 1; ll <= 15; ++ll) {
            const uint64_t m = __ballot(l == (uint32_t)ll);
            if (l == (uint32_t)ll) at = cur[ll] + (uint32_t)__popcll(m & ((1ull << lane) - 1));
            cur[ll] += (uint32_t)__popcll(m);
        }
        if (l) sorted[at] = (E)(DIST ? dist_entry(s, 0) : lit_entry(s, 0));
    }
    __syncthreads();
    // LUT: the TB bits x (first input bit lowest) decode to the code of length l
    // when lj[l-1] <= rev < lj[l], rev = x reversed and left-justified to 15 bits
    for (uint32_t x = lane; x < (1u << TB); x += 64) {
        const uint32_t rev = (__brev(x) >> (32 - TB)) << (15 - TB);
        uint32_t l = 1;
#pragma unroll
        for (int ll = 1; ll < TB; ++ll) l += rev >= ljv[ll] ? 1u : 0u;
        uint32_t e = 0;
        if (rev < lt.lj[l]) e = (uint32_t)sorted[(uint32_t)((int)(rev >> (15 - l)) + (int)lt.bias[l])] | l;
        if (DIST) static_cast<uint32_t*>(lutv)[x] = e;
        else static_cast<uint16_t*>(lutv)[x] = (uint16_t)e;
    }
    __syncthreads();
    return true;
}

template <int LB>
struct WaveLds {
    uint16_t lit[1 << LB];     // lit_entry
    uint32_t dist[1 << WDLB];  // dist_entry
    uint16_t pre[128];         // precode LUT (7 bits): (symbol << 4) | length
    LongTab ltL, ltD;          // codes longer than the LUTs
    uint16_t sortL[288];  // lit_entry(symbol, 0) in (length, symbol) order
    uint32_t sortD[32];   // dist_entry(symbol, 0) in (length, symbol) order
    uint8_t lens[320];       // code lengths: literal/length, then distance
    uint8_t cl[32];          // precode lengths
    uint32_t c32[16];        // code-length counts (table build)
    uint16_t cnt[16], tnext[16], toffs[16];  // precode table build
    union {
        // per lane, for its window's first kPMS crossings: the first symbol start
        // at or past the crossing (from the window start; 0xFFFF: not reached),
        // and the tokens its walk produced before it; crossing-major (crossing
        // m of lane k at X[m][k]) so that the lanes' records of one crossing
        // fall in distinct banks
        struct {
            uint16_t X[kPMS][64];
            uint16_t T[kPMS][64];
        };
        // once a round's ranges are known: each lane's tokens on their way to
        // memory, slot-major (slot j of lane k at TB[j][k]: the lanes' accesses
        // to one slot fall in distinct banks)
        uint32_t TB[kPMS][64];
    };
    WaveHand h;
    SerialState ss;
};

template <int LB>
int wave_lds_bytes() {
    return (int)sizeof(WaveLds<LB>);
}

enum : uint32_t { WK_WALK = 0, WK_EOB = 1, WK_ERR = 2 };
enum : uint32_t { WT_END = 64, WT_EOB = 65, WT_ERR = 66 };

template <int LB>
// 4 waves per SIMD (<= 128 VGPRs): the rounds need about 64; lane 0's header
// code, which needs more, spills instead
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_tokens_wave(const uint64_t* __restrict__ in64, const SegDesc* __restrict__ segs,
                                                    SegResult* __restrict__ res, MemberRec* __restrict__ mrec,
                                                    const uint32_t* __restrict__ work, uint32_t n_work,
                                                    uint32_t* __restrict__ counter, uint32_t* __restrict__ toks,
                                                    uint8_t* __restrict__ wscratch) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    WaveLds<LB>& L = *reinterpret_cast<WaveLds<LB>*>(lds_raw);
    constexpr uint32_t LMASK = (1u << LB) - 1;
    constexpr uint32_t DMASK = (1u << WDLB) - 1;
    const uint32_t lane = threadIdx.x;
    uint8_t* my_scr = wscratch + (size_t)blockIdx.x * kWaveScratchBytes;
    uint8_t* const lens = L.lens;
    (void)my_scr;

    if (lane == 0) {
        L.ss = SerialState{};
        L.ss.state = S_FETCH;
        L.ss.parW = kPWMax;
    }

#define WS_REFILL()                       \
    do {                                  \
        bitbuf = bits64_at(in64, pos);    \
        asm volatile("" : "+v"(bitbuf));  \
        bitcnt = 64;                      \
    } while (0)
#define WS_REFILL_W()                                                               \
    do {                                                                            \
        const uint64_t _q = pos >> 6;                                               \
        if (_q != ci) {                                                             \
            uint32_t _st = 3;                                                       \
            if (_q > ci && _q - ci < 3) _st = (uint32_t)(_q - ci);                   \
            ci = _q;                                                                \
            for (uint32_t _k = _st; _k > 0; --_k) {                                 \
                C0 = C1;                                                            \
                C1 = C2;                                                            \
                C2 = in64[ci + 3 - _k];                                             \
            }                                                                       \
        }                                                                           \
        bitbuf = funnel_lo(C0, C1, (uint32_t)(pos & 63));                           \
        bitcnt = 64;                                                                \
    } while (0)
#define WS_CONSUME(n)            \
    do {                         \
        const uint32_t _n = (n); \
        bitbuf >>= _n;           \
        bitcnt -= _n;            \
        pos += _n;               \
    } while (0)
#define WS_FAIL(st, det)    \
    do {                    \
        status = (st);      \
        detail = (det);     \
        state = S_FETCH;    \
        goto finish;        \
    } while (0)
#define WS_GETBITS(dst, n)                                          \
    do {                                                            \
        const uint32_t _k = (n);                                    \
        if (pos + _k > nbits) WS_FAIL(ST_EOF, DET_NONE);            \
        if (bitcnt < _k) WS_REFILL();                               \
        (dst) = _k ? (uint32_t)(bitbuf & ((1ull << _k) - 1)) : 0u; \
        WS_CONSUME(_k);                                             \
    } while (0)
#define WS_EMIT(t)                                \
    do {                                          \
        if (n_tok >= tok_cap) counting = 1;       \
        if (!counting) tk[n_tok] = (t);           \
        ++n_tok;                                  \
    } while (0)
#define WS_FLUSH_LITS()                                          \
    do {                                                         \
        if (lit_n) {                                             \
            WS_EMIT((TOK_LIT << 30) | (lit_n << 24) | lit_acc);  \
            lit_n = 0;                                           \
            lit_acc = 0;                                         \
        }                                                        \
    } while (0)
// one symbol of lane 0's serial decode (the lane-per-segment kernel's symbol step)
#define WS_SYMBOL(CHK)                                                                           \
    {                                                                                            \
        uint32_t e = L.lit[bitbuf & LMASK];                                                      \
        uint32_t l = e & 15u;                                                                    \
        if (l == 0) {                                                                            \
            if ((CHK) && pos + 15 > nbits) WS_FAIL(ST_EOF, DET_NONE);                            \
            e = wave_long<LB>(bitbuf, L.ltL, L.sortL);                                           \
            if (!e) WS_FAIL(ST_INVALID, DET_NO_CODE);                                            \
            l = e & 15u;                                                                         \
        }                                                                                        \
        if ((CHK) && pos + l > nbits) WS_FAIL(ST_EOF, DET_NONE);                                 \
        WS_CONSUME(l);                                                                           \
        if (!(e & 16u)) {                                                                        \
            lit_acc |= (e >> 8) << (8 * lit_n);                                                  \
            if (++lit_n == 3) {                                                                  \
                WS_EMIT((TOK_LIT << 30) | (3u << 24) | lit_acc);                                 \
                lit_n = 0;                                                                       \
                lit_acc = 0;                                                                     \
            }                                                                                    \
            ++out_n;                                                                             \
            continue;                                                                            \
        }                                                                                        \
        const uint32_t lx = (e >> 5) & 7u;                                                       \
        if (lx == 7) {                                                                           \
            state = S_BLOCK_END;                                                                 \
            break;                                                                               \
        }                                                                                        \
        if (lx == 6) WS_FAIL(ST_INVALID, DET_LITLEN_RANGE);                                      \
        if ((CHK) && pos + lx > nbits) WS_FAIL(ST_EOF, DET_NONE);                                \
        const uint32_t length = (e >> 8) + 3u + (uint32_t)(bitbuf & ((1ull << lx) - 1));         \
        WS_CONSUME(lx);                                                                          \
        uint32_t de = L.dist[bitbuf & DMASK];                                                    \
        uint32_t dl = de & 15u;                                                                  \
        if (dl == 0) {                                                                           \
            if ((CHK) && pos + 15 > nbits) WS_FAIL(ST_EOF, DET_NONE);                            \
            de = wave_long<WDLB>(bitbuf, L.ltD, L.sortD);                                        \
            if (!de) WS_FAIL(ST_INVALID, DET_NO_CODE);                                           \
            dl = de & 15u;                                                                       \
        }                                                                                        \
        if ((CHK) && pos + dl > nbits) WS_FAIL(ST_EOF, DET_NONE);                                \
        WS_CONSUME(dl);                                                                          \
        const uint32_t dx = (de >> 4) & 15u;                                                     \
        if ((CHK) && pos + dx > nbits) WS_FAIL(ST_EOF, DET_NONE);                                \
        const uint32_t distance = (de >> 8) + (uint32_t)(bitbuf & ((1ull << dx) - 1));           \
        WS_CONSUME(dx);                                                                          \
        if (distance > out_n - member_base && !spec) WS_FAIL(ST_INVALID, DET_DIST_FAR);          \
        WS_FLUSH_LITS();                                                                         \
        WS_EMIT((TOK_MATCH << 30) | ((length - 3) << 16) | (distance - 1));                      \
        out_n += length;                                                                         \
    }

    while (true) {
        // ================= lane 0: the serial decode =================
        if (lane == 0) {
            SerialState& S = L.ss;
            int state = (int)S.state;
            uint32_t seg = S.seg;
            uint64_t pos = S.pos, hint = S.hint, seg_byte = S.seg_byte, nbits = S.nbits, tk_base = S.tk_base;
            uint64_t body_start = S.body_start;
            uint32_t tok_cap = S.tok_cap, n_tok = S.n_tok, out_n = S.out_n, member_base = S.member_base;
            uint32_t spec = S.spec, mflags = S.mflags, bfinal = S.bfinal, n_mb = S.n_mb, mb_base = S.mb_base;
            uint32_t mb_cap = S.mb_cap, fixed_seq = S.fixed_seq, counting = S.counting;
            uint32_t force_serial = S.force_serial, parW = S.parW;
            uint32_t status = S.status, detail = S.detail, final_block = S.final_block;
            uint32_t s_hlit = S.hlit, s_hdist = S.hdist;
            uint32_t* tk = toks + tk_base;
            uint32_t lit_acc = 0, lit_n = 0;
            uint64_t bitbuf = 0, ci = ~0ull, C0 = 0, C1 = 0, C2 = 0;
            uint32_t bitcnt = 0;
            WS_REFILL();
            uint32_t cmd = 0;
            while (true) {
                if (state == S_FINISH) {
                    // the wave stopped the segment (status set)
                    state = S_FETCH;
                    goto finish;
                }
                if (state == S_FETCH) {
                    const uint32_t w = atomicAdd(counter, 1u);
                    if (w >= n_work) {
                        state = S_EXIT;
                    } else {
                        seg = work[w];
                        const SegDesc d = segs[seg];
                        pos = d.start_bit;
                        seg_byte = d.start_bit >> 3;
                        hint = d.hint_bit;
                        counting = 0;
                        nbits = d.end_byte * 8;
                        tk_base = d.tok_base;
                        tk = toks + d.tok_base;
                        tok_cap = (uint32_t)d.tok_cap;
                        ci = ~0ull;
                        n_tok = 0;
                        lit_n = 0;
                        lit_acc = 0;
                        out_n = 0;
                        member_base = 0;
                        spec = d.flags & SEG_SPEC;
                        mflags = d.flags & (SEG_SINGLE_MEMBER | SEG_BLOCK_ENDS);
                        mb_base = d.mb_base;
                        mb_cap = d.mb_cap;
                        n_mb = 0;
                        status = ST_OK;
                        detail = DET_NONE;
                        final_block = 0;
                        bfinal = 0;
                        parW = kPWMax;
                        if (pos > nbits) {
                            status = ST_EOF;
                            goto finish;
                        }
                        WS_REFILL();
                        state = S_BLOCK_START;
                    }
                }
                if (state == S_EXIT) break;

                if (state == S_SYM) {
                    if (!force_serial && pos + 64ull * kPWMin + 160 <= nbits) {
                        cmd = 1;  // the wave decodes this body
                        break;
                    }
                    // serial decode: near the end of the input, or on from a
                    // range where the wave met an error
                    for (int it = 0; it < 64; ++it) {
                        if (bitcnt < 48) WS_REFILL_W();
                        if (pos + 64 <= nbits) {
                            WS_SYMBOL(0)
                        } else {
                            WS_SYMBOL(1)
                        }
                    }
                } else if (state == S_BLOCK_START) {
                    force_serial = 0;
                    if (pos >= hint) {
                        state = S_FETCH;
                        goto finish;
                    }
                    uint32_t v, bt;
                    WS_GETBITS(v, 1);
                    bfinal = v;
                    WS_GETBITS(bt, 2);
                    if (bt == 0) {
                        pos = (pos + 7) & ~7ull;
                        WS_REFILL();
                        uint32_t ln, nln;
                        WS_GETBITS(ln, 16);
                        WS_GETBITS(nln, 16);
                        if ((ln ^ nln) != 0xffffu) WS_FAIL(ST_INVALID, DET_STORED_NLEN);
                        if (pos + 8ull * ln > nbits) WS_FAIL(ST_EOF, DET_NONE);
                        if (ln) {
                            const uint64_t off = (pos >> 3) - seg_byte;
                            if (off + ln >= (1ull << 30) || out_n + ln >= kMaxSegCells) WS_FAIL(ST_TOO_BIG, DET_NONE);
                            WS_FLUSH_LITS();
                            for (uint32_t k = 0; k < ln; k += 256) {
                                WS_EMIT((TOK_STORED << 30) | min(ln - k, 256u));
                                WS_EMIT((TOK_OFFSET << 30) | (uint32_t)(off + k));
                            }
                            out_n += ln;
                            pos += 8ull * ln;
                        }
                        WS_REFILL();
                        state = S_BLOCK_END;
                    } else if (bt == 1) {
                        for (int s = 0; s < 288; ++s) lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
                        for (int s = 0; s < 30; ++s) lens[288 + s] = 5;
                        s_hlit = 288;
                        s_hdist = 30;
                        body_start = pos;
                        state = S_SYM;
                        cmd = 2;  // the wave builds the tables
                        break;
                    } else if (bt == 2) {
                        uint32_t a, b2, c2;
                        WS_GETBITS(a, 5);
                        WS_GETBITS(b2, 5);
                        WS_GETBITS(c2, 4);
                        const int hlit = (int)a + 257, hdist = (int)b2 + 1, hclen = (int)c2 + 4;
                        if (hlit > 286 || hdist > 30) WS_FAIL(ST_INVALID, DET_HLIT_HDIST);
                        uint8_t* cl = L.cl;
                        for (int i = 0; i < 19; ++i) cl[i] = 0;
                        for (int i = 0; i < hclen; ++i) {
                            uint32_t x;
                            WS_GETBITS(x, 3);
                            cl[c_cl_order[i]] = (uint8_t)x;
                        }
                        if (!build_table(cl, 19, L.pre, 7, L.cnt, L.tnext, L.toffs, nullptr, (uint16_t*)nullptr, 0,
                                         (uint32_t*)nullptr))
                            WS_FAIL(ST_INVALID, DET_OVERSUBSCRIBED);
                        const int total = hlit + hdist;
                        int i = 0;
                        uint8_t last = 0;
                        while (i < total) {
                            if (pos + 10 > nbits) WS_FAIL(ST_EOF, DET_NONE);
                            if (bitcnt < 32) WS_REFILL();
                            const uint32_t e = L.pre[bitbuf & 127u];
                            const uint32_t l = e & 15u;
                            if (l == 0) {
                                if (pos + 15 > nbits) WS_FAIL(ST_EOF, DET_NONE);
                                WS_FAIL(ST_INVALID, DET_NO_CODE);
                            }
                            WS_CONSUME(l);
                            const uint32_t sym = e >> 4;
                            if (sym <= 15) {
                                last = (uint8_t)sym;
                                lens[i++] = last;
                            } else if (sym == 16) {
                                if (i == 0) WS_FAIL(ST_INVALID, DET_CODE16_FIRST);
                                uint32_t x;
                                WS_GETBITS(x, 2);
                                const int n = (int)x + 3;
                                if (i + n > total) WS_FAIL(ST_INVALID, DET_CODE16_OVERRUN);
                                for (int j = 0; j < n; ++j) lens[i + j] = last;
                                i += n;
                            } else if (sym == 17) {
                                uint32_t x;
                                WS_GETBITS(x, 3);
                                const int n = (int)x + 3;
                                if (i + n > total) WS_FAIL(ST_INVALID, DET_CODE17_OVERRUN);
                                for (int j = 0; j < n; ++j) lens[i + j] = 0;
                                last = 0;
                                i += n;
                            } else {
                                uint32_t x;
                                WS_GETBITS(x, 7);
                                const int n = (int)x + 11;
                                if (i + n > total) WS_FAIL(ST_INVALID, DET_CODE18_OVERRUN);
                                for (int j = 0; j < n; ++j) lens[i + j] = 0;
                                last = 0;
                                i += n;
                            }
                        }
                        if (lens[256] == 0) WS_FAIL(ST_INVALID, DET_EOB_ZERO);
                        s_hlit = (uint32_t)hlit;
                        s_hdist = (uint32_t)hdist;
                        body_start = pos;
                        state = S_SYM;
                        cmd = 2;  // the wave builds the tables (an oversubscribed code fails there)
                        break;
                    } else {
                        WS_FAIL(ST_INVALID, DET_RESERVED_BTYPE);
                    }
                } else if (state == S_BLOCK_END) {
                    if (!bfinal) {
                        if (mflags & SEG_BLOCK_ENDS) {
                            if (n_mb >= mb_cap) WS_FAIL(ST_MB_OVERFLOW, DET_NONE);
                            MemberRec m;
                            m.out_off = out_n;
                            m.trailer_byte = MB_BLOCK_END;
                            m.crc = 0;
                            m.isize = 0;
                            mrec[mb_base + n_mb] = m;
                            ++n_mb;
                        }
                        state = S_BLOCK_START;
                    } else {
                        pos = (pos + 7) & ~7ull;
                        const uint64_t tb = pos >> 3;
                        if (tb + 8 > (nbits >> 3)) WS_FAIL(ST_TRUNCATED, DET_NONE);
                        const uint64_t w = bits64_at(in64, pos);
                        if (n_mb >= mb_cap) WS_FAIL(ST_MB_OVERFLOW, DET_NONE);
                        MemberRec m;
                        m.out_off = out_n;
                        m.trailer_byte = tb;
                        m.crc = (uint32_t)w;
                        m.isize = (uint32_t)(w >> 32);
                        mrec[mb_base + n_mb] = m;
                        ++n_mb;
                        const uint64_t after = tb + 8;
                        pos = after * 8;
                        if (after >= (nbits >> 3)) {
                            final_block = 1;
                            state = S_FETCH;
                            goto finish;
                        }
                        if (mflags & SEG_SINGLE_MEMBER) {
                            state = S_FETCH;
                            goto finish;
                        }
                        uint64_t hl = 0;
                        const int hs = parse_gz_header(in64, nbits >> 3, after, &hl);
                        if (hs != ST_OK) WS_FAIL((uint32_t)hs, DET_NONE);
                        pos = (after + hl) * 8;
                        WS_REFILL();
                        member_base = out_n;
                        spec = 0;
                        state = S_BLOCK_START;
                    }
                }
                continue;
            finish : {
                WS_FLUSH_LITS();
                if (counting && (status == ST_OK || (mflags & SEG_BLOCK_ENDS))) status = ST_OVERFLOW;
                SegResult r;
                r.n_cells = out_n;
                r.end_bit = pos;
                r.status = status;
                r.detail = detail;
                r.final_block = final_block;
                r.n_mb = n_mb;
                r.n_tok = n_tok;
                r.pad = 0;
                res[seg] = r;
            }
            }
            // hand the body to the wave (pending literals close first, so the
            // token count is exact)
            if (cmd) WS_FLUSH_LITS();
            S.state = (uint32_t)state;
            S.seg = seg;
            S.pos = pos;
            S.hint = hint;
            S.seg_byte = seg_byte;
            S.nbits = nbits;
            S.tk_base = tk_base;
            S.body_start = body_start;
            S.tok_cap = tok_cap;
            S.n_tok = n_tok;
            S.out_n = out_n;
            S.member_base = member_base;
            S.spec = spec;
            S.mflags = mflags;
            S.bfinal = bfinal;
            S.n_mb = n_mb;
            S.mb_base = mb_base;
            S.mb_cap = mb_cap;
            S.fixed_seq = fixed_seq;
            S.counting = counting;
            S.force_serial = force_serial;
            S.parW = parW;
            S.status = status;
            S.detail = detail;
            S.final_block = final_block;
            S.hlit = s_hlit;
            S.hdist = s_hdist;
            if (cmd) {
                L.h.pos = pos;
                L.h.nbits = nbits;
                L.h.tk_base = tk_base;
                L.h.n_tok = n_tok;
                L.h.out_n = out_n;
                L.h.tok_cap = tok_cap;
                L.h.member_base = member_base;
                L.h.spec = spec;
                L.h.counting = counting;
                L.h.fixed_seq = fixed_seq;
                L.h.W = parW;
                L.h.hlit = s_hlit;
                L.h.hdist = s_hdist;
            }
            L.h.cmd = cmd;
        }
        __syncthreads();
        const uint32_t wcmd = uni32(L.h.cmd);
        if (wcmd == 0) break;
        bool tables_ok = true;
        if (wcmd == 2) {
            // the block's Huffman tables, by the whole wave
            const uint32_t hl = uni32(L.h.hlit), hd = uni32(L.h.hdist);
            tables_ok = wave_build<LB, false>(L.lens, hl, L.lit, L.sortL, L.ltL, L.c32, lane) &&
                        wave_build<WDLB, true>(L.lens + hl, hd, L.dist, L.sortD, L.ltD, L.c32, lane);
        }

        // ================= the wave: rounds over the block body =================
        {
            uint64_t P = uni64(L.h.pos);
            const uint64_t nb_end = uni64(L.h.nbits);
            uint32_t* const tkw = toks + uni64(L.h.tk_base);
            uint32_t w_ntok = uni32(L.h.n_tok), w_outn = uni32(L.h.out_n);
            const uint32_t w_cap = uni32(L.h.tok_cap), w_mbase = uni32(L.h.member_base);
            const uint32_t w_spec = uni32(L.h.spec);
            uint32_t w_counting = uni32(L.h.counting);
            uint32_t Wwant = uni32(L.h.W);
            uint32_t w_next = tables_ok ? (uint32_t)S_SYM : (uint32_t)S_FINISH;
            const uint32_t k = lane;
            uint16_t* const myX = &L.X[0][k];  // crossing m at myX[64 * m]
            uint16_t* const myT = &L.T[0][k];

            // Lane walk state.  Positions r are bits after P.  The lane's input
            // window D0..D2 holds the words wq..wq+2 after P's word; a unit
            // moves r by at most 48 bits, so before each unit the window moves
            // by at most one word, and the word two ahead is read then and
            // first used one word later.  The loop bodies are branch-free
            // (every lane runs every instruction, state changes are selects):
            // with a branch per lane the compiler copies the loop-carried
            // registers, the pending input word among them, and has to wait for
            // it.  Only the rare long-code decode and the stores are
            // predicated.
            uint64_t bb = 0, D0 = 0, D1 = 0, D2 = 0;
            uint32_t r = 0, wq = 0;
            uint32_t ukind = 0, usym = 0, ulen = 0, udist = 0, un = 0;
#define WP_START()                                  \
    do {                                            \
        const uint32_t _a = Pb + r;                 \
        wq = _a >> 6;                               \
        D0 = inP[wq];                               \
        D1 = inP[wq + 1];                           \
        D2 = inP[wq + 2];                           \
    } while (0)
#define WP_WIN()                                    \
    do {                                            \
        const uint32_t _a = Pb + r;                 \
        const uint32_t _q = _a >> 6;                \
        const bool _adv = _q != wq;                 \
        D0 = _adv ? D1 : D0;                        \
        D1 = _adv ? D2 : D1;                        \
        /* the old D2 is dead here: the load below  \
           may reuse its register */                \
        asm volatile("" : "+v"(D1) : : "memory");   \
        if (_adv) D2 = inP[_q + 2];                 \
        wq = _q;                                    \
        bb = funnel_lo(D0, D1, _a & 63u);           \
    } while (0)
// One unit at r (in bb): a literal (ukind 0, usym), a match (1, ulen, udist),
// the end of block (2) or an invalid code (3), un bits long.  Both the literal
// and the match fields are computed; the distance LUT is read for every lane.
#define WP_DEC(ACT)                                                                                     \
    do {                                                                                                \
        uint32_t _e = L.lit[(uint32_t)bb & LMASK];                                                      \
        if ((ACT) && (_e & 15u) == 0) _e = wave_long<LB>(bb, L.ltL, L.sortL);                          \
        const uint32_t _l = _e & 15u;                                                                   \
        const uint32_t _x = (_e >> 5) & 7u;                                                             \
        const bool _oth = (_e & 16u) != 0;                                                              \
        const bool _len = _oth && _x < 6;                                                               \
        const uint32_t _xl = _len ? _x : 0u;                                                            \
        ulen = (_e >> 8) + 3u + ((uint32_t)(bb >> _l) & ((1u << _xl) - 1u));                            \
        const uint32_t _o = _l + _xl;                                                                   \
        const uint64_t _b2 = bb >> _o;                                                                  \
        uint32_t _d = L.dist[(uint32_t)_b2 & DMASK];                                                    \
        if ((ACT) && _len && (_d & 15u) == 0) _d = wave_long<WDLB>(_b2, L.ltD, L.sortD);               \
        const uint32_t _dl = _d & 15u;                                                                  \
        const uint32_t _dx = (_d >> 4) & 15u;                                                           \
        udist = (_d >> 8) + ((uint32_t)(_b2 >> _dl) & ((1u << _dx) - 1u));                              \
        usym = _e >> 8;                                                                                 \
        un = _len ? _o + _dl + _dx : _l;                                                                \
        ukind = _l == 0 ? 3u : !_oth ? 0u : _len ? (_dl ? 1u : 3u) : (_x == 7 ? 2u : 3u);               \
    } while (0)
// token counting of a walk (literal tokens close at 3 and before a match or
// the end of block)
#define WP_COUNT(A)                                                                                     \
    do {                                                                                                \
        const uint32_t _ln1 = litn + 1;                                                                 \
        const bool _lit = ukind == 0;                                                                   \
        tok += (A) ? (_lit ? (_ln1 == 3 ? 1u : 0u) : (litn != 0 ? 1u : 0u) + (ukind == 1 ? 1u : 0u)) : 0u; \
        litn = (A) ? (_lit ? (_ln1 == 3 ? 0u : _ln1) : 0u) : litn;                                      \
        r += (A) ? un : 0u;                                                                             \
    } while (0)

            while (tables_ok) {
                // the window: as asked, shrunk so that every read stays inside the input
                const uint64_t room = nb_end > P + 160 ? (nb_end - P - 160) / 64 : 0;
                const uint32_t Wr = (uint32_t)((uint64_t)Wwant < room ? (uint64_t)Wwant : room) & ~(kPB - 1);
                if (Wr < kPWMin) {
                    w_next = S_SYM;
                    break;
                }
                const uint32_t NMr = Wr / kPB;
                const uint32_t RW = 64u * Wr;
                const uint64_t* const inP = in64 + (P >> 6);
                const uint32_t Pb = (uint32_t)(P & 63);
                // ---- 1. own window ----
                const uint32_t g = k * Wr;
                r = g;
                WP_START();
                uint32_t nbd = g, m = 0, tok = 0, litn = 0, kind = WK_WALK;
                bool act = true;
                while (__ballot(act)) {
                    // a crossing at this unit's start: literal tokens close here
                    const bool cr = act && r >= nbd;
                    tok += (cr && litn != 0) ? 1u : 0u;
                    litn = cr ? 0u : litn;
                    if (cr && m < kPMS) {
                        myX[64u * m] = (uint16_t)(r - g);
                        myT[64u * m] = (uint16_t)tok;
                    }
                    act = act && !(cr && m == NMr);
                    m += cr ? 1u : 0u;
                    nbd += cr ? kPB : 0u;
                    WP_WIN();
                    WP_DEC(act);
                    WP_COUNT(act);
                    kind = (act && ukind >= 2) ? (ukind == 2 ? WK_EOB : WK_ERR) : kind;
                    act = act && ukind < 2;
                }
                // crossings this walk did not reach hold no position
                for (uint32_t mm = kind == WK_WALK ? NMr + 1 : m; mm < kPMS; ++mm) myX[64u * mm] = 0xFFFFu;
                __syncthreads();
                // ---- 2. on into the next windows until a crossing agrees ----
                uint32_t tgt = kind == WK_EOB ? WT_EOB : kind == WK_ERR ? WT_ERR : WT_END, tm = 0;
                {
                    uint32_t j = k + 1, mm = 0, bnd = (k + 1) * Wr;
                    act = kind == WK_WALK && k < 63;
                    while (__ballot(act)) {
                        const bool cr = act && r >= bnd;
                        const bool end = cr && j == 64;
                        const bool cr2 = cr && !end;
                        // literals close at every crossing, the round's end included,
                        // as they do in step 4
                        tok += (cr && litn != 0) ? 1u : 0u;
                        litn = cr ? 0u : litn;
                        const uint32_t xj = (uint32_t)L.X[mm < kPMS ? mm : 0u][j < 64 ? j : 63u];
                        const bool hit = cr2 && mm < kPMS && xj == r - j * Wr;
                        tgt = end ? WT_END : hit ? j : tgt;
                        tm = hit ? mm : tm;
                        const bool mv = cr2 && !hit;
                        bnd += mv ? kPB : 0u;
                        const bool wrap = mv && mm + 1 == NMr;
                        mm = mv ? (wrap ? 0u : mm + 1) : mm;
                        j += wrap ? 1u : 0u;
                        act = act && !end && !hit;
                        WP_WIN();
                        WP_DEC(act);
                        WP_COUNT(act);
                        tgt = (act && ukind >= 2) ? (ukind == 2 ? WT_EOB : WT_ERR) : tgt;
                        act = act && ukind < 2;
                    }
                }
                const uint32_t r_end = r;  // where this lane's walk ended
                // ---- 3. the true lanes, from lane 0 on ----
                uint64_t valid = 0;
                uint32_t c = 0, sm = 0, mysm = 0, last = 0, last_t = 0;
                while (true) {
                    valid |= 1ull << c;
                    if (k == c) mysm = sm;
                    const uint32_t t = uni32((uint32_t)__builtin_amdgcn_readlane((int)tgt, (int)c));
                    if (t >= 64) {
                        last = c;
                        last_t = t;
                        break;
                    }
                    sm = uni32((uint32_t)__builtin_amdgcn_readlane((int)tm, (int)c));
                    c = t;
                }
                const uint64_t committed = last_t == WT_ERR ? (valid & ((1ull << last) - 1)) : valid;
                const bool com = (committed >> k) & 1ull;
                const uint32_t start = g + (uint32_t)myX[64u * mysm];  // the lane's range starts here (if true)
                const uint32_t ntok = com ? tok - (uint32_t)myT[64u * mysm] : 0u;
                uint32_t it = ntok;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const uint32_t o1 = (uint32_t)__shfl_up((int)it, d, 64);
                    if ((int)k >= d) it += o1;
                }
                const uint32_t Tr = uni32((uint32_t)__builtin_amdgcn_readlane((int)it, 63));
                if (!w_counting && (uint64_t)w_ntok + Tr > w_cap) w_counting = 1;
                // ---- 4. the true lanes decode their ranges again and store ----
                uint32_t bad = 0, emitted = 0, lc = 0;
                int32_t reach = 0;  // largest distance - cells before it: how far before the range a match reads
                {
                    const uint32_t stop_at = tgt < 64 ? tgt * Wr + (uint32_t)L.X[tm][tgt] : 0xFFFFFFFFu;
                    uint32_t nb2 = g + (mysm + 1) * kPB;
                    const uint32_t tbase = w_ntok + (it - ntok);
                    uint32_t ln = 0, lacc = 0;
                    // The crossing table has been read; its space now buffers each
                    // lane's tokens (at most 15).  They go to memory together, in
                    // 16-byte groups, when some lane holds 14 or more: stores in
                    // bursts instead of one every few units, so that the input
                    // reads of the walk seldom wait behind them.
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    uint32_t* const tbuf = &L.TB[0][k];  // slot j at tbuf[64 * j]
                    uint32_t nbf = 0, flushed = 0;  // tokens in the buffer; tokens stored
#define WP_PUT(t, V)                 \
    do {                             \
        tbuf[64u * (nbf & 15u)] = (t); \
        nbf += (V) ? 1u : 0u;        \
        emitted += (V) ? 1u : 0u;    \
    } while (0)
// stores the buffered tokens: single tokens up to a 16-byte boundary, then
// groups of four; fewer than four stay behind unless ALL
#define WP_FLUSH(ALL)                                                                      \
    do {                                                                                   \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                             \
        __builtin_amdgcn_wave_barrier();                                                   \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                             \
        const uint32_t _i0 = tbase + flushed;                                              \
        const uint32_t _al = min((4u - (_i0 & 3u)) & 3u, nbf);                             \
        uint32_t _j = 0;                                                                   \
        for (; _j < _al; ++_j)                                                             \
            if (!w_counting && _i0 + _j < w_cap) tkw[_i0 + _j] = tbuf[64u * _j];           \
        for (; _j + 4 <= nbf; _j += 4) {                                                   \
            const uint4 _v = make_uint4(tbuf[64u * _j], tbuf[64u * (_j + 1)], tbuf[64u * (_j + 2)], \
                                        tbuf[64u * (_j + 3)]);                             \
            if (!w_counting && _i0 + _j + 3 < w_cap) *reinterpret_cast<uint4*>(tkw + _i0 + _j) = _v; \
        }                                                                                  \
        if (ALL)                                                                           \
            for (; _j < nbf; ++_j)                                                         \
                if (!w_counting && _i0 + _j < w_cap) tkw[_i0 + _j] = tbuf[64u * _j];       \
        const uint32_t _rest = nbf - _j;                                                   \
        for (uint32_t _m = 0; _m < _rest; ++_m) tbuf[64u * _m] = tbuf[64u * (_j + _m)];    \
        flushed += _j;                                                                     \
        nbf = _rest;                                                                       \
    } while (0)
                    if (com) {
                        r = start;
                        WP_START();
                    }
                    act = com;
                    while (__ballot(act)) {
                        // a crossing: the pending literals close (token pa)
                        const bool cr = act && r >= nb2;
                        const bool hpa0 = cr && ln != 0;
                        uint32_t pa = (TOK_LIT << 30) | (ln << 24) | lacc;
                        ln = cr ? 0u : ln;
                        lacc = cr ? 0u : lacc;
                        nb2 += cr ? kPB : 0u;
                        act = act && !(r == stop_at || (tgt == WT_END && r >= RW));
                        WP_WIN();
                        WP_DEC(act);
                        const bool lit = act && ukind == 0, mat = act && ukind == 1, eob = act && ukind == 2;
                        // literals close before a match or the end of block (pa; never
                        // together with a crossing's, which left none pending)
                        const bool fl = (mat || eob) && ln != 0;
                        const uint32_t lacc1 = lacc | (usym << (8 * ln));
                        const bool lit3 = lit && ln == 2;
                        const uint32_t pb = lit3 ? (TOK_LIT << 30) | (3u << 24) | lacc1
                                                 : (TOK_MATCH << 30) | ((ulen - 3) << 16) | (udist - 1);
                        const bool hpb = lit3 || mat;
                        reach = mat ? max(reach, (int32_t)udist - (int32_t)lc) : reach;
                        lc += lit ? 1u : mat ? ulen : 0u;
                        ln = lit ? (lit3 ? 0u : ln + 1) : (act ? 0u : ln);
                        lacc = lit ? (lit3 ? 0u : lacc1) : (act ? 0u : lacc);
                        bad |= (act && ukind == 3) ? 1u : 0u;
                        r += act ? un : 0u;
                        act = lit || mat;
                        if (act && r > RW + 64) {  // past every range end: cannot happen on a true walk
                            bad = 1;
                            act = false;
                        }
                        WP_PUT(pa, hpa0 || fl);
                        WP_PUT(pb, hpb);
                        if (__ballot(nbf >= 14u)) WP_FLUSH(false);  // a unit adds at most 2: never past 15
                    }
                    WP_FLUSH(true);
#undef WP_FLUSH
#undef WP_PUT
                    if (com && !bad && emitted != ntok) bad = 1;
                }
                // cells: a prefix sum over the ranges; a back-reference before the
                // member's start fails its range (the serial decode from that range
                // then stops at the symbol, fast_inflate.rs:1704-1715)
                uint32_t ic = lc;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const uint32_t o2 = (uint32_t)__shfl_up((int)ic, d, 64);
                    if ((int)k >= d) ic += o2;
                }
                const uint32_t Cr = uni32((uint32_t)__builtin_amdgcn_readlane((int)ic, 63));
                const uint32_t cbase = w_outn + (ic - lc);
                if (com && !w_spec && reach > 0 && (uint32_t)reach > cbase - w_mbase) bad = 1;
                const uint64_t badm = __ballot(com && bad);
                if (badm) {
                    // lane 0 decodes on from the first failing range
                    const uint32_t f = (uint32_t)__builtin_ctzll(badm);
                    w_ntok += uni32((uint32_t)__builtin_amdgcn_readlane((int)(it - ntok), (int)f));
                    w_outn += uni32((uint32_t)__builtin_amdgcn_readlane((int)(ic - lc), (int)f));
                    P += uni32((uint32_t)__builtin_amdgcn_readlane((int)start, (int)f));
                    w_next = S_SYM;
                    break;
                }
                if ((uint64_t)w_outn + Cr >= kMaxSegCells) {
                    w_next = S_EXIT;
                    break;
                }
                w_ntok += Tr;
                w_outn += Cr;
                if (last_t == WT_ERR) {
                    P += uni32((uint32_t)__builtin_amdgcn_readlane((int)start, (int)last));
                    w_next = S_SYM;
                    break;
                }
                P += uni32((uint32_t)__builtin_amdgcn_readlane((int)r_end, (int)last));
                if (last_t == WT_EOB) {
                    w_next = S_BLOCK_END;
                    break;
                }
                // the body goes on past this round: the windows were sized from
                // the body before, which may have been much shorter
                Wwant = min(2u * Wwant, kPWMax);
                __syncthreads();  // the next round rewrites the crossing table
            }
#undef WP_COUNT
#undef WP_DEC
#undef WP_WIN
#undef WP_START
            if (lane == 0) {
                L.h.pos = P;
                L.h.n_tok = w_ntok;
                L.h.out_n = w_outn;
                L.h.counting = w_counting;
                L.h.next = w_next;
            }
        }
        __syncthreads();
        if (lane == 0) {
            // lane 0 goes on from where the wave stopped
            SerialState& S = L.ss;
            const uint32_t nx = L.h.next;
            S.pos = L.h.pos;
            S.n_tok = L.h.n_tok;
            S.out_n = L.h.out_n;
            S.counting = L.h.counting;
            if (nx == S_BLOCK_END) {
                // windows for the next body: about this one's size over 64 lanes
                const uint64_t body = S.pos - S.body_start;
                const uint64_t w = ((body / 60 + kPB - 1) / kPB) * kPB;
                S.parW = (uint32_t)(w < kPWMin ? kPWMin : w > kPWMax ? kPWMax : w);
                S.state = S_BLOCK_END;
            } else if (nx == S_SYM) {
                S.force_serial = 1;  // the rest of this body: serially, with every check
                S.state = S_SYM;
            } else if (nx == S_FINISH) {
                S.status = ST_INVALID;  // a table of the block is oversubscribed
                S.detail = DET_OVERSUBSCRIBED;
                S.state = S_FINISH;
            } else {
                S.status = ST_TOO_BIG;
                S.detail = DET_NONE;
                S.state = S_FINISH;
            }
        }
        __syncthreads();
    }
#undef WS_SYMBOL
#undef WS_REFILL_W
#undef WS_REFILL
#undef WS_CONSUME
#undef WS_FAIL
#undef WS_GETBITS
#undef WS_EMIT
#undef WS_FLUSH_LITS
}

int inflate_wave_lds_bytes(int lut_bits) {
    switch (lut_bits) {
        case 9: return wave_lds_bytes<9>();
        case 11: return wave_lds_bytes<11>();
        default: return wave_lds_bytes<10>();
    }
}

int inflate_wave_blocks_per_cu(int lut_bits) {
    const size_t shm = (size_t)inflate_wave_lds_bytes(lut_bits);
    int nb = 0;
    hipError_t e;
    switch (lut_bits) {
        case 9: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tokens_wave<9>, 64, shm); break;
        case 11: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tokens_wave<11>, 64, shm); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tokens_wave<10>, 64, shm); break;
    }
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_tokens_wave(const uint64_t* d_in64, const SegDesc* d_segs, SegResult* d_res, MemberRec* d_mrec,
                              const uint32_t* d_work, uint32_t n_work, uint32_t* d_counter, uint32_t* d_toks,
                              uint8_t* d_scratch, const InflateGeom& g, hipStream_t s) {
    if (n_work == 0) return hipSuccess;
    const size_t shm = (size_t)inflate_wave_lds_bytes(g.lut_bits);
    switch (g.lut_bits) {
        case 9:
            hipLaunchKernelGGL(k_tokens_wave<9>, dim3(g.blocks), dim3(64), shm, s, d_in64, d_segs, d_res, d_mrec,
                               d_work, n_work, d_counter, d_toks, d_scratch);
            break;
        case 11:
            hipLaunchKernelGGL(k_tokens_wave<11>, dim3(g.blocks), dim3(64), shm, s, d_in64, d_segs, d_res, d_mrec,
                               d_work, n_work, d_counter, d_toks, d_scratch);
            break;
        default:
            hipLaunchKernelGGL(k_tokens_wave<10>, dim3(g.blocks), dim3(64), shm, s, d_in64, d_segs, d_res, d_mrec,
                               d_work, n_work, d_counter, d_toks, d_scratch);
            break;
    }
    return hipGetLastError();
}

// ===========================================================================
// Expansion: one wavefront per segment turns its tokens into cells (or, for
// members that cannot hold markers, bytes).  Lane k of a batch owns token k:
// a wave prefix sum over the tokens' lengths places them (at most 64 tokens
// and kBatchCells cells per batch), and each lane writes its token's cells
// into the wave's LDS ring of recent output (kRing cells) in one pass.
// Literals and stored bytes are final; a match cell j is cell j - d (by
// periodicity, for d < 8, a cell before the match's start), which is a marker
// 0x8000 | (offset before the segment) when it precedes the segment
// (fast_inflate.rs:2209-2222), a final cell when it precedes the batch (read
// from the ring, or from the stored output when older than the ring), and
// otherwise a cell of this batch: then the cell holds a link to it, and the
// links are resolved in cell order afterwards.  Complete 64-cell lines then
// go from the ring to memory as 16-byte stores.
// ===========================================================================
constexpr uint32_t kRing = 2048;        // cells of recent output per wave, in LDS (8 waves per SIMD)
constexpr uint32_t kBatchCells = 1024;  // cells per batch (<= kRing / 2)
constexpr uint32_t kLink = 0x4000u;     // a ring cell that names the ring slot of its value's cell

__device__ __forceinline__ uint32_t tok_len(uint32_t w) {
    const uint32_t ty = w >> 30;
    return ty == TOK_LIT ? (w >> 24) & 3u : ty == TOK_MATCH ? ((w >> 16) & 0xffu) + 3u : ty == TOK_STORED ? (w & 0xffffu) : 0u;
}

__global__ __launch_bounds__(64) void k_expand(const uint32_t* __restrict__ toks, const ExpandItem* __restrict__ items,
                                               uint32_t n_items, uint32_t* __restrict__ counter,
                                               const uint8_t* __restrict__ in8, uint16_t* __restrict__ cells,
                                               uint8_t* __restrict__ out) {
    // the ring, then 8 cells that absorb the unconditional stores of cells past a
    // token's end (a select of the address instead of a branch per cell)
    __shared__ __align__(16) uint16_t ring[kRing + 8];
    constexpr uint32_t RM = kRing - 1;
    const int lane = (int)threadIdx.x;
    while (true) {
        uint32_t wi = 0;
        if (lane == 0) wi = atomicAdd(counter, 1u);
        wi = uni32((uint32_t)__shfl((int)wi, 0, 64));
        if (wi >= n_items) break;
        const uint64_t tok_base = uni64(items[wi].tok_base);
        const uint64_t dst = uni64(items[wi].dst);
        const uint64_t in_byte = uni64(items[wi].in_byte);
        const uint32_t n_cells = uni32((uint32_t)items[wi].n_cells);
        const uint32_t n_tok = uni32(items[wi].n_tok);
        const bool u8 = (uni32(items[wi].flags) & EXP_U8) != 0;
        const uint32_t* tk = toks + tok_base;
        uint16_t* c16 = cells + dst;
        uint8_t* o8 = out + dst;
        const uint8_t* src_in = in8 + in_byte;
        uint32_t P0 = 0;       // first cell of the batch
        uint32_t flushed = 0;  // cells stored to memory (a multiple of 64 until the end)
        // the tokens of the next batch are loaded while this one is expanded
        uint32_t wn = (uint32_t)lane < n_tok ? tk[lane] : (TOK_OFFSET << 30);
        for (uint32_t t0 = 0; t0 < n_tok;) {
            const uint32_t i = t0 + (uint32_t)lane;
            const uint32_t w = wn;
            const uint32_t ty = w >> 30;
            const uint32_t L = tok_len(w);
            uint32_t incl = L;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = (uint32_t)__shfl_up((int)incl, d, 64);
                if (lane >= d) incl += o;
            }
            // the tokens of this batch: a prefix of the lanes (every token is
            // at most 258 cells, so the first always fits)
            const uint32_t nb = (uint32_t)__popcll(__ballot(incl <= kBatchCells && i < n_tok));
            const uint32_t T = uni32((uint32_t)__builtin_amdgcn_readlane((int)incl, (int)nb - 1));
            {
                const uint32_t n2 = t0 + nb + (uint32_t)lane;  // lane's token of the next batch
                const uint32_t keep = (uint32_t)__shfl((int)w, (int)min((uint32_t)lane + nb, 63u), 64);
                // tokens already loaded move down; the rest are loaded now
                wn = (uint32_t)lane + nb < 64 ? keep : (n2 < n_tok ? tk[n2] : (TOK_OFFSET << 30));
            }
            const uint32_t S = P0 + incl - L;  // this token's first cell
            uint32_t aux = (uint32_t)__shfl((int)w, lane < 63 ? lane + 1 : 63, 64);
            if (ty == TOK_STORED && lane == 63 && i + 1 < n_tok) aux = tk[i + 1];
            const bool mine = (uint32_t)lane < nb && L != 0;
            // long matches and stored pieces are copied by the whole wave
            const bool longtok = L > 32 && (ty == TOK_MATCH || ty == TOK_STORED);
            const uint32_t ring_lo = P0 + T > kRing ? P0 + T - kRing : 0u;  // older cells are no longer in the ring
            // The value of cell sj, a copy of cell q = sj - dd: a marker when q
            // precedes the segment, a link to cell q when it lies in this batch
            // (kLink | ring slot, resolved below), the cell itself when it is
            // still in the ring, else the stored cell (read only when ld).
            auto cell = [&](uint32_t sj, uint32_t dd, bool ld) -> uint32_t {
                const uint32_t q = sj - dd;  // wraps when the source precedes the segment
                const bool mk = sj < dd;
                uint32_t x = ring[q & RM];
                x = q >= P0 ? (kLink | (q & RM)) : x;
                x = mk ? (0x8000u | (dd - sj - 1u)) : x;
                if (!mk && q < ring_lo && ld) x = u8 ? (uint32_t)o8[q] : (uint32_t)c16[q];
                return x;
            };
            // Every token is written once.  Literals and stored bytes are final;
            // a match's cells are final when its source precedes the batch and
            // links otherwise (32-bit cell arithmetic: a segment has < 2^32 cells).
            const uint32_t d = (w & 0x7fffu) + 1u;
            if (mine && !longtok) {
                if (ty == TOK_LIT) {
#pragma unroll
                    for (uint32_t j = 0; j < 3; ++j) ring[j < L ? ((S + j) & RM) : kRing + j] = (uint16_t)((w >> (8 * j)) & 0xffu);
                } else if (ty == TOK_STORED) {
                    const uint8_t* sp = src_in + (aux & 0x3fffffffu);
                    for (uint32_t j = 0; j < L; ++j) ring[(S + j) & RM] = sp[j];
                } else if (d >= 8) {
                    // cell j is cell j - d: a chunk of 8 never reads its own cells.
                    // A chunk whose source lies wholly before the segment
                    // (markers), in the batch (links), in the ring before the
                    // batch, or below the ring (stored) needs no per-cell cases.
                    for (uint32_t j0 = 0; j0 < L; j0 += 8) {
                        uint32_t v[8];
                        const uint32_t sj0 = S + j0, q = sj0 - d;
                        if (sj0 + 8 <= d) {
#pragma unroll
                            for (int u = 0; u < 8; ++u) v[u] = 0x8000u | (d - sj0 - (uint32_t)u - 1u);
                        } else if (sj0 >= d && q >= P0) {
#pragma unroll
                            for (int u = 0; u < 8; ++u) v[u] = kLink | ((q + (uint32_t)u) & RM);
                        } else if (sj0 >= d && q >= ring_lo && q + 8 <= P0) {
                            const uint32_t r = q & RM;
                            if (r + 8 <= kRing) {
#pragma unroll
                                for (int u = 0; u < 8; ++u) v[u] = ring[r + (uint32_t)u];
                            } else {
#pragma unroll
                                for (int u = 0; u < 8; ++u) v[u] = ring[(r + (uint32_t)u) & RM];
                            }
                        } else if (sj0 >= d && q + 8 <= ring_lo) {
                            if (u8) {
                                uint2 b;
                                __builtin_memcpy(&b, o8 + q, 8);
#pragma unroll
                                for (int u = 0; u < 8; ++u) v[u] = ((u < 4 ? b.x : b.y) >> (8 * (u & 3))) & 0xffu;
                            } else {
                                const uint4 c = load_cells8(c16 + q);
                                const uint32_t cw[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
                                for (int u = 0; u < 8; ++u) v[u] = (cw[u >> 1] >> (16 * (u & 1))) & 0xffffu;
                            }
                        } else {
#pragma unroll
                            for (int u = 0; u < 8; ++u) v[u] = cell(sj0 + (uint32_t)u, d, j0 + (uint32_t)u < L);
                        }
                        const uint32_t t0r = sj0 & RM, nrem = L - j0;
                        if (t0r + 8 <= kRing) {
#pragma unroll
                            for (int u = 0; u < 8; ++u)
                                ring[(uint32_t)u < nrem ? t0r + (uint32_t)u : kRing + (uint32_t)u] = (uint16_t)v[u];
                        } else {
#pragma unroll
                            for (int u = 0; u < 8; ++u)
                                ring[(uint32_t)u < nrem ? ((sj0 + (uint32_t)u) & RM) : kRing + (uint32_t)u] = (uint16_t)v[u];
                        }
                    }
                } else {
                    // d < 8: the source repeats with period d
                    uint32_t jm = 0;  // j mod d
                    for (uint32_t j0 = 0; j0 < L; j0 += 8) {
                        uint32_t v[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            v[u] = cell(S + jm, d, j0 + (uint32_t)u < L);
                            jm = (jm + 1 == d) ? 0u : jm + 1;
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            ring[j0 + (uint32_t)u < L ? ((S + j0 + (uint32_t)u) & RM) : kRing + (uint32_t)u] = (uint16_t)v[u];
                    }
                }
            }
            // long tokens: the 64 lanes copy each together, 64 cells at a time
            // (cell j comes from the cell j mod d after the start of its source).
            // With several in the batch (runs, short-period repeats) they go in
            // token order and their cells are final: an in-batch source cell is
            // read after the tokens before it are written, following the links
            // of short matches back to a final cell, so a batch of long matches
            // leaves no links to resolve.  A single one writes links (no waits).
            const uint64_t lmask = __ballot(mine && longtok);
            const bool lfinal = __popcll(lmask) >= 2;
            if (lfinal) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            for (uint64_t lm = lmask; lm; lm &= lm - 1) {
                const int k = (int)__builtin_ctzll(lm);
                const uint32_t kw = uni32((uint32_t)__builtin_amdgcn_readlane((int)w, k));
                const uint32_t kL = tok_len(kw);
                const uint32_t kS = uni32((uint32_t)__builtin_amdgcn_readlane((int)S, k));
                if ((kw >> 30) == TOK_STORED) {
                    const uint8_t* sp = src_in + (uni32((uint32_t)__builtin_amdgcn_readlane((int)aux, k)) & 0x3fffffffu);
                    for (uint32_t j = (uint32_t)lane; j < kL; j += 64) ring[(kS + j) & RM] = sp[j];
                } else {
                    const uint32_t kd = (kw & 0x7fffu) + 1u;
                    for (uint32_t j = (uint32_t)lane; j < kL; j += 64) {
                        const uint32_t jm = kd == 1 ? 0u : (j < kd ? j : j % kd);
                        uint32_t x = cell(kS + jm, kd, true);
                        // a link names an earlier cell of the batch: the chain ends
                        if (lfinal)
                            while ((x & 0xc000u) == kLink) x = ring[x & RM];
                        ring[(kS + j) & RM] = (uint16_t)x;
                    }
                }
                if (lfinal) {
                    // the next long token may read these cells
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // Links, in cell order, 64 cells at a time: a link points to an
            // earlier cell, so every cell of the rows before is final; within a
            // row each pass replaces a link by what its target holds (the final
            // value or that target's own link, further back), so a chain of n
            // links in a row takes about log2(n) passes.
            const uint64_t lkm = __ballot(mine && !(longtok && lfinal) && ty == TOK_MATCH && S + L > d + P0);  // its last source cell is in the batch
            if (lkm) {
                const uint32_t lo = uni32((uint32_t)__builtin_amdgcn_readlane((int)S, (int)__builtin_ctzll(lkm)));
                const uint32_t hi = uni32((uint32_t)__builtin_amdgcn_readlane((int)(S + L), 63 - (int)__builtin_clzll(lkm)));
                for (uint32_t c0 = lo; c0 < hi; c0 += 64) {
                    const uint32_t c = c0 + (uint32_t)lane;
                    uint32_t v = ring[c & RM];
                    bool lk = c < hi && (v & 0xc000u) == kLink;
                    while (__ballot(lk)) {
                        const uint32_t x = ring[v & RM];
                        v = lk ? x : v;
                        const bool lk2 = lk && (x & 0xc000u) == kLink;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        if (lk) ring[c & RM] = (uint16_t)v;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        lk = lk2;
                    }
                }
            }
            // store the complete lines (all of the rest after the last batch)
            const uint32_t end = P0 + T;
            const bool last = t0 + nb >= n_tok;
            const uint32_t upto = last ? min(end, n_cells) : (end & ~63u);
            for (uint32_t c = flushed + 8u * (uint32_t)lane; c < upto; c += 512) {
                const uint32_t r = c & RM;  // 8 cells never wrap: c and kRing are multiples of 8
                if (c + 8 <= upto) {
                    const uint4 g = *reinterpret_cast<const uint4*>(ring + r);
                    if (u8) {
                        const uint32_t lo = __builtin_amdgcn_perm(g.y, g.x, 0x06040200u);
                        const uint32_t hi = __builtin_amdgcn_perm(g.w, g.z, 0x06040200u);
                        *reinterpret_cast<uint2*>(o8 + c) = make_uint2(lo, hi);
                    } else {
                        store_cells8(c16 + c, g);  // a piece of a long item may start at any cell
                    }
                } else {
                    for (uint32_t k = c; k < upto; ++k) {
                        if (u8) o8[k] = (uint8_t)ring[k & RM];
                        else c16[k] = ring[k & RM];
                    }
                }
            }
            flushed = upto;
            P0 = end;
            t0 += nb;
        }
        // the next item reuses the ring: its first reads come after its own writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

hipError_t launch_expand(const uint32_t* d_toks, const ExpandItem* d_items, uint32_t n_items, uint32_t* d_counter,
                         const uint8_t* d_in8, uint16_t* d_cells, uint8_t* d_out, int n_cu, hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    // persistent single-wave workgroups (16 per CU; two waves per workgroup,
    // 32 per CU, measured slower: the kernel is instruction-bound)
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_expand, 64, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    const uint32_t blocks = std::min<uint32_t>((uint32_t)per_cu * (uint32_t)n_cu, n_items);
    hipLaunchKernelGGL(k_expand, dim3(blocks), dim3(64), 0, s, d_toks, d_items, n_items, d_counter, d_in8, d_cells,
                       d_out);
    return hipGetLastError();
}

// Cut points of long expansion items: one wavefront per item walks its
// tokens with a prefix sum over their lengths; for every m >= 1 with
// m * piece < n_cells, the token holding cell m * piece writes
// cuts[first[b] + m - 1] = (token index, its first cell).  Tokens cover the
// item's cells without gaps, so each record is written exactly once (offset
// tokens hold no cells and are never a cut).
__global__ __launch_bounds__(64) void k_expand_cuts(const uint32_t* __restrict__ toks,
                                                    const ExpandItem* __restrict__ items, const uint32_t* __restrict__ first,
                                                    uint32_t n, uint32_t piece, uint2* __restrict__ cuts) {
    const uint32_t b = blockIdx.x;
    if (b >= n) return;
    const int lane = (int)threadIdx.x;
    const uint32_t* tk = toks + items[b].tok_base;
    const uint32_t n_tok = items[b].n_tok;
    const uint64_t n_cells = items[b].n_cells;
    uint2* out = cuts + first[b];
    uint64_t base = 0;
    for (uint32_t t0 = 0; t0 < n_tok; t0 += 64) {
        const uint32_t i = t0 + (uint32_t)lane;
        const uint32_t L = i < n_tok ? tok_len(tk[i]) : 0u;
        uint32_t incl = L;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)incl, d, 64);
            if (lane >= d) incl += o;
        }
        const uint64_t s0 = base + incl - L, e0 = s0 + L;
        if (L != 0 && e0 <= n_cells) {
            uint64_t m = (s0 + piece - 1) / piece;
            if (m == 0) m = 1;
            for (; m * piece < e0; ++m) out[m - 1] = make_uint2(i, (uint32_t)s0);
        }
        base += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    }
}

hipError_t launch_expand_cuts(const uint32_t* d_toks, const ExpandItem* d_items, const uint32_t* d_first, uint32_t n,
                              uint32_t piece, uint2* d_cuts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_expand_cuts, dim3(n), dim3(64), 0, s, d_toks, d_items, d_first, n, piece, d_cuts);
    return hipGetLastError();
}

// ===========================================================================
// Tail chain.  A segment's markers can only reference the 32 Ki cells before
// it, which always lie in the last min(len, 32 Ki) cells of earlier segments
// (their "tail zones").  The reference carries that window serially from chunk
// to chunk (build_prev_tail_fast, pipeline.rs:1043-1078).  Here it is a
// three-phase scan over blocks of G consecutive segments:
//
//   1. k_tail_block (one workgroup per block, in parallel): sweep the block's
//      segments with the window held in LDS in *symbolic* form -- a cell is a
//      literal byte or a marker relative to the block's entry window (the 32 Ki
//      bytes before the block).  The window starts as the identity (cell i =
//      marker for entry byte i); each segment's tail cells are rewritten in
//      place to symbolic form and the final window is the block's exit tail.
//   2. entry windows: entry of block B+1 = exit tail of B resolved against
//      the entry of B.  The composition is associative, so it runs as a
//      three-level scan: k_tail_compose folds groups of ~sqrt(nb) blocks in
//      parallel, one k_tail_entry workgroup walks the group windows, and
//      k_tail_entry expands every group in parallel.
//   3. k_tail_patch (parallel): every symbolic marker in a tail zone becomes
//      the literal byte of its block's entry window.
//
// After phase 3 every tail-zone cell is a literal, so k_resolve needs one
// gather per marker.  Window cells carry 0x100 for bytes before the start of
// the output (a marker reaching there is the reference's "marker references
// bytes outside the available prefix window").  Several files can share one
// chain: a segment flagged `first` starts a new output, and the window is
// reset to 0x100 in front of it, so no marker reaches into another file.
// ===========================================================================
constexpr uint32_t WIN = 32768;

__device__ __forceinline__ bool marker_target(const ChainSeg* __restrict__ cs, uint32_t c, uint32_t off,
                                              uint32_t* s_out, uint64_t* local_out) {
    uint64_t rem = (uint64_t)off + 1;
    int64_t s = (int64_t)c - 1;
    while (s >= 0 && rem > cs[s].len) {
        if (cs[s].first) return false;  // before the start of this file's output
        rem -= cs[s].len;
        --s;
    }
    if (s < 0) return false;
    *s_out = (uint32_t)s;
    *local_out = cs[s].len - rem;
    return true;
}


// Two packed cells with each marker replaced by the window cell it names
// (tab[WIN - 1 - offset]); literal cells stay.
__device__ __forceinline__ uint32_t sym2(uint32_t w, const uint16_t* tab) {
    uint32_t lo = w & 0xffffu, hi = w >> 16;
    if (lo & 0x8000u) lo = tab[WIN - 1 - (lo & 0x7fffu)];
    if (hi & 0x8000u) hi = tab[WIN - 1 - (hi & 0x7fffu)];
    return lo | (hi << 16);
}

__global__ __launch_bounds__(1024) void k_tail_block(const ChainSeg* __restrict__ cs, uint32_t n_seg, uint32_t G,
                                                     uint16_t* __restrict__ cells, uint16_t* __restrict__ exit_tail) {
    __shared__ __align__(16) uint16_t win[WIN];
    const uint32_t B = blockIdx.x, tid = threadIdx.x;
    const uint32_t first = B * G, end = min(n_seg, first + G);
    const bool fresh = first < n_seg && cs[first].first;
    for (uint32_t i = tid; i < WIN; i += 1024) win[i] = fresh ? (uint16_t)0x100 : (uint16_t)(0x8000u | (WIN - 1 - i));
    __syncthreads();
    for (uint32_t s = first; s < end; ++s) {
        if (cs[s].first && s != first) {  // a new file: nothing before it is reachable
            for (uint32_t i = tid; i < WIN; i += 1024) win[i] = 0x100;
            __syncthreads();
        }
        const uint64_t len = cs[s].len;
        const uint32_t tl = (uint32_t)(len < WIN ? len : WIN);
        if (tl == 0) continue;
        const uint64_t tz0 = cs[s].cells_base + len - tl;  // first tail cell
        if (tl == WIN) {
            // the whole window is this segment's tail: four runs of 8 cells
            // per thread, all four loads in flight before the first lookup
            uint4 q[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) q[r] = load_cells8(cells + tz0 + 8u * (tid + 1024u * r));
#pragma unroll
            for (int r = 0; r < 4; ++r) q[r] = make_uint4(sym2(q[r].x, win), sym2(q[r].y, win), sym2(q[r].z, win),
                                                          sym2(q[r].w, win));
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t i = 8u * (tid + 1024u * r);
                *reinterpret_cast<uint4*>(win + i) = q[r];
                store_cells8(cells + tz0 + i, q[r]);
            }
            __syncthreads();
            continue;
        }
        const uint32_t keep = WIN - tl;  // window cells shifted down
        uint16_t v[WIN / 1024];
        // loads first (all in flight), then the lookups
#pragma unroll
        for (int r = 0; r < (int)(WIN / 1024); ++r) {
            const uint32_t i = tid + 1024u * r;
            v[r] = i >= keep ? cells[tz0 + (i - keep)] : (uint16_t)0;
        }
#pragma unroll
        for (int r = 0; r < (int)(WIN / 1024); ++r) {
            const uint32_t i = tid + 1024u * r;
            if (i < keep)
                v[r] = win[i + tl];
            else if (v[r] & 0x8000u)
                v[r] = win[WIN - 1 - (v[r] & 0x7fffu)];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < (int)(WIN / 1024); ++r) {
            const uint32_t i = tid + 1024u * r;
            win[i] = v[r];
            if (i >= keep) cells[tz0 + (i - keep)] = v[r];
        }
        __syncthreads();
    }
    uint4* out = reinterpret_cast<uint4*>(exit_tail + (size_t)B * WIN);
    const uint4* w4 = reinterpret_cast<const uint4*>(win);
    for (uint32_t i = tid; i < WIN / 8; i += 1024) out[i] = w4[i];
}

// Composition of symbolic windows: out[i] = X[i] when it is a literal (or
// 0x100), else Y's cell for the byte it names.  Both windows describe the 32 Ki
// bytes before some point; Y is relative to the point where X's markers refer.
// k_tail_compose: each workgroup folds the exit windows of K consecutive
// blocks into one window relative to the entry of the first of them.
__global__ __launch_bounds__(1024) void k_tail_compose(const uint16_t* __restrict__ exit_tail, uint32_t n_blocks,
                                                       uint32_t K, uint16_t* __restrict__ group_exit) {
    __shared__ __align__(16) uint16_t acc[WIN];
    const uint32_t tid = threadIdx.x, g = blockIdx.x;
    const uint32_t first = g * K, end = min(n_blocks, first + K);
    for (uint32_t i = tid; i < WIN; i += 1024) acc[i] = (uint16_t)(0x8000u | (WIN - 1 - i));  // identity
    __syncthreads();
    uint4* a4 = reinterpret_cast<uint4*>(acc);
    for (uint32_t B = first; B < end; ++B) {
        const uint4* ex = reinterpret_cast<const uint4*>(exit_tail + (size_t)B * WIN);
        uint4 q[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r] = ex[tid + 1024u * r];
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r] = make_uint4(sym2(q[r].x, acc), sym2(q[r].y, acc), sym2(q[r].z, acc),
                                                      sym2(q[r].w, acc));
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r) a4[tid + 1024u * r] = q[r];
        __syncthreads();
    }
    uint4* out = reinterpret_cast<uint4*>(group_exit + (size_t)g * WIN);
    for (uint32_t i = tid; i < WIN / 8; i += 1024) out[i] = a4[i];
}

// k_tail_entry: workgroup g walks blocks [gK, gK + K) in order, starting from
// init[g] (or "before the output start" when init is null): the entry window
// of each block, then the next entry = that block's exit resolved against it.
__global__ __launch_bounds__(1024) void k_tail_entry(const uint16_t* __restrict__ exit_tail, uint32_t n_blocks,
                                                     uint32_t K, const uint16_t* __restrict__ init,
                                                     uint16_t* __restrict__ entry) {
    __shared__ __align__(16) uint16_t ent[WIN];
    const uint32_t tid = threadIdx.x, g = blockIdx.x;
    const uint32_t first = g * K, end = min(n_blocks, first + K);
    for (uint32_t i = tid; i < WIN; i += 1024) ent[i] = init ? init[(size_t)g * WIN + i] : (uint16_t)0x100;
    __syncthreads();
    uint4* e4 = reinterpret_cast<uint4*>(ent);
    for (uint32_t B = first; B < end; ++B) {
        const uint4* ex = reinterpret_cast<const uint4*>(exit_tail + (size_t)B * WIN);
        uint4* en = reinterpret_cast<uint4*>(entry + (size_t)B * WIN);
        uint4 q[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r] = ex[tid + 1024u * r];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            en[tid + 1024u * r] = e4[tid + 1024u * r];
            q[r] = make_uint4(sym2(q[r].x, ent), sym2(q[r].y, ent), sym2(q[r].z, ent), sym2(q[r].w, ent));
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; ++r) e4[tid + 1024u * r] = q[r];
        __syncthreads();
    }
}

// Two packed tail cells made literal: a marker (relative to the block entry)
// becomes the entry window's byte; a 0x100 cell (a byte before the output
// start) sets `bad`.  The entry cell is read whatever the cell holds (the
// index is below WIN either way), so a run's reads are all in flight at once.
__device__ __forceinline__ uint32_t lit2(uint32_t w, const uint16_t* __restrict__ en, bool& bad) {
    uint32_t c[2] = {w & 0xffffu, w >> 16};
    const uint32_t e[2] = {en[WIN - 1 - (c[0] & 0x7fffu)], en[WIN - 1 - (c[1] & 0x7fffu)]};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t b = (c[h] & 0x8000u) ? e[h] : c[h];
        bad |= (b & 0x100u) != 0;
        c[h] = b & 0xffu;
    }
    return c[0] | (c[1] << 16);
}

__global__ __launch_bounds__(256) void k_tail_patch(const ChainSeg* __restrict__ cs, uint32_t n_seg, uint32_t G,
                                                    const uint16_t* __restrict__ entry, uint16_t* __restrict__ cells,
                                                    uint32_t* __restrict__ err) {
    const uint32_t s = blockIdx.x;
    if (s >= n_seg) return;
    const uint64_t len = cs[s].len;
    const uint32_t tl = (uint32_t)(len < WIN ? len : WIN);
    const uint64_t tz0 = cs[s].cells_base + len - tl;
    const uint16_t* en = entry + (size_t)(s / G) * WIN;
    if (tl == WIN) {
        // sixteen runs of 8 cells per thread, all loads in flight before the
        // first test; runs without a marker or 0x100 cell are not rewritten
        bool bad = false;
        uint4 q[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) q[r] = load_cells8(cells + tz0 + 8u * (threadIdx.x + 256u * r));
#pragma unroll
        for (int r = 0; r < 16; ++r)
            if ((q[r].x | q[r].y | q[r].z | q[r].w) & 0x81008100u) {
                q[r] = make_uint4(lit2(q[r].x, en, bad), lit2(q[r].y, en, bad), lit2(q[r].z, en, bad),
                                  lit2(q[r].w, en, bad));
                store_cells8(cells + tz0 + 8u * (threadIdx.x + 256u * r), q[r]);
            }
        if (bad) err[s] = 1u;
        return;
    }
    for (uint32_t k = threadIdx.x; k < tl; k += 256) {
        const uint16_t c = cells[tz0 + k];
        if (c & 0x8100u) {
            // a marker still relative to the block entry, or a byte before the output start
            const uint16_t b = (c & 0x8000u) ? en[WIN - 1 - (c & 0x7fffu)] : c;
            if (b & 0x100u) err[s] = 1u;
            cells[tz0 + k] = (uint16_t)(b & 0xffu);
        }
    }
}

// Blocks per group of the entry scan (about the square root of the block count).
uint32_t tail_group(uint32_t nb) {
    uint32_t K = 1;
    while (K * K < nb) ++K;
    return K;
}

size_t tail_scratch_bytes(uint32_t n_seg, uint32_t G) {
    const uint32_t nb = (n_seg + G - 1) / G;
    const uint32_t K = tail_group(nb), ng = (nb + K - 1) / K;
    return ((size_t)2 * nb + 2 * (size_t)ng) * WIN * sizeof(uint16_t);
}

hipError_t launch_tail(const ChainSeg* d_cs, uint32_t n_seg, uint32_t G, uint16_t* d_cells, uint16_t* d_scratch,
                       uint32_t* d_err, hipStream_t s) {
    if (n_seg == 0) return hipSuccess;
    const uint32_t nb = (n_seg + G - 1) / G;
    const uint32_t K = tail_group(nb), ng = (nb + K - 1) / K;
    uint16_t* d_exit = d_scratch;
    uint16_t* d_entry = d_exit + (size_t)nb * WIN;
    uint16_t* d_gexit = d_entry + (size_t)nb * WIN;
    uint16_t* d_gentry = d_gexit + (size_t)ng * WIN;
    hipLaunchKernelGGL(k_tail_block, dim3(nb), dim3(1024), 0, s, d_cs, n_seg, G, d_cells, d_exit);
    // entry windows: groups of K blocks folded in parallel, a short serial pass
    // over the ng group windows, then every group expanded in parallel
    hipLaunchKernelGGL(k_tail_compose, dim3(ng), dim3(1024), 0, s, d_exit, nb, K, d_gexit);
    hipLaunchKernelGGL(k_tail_entry, dim3(1), dim3(1024), 0, s, d_gexit, ng, ng, (const uint16_t*)nullptr, d_gentry);
    hipLaunchKernelGGL(k_tail_entry, dim3(ng), dim3(1024), 0, s, d_exit, nb, K, (const uint16_t*)d_gentry, d_entry);
    hipLaunchKernelGGL(k_tail_patch, dim3(n_seg), dim3(256), 0, s, d_cs, n_seg, G, d_entry, d_cells, d_err);
    return hipGetLastError();
}

// ===========================================================================
// Resolve + narrow.  Each thread owns 16-byte output chunks (16-byte aligned
// in output coordinates).  The 16 cells behind a chunk start at a phase that
// is constant over the tile (and so over the wave): three aligned 16-byte
// cell loads and a wave-uniform funnel give them.  Without markers the low
// bytes are packed with v_perm; marker cells read the (already patched,
// literal) cell they copy from the preceding tail zone.  Chunks that straddle
// a tile edge are written byte by byte.
// ===========================================================================
__device__ __forceinline__ uint32_t resolve_cell(const ChainSeg* __restrict__ cs, const uint16_t* __restrict__ cells,
                                                 uint32_t seg, uint32_t cell, uint32_t* __restrict__ err) {
    if (cell & 0x8000u) {
        uint32_t s;
        uint64_t local;
        if (marker_target(cs, seg, cell & 0x7fffu, &s, &local)) return cells[cs[s].cells_base + local] & 0xffu;
        err[seg] = 1u;
        return 0u;
    }
    return cell & 0xffu;
}

constexpr int RESOLVE_THREADS = 512;
constexpr int WIN_PARTS = 64;  // chain segments a window may be gathered from

__global__ __launch_bounds__(RESOLVE_THREADS) void k_resolve(const ResolveTile* __restrict__ tiles,
                                                             const ChainSeg* __restrict__ cs,
                                                             const uint16_t* __restrict__ cells,
                                                             uint8_t* __restrict__ out, uint32_t* __restrict__ err) {
    // byte that marker offset m names: win[WIN - 1 - m], valid for m < s_avail
    __shared__ uint8_t win[WIN];
    __shared__ uint64_t part_src[WIN_PARTS];
    __shared__ uint32_t part_dst[WIN_PARTS], part_n[WIN_PARTS];
    __shared__ uint32_t s_parts, s_avail;
    const ResolveTile t = tiles[blockIdx.x];
    const bool spec = cs[t.seg].spec != 0;
    if (spec) {
        // the 32 Ki bytes before the segment: tail zones (already literal) of
        // the chain segments before it, back to the start of its file
        if (threadIdx.x == 0) {
            uint32_t rem = WIN, np = 0;
            for (int64_t s = (int64_t)t.seg - 1; s >= 0 && rem; --s) {
                const uint64_t len = cs[s].len;
                const uint32_t take = (uint32_t)(len < rem ? len : rem);
                if (take) {
                    if (np == WIN_PARTS) break;  // fragmented window: markers use the chain walk
                    part_src[np] = cs[s].cells_base + len - take;
                    part_dst[np] = rem - take;
                    part_n[np] = take;
                    ++np;
                    rem -= take;
                }
                if (cs[s].first) break;
            }
            s_parts = (rem && np == WIN_PARTS) ? ~0u : np;
            s_avail = WIN - rem;
        }
        __syncthreads();
        const uint32_t np = s_parts;
        if (np != ~0u)
            for (uint32_t q = 0; q < np; ++q) {
                const uint64_t src = part_src[q];
                const uint32_t dst = part_dst[q], n = part_n[q];
                // 8 cells per load (the window is read once per tile)
                for (uint32_t i = 8u * threadIdx.x; i < n; i += 8u * RESOLVE_THREADS) {
                    if (i + 8 <= n) {
                        const uint4 g = load_cells8(cells + src + i);
                        const uint32_t lo = __builtin_amdgcn_perm(g.y, g.x, 0x06040200u);
                        const uint32_t hi = __builtin_amdgcn_perm(g.w, g.z, 0x06040200u);
#pragma unroll
                        for (uint32_t b = 0; b < 8; ++b) win[dst + i + b] = (uint8_t)((b < 4 ? lo : hi) >> (8 * (b & 3)));
                    } else {
                        for (uint32_t k = i; k < n; ++k) win[dst + k] = (uint8_t)cells[src + k];
                    }
                }
            }
        __syncthreads();
    }
    const bool lds_win = spec && s_parts != ~0u;
    const uint32_t avail = spec ? s_avail : 0u;
    const uint64_t ob = t.out_off, oe = t.out_off + t.n;
    // cell index of output byte p is delta + p
    const int64_t delta = (int64_t)t.cell_addr - (int64_t)ob;
    const uint32_t phase = (uint32_t)((uint64_t)delta & 7u);  // cells, within a 16-byte cell group
    const uint32_t dsh = phase >> 1, odd = phase & 1u;
    const uint4* c128 = reinterpret_cast<const uint4*>(cells);
    // two steps per iteration: the next step's loads are issued before this one's store
#pragma unroll 2
    for (uint64_t p0 = (ob & ~15ull) + 16ull * threadIdx.x; p0 < oe; p0 += 16ull * RESOLVE_THREADS) {
        const int64_t q = (delta + (int64_t)p0 - (int64_t)phase) >> 3;  // first 16-byte cell group
        // a group before the buffer start only holds cells of bytes before the tile
        const uint4 g0 = c128[q < 0 ? 0 : q], g1 = c128[q + 1 < 0 ? 0 : q + 1], g2 = c128[q + 2 < 0 ? 0 : q + 2];
        const uint32_t w[12] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w, g2.x, g2.y, g2.z, g2.w};
        uint32_t o[8];
        // wave-uniform selection of the funnel (phase is constant over the tile)
#define RRG_FUNNEL(D)                                                                            \
    _Pragma("unroll") for (int j = 0; j < 8; ++j) o[j] = odd ? ((w[j + (D)] >> 16) | (w[j + (D) + 1] << 16)) \
                                                             : w[j + (D)];
        switch (dsh) {
            case 0: RRG_FUNNEL(0) break;
            case 1: RRG_FUNNEL(1) break;
            case 2: RRG_FUNNEL(2) break;
            default: RRG_FUNNEL(3) break;
        }
#undef RRG_FUNNEL
        uint32_t b[4];
        const uint32_t any = (o[0] | o[1] | o[2] | o[3] | o[4] | o[5] | o[6] | o[7]) & 0x80008000u;
        if (!any) {
#pragma unroll
            for (int m = 0; m < 4; ++m) b[m] = __builtin_amdgcn_perm(o[2 * m + 1], o[2 * m], 0x06040200u);
        } else {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                uint32_t v = 0;
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const uint32_t cell = (o[2 * m + (h >> 1)] >> (16 * (h & 1))) & 0xffffu;
                    const uint64_t p = p0 + (uint64_t)(4 * m + h);
                    uint32_t byte = cell & 0xffu;
                    if ((cell & 0x8000u) && p >= ob && p < oe) {
                        const uint32_t off = cell & 0x7fffu;
                        if (lds_win) {
                            if (off < avail) {
                                byte = win[WIN - 1 - off];
                            } else {
                                err[t.seg] = 1u;
                                byte = 0;
                            }
                        } else {
                            byte = resolve_cell(cs, cells, t.seg, cell, err);
                        }
                    }
                    v |= byte << (8 * h);
                }
                b[m] = v;
            }
        }
        if (p0 >= ob && p0 + 16 <= oe) {
            *reinterpret_cast<uint4*>(out + p0) = make_uint4(b[0], b[1], b[2], b[3]);
        } else {
            for (int i = 0; i < 16; ++i) {
                const uint64_t p = p0 + (uint64_t)i;
                if (p >= ob && p < oe) out[p] = (uint8_t)(b[i >> 2] >> (8 * (i & 3)));
            }
        }
    }
}

hipError_t launch_resolve(const ResolveTile* d_tiles, uint32_t n_tiles, const ChainSeg* d_cs, const uint16_t* d_cells,
                          uint8_t* d_out, uint32_t* d_err, hipStream_t s) {
    if (n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_resolve, dim3(n_tiles), dim3(RESOLVE_THREADS), 0, s, d_tiles, d_cs, d_cells, d_out, d_err);
    return hipGetLastError();
}

// ===========================================================================
// CRC-32 of pieces (raw: register starts at 0, no final inversion; the host
// folds pieces into member CRCs with GF(2) shifts).  256 threads per piece.
// The piece is read as aligned 8-byte words (bytes before the piece masked to
// zero, which a zero register ignores) and the words are dealt round-robin:
// thread t takes every 256th word, so a wave reads 512 contiguous bytes per
// step.  Thread t's words are 2 KiB apart: acc <- acc * x^(8*2048) + crc(word),
// the multiplication by four 256-entry tables; crc(word) by slice-by-8 tables.
// The 256 partial values are folded in a tree with x^(64*2^k); the last
// (unaligned) bytes of the piece follow byte by byte.
// ===========================================================================
// GF(2) constants of the fold, passed by value: shift[k] = x^(64*2^k) mod P
// (2^k words of 8 bytes), step = x^(8*2048) mod P (256 words), reflected
struct CrcConsts {
    uint32_t shift[8];
    uint32_t step;
};

__device__ __forceinline__ uint32_t gf2_mul(uint32_t a, uint32_t b) {
    // a*b mod P over GF(2), reflected bit order (bit 31 = x^0)
    uint32_t prod = 0;
#pragma unroll 4
    for (int i = 0; i < 32; ++i) {
        prod ^= (0u - ((a >> (31 - i)) & 1u)) & b;
        b = (b >> 1) ^ (0xEDB88320u & (0u - (b & 1u)));
    }
    return prod;
}

__global__ __launch_bounds__(256) void k_crc_pieces(const uint8_t* __restrict__ out,
                                                    const uint64_t* __restrict__ piece_off,
                                                    const uint32_t* __restrict__ piece_len, uint32_t* __restrict__ crc,
                                                    const CrcConsts K) {
    __shared__ uint32_t T[8][256];  // slice-by-8
    __shared__ uint32_t M[4][256];  // v -> v * x^(8*2048), by bytes of v
    __shared__ uint32_t part[256];
    const uint32_t tid = threadIdx.x;
    {
        uint32_t c = tid;
        for (int k = 0; k < 8; ++k) c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
        T[0][tid] = c;
        const uint32_t X = K.step;
#pragma unroll
        for (int j = 0; j < 4; ++j) M[j][tid] = gf2_mul(X, tid << (8 * j));
    }
    __syncthreads();
    for (int s = 1; s < 8; ++s) {
        const uint32_t prev = T[s - 1][tid];
        T[s][tid] = (prev >> 8) ^ T[0][prev & 0xffu];
        __syncthreads();
    }
    const uint64_t start = piece_off[blockIdx.x];
    const uint64_t end = start + piece_len[blockIdx.x];
    const uint64_t a0 = start & ~7ull, E = end & ~7ull;
    const uint64_t N = E > start ? (E - a0) >> 3 : 0;  // aligned words
    const uint64_t tail = E > start ? E : start;       // bytes [tail, end) follow one by one
    const uint64_t* w64 = reinterpret_cast<const uint64_t*>(out) + (a0 >> 3);
    uint32_t acc = 0;
    const uint64_t r0 = 255 - tid;  // this thread's words: j = N - 1 - (r0 + 256 m)
    if (N > r0) {
        for (int64_t m = (int64_t)((N - 1 - r0) >> 8); m >= 0; --m) {
            const uint64_t j = N - 1 - (r0 + 256ull * (uint64_t)m);
            uint64_t v = w64[j];
            if (j == 0) v &= ~0ull << (8 * (start & 7));
            const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
            const uint32_t cw = T[7][lo & 0xffu] ^ T[6][(lo >> 8) & 0xffu] ^ T[5][(lo >> 16) & 0xffu] ^
                                T[4][lo >> 24] ^ T[3][hi & 0xffu] ^ T[2][(hi >> 8) & 0xffu] ^
                                T[1][(hi >> 16) & 0xffu] ^ T[0][hi >> 24];
            acc = M[0][acc & 0xffu] ^ M[1][(acc >> 8) & 0xffu] ^ M[2][(acc >> 16) & 0xffu] ^ M[3][acc >> 24] ^ cw;
        }
    }
    part[tid] = acc;
    __syncthreads();
    for (int k = 0; k < 8; ++k) {
        const uint32_t span = 1u << k;
        if ((tid & ((span << 1) - 1)) == 0) part[tid] = gf2_mul(K.shift[k], part[tid]) ^ part[tid + span];
        __syncthreads();
    }
    if (tid == 0) {
        uint32_t c = part[0];
        for (uint64_t b = tail; b < end; ++b) c = (c >> 8) ^ T[0][(c ^ out[b]) & 0xffu];
        crc[blockIdx.x] = c;
    }
}

// ===========================================================================
// Member reach check of the single-worker paths.  There a back-reference
// before the start of its member is an error (fast_inflate.rs:1704-1715,
// inflate.rs:339-346), but a segment that starts inside a member turns it
// into a marker like any other.  The first cell of a listed segment whose
// marker offset is at or beyond the segment's reach is that back-reference's
// first output byte.  One workgroup per segment, 8 cells per thread step.
// ===========================================================================
__global__ __launch_bounds__(256) void k_marker_reach(const ReachScan* __restrict__ items, uint32_t n,
                                                      const uint16_t* __restrict__ cells, uint32_t* __restrict__ first) {
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const ReachScan it = items[i];
        const uint4* c128 = reinterpret_cast<const uint4*>(cells + it.cells_base);
        const uint32_t groups = (it.len + 7u) >> 3;
        uint32_t best = ~0u;
        for (uint32_t g = threadIdx.x; g < groups; g += 256u) {
            const uint4 v = c128[g];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t c = (w[k >> 1] >> (16 * (k & 1))) & 0xffffu;
                const uint32_t idx = 8u * g + (uint32_t)k;
                if (idx < it.len && (c & 0x8000u) && (c & 0x7fffu) >= it.reach && idx < best) best = idx;
            }
            if (best != ~0u) break;  // later groups of this thread only hold later cells
        }
        if (best != ~0u) atomicMin(first + i, best);
    }
}

hipError_t launch_marker_reach(const ReachScan* d_items, uint32_t n, const uint16_t* d_cells, uint32_t* d_first,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_marker_reach, dim3(n < 4096u ? n : 4096u), dim3(256), 0, s, d_items, n, d_cells, d_first);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_widen(const uint8_t* __restrict__ src, uint16_t* __restrict__ cells,
                                               uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        cells[i] = src[i];
}

hipError_t launch_widen(const uint8_t* d_src, uint16_t* d_cells, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_widen, dim3((uint32_t)(blocks < 1024 ? blocks : 1024)), dim3(256), 0, s, d_src, d_cells, n);
    return hipGetLastError();
}

// The fold constants, computed once (thread-safe static initialisation).
static CrcConsts make_crc_consts() {
    auto mul = [](uint32_t a, uint32_t b) {
        uint32_t prod = 0;
        for (int i = 0; i < 32; ++i) {
            if (a & (0x80000000u >> i)) prod ^= b;
            b = (b & 1u) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
        }
        return prod;
    };
    CrcConsts K{};
    uint32_t x = 0x40000000u;                   // x^1
    for (int i = 0; i < 6; ++i) x = mul(x, x);  // x^64: one 8-byte word
    for (int k = 0; k < 8; ++k) {
        K.shift[k] = x;
        x = mul(x, x);
    }
    K.step = x;  // x^(64*256) = x^(8*2048)
    return K;
}

hipError_t launch_crc_pieces(const uint8_t* d_out, const uint64_t* d_piece_off, const uint32_t* d_piece_len,
                             uint32_t n_pieces, uint32_t* d_crc, hipStream_t s) {
    if (n_pieces == 0) return hipSuccess;
    static const CrcConsts K = make_crc_consts();
    hipLaunchKernelGGL(k_crc_pieces, dim3(n_pieces), dim3(256), 0, s, d_out, d_piece_off, d_piece_len, d_crc, K);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The post-expansion kernels as one graph.  Kernel nodes are described by
// their function, grid and argument pointers; the arguments are copied when
// the node is set, so locals will do.
// ---------------------------------------------------------------------------
void PostGraphs::release() {
    for (int v = 0; v < 8; ++v) {
        if (exec[v]) (void)hipGraphExecDestroy(exec[v]);
        if (graph[v]) (void)hipGraphDestroy(graph[v]);
        exec[v] = nullptr;
        graph[v] = nullptr;
        n_nodes[v] = 0;
    }
}

namespace {

struct NodeArgs {
    // the argument values of every node of one launch, alive while the nodes are set
    const ChainSeg* cs;
    uint32_t n_seg, G, nb, K, ng, one;
    uint16_t *cells, *exit_w, *entry_w, *gexit_w, *gentry_w;
    const uint16_t* null16;
    uint32_t* err;
    const ResolveTile* tiles;
    uint8_t* out;
    const uint64_t* poff;
    const uint32_t* plen;
    uint32_t* pcrc;
    CrcConsts K_crc;
    void* a_block[5];
    void* a_compose[4];
    void* a_entry1[5];
    void* a_entry2[5];
    void* a_patch[6];
    void* a_resolve[5];
    void* a_crc[5];
};

int post_nodes(const PostLaunch& p, NodeArgs& A, hipKernelNodeParams* kp) {
    int n = 0;
    auto node = [&](const void* f, uint32_t grid, uint32_t block, void** args) {
        hipKernelNodeParams q{};
        q.func = const_cast<void*>(f);
        q.gridDim = dim3(grid);
        q.blockDim = dim3(block);
        q.sharedMemBytes = 0;
        q.kernelParams = args;
        q.extra = nullptr;
        kp[n++] = q;
    };
    if (p.tail) {
        A.cs = p.cs;
        A.n_seg = p.n_seg;
        A.G = p.G;
        A.nb = (p.n_seg + p.G - 1) / p.G;
        A.K = tail_group(A.nb);
        A.ng = (A.nb + A.K - 1) / A.K;
        A.one = 1;
        A.cells = p.cells;
        A.exit_w = p.scratch;
        A.entry_w = A.exit_w + (size_t)A.nb * WIN;
        A.gexit_w = A.entry_w + (size_t)A.nb * WIN;
        A.gentry_w = A.gexit_w + (size_t)A.ng * WIN;
        A.null16 = nullptr;
        A.err = p.err;
        void* ab[] = {&A.cs, &A.n_seg, &A.G, &A.cells, &A.exit_w};
        std::memcpy(A.a_block, ab, sizeof ab);
        void* ac[] = {&A.exit_w, &A.nb, &A.K, &A.gexit_w};
        std::memcpy(A.a_compose, ac, sizeof ac);
        void* ae1[] = {&A.gexit_w, &A.ng, &A.ng, &A.null16, &A.gentry_w};
        std::memcpy(A.a_entry1, ae1, sizeof ae1);
        void* ae2[] = {&A.exit_w, &A.nb, &A.K, &A.gentry_w, &A.entry_w};
        std::memcpy(A.a_entry2, ae2, sizeof ae2);
        void* ap[] = {&A.cs, &A.n_seg, &A.G, &A.entry_w, &A.cells, &A.err};
        std::memcpy(A.a_patch, ap, sizeof ap);
        node((const void*)k_tail_block, A.nb, 1024, A.a_block);
        node((const void*)k_tail_compose, A.ng, 1024, A.a_compose);
        node((const void*)k_tail_entry, 1, 1024, A.a_entry1);
        node((const void*)k_tail_entry, A.ng, 1024, A.a_entry2);
        node((const void*)k_tail_patch, p.n_seg, 256, A.a_patch);
    }
    if (p.n_tiles) {
        A.tiles = p.tiles;
        A.cs = p.cs;
        A.cells = p.cells;
        A.out = p.out;
        A.err = p.err;
        void* ar[] = {&A.tiles, &A.cs, &A.cells, &A.out, &A.err};
        std::memcpy(A.a_resolve, ar, sizeof ar);
        node((const void*)k_resolve, p.n_tiles, RESOLVE_THREADS, A.a_resolve);
    }
    if (p.n_pieces) {
        static const CrcConsts Kc = make_crc_consts();
        A.K_crc = Kc;
        A.out = p.out;
        A.poff = p.poff;
        A.plen = p.plen;
        A.pcrc = p.pcrc;
        void* aq[] = {&A.out, &A.poff, &A.plen, &A.pcrc, &A.K_crc};
        std::memcpy(A.a_crc, aq, sizeof aq);
        node((const void*)k_crc_pieces, p.n_pieces, 256, A.a_crc);
    }
    return n;
}

}  // namespace

hipError_t launch_post(const PostLaunch& p, PostGraphs* g, hipStream_t s) {
    if (!g) {
        hipError_t e = p.tail ? launch_tail(p.cs, p.n_seg, p.G, p.cells, p.scratch, p.err, s) : hipSuccess;
        if (e == hipSuccess) e = launch_resolve(p.tiles, p.n_tiles, p.cs, p.cells, p.out, p.err, s);
        if (e == hipSuccess) e = launch_crc_pieces(p.out, p.poff, p.plen, p.n_pieces, p.pcrc, s);
        return e;
    }
    const int v = (p.tail ? 1 : 0) | (p.n_tiles ? 2 : 0) | (p.n_pieces ? 4 : 0);
    if (!v) return hipSuccess;
    NodeArgs A{};
    hipKernelNodeParams kp[7];
    const int n = post_nodes(p, A, kp);
    hipError_t e = hipSuccess;
    if (!g->exec[v]) {
        // built once: the nodes in a chain, each after the one before
        e = hipGraphCreate(&g->graph[v], 0);
        for (int k = 0; k < n && e == hipSuccess; ++k)
            e = hipGraphAddKernelNode(&g->node[v][k], g->graph[v], k ? &g->node[v][k - 1] : nullptr, k ? 1 : 0,
                                      &kp[k]);
        if (e == hipSuccess) e = hipGraphInstantiate(&g->exec[v], g->graph[v], nullptr, nullptr, 0);
        if (e != hipSuccess) {
            if (g->graph[v]) (void)hipGraphDestroy(g->graph[v]);
            g->graph[v] = nullptr;
            g->exec[v] = nullptr;
            return e;
        }
        g->n_nodes[v] = n;
    } else {
        for (int k = 0; k < n && e == hipSuccess; ++k) e = hipGraphExecKernelNodeSetParams(g->exec[v], g->node[v][k], &kp[k]);
        if (e != hipSuccess) return e;
    }
    return hipGraphLaunch(g->exec[v], s);
}

}  // namespace rrg
