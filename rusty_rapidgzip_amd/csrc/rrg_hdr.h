// Strict dynamic-header check for one candidate position, written for one
// lane at a time: every lane of a wave runs it on its own candidate, with its
// own bit reader and precode tables held in registers.  The finder runs this
// over a batch of candidates and passes only the survivors to its
// wave-cooperative verification, which repeats the check and adds the trial
// decode; a position passes here exactly when that strict header parse
// succeeds (block_finder.rs:196-262).
//
// The same source compiles for the host (inline, no HIP headers), where the
// CPU tests compare it with the oracle position by position.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RRG_HD __host__ __device__ __forceinline__
#else
#define RRG_HD inline
#endif

namespace rrg {

// the 7 bits of x, reversed (the first stream bit becomes the most significant)
RRG_HD uint32_t rev7(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bitreverse32(x) >> 25;
#else
    x = ((x & 0x55u) << 1) | ((x >> 1) & 0x55u);
    x = ((x & 0x33u) << 2) | ((x >> 2) & 0x33u);
    x = ((x & 0x0fu) << 4) | ((x >> 4) & 0x0fu);
    return (x & 0xffu) >> 1;
#endif
}

RRG_HD uint32_t popcount32(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

// Cheap stage (block_finder.rs:196-213): header fields + strict precode Kraft.
// `lo`/`hi` hold the 128 bits at p; branch-free over the 19 possible fields.
RRG_HD bool precode_ok(uint64_t lo, uint64_t hi, uint64_t p, uint64_t nbits) {
    const uint32_t hclen = 4u + (uint32_t)((lo >> 13) & 15u);
    if (p + 17 + 3ull * hclen > nbits) return false;
    const uint64_t pc = (lo >> 17) | (hi << 47);  // precode lengths, 3 bits each
    uint32_t sum = 0, nsym = 0;
#pragma unroll
    for (uint32_t i = 0; i < 19; ++i) {
        uint32_t l = (uint32_t)(pc >> (3 * i)) & 7u;
        l = i < hclen ? l : 0u;
        sum += (128u >> l) & (0u - (uint32_t)(l != 0));
        nsym += (uint32_t)(l != 0);
    }
    return nsym == 1 || (nsym >= 2 && sum == 128u);
}

// The same test with a table of three code lengths per entry: entry k holds
// the Kraft sum (units of 2^-7, bits 0-15) and the count (bits 16-31) of the
// nonzero 3-bit lengths packed in k.  Seven lookups cover the 19 lengths
// (the lengths past hclen are masked to zero first).
RRG_HD uint32_t precode3_entry(uint32_t k) {
    uint32_t sum = 0, n = 0;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const uint32_t l = (k >> (3 * f)) & 7u;
        if (l) {
            sum += 128u >> l;
            ++n;
        }
    }
    return sum | (n << 16);
}
RRG_HD bool precode_ok_lut(const uint32_t* __restrict__ tab, uint64_t lo, uint64_t hi, uint64_t p,
                                               uint64_t nbits) {
    const uint32_t hclen = 4u + (uint32_t)((lo >> 13) & 15u);
    if (p + 17 + 3ull * hclen > nbits) return false;
    const uint64_t pc = ((lo >> 17) | (hi << 47)) & ((1ull << (3 * hclen)) - 1ull);
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i) acc += tab[(uint32_t)(pc >> (9 * i)) & 511u];
    const uint32_t sum = acc & 0xffffu, n = acc >> 16;
    return n == 1u || (n >= 2u && sum == 128u);
}

// `fetch(q)` returns 64-bit word q of the stream (zero past its end); bits
// are little-endian within words.  `p` is the block's first bit (BFINAL);
// BFINAL and BTYPE are the caller's.  `nbits` bounds every read as the
// reference's bit reader does: the code-length probe needs 10 bits present,
// and each field its own width.
template <class Fetch>
RRG_HD bool strict_header_lane(const Fetch& fetch, uint64_t nbits, uint64_t p) {
    if (p + 17 > nbits) return false;
    // words cw, cw + 1 are the current pair; cw + 2 is loaded ahead
    uint64_t cw = p >> 6, pos = p;
    uint64_t a = fetch(cw), b = fetch(cw + 1), c = fetch(cw + 2);
    const uint64_t w = (pos & 63) ? (a >> (pos & 63)) | (b << (64 - (pos & 63))) : a;
    const uint32_t hlit = 257u + (uint32_t)((w >> 3) & 31u);
    const uint32_t hdist = 1u + (uint32_t)((w >> 8) & 31u);
    const uint32_t hclen = 4u + (uint32_t)((w >> 13) & 15u);
    if (hlit > 286u || hdist > 30u || p + 17 + 3ull * hclen > nbits) return false;
    auto skip = [&](uint32_t n) __attribute__((always_inline)) {
        pos += n;
        while ((pos >> 6) != cw) {
            ++cw;
            a = b;
            b = c;
            c = fetch(cw + 2);
        }
    };
    auto peek = [&]() __attribute__((always_inline)) -> uint64_t {
        const uint32_t sh = (uint32_t)(pos & 63);
        return sh ? (a >> sh) | (b << (64 - sh)) : a;
    };
    skip(17);
    const uint64_t pc = peek() & ((1ull << (3 * hclen)) - 1);  // precode lengths in header order
    skip(3 * hclen);
    // precode lengths in symbol order; counts per length (8-bit fields);
    // strict Kraft (a single code is accepted, as the reference does)
    constexpr uint8_t kPos[19] = {3, 17, 15, 13, 11, 9, 7, 5, 4, 6, 8, 10, 12, 14, 16, 18, 0, 1, 2};
    uint64_t ls = 0, cnt = 0;
    uint32_t nsym = 0, kraft = 0;
#pragma unroll
    for (uint32_t s = 0; s < 19; ++s) {
        const uint32_t L = (uint32_t)(pc >> (3 * kPos[s])) & 7u;
        ls |= (uint64_t)L << (3 * s);
        cnt += 1ull << (8 * L);
        nsym += L != 0;
        kraft += L ? 128u >> L : 0u;
    }
    if (!(nsym == 1 || (nsym >= 2 && kraft == 128u))) return false;
    // canonical code (RFC 1951 3.2.2), left-justified to 7 bits: a code of
    // length l is below lim[l] and not below lim[l - 1]; its index in the
    // length-sorted symbol list is base[l] + (code >> (7 - l)) (mod 256).
    // lim[1..7] sit in bytes 1..7 of limlo/limhi (byte 0 is zero).
    uint32_t limlo = 0, limhi = 0;
    uint64_t base = 0, nxt = 0;
    {
        uint32_t code = 0, off = 0;
#pragma unroll
        for (uint32_t l = 1; l <= 7; ++l) {
            const uint32_t n = (uint32_t)(cnt >> (8 * l)) & 255u;
            const uint32_t lim = (code + n) << (7 - l);
            if (l < 4)
                limlo |= lim << (8 * l);
            else
                limhi |= lim << (8 * (l - 4));
            base |= (uint64_t)((off - code) & 255u) << (8 * l);
            nxt |= (uint64_t)off << (8 * l);
            off += n;
            code = (code + n) << 1;
        }
    }
    uint64_t sa = 0, sb = 0;  // sorted symbols, 5 bits each: list entries 0..11 in sa, 12..18 in sb
#pragma unroll
    for (uint32_t s = 0; s < 19; ++s) {
        const uint32_t L = (uint32_t)(ls >> (3 * s)) & 7u;
        if (L) {
            const uint32_t k = (uint32_t)(nxt >> (8 * L)) & 255u;
            nxt += 1ull << (8 * L);
            if (k < 12)
                sa |= (uint64_t)s << (5 * k);
            else
                sb |= (uint64_t)s << (5 * (k - 12));
        }
    }
    const uint32_t total = hlit + hdist;
    uint32_t i = 0, last = 0, sumL = 0, nL = 0, sumD = 0, nD = 0, len256 = 0;
    while (i < total) {
        if (pos + 10 > nbits) return false;
        const uint64_t x = peek();
        const uint32_t r = rev7((uint32_t)x & 127u);
        // l = 1 + #{k : r >= lim[k]}, the seven compares at once: in each byte
        // (r + 128) - lim[k] keeps bit 7 exactly when r >= lim[k] (lim <= 128,
        // so no byte borrows from the next)
        const uint32_t rb = (r | 128u) * 0x01010101u;
        const uint32_t ge = ((rb & 0xFFFFFF00u) - limlo) & 0x80808000u;
        const uint32_t gh = (rb - limhi) & 0x80808080u;
        const uint32_t l = 1u + (uint32_t)(popcount32(ge) + popcount32(gh));
        if (l > 7) return false;  // an unused pattern of an incomplete (single-code) precode
        const uint32_t k = (((uint32_t)(base >> (8 * l)) + (r >> (7 - l))) & 255u);
        const uint32_t sym = (uint32_t)((k < 12 ? sa >> (5 * k) : sb >> (5 * (k - 12))) & 31u);
        // the four kinds by selects: a length (n = 1), 16 (repeat the last, 3-6
        // times, 2 extra bits), 17 (3-10 zeros, 3 bits), 18 (11-138 zeros, 7 bits)
        const bool is_len = sym <= 15, is16 = sym == 16, is17 = sym == 17;
        const uint32_t eb = is_len ? 0u : is16 ? 2u : is17 ? 3u : 7u;
        const uint32_t n = (is_len ? 1u : is16 || is17 ? 3u : 11u) + ((uint32_t)(x >> l) & ((1u << eb) - 1u));
        const uint32_t L = is_len ? sym : is16 ? last : 0u;
        const uint32_t adv = l + eb;
        if (is16 && i == 0) return false;
        if (pos + adv > nbits) return false;
        if (i + n > total) return false;
        last = L;
        skip(adv);
        {
            const uint32_t nl = L ? (i < hlit ? (n < hlit - i ? n : hlit - i) : 0u) : 0u;
            const uint32_t nd = L ? n - nl : 0u;
            const uint32_t sh = 15u - (L ? L : 15u);
            sumL += nl << sh;
            nL += nl;
            sumD += nd << sh;
            nD += nd;
        }
        if (i <= 256u && 256u < i + n) len256 = L;
        const bool crossed = i < hlit && i + n >= hlit;
        i += n;
        // oversubscription only grows; the literal/length code is settled
        // once its lengths are all read
        if ((nL >= 2 && sumL > 32768u) || (nD >= 2 && sumD > 32768u)) return false;
        if (crossed && (len256 == 0 || (nL >= 2 && sumL != 32768u))) return false;
    }
    if (len256 == 0) return false;
    return !((nL >= 2 && sumL != 32768u) || (nD >= 2 && sumD != 32768u));
}

}  // namespace rrg
