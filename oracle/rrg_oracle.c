/*
 * rrg_oracle.c -- CPU restatement of rusty_rapidgzip's gzip/DEFLATE decode path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU path and the "port" CPU
 * baseline.  Nothing in the product links or calls this file.
 *
 * Citations are relative to crates/rusty-rapidgzip/src/ of the reference.
 *
 * Semantics kept from the reference on purpose:
 *  - Huffman tables: 10-bit LUT + canonical long-code list (deflate/huffman.rs:83,
 *    215-325); incomplete codes accepted except in the block finder's strict mode
 *    (huffman.rs:245-247).
 *  - Symbol decode in headers and in the block finder peeks 10 bits (15 for long
 *    codes) and fails when fewer remain (huffman.rs:346-389, bitreader.rs:150-156).
 *  - In the block-data hot loop a read fails only when it consumes bits past the
 *    end of the input.  (The reference's hot loop also demands 48 buffered bits
 *    near the end, fast_inflate.rs:2065-2080; that only changes which error a
 *    truncated stream reports, never the bytes of a stream that decodes.)
 *  - Speculative decode emits u16 cells for the chunk's first member, markers
 *    0x8000|prefix_offset for back-references before the chunk start
 *    (fast_inflate.rs:2209-2222); later members decode isolated (a back-reference
 *    before the member start is an error, like libdeflate at
 *    libdeflate_ffi.rs:64-123).
 */
#include "rrg_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------------ */
/* RFC 1951 tables (deflate/tables.rs:1-49)                                   */
/* ------------------------------------------------------------------------ */
static const uint8_t CL_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
static const uint16_t LEN_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
static const uint8_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                      2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
static const uint16_t DIST_BASE[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,  25,
                                       33,   49,   65,   97,   129,  193,   257,   385,   513, 769,
                                       1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
static const uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                       6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

/* ------------------------------------------------------------------------ */
/* Bit access (deflate/bitreader.rs): LSB-first, absolute bit positions.      */
/* ------------------------------------------------------------------------ */
static inline uint64_t bits_at(const uint8_t* d, size_t len, uint64_t pos) {
    /* >= 57 valid bits starting at pos, zero beyond the end of the input */
    size_t byte = (size_t)(pos >> 3);
    uint64_t w = 0;
    if (byte + 8 <= len) {
        memcpy(&w, d + byte, 8);
    } else {
        for (size_t i = 0; i < 8 && byte + i < len; i++) w |= (uint64_t)d[byte + i] << (8 * i);
    }
    return w >> (pos & 7);
}

typedef struct {
    const uint8_t* d;
    size_t len;
    uint64_t nbits;
    uint64_t pos;
} BR;

/* bitreader.rs:124-134 read(n): fails iff fewer than n bits remain */
static inline int br_read(BR* b, unsigned n, uint32_t* v) {
    if (n == 0) {
        *v = 0;
        return ORC_OK;
    }
    if (b->pos + n > b->nbits) return ORC_E_EOF;
    *v = (uint32_t)(bits_at(b->d, b->len, b->pos) & ((1ull << n) - 1));
    b->pos += n;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Canonical Huffman decoder (deflate/huffman.rs:131-428)                     */
/* ------------------------------------------------------------------------ */
#define LUT_BITS 10
#define LUT_SIZE (1 << LUT_BITS)
#define HUFFDEC_LITERAL (1u << 31)
#define HUFFDEC_EXCEPTIONAL (1u << 15)
enum { ENC_SIMPLE, ENC_LITLEN, ENC_DIST };

typedef struct {
    uint32_t lut[LUT_SIZE];
    int nlong;
    uint32_t lrev[288];
    uint8_t llen[288];
    uint32_t lent[288];
} Huff;

/* huffman.rs:393-428 */
static uint32_t encode_entry(unsigned sym, unsigned len, int enc) {
    if (enc == ENC_SIMPLE) return (sym << 16) | len;
    if (enc == ENC_LITLEN) {
        if (sym < 256) return HUFFDEC_LITERAL | (sym << 16) | len;
        if (sym == 256) return HUFFDEC_EXCEPTIONAL | len;
        if (sym <= 285) {
            unsigned li = sym - 257;
            return ((uint32_t)LEN_BASE[li] << 16) | ((uint32_t)LEN_EXTRA[li] << 8) | len;
        }
        return HUFFDEC_EXCEPTIONAL | (sym << 16) | len;
    }
    if (sym <= 29) return ((uint32_t)DIST_BASE[sym] << 16) | ((uint32_t)DIST_EXTRA[sym] << 8) | len;
    return HUFFDEC_EXCEPTIONAL | len;
}

/* huffman.rs:215-325 build_into.  Returns ORC_OK or ORC_E_INVALID. */
static int huff_build(Huff* h, const uint8_t* lengths, int n, int strict, int enc, const char** msg) {
    memset(h->lut, 0, sizeof h->lut);
    h->nlong = 0;
    uint32_t count[16] = {0};
    for (int i = 0; i < n; i++) {
        if (lengths[i] > 15) {
            *msg = "huffman code length > 15";
            return ORC_E_INVALID;
        }
        count[lengths[i]]++;
    }
    count[0] = 0;
    uint64_t total = 0;
    uint32_t nsym = 0;
    for (int l = 1; l <= 15; l++) {
        total += (uint64_t)count[l] << (15 - l);
        nsym += count[l];
    }
    if (nsym >= 2 && total > (1u << 15)) {
        *msg = "oversubscribed huffman tree";
        return ORC_E_INVALID;
    }
    if (strict && nsym >= 2 && total != (1u << 15)) {
        *msg = "incomplete huffman tree (strict)";
        return ORC_E_INVALID;
    }
    /* symbols sorted by (length, symbol) */
    uint16_t sorted[288];
    uint32_t off[17] = {0};
    for (int l = 1; l <= 15; l++) off[l + 1] = off[l] + count[l];
    for (int s = 0; s < n; s++)
        if (lengths[s]) sorted[off[lengths[s]]++] = (uint16_t)s;
    /* canonical codes in order, kept bit-reversed (LSB-first index) */
    uint32_t code = 0;
    int k = 0;
    for (int l = 1; l <= 15; l++) {
        code <<= 1;
        for (uint32_t i = 0; i < count[l]; i++, k++) {
            uint32_t rev = 0;
            for (int b = 0; b < l; b++) rev |= ((code >> b) & 1u) << (l - 1 - b);
            uint32_t e = encode_entry(sorted[k], (unsigned)l, enc);
            if (l <= LUT_BITS) {
                for (uint32_t idx = rev; idx < LUT_SIZE; idx += 1u << l) h->lut[idx] = e;
            } else {
                h->lrev[h->nlong] = rev;
                h->llen[h->nlong] = (uint8_t)l;
                h->lent[h->nlong] = e;
                h->nlong++;
            }
            code++;
        }
    }
    return ORC_OK;
}

/* huffman.rs:360-389 lookup(): peek(10) must succeed, long codes peek(15). */
static int huff_lookup(const Huff* h, BR* b, uint32_t* entry, const char** msg) {
    if (b->pos + LUT_BITS > b->nbits) return ORC_E_EOF;
    uint64_t w = bits_at(b->d, b->len, b->pos);
    uint32_t e = h->lut[w & (LUT_SIZE - 1)];
    if (e & 0xff) {
        b->pos += e & 0xff;
        *entry = e;
        return ORC_OK;
    }
    if (b->pos + 15 > b->nbits) return ORC_E_EOF;
    for (int i = 0; i < h->nlong; i++) {
        uint32_t mask = (1u << h->llen[i]) - 1;
        if ((w & mask) == h->lrev[i]) {
            b->pos += h->llen[i];
            *entry = h->lent[i];
            return ORC_OK;
        }
    }
    *msg = "no matching huffman code";
    return ORC_E_INVALID;
}

/* Hot-loop lookup (fast_inflate.rs:2106-2122): LUT on zero-padded bits; a code
 * that would extend past the end of input is an EOF; long codes peek(15). */
static inline int huff_lookup_hot(const Huff* h, BR* b, uint32_t* entry, const char** msg) {
    uint64_t w = bits_at(b->d, b->len, b->pos);
    uint32_t e = h->lut[w & (LUT_SIZE - 1)];
    unsigned l = e & 0xff;
    if (l) {
        if (b->pos + l > b->nbits) return ORC_E_EOF;
        b->pos += l;
        *entry = e;
        return ORC_OK;
    }
    if (b->pos + 15 > b->nbits) return ORC_E_EOF;
    for (int i = 0; i < h->nlong; i++) {
        uint32_t mask = (1u << h->llen[i]) - 1;
        if ((w & mask) == h->lrev[i]) {
            b->pos += h->llen[i];
            *entry = h->lent[i];
            return ORC_OK;
        }
    }
    *msg = "no matching huffman code";
    return ORC_E_INVALID;
}

/* Dynamic header (inflate.rs:74-145; strict variant block_finder.rs:196-262).
 * After the BFINAL/BTYPE bits.  Builds lit (LitLen or Simple) and dist. */
static int read_dynamic_header(BR* b, Huff* cl, Huff* lit, Huff* dist, int strict, const char** msg) {
    uint32_t v;
    int r;
    if ((r = br_read(b, 5, &v))) return r;
    unsigned hlit = v + 257;
    if ((r = br_read(b, 5, &v))) return r;
    unsigned hdist = v + 1;
    if ((r = br_read(b, 4, &v))) return r;
    unsigned hclen = v + 4;
    if (hlit > 286 || hdist > 30) {
        *msg = "dynamic block: HLIT/HDIST out of range";
        return ORC_E_INVALID;
    }
    uint8_t cll[19] = {0};
    for (unsigned i = 0; i < hclen; i++) {
        if ((r = br_read(b, 3, &v))) return r;
        cll[CL_ORDER[i]] = (uint8_t)v;
    }
    if ((r = huff_build(cl, cll, 19, strict, ENC_SIMPLE, msg))) return r;
    uint8_t lens[316];
    memset(lens, 0, sizeof lens);
    unsigned total = hlit + hdist, i = 0;
    while (i < total) {
        uint32_t e;
        if ((r = huff_lookup(cl, b, &e, msg))) return r;
        unsigned sym = e >> 16;
        if (sym <= 15) {
            lens[i++] = (uint8_t)sym;
        } else if (sym == 16) {
            if (i == 0) {
                *msg = "code 16 with no previous";
                return ORC_E_INVALID;
            }
            if ((r = br_read(b, 2, &v))) return r;
            unsigned n = v + 3;
            if (i + n > total) {
                *msg = "code 16 overruns lengths";
                return ORC_E_INVALID;
            }
            uint8_t prev = lens[i - 1];
            for (unsigned j = 0; j < n; j++) lens[i + j] = prev;
            i += n;
        } else if (sym == 17) {
            if ((r = br_read(b, 3, &v))) return r;
            unsigned n = v + 3;
            if (i + n > total) {
                *msg = "code 17 overruns lengths";
                return ORC_E_INVALID;
            }
            i += n;
        } else if (sym == 18) {
            if ((r = br_read(b, 7, &v))) return r;
            unsigned n = v + 11;
            if (i + n > total) {
                *msg = "code 18 overruns lengths";
                return ORC_E_INVALID;
            }
            i += n;
        } else {
            *msg = "bad code-length symbol";
            return ORC_E_INVALID;
        }
    }
    if (lens[256] == 0) {
        *msg = "EOB symbol has zero length";
        return ORC_E_INVALID;
    }
    if ((r = huff_build(lit, lens, (int)hlit, strict, strict ? ENC_SIMPLE : ENC_LITLEN, msg))) return r;
    if ((r = huff_build(dist, lens + hlit, (int)hdist, strict, strict ? ENC_SIMPLE : ENC_DIST, msg))) return r;
    return ORC_OK;
}

static void fixed_tables(Huff* lit, Huff* dist) {
    /* tables.rs:34-49 */
    uint8_t l[288], d[30];
    const char* m;
    for (int i = 0; i < 144; i++) l[i] = 8;
    for (int i = 144; i < 256; i++) l[i] = 9;
    for (int i = 256; i < 280; i++) l[i] = 7;
    for (int i = 280; i < 288; i++) l[i] = 8;
    for (int i = 0; i < 30; i++) d[i] = 5;
    huff_build(lit, l, 288, 0, ENC_LITLEN, &m);
    huff_build(dist, d, 30, 0, ENC_DIST, &m);
}

/* ------------------------------------------------------------------------ */
/* Block finder (deflate/block_finder.rs:57-301)                               */
/* ------------------------------------------------------------------------ */
static int is_dynamic_candidate(uint32_t bits) {
    if (bits & 1) return 0;
    if (((bits >> 1) & 3) != 2) return 0;
    if (((bits >> 3) & 31) > 29) return 0;
    if (((bits >> 8) & 31) > 29) return 0;
    return 1;
}

static uint8_t NEXT_LUT[1 << 13];
static pthread_once_t next_lut_once = PTHREAD_ONCE_INIT;
static void build_next_lut(void) {
    for (uint32_t i = 0; i < (1u << 13); i++) {
        uint32_t bits = i;
        uint8_t step = 0;
        while (step < 13) {
            if (is_dynamic_candidate(bits)) break;
            bits >>= 1;
            step++;
        }
        NEXT_LUT[i] = step;
    }
}
uint8_t orc_next_lut(uint32_t w) {
    pthread_once(&next_lut_once, build_next_lut);
    return NEXT_LUT[w & 0x1fff];
}

/* block_finder.rs:268-301 */
static int trial_decode(BR* b, const Huff* lit, const Huff* dist) {
    const char* msg;
    for (int k = 0; k < 64; k++) {
        uint32_t e, v;
        if (huff_lookup(lit, b, &e, &msg)) return 0;
        unsigned sym = e >> 16;
        if (sym <= 255) continue;
        if (sym == 256) return 1;
        if (sym <= 285) {
            if (br_read(b, LEN_EXTRA[sym - 257], &v)) return 0;
            if (huff_lookup(dist, b, &e, &msg)) return 0;
            unsigned ds = e >> 16;
            if (ds >= 30) return 0;
            if (br_read(b, DIST_EXTRA[ds], &v)) return 0;
        } else {
            return 0;
        }
    }
    return 1;
}

/* block_finder.rs:175-189 */
static int verify_candidate(const uint8_t* data, size_t len, uint64_t pos, Huff* pool) {
    BR b = {data, len, (uint64_t)len * 8, pos};
    uint32_t v;
    const char* msg;
    if (br_read(&b, 3, &v)) return 0;
    if (read_dynamic_header(&b, &pool[0], &pool[1], &pool[2], 1, &msg)) return 0;
    return trial_decode(&b, &pool[1], &pool[2]);
}

/* block_finder.rs:196-262: the strict dynamic header alone (BFINAL/BTYPE at
 * `pos` are skipped, not checked).  1 = the header parses in strict mode. */
int orc_strict_header_ok(const uint8_t* data, size_t len, uint64_t pos) {
    Huff* pool = (Huff*)malloc(3 * sizeof(Huff));
    BR b = {data, len, (uint64_t)len * 8, pos};
    uint32_t v;
    const char* msg;
    int ok = !br_read(&b, 3, &v) && !read_dynamic_header(&b, &pool[0], &pool[1], &pool[2], 1, &msg);
    free(pool);
    return ok;
}

/* block_finder.rs:109-141 */
int orc_find_next_dynamic_block(const uint8_t* data, size_t len, uint64_t from_bit, uint64_t until_bit,
                                uint64_t* out) {
    pthread_once(&next_lut_once, build_next_lut);
    uint64_t total = (uint64_t)len * 8;
    uint64_t limit = until_bit < total ? until_bit : total;
    if (from_bit >= limit) return 0;
    Huff* pool = (Huff*)malloc(3 * sizeof(Huff));
    int found = 0;
    uint64_t pos = from_bit;
    while (pos + 13 <= limit) {
        uint32_t w = (uint32_t)(bits_at(data, len, pos) & 0x1fff);
        uint8_t step = NEXT_LUT[w];
        if (step == 0) {
            if (verify_candidate(data, len, pos, pool)) {
                *out = pos;
                found = 1;
                break;
            }
            pos += 1;
        } else {
            pos += step;
        }
    }
    free(pool);
    return found;
}

/* ------------------------------------------------------------------------ */
/* CRC-32 (IEEE, reflected 0xEDB88320), combine by GF(2) polynomial products   */
/* ------------------------------------------------------------------------ */
static uint32_t CRC_T[8][256];
static uint32_t X2N[64]; /* x^(2^k) mod P, reflected */
static pthread_once_t crc_once = PTHREAD_ONCE_INIT;

/* a*b mod P over GF(2), reflected bit order (bit 31 = x^0). */
static uint32_t gf2_mulmod(uint32_t a, uint32_t b) {
    uint32_t prod = 0;
    for (int i = 0; i < 32; i++) {
        if (a & (0x80000000u >> i)) prod ^= b;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
    }
    return prod;
}
static void crc_init(void) {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
        CRC_T[0][n] = c;
    }
    for (uint32_t n = 0; n < 256; n++)
        for (int s = 1; s < 8; s++) CRC_T[s][n] = (CRC_T[s - 1][n] >> 8) ^ CRC_T[0][CRC_T[s - 1][n] & 0xff];
    X2N[0] = 0x40000000u; /* x^1 */
    for (int k = 1; k < 64; k++) X2N[k] = gf2_mulmod(X2N[k - 1], X2N[k - 1]);
}
uint32_t orc_crc32(uint32_t crc, const uint8_t* p, size_t n) {
    pthread_once(&crc_once, crc_init);
    uint32_t c = ~crc;
    while (n >= 8) {
        uint32_t lo, hi;
        memcpy(&lo, p, 4);
        memcpy(&hi, p + 4, 4);
        lo ^= c;
        c = CRC_T[7][lo & 0xff] ^ CRC_T[6][(lo >> 8) & 0xff] ^ CRC_T[5][(lo >> 16) & 0xff] ^ CRC_T[4][lo >> 24] ^
            CRC_T[3][hi & 0xff] ^ CRC_T[2][(hi >> 8) & 0xff] ^ CRC_T[1][(hi >> 16) & 0xff] ^ CRC_T[0][hi >> 24];
        p += 8;
        n -= 8;
    }
    while (n--) c = (c >> 8) ^ CRC_T[0][(c ^ *p++) & 0xff];
    return ~c;
}
/* crc(A||B) = (x^(8|B|) * crc(A)) mod P  xor  crc(B) */
uint32_t orc_crc32_combine(uint32_t a, uint32_t b, uint64_t len_b) {
    pthread_once(&crc_once, crc_init);
    uint32_t x = 0x80000000u; /* 1 */
    uint64_t n = len_b;
    int k = 3; /* multiply by x^(8n) = prod over set bits of n of x^(2^(k+3)) */
    while (n) {
        if (n & 1) x = gf2_mulmod(X2N[k], x);
        n >>= 1;
        k++;
    }
    return gf2_mulmod(x, a) ^ b;
}

/* ------------------------------------------------------------------------ */
/* gzip framing (gzip.rs:252-351)                                             */
/* ------------------------------------------------------------------------ */
long orc_parse_header(const uint8_t* p, size_t n) {
    if (n < 10) return -ORC_E_TRUNCATED;
    if (p[0] != 0x1f || p[1] != 0x8b) return -ORC_E_BADMAGIC;
    if (p[2] != 8) return -ORC_E_METHOD;
    uint8_t flg = p[3];
    if (flg & 0xe0) return -ORC_E_RESERVED;
    size_t pos = 10;
    if (flg & 4) {
        if (pos + 2 > n) return -ORC_E_TRUNCATED;
        size_t xlen = p[pos] | ((size_t)p[pos + 1] << 8);
        pos += 2 + xlen;
        if (pos > n) return -ORC_E_TRUNCATED;
    }
    for (int f = 8; f <= 16; f <<= 1) {
        if (flg & f) {
            if (pos > n) return -ORC_E_TRUNCATED;
            const uint8_t* z = memchr(p + pos, 0, n - pos);
            if (!z) return -ORC_E_TRUNCATED;
            pos = (size_t)(z - p) + 1;
        }
    }
    if (flg & 2) {
        if (pos + 2 > n) return -ORC_E_TRUNCATED;
        pos += 2;
    }
    return (long)pos;
}

uint32_t orc_parse_bgzf_block_size(const uint8_t* p, size_t n) {
    if (n < 12 || p[0] != 0x1f || p[1] != 0x8b || p[2] != 8) return 0;
    if (!(p[3] & 4)) return 0;
    size_t xlen = p[10] | ((size_t)p[11] << 8);
    if (12 + xlen > n) return 0;
    size_t q = 12, end = 12 + xlen;
    while (q + 4 <= end) {
        size_t slen = p[q + 2] | ((size_t)p[q + 3] << 8);
        if (q + 4 + slen > end) return 0;
        if (p[q] == 'B' && p[q + 1] == 'C' && slen == 2) return (uint32_t)(p[q + 4] | (p[q + 5] << 8)) + 1;
        q += 4 + slen;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Textbook inflate (safe_inflate.rs:278-644, puff.c semantics)               */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint16_t count[16];
    uint16_t symbol[288];
} Puff;

static int puff_construct(Puff* h, const uint8_t* len, int n) {
    memset(h->count, 0, sizeof h->count);
    for (int i = 0; i < n; i++) {
        if (len[i] > 15) return -1;
        h->count[len[i]]++;
    }
    if (h->count[0] == n) return 0;
    int left = 1;
    for (int l = 1; l <= 15; l++) {
        left = (left << 1) - h->count[l];
        if (left < 0) return -1;
    }
    uint16_t offs[16];
    offs[1] = 0;
    for (int l = 1; l < 15; l++) offs[l + 1] = offs[l] + h->count[l];
    for (int s = 0; s < n; s++)
        if (len[s]) h->symbol[offs[len[s]]++] = (uint16_t)s;
    return left;
}

static int puff_decode(BR* b, const Puff* h, int* sym) {
    int code = 0, first = 0, index = 0;
    for (int l = 1; l <= 15; l++) {
        uint32_t bit;
        if (br_read(b, 1, &bit)) return ORC_E_EOF;
        code |= (int)bit;
        int count = h->count[l];
        if (code - count < first) {
            *sym = h->symbol[index + (code - first)];
            return ORC_OK;
        }
        index += count;
        first += count;
        first <<= 1;
        code <<= 1;
    }
    return ORC_E_INVALID;
}

static int puff_codes(BR* b, uint8_t* out, size_t cap, size_t* n, const Puff* lc, const Puff* dc, const char** msg) {
    for (;;) {
        int sym, r;
        if ((r = puff_decode(b, lc, &sym))) {
            *msg = "invalid huffman code";
            return r;
        }
        if (sym < 256) {
            if (*n >= cap) return ORC_E_NOMEM;
            out[(*n)++] = (uint8_t)sym;
        } else if (sym == 256) {
            return ORC_OK;
        } else {
            sym -= 257;
            if (sym >= 29) {
                *msg = "length symbol out of range";
                return ORC_E_INVALID;
            }
            uint32_t v;
            if (br_read(b, LEN_EXTRA[sym], &v)) return ORC_E_EOF;
            size_t length = LEN_BASE[sym] + v;
            int ds;
            if ((r = puff_decode(b, dc, &ds))) {
                *msg = "invalid huffman code";
                return r;
            }
            if (ds >= 30) {
                *msg = "distance symbol out of range";
                return ORC_E_INVALID;
            }
            if (br_read(b, DIST_EXTRA[ds], &v)) return ORC_E_EOF;
            size_t dist = DIST_BASE[ds] + v;
            if (dist > *n) {
                *msg = "distance out of bounds";
                return ORC_E_INVALID;
            }
            if (*n + length > cap) return ORC_E_NOMEM;
            for (size_t k = 0; k < length; k++, (*n)++) out[*n] = out[*n - dist];
        }
    }
}

int orc_safe_inflate(const uint8_t* in, size_t in_len, uint8_t* out, size_t cap, size_t* out_len,
                     size_t* consumed) {
    BR b = {in, in_len, (uint64_t)in_len * 8, 0};
    size_t n = 0;
    const char* msg = 0;
    Puff lc, dc;
    int r = ORC_OK;
    for (;;) {
        uint32_t last, type;
        if ((r = br_read(&b, 1, &last))) break;
        if ((r = br_read(&b, 2, &type))) break;
        if (type == 0) {
            b.pos = (b.pos + 7) & ~7ull;
            size_t p = (size_t)(b.pos >> 3);
            if (p + 4 > in_len) {
                r = ORC_E_EOF;
                break;
            }
            unsigned len = in[p] | (in[p + 1] << 8), nlen = in[p + 2] | (in[p + 3] << 8);
            if ((len ^ nlen) != 0xffff) {
                r = ORC_E_INVALID;
                break;
            }
            p += 4;
            if (p + len > in_len) {
                r = ORC_E_EOF;
                break;
            }
            if (n + len > cap) {
                r = ORC_E_NOMEM;
                break;
            }
            memcpy(out + n, in + p, len);
            n += len;
            b.pos = (uint64_t)(p + len) * 8;
        } else if (type == 1) {
            uint8_t l[288];
            for (int i = 0; i < 144; i++) l[i] = 8;
            for (int i = 144; i < 256; i++) l[i] = 9;
            for (int i = 256; i < 280; i++) l[i] = 7;
            for (int i = 280; i < 288; i++) l[i] = 8;
            puff_construct(&lc, l, 288);
            for (int i = 0; i < 30; i++) l[i] = 5;
            puff_construct(&dc, l, 30);
            if ((r = puff_codes(&b, out, cap, &n, &lc, &dc, &msg))) break;
        } else if (type == 2) {
            uint32_t v;
            if ((r = br_read(&b, 5, &v))) break;
            unsigned nlen = v + 257;
            if ((r = br_read(&b, 5, &v))) break;
            unsigned ndist = v + 1;
            if ((r = br_read(&b, 4, &v))) break;
            unsigned ncode = v + 4;
            if (nlen > 286 || ndist > 30) {
                r = ORC_E_INVALID;
                break;
            }
            uint8_t lengths[320];
            memset(lengths, 0, sizeof lengths);
            for (unsigned i = 0; i < ncode; i++) {
                if ((r = br_read(&b, 3, &v))) break;
                lengths[CL_ORDER[i]] = (uint8_t)v;
            }
            if (r) break;
            if (puff_construct(&lc, lengths, 19) < 0) {
                r = ORC_E_INVALID;
                break;
            }
            unsigned idx = 0;
            while (idx < nlen + ndist) {
                int sym;
                if ((r = puff_decode(&b, &lc, &sym))) break;
                if (sym < 16) {
                    lengths[idx++] = (uint8_t)sym;
                } else {
                    uint8_t lenv = 0;
                    unsigned rep;
                    if (sym == 16) {
                        if (idx == 0) {
                            r = ORC_E_INVALID;
                            break;
                        }
                        lenv = lengths[idx - 1];
                        if ((r = br_read(&b, 2, &v))) break;
                        rep = 3 + v;
                    } else if (sym == 17) {
                        if ((r = br_read(&b, 3, &v))) break;
                        rep = 3 + v;
                    } else {
                        if ((r = br_read(&b, 7, &v))) break;
                        rep = 11 + v;
                    }
                    if (idx + rep > nlen + ndist) {
                        r = ORC_E_INVALID;
                        break;
                    }
                    while (rep--) lengths[idx++] = lenv;
                }
            }
            if (r) break;
            if (lengths[256] == 0) {
                r = ORC_E_INVALID;
                break;
            }
            if (puff_construct(&lc, lengths, (int)nlen) < 0 || puff_construct(&dc, lengths + nlen, (int)ndist) < 0) {
                r = ORC_E_INVALID;
                break;
            }
            if ((r = puff_codes(&b, out, cap, &n, &lc, &dc, &msg))) break;
        } else {
            r = ORC_E_INVALID;
            break;
        }
        if (last) break;
    }
    *out_len = n;
    *consumed = (size_t)((b.pos + 7) >> 3);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Speculative chunk decode (pipeline.rs:1080-1185, fast_inflate.rs:1794-2234) */
/* ------------------------------------------------------------------------ */
static int chunk_reserve(orc_chunk* c, size_t extra) {
    if (c->n + extra <= c->cap) return 0;
    size_t nc = c->cap ? c->cap : 1 << 16;
    while (nc < c->n + extra) nc *= 2;
    uint16_t* p = (uint16_t*)realloc(c->cells, nc * sizeof(uint16_t));
    if (!p) return -1;
    c->cells = p;
    c->cap = nc;
    return 0;
}

/* One compressed block into cells.  member_base: cell index where the member
 * began; spec: this is the chunk's first member (markers allowed). */
static int decode_huff_block(BR* b, orc_chunk* c, const Huff* lit, const Huff* dist, size_t member_base, int spec,
                             const char** msg) {
    for (;;) {
        uint32_t e, v;
        int r;
        if ((r = huff_lookup_hot(lit, b, &e, msg))) return r;
        if (e & HUFFDEC_LITERAL) {
            if (chunk_reserve(c, 1)) return ORC_E_NOMEM;
            c->cells[c->n++] = (uint16_t)((e >> 16) & 0xff);
            continue;
        }
        if (e & HUFFDEC_EXCEPTIONAL) {
            if ((e >> 16) == 0) return ORC_OK;
            *msg = "literal/length symbol out of range";
            return ORC_E_INVALID;
        }
        if ((r = br_read(b, (e >> 8) & 0x1f, &v))) return r;
        size_t length = ((e >> 16) & 0x1ff) + v;
        uint32_t de;
        if ((r = huff_lookup_hot(dist, b, &de, msg))) return r;
        if (de & HUFFDEC_EXCEPTIONAL) {
            *msg = "distance symbol out of range";
            return ORC_E_INVALID;
        }
        if ((r = br_read(b, (de >> 8) & 0x1f, &v))) return r;
        size_t distance = (de >> 16) + v;
        if (chunk_reserve(c, length)) return ORC_E_NOMEM;
        size_t cur = c->n - member_base;
        uint16_t* o = c->cells;
        if (distance > cur) {
            if (!spec) {
                *msg = "back-reference distance out of bounds"; /* fast_inflate.rs:1711-1715 */
                return ORC_E_INVALID;
            }
            /* fast_inflate.rs:2209-2222 */
            size_t pre = distance - cur < length ? distance - cur : length;
            for (size_t k = 0; k < pre; k++) o[c->n + k] = (uint16_t)(0x8000u | (distance - cur - k - 1));
            c->n += pre;
            for (size_t k = pre; k < length; k++, c->n++) o[c->n] = o[c->n - distance];
        } else {
            for (size_t k = 0; k < length; k++, c->n++) o[c->n] = o[c->n - distance];
        }
    }
}

static int decode_block(BR* b, orc_chunk* c, Huff* tabs, size_t member_base, int spec, int* bfinal,
                        const char** msg) {
    uint32_t f, t;
    int r;
    if ((r = br_read(b, 1, &f))) return r;
    if ((r = br_read(b, 2, &t))) return r;
    *bfinal = (int)f;
    if (t == 0) { /* fast_inflate.rs:1937-1963 */
        b->pos = (b->pos + 7) & ~7ull;
        uint32_t len, nlen;
        if ((r = br_read(b, 16, &len))) return r;
        if ((r = br_read(b, 16, &nlen))) return r;
        if ((len ^ nlen) != 0xffff) {
            *msg = "stored block: LEN/NLEN mismatch";
            return ORC_E_INVALID;
        }
        if (b->pos + 8ull * len > b->nbits) return ORC_E_EOF;
        if (chunk_reserve(c, len)) return ORC_E_NOMEM;
        const uint8_t* src = b->d + (b->pos >> 3);
        for (uint32_t i = 0; i < len; i++) c->cells[c->n++] = src[i];
        b->pos += 8ull * len;
        return ORC_OK;
    }
    if (t == 1) {
        fixed_tables(&tabs[1], &tabs[2]);
        return decode_huff_block(b, c, &tabs[1], &tabs[2], member_base, spec, msg);
    }
    if (t == 2) {
        if ((r = read_dynamic_header(b, &tabs[0], &tabs[1], &tabs[2], 0, msg))) return r;
        return decode_huff_block(b, c, &tabs[1], &tabs[2], member_base, spec, msg);
    }
    *msg = "reserved block type";
    return ORC_E_INVALID;
}

static int push_mb(orc_chunk* c, orc_member_boundary m) {
    if (c->n_mb == c->cap_mb) {
        size_t nc = c->cap_mb ? 2 * c->cap_mb : 8;
        orc_member_boundary* p = (orc_member_boundary*)realloc(c->mb, nc * sizeof *p);
        if (!p) return -1;
        c->mb = p;
        c->cap_mb = nc;
    }
    c->mb[c->n_mb++] = m;
    return 0;
}

int orc_decode_chunk(const uint8_t* body, size_t len, uint64_t start_bit, uint64_t end_hint, orc_chunk* c) {
    c->n = 0;
    c->n_mb = 0;
    c->final_block = 0;
    c->start_bit = start_bit;
    c->end_hint = end_hint;
    c->msg = 0;
    c->err = ORC_OK;
    BR b = {body, len, (uint64_t)len * 8, start_bit};
    if (start_bit > b.nbits) return c->err = ORC_E_EOF;
    Huff* tabs = (Huff*)malloc(3 * sizeof(Huff));
    int first = 1, r = ORC_OK;
    size_t member_base = 0;
    for (;;) {
        if (b.pos >= end_hint) break;
        int bfinal = 0;
        for (;;) { /* blocks of one member */
            if ((r = decode_block(&b, c, tabs, member_base, first, &bfinal, &c->msg))) goto done;
            if (bfinal || b.pos >= end_hint) break;
        }
        first = 0;
        if (!bfinal) break;
        b.pos = (b.pos + 7) & ~7ull;
        size_t tb = (size_t)(b.pos >> 3);
        if (tb + 8 > len) {
            r = ORC_E_TRUNCATED;
            goto done;
        }
        orc_member_boundary m;
        m.out_off = c->n;
        m.crc = (uint32_t)body[tb] | ((uint32_t)body[tb + 1] << 8) | ((uint32_t)body[tb + 2] << 16) |
                ((uint32_t)body[tb + 3] << 24);
        m.isize = (uint32_t)body[tb + 4] | ((uint32_t)body[tb + 5] << 8) | ((uint32_t)body[tb + 6] << 16) |
                  ((uint32_t)body[tb + 7] << 24);
        m.trailer_byte = tb;
        if (push_mb(c, m)) {
            r = ORC_E_NOMEM;
            goto done;
        }
        size_t after = tb + 8;
        if (after >= len) {
            c->final_block = 1;
            b.pos = (uint64_t)after * 8;
            break;
        }
        long hl = orc_parse_header(body + after, len - after);
        if (hl < 0) {
            r = (int)-hl;
            goto done;
        }
        b.pos = (uint64_t)(after + (size_t)hl) * 8;
        member_base = c->n;
    }
done:
    free(tabs);
    c->actual_end_bit = b.pos;
    return c->err = r;
}

void orc_chunk_free(orc_chunk* c) {
    free(c->cells);
    free(c->mb);
    memset(c, 0, sizeof *c);
}

int orc_resolve_cells(const uint16_t* cells, size_t n, const uint8_t* tail, size_t tl, uint8_t* out) {
    for (size_t i = 0; i < n; i++) {
        uint16_t c = cells[i];
        if (c & 0x8000) {
            size_t off = c & 0x7fff;
            if (off >= tl) return ORC_E_INVALID;
            out[i] = tail[tl - 1 - off];
        } else {
            out[i] = (uint8_t)c;
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Parallel read_gz restatement                                                */
/* ------------------------------------------------------------------------ */
typedef struct {
    void (*fn)(void* ctx, size_t i);
    void* ctx;
    size_t n;
    atomic_size_t next;
} Job;

static void* job_worker(void* arg) {
    Job* j = (Job*)arg;
    for (;;) {
        size_t i = atomic_fetch_add(&j->next, 1);
        if (i >= j->n) break;
        j->fn(j->ctx, i);
    }
    return 0;
}

static void parallel_for(size_t n, int nthreads, void (*fn)(void*, size_t), void* ctx) {
    Job j;
    j.fn = fn;
    j.ctx = ctx;
    j.n = n;
    atomic_init(&j.next, 0);
    if (nthreads <= 1 || n <= 1) {
        job_worker(&j);
        return;
    }
    int t = nthreads < (int)n ? nthreads : (int)n;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)t);
    for (int i = 0; i < t; i++) pthread_create(&th[i], 0, job_worker, &j);
    for (int i = 0; i < t; i++) pthread_join(th[i], 0);
    free(th);
}

typedef struct {
    const uint8_t* body;
    size_t len;
    const uint64_t* targets;
    uint64_t* found;
    uint64_t total_bits;
} FindCtx;

static void find_one(void* ctx, size_t i) {
    FindCtx* f = (FindCtx*)ctx;
    uint64_t r;
    f->found[i] = orc_find_next_dynamic_block(f->body, f->len, f->targets[i], f->total_bits, &r) ? r : UINT64_MAX;
}

/* pipeline.rs:489-544 */
size_t orc_chunk_boundaries(const uint8_t* body, size_t len, size_t chunk_size, int nthreads, uint64_t* out,
                            size_t cap) {
    uint64_t total = (uint64_t)len * 8;
    uint64_t cb = (uint64_t)(chunk_size > 65536 ? chunk_size : 65536) * 8;
    size_t nt = 0;
    for (uint64_t c = cb; c < total; c += cb) nt++;
    uint64_t* targets = (uint64_t*)malloc((nt + 1) * sizeof(uint64_t));
    uint64_t* found = (uint64_t*)malloc((nt + 1) * sizeof(uint64_t));
    for (size_t i = 0; i < nt; i++) targets[i] = cb * (i + 1);
    FindCtx f = {body, len, targets, found, total};
    parallel_for(nt, nthreads, find_one, &f);
    size_t n = 0;
    if (cap) out[n++] = 0;
    for (size_t i = 0; i < nt; i++)
        if (found[i] != UINT64_MAX && found[i] > out[n - 1] && n < cap) out[n++] = found[i];
    free(targets);
    free(found);
    return n;
}

typedef struct {
    const uint8_t* body;
    size_t len;
    const uint64_t* bounds;
    size_t nb;
    orc_chunk* chunks;
} DecCtx;

static void decode_one(void* ctx, size_t i) {
    DecCtx* d = (DecCtx*)ctx;
    uint64_t hint = i + 1 < d->nb ? d->bounds[i + 1] : UINT64_MAX;
    orc_decode_chunk(d->body, d->len, d->bounds[i], hint, &d->chunks[i]);
}

typedef struct {
    orc_chunk* chunks;
    uint8_t** tails;
    size_t* tail_len;
    uint8_t** bytes;
    int* rerr;
    uint32_t** piece_crc; /* per chunk: n_mb + 1 pieces */
} ResCtx;

static void resolve_one(void* ctx, size_t i) {
    ResCtx* r = (ResCtx*)ctx;
    orc_chunk* c = &r->chunks[i];
    r->bytes[i] = (uint8_t*)malloc(c->n ? c->n : 1);
    r->rerr[i] = orc_resolve_cells(c->cells, c->n, r->tails[i], r->tail_len[i], r->bytes[i]);
    r->piece_crc[i] = (uint32_t*)malloc((c->n_mb + 1) * sizeof(uint32_t));
    size_t prev = 0;
    for (size_t k = 0; k <= c->n_mb; k++) {
        size_t end = k < c->n_mb ? c->mb[k].out_off : c->n;
        r->piece_crc[i][k] = orc_crc32(0, r->bytes[i] + prev, end - prev);
        prev = end;
    }
}

static void set_err(orc_error* e, int kind, const char* msg) {
    if (!e) return;
    memset(e, 0, sizeof *e);
    e->kind = kind;
    e->msg = msg;
}

/* pipeline.rs:468-1041 (parallel_decode_member).  *consumed_chunks etc are
 * accumulated into st. */
static int decode_member_path(const uint8_t* body, size_t len, size_t chunk_size, int nthreads, orc_sink_cb cb,
                              void* user, orc_stats* st, orc_error* err) {
    size_t cap = (size_t)(((uint64_t)len * 8) / ((uint64_t)(chunk_size > 65536 ? chunk_size : 65536) * 8)) + 2;
    uint64_t* bounds = (uint64_t*)malloc(cap * sizeof(uint64_t));
    size_t nb = orc_chunk_boundaries(body, len, chunk_size, nthreads, bounds, cap);
    orc_chunk* chunks = (orc_chunk*)calloc(nb, sizeof(orc_chunk));
    DecCtx dc = {body, len, bounds, nb, chunks};
    parallel_for(nb, nthreads, decode_one, &dc);

    /* Stage A (pipeline.rs:742-856): anchored chain, re-decode, tails. */
    uint8_t** tails = (uint8_t**)calloc(nb, sizeof(uint8_t*));
    size_t* tail_len = (size_t*)calloc(nb, sizeof(size_t));
    uint64_t expected = 0;
    size_t nproc = 0;
    int rc = ORC_OK;
    uint8_t* prev_tail = (uint8_t*)malloc(32768);
    size_t prev_len = 0;
    for (size_t i = 0; i < nb; i++) {
        orc_chunk* c = &chunks[i];
        if (c->err != ORC_OK || c->start_bit != expected) {
            st->speculation_failures++;
            uint64_t hint = c->end_hint;
            orc_chunk_free(c);
            orc_decode_chunk(body, len, expected, hint, c);
            if (c->err) {
                set_err(err, c->err, c->msg);
                rc = c->err;
                nproc = i; /* chunks before i are still emitted */
                break;
            }
        }
        expected = c->actual_end_bit;
        tails[i] = (uint8_t*)malloc(32768);
        memcpy(tails[i], prev_tail, prev_len);
        tail_len[i] = prev_len;
        /* build_prev_tail_fast (pipeline.rs:1045-1078) */
        uint8_t nt[32768];
        size_t W = 32768, n = c->n, ts = n > W ? n - W : 0, tl = n - ts;
        for (size_t k = 0; k < tl; k++) {
            uint16_t cell = c->cells[ts + k];
            if (cell & 0x8000) {
                size_t off = cell & 0x7fff;
                nt[k] = off < prev_len ? prev_tail[prev_len - 1 - off] : 0;
            } else {
                nt[k] = (uint8_t)cell;
            }
        }
        if (n < W && prev_len) {
            size_t need = W - n, carry = prev_len < need ? prev_len : need;
            memmove(prev_tail, prev_tail + prev_len - carry, carry);
            memcpy(prev_tail + carry, nt, tl);
            prev_len = carry + tl;
        } else {
            memcpy(prev_tail, nt, tl);
            prev_len = tl;
        }
        nproc = i + 1;
        if (c->final_block) break;
    }
    free(prev_tail);

    /* Resolve pool + CRC pieces (pipeline.rs:676-739). */
    uint8_t** bytes = (uint8_t**)calloc(nb, sizeof(uint8_t*));
    int* rerr = (int*)calloc(nb, sizeof(int));
    uint32_t** pcrc = (uint32_t**)calloc(nb, sizeof(uint32_t*));
    ResCtx rcx = {chunks, tails, tail_len, bytes, rerr, pcrc};
    parallel_for(nproc, nthreads, resolve_one, &rcx);

    /* Output stage (pipeline.rs:858-952). */
    uint32_t cur_crc = 0;
    uint64_t cur_len = 0;
    uint32_t member = 0;
    int saw_final = 0, closed = 0;
    for (size_t i = 0; i < nproc && rc == ORC_OK; i++) {
        orc_chunk* c = &chunks[i];
        if (rerr[i]) {
            set_err(err, ORC_E_INVALID, "marker references bytes outside the available prefix window");
            rc = ORC_E_INVALID;
            break;
        }
        size_t prev = 0;
        for (size_t k = 0; k < c->n_mb; k++) {
            size_t plen = c->mb[k].out_off - prev;
            cur_crc = orc_crc32_combine(cur_crc, pcrc[i][k], plen);
            cur_len += plen;
            prev = c->mb[k].out_off;
            if (cur_crc != c->mb[k].crc) {
                set_err(err, ORC_E_CRC, "crc mismatch");
                if (err) {
                    err->expected = c->mb[k].crc;
                    err->got32 = cur_crc;
                    err->member = member;
                    err->got64 = cur_len;
                    err->trailer_byte = c->mb[k].trailer_byte;
                }
                rc = ORC_E_CRC;
                break;
            }
            if ((uint32_t)cur_len != c->mb[k].isize) {
                set_err(err, ORC_E_ISIZE, "isize mismatch");
                if (err) {
                    err->expected = c->mb[k].isize;
                    err->got64 = cur_len;
                    err->member = member;
                    err->trailer_byte = c->mb[k].trailer_byte;
                }
                rc = ORC_E_ISIZE;
                break;
            }
            member++;
            cur_crc = 0;
            cur_len = 0;
        }
        if (rc) break;
        size_t tl = c->n - prev;
        cur_crc = orc_crc32_combine(cur_crc, pcrc[i][c->n_mb], tl);
        cur_len += tl;
        st->uncompressed_bytes += c->n;
        if (cb && cb(user, bytes[i], c->n)) { /* sink closed: Ok (pipeline.rs:1028-1030) */
            closed = 1;
            break;
        }
        st->chunks_decoded++;
        if (c->final_block) {
            saw_final = 1;
            break;
        }
    }
    if (rc == ORC_OK && !saw_final && !closed) {
        set_err(err, ORC_E_IO, "parallel decode produced no final block");
        rc = ORC_E_IO;
    }
    for (size_t i = 0; i < nb; i++) {
        orc_chunk_free(&chunks[i]);
        free(tails[i]);
        free(bytes[i]);
        free(pcrc[i]);
    }
    free(chunks);
    free(tails);
    free(tail_len);
    free(bytes);
    free(rerr);
    free(pcrc);
    free(bounds);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Single-worker paths.                                                        */
/* ------------------------------------------------------------------------ */
/* serial_decode_member (pipeline.rs:1211-1345) when streaming == 0: blocks
 * decode one after another with a real 32 KiB window; after a non-final block,
 * once chunk_size + 32 KiB bytes are pending, all but the last 32 KiB are sent;
 * a member's final block sends the rest, before its trailer is checked.
 * file[body_off..] is the first member's deflate body; trailer offsets in
 * errors are file offsets (pipeline.rs:1318,1326).
 *
 * read_gz_streaming (lib.rs:244-361) when streaming != 0: file[0..] starts at
 * the first member's header; after every non-final block all but the last
 * 32 KiB are sent (lib.rs:286-303), an input that ends inside a block or
 * trailer is Truncated (lib.rs:200-205, 188-190), trailer offsets in errors
 * are 0 (lib.rs:323-340), an empty input is Truncated (lib.rs:355-357), and a
 * closed sink does not stop the decode (lib.rs:296-298). */
static int serial_member_path(const uint8_t* file, size_t len, size_t body_off, size_t chunk_size, int streaming,
                              orc_sink_cb cb, void* user, orc_stats* st, orc_error* err) {
    const size_t W = 32768;
    const size_t T = chunk_size > 65536 ? chunk_size : 65536;
    orc_chunk c;
    memset(&c, 0, sizeof c);
    Huff* tabs = (Huff*)malloc(3 * sizeof(Huff));
    uint8_t* bytes = NULL;
    size_t bytes_cap = 0;
    int rc = ORC_OK, closed = 0;
    uint32_t member = 0;
    size_t pos = body_off; /* streaming: next header; else: next body */
    int need_header = streaming;
    if (streaming && len == 0) {
        set_err(err, ORC_E_TRUNCATED, "truncated gzip stream");
        rc = ORC_E_TRUNCATED;
        goto out;
    }
    for (;;) {
        if (need_header) {
            long hl = orc_parse_header(file + pos, len - pos);
            if (hl < 0) {
                set_err(err, (int)-hl, "bad gzip header");
                rc = (int)-hl;
                goto out;
            }
            pos += (size_t)hl;
        }
        need_header = 1;
        BR b = {file, len, (uint64_t)len * 8, (uint64_t)pos * 8};
        uint32_t crc = 0;
        uint64_t muc = 0;
        c.n = 0;
        for (;;) {
            int bfinal = 0;
            const char* msg = NULL;
            int r = decode_block(&b, &c, tabs, 0, 0, &bfinal, &msg);
            if (r) {
                if (streaming && r == ORC_E_EOF) {
                    set_err(err, ORC_E_TRUNCATED, "truncated gzip stream");
                    rc = ORC_E_TRUNCATED;
                } else {
                    set_err(err, r, msg ? msg : "deflate error");
                    rc = r;
                }
                goto out;
            }
            size_t emit = 0; /* cells sent now, from the front */
            if (bfinal) emit = c.n;
            else if (streaming ? c.n > W : c.n >= T + W) emit = c.n - W;
            if (emit) {
                if (emit > bytes_cap) {
                    free(bytes);
                    bytes_cap = emit;
                    bytes = (uint8_t*)malloc(bytes_cap);
                }
                for (size_t k = 0; k < emit; k++) bytes[k] = (uint8_t)c.cells[k];
                crc = orc_crc32(crc, bytes, emit);
                muc += emit;
                if (!closed && cb && cb(user, bytes, emit)) {
                    if (!streaming) goto out; /* Ok with what was sent (pipeline.rs:1264-1266) */
                    closed = 1;
                }
                st->uncompressed_bytes += emit;
                st->chunks_decoded++;
                /* keep the window: the remaining cells move to the front; a
                 * back-reference can never reach past them (distance <= 32 KiB) */
                memmove(c.cells, c.cells + emit, (c.n - emit) * sizeof(uint16_t));
                c.n -= emit;
            }
            if (bfinal) break;
        }
        b.pos = (b.pos + 7) & ~7ull;
        size_t tb = (size_t)(b.pos >> 3);
        if (tb + 8 > len) {
            set_err(err, ORC_E_TRUNCATED, "truncated gzip stream");
            rc = ORC_E_TRUNCATED;
            goto out;
        }
        const uint32_t want_crc = (uint32_t)file[tb] | ((uint32_t)file[tb + 1] << 8) |
                                  ((uint32_t)file[tb + 2] << 16) | ((uint32_t)file[tb + 3] << 24);
        const uint32_t want_isize = (uint32_t)file[tb + 4] | ((uint32_t)file[tb + 5] << 8) |
                                    ((uint32_t)file[tb + 6] << 16) | ((uint32_t)file[tb + 7] << 24);
        if (crc != want_crc) {
            set_err(err, ORC_E_CRC, "crc mismatch");
            if (err) {
                err->expected = want_crc;
                err->got32 = crc;
                err->member = member;
                err->got64 = muc;
                err->trailer_byte = streaming ? 0 : tb;
            }
            rc = ORC_E_CRC;
            goto out;
        }
        if ((uint32_t)muc != want_isize) {
            set_err(err, ORC_E_ISIZE, "isize mismatch");
            if (err) {
                err->expected = want_isize;
                err->got64 = muc;
                err->member = member;
                err->trailer_byte = streaming ? 0 : tb;
            }
            rc = ORC_E_ISIZE;
            goto out;
        }
        member++;
        pos = tb + 8;
        if (pos >= len) break;
    }
out:
    free(bytes);
    free(tabs);
    orc_chunk_free(&c);
    return rc;
}

typedef struct {
    const uint8_t* file;
    const size_t* ms;
    const size_t* me;
    size_t first;
    uint8_t** out;
    size_t* olen;
    int* oerr;
    size_t* mstart; /* per batch: first member index */
    size_t* mcount;
} BgzfCtx;

/* one BGZF batch: pipeline.rs:1394-1474 with gzip.rs:146-190 per member */
static void bgzf_batch(void* ctx, size_t bi) {
    BgzfCtx* b = (BgzfCtx*)ctx;
    size_t total = 0;
    for (size_t k = 0; k < b->mcount[bi]; k++) {
        size_t m = b->mstart[bi] + k;
        const uint8_t* p = b->file + b->ms[m];
        size_t n = b->me[m] - b->ms[m];
        if (n >= 18) total += (size_t)p[n - 4] | ((size_t)p[n - 3] << 8) | ((size_t)p[n - 2] << 16) | ((size_t)p[n - 1] << 24);
    }
    uint8_t* out = (uint8_t*)malloc(total ? total : 1);
    size_t pos = 0;
    int e = ORC_OK;
    for (size_t k = 0; k < b->mcount[bi] && !e; k++) {
        size_t m = b->mstart[bi] + k;
        const uint8_t* p = b->file + b->ms[m];
        size_t n = b->me[m] - b->ms[m];
        if (n < 18) {
            e = ORC_E_INVALID;
            break;
        }
        uint32_t isz = (uint32_t)p[n - 4] | ((uint32_t)p[n - 3] << 8) | ((uint32_t)p[n - 2] << 16) | ((uint32_t)p[n - 1] << 24);
        long hl = orc_parse_header(p, n);
        if (hl < 0) {
            e = ORC_E_INVALID;
            break;
        }
        size_t made = 0, used = 0;
        int r = orc_safe_inflate(p + hl, n - (size_t)hl, out + pos, isz, &made, &used);
        if (r || (size_t)hl + used + 8 > n) {
            e = ORC_E_INVALID;
            break;
        }
        const uint8_t* t = p + hl + used;
        uint32_t crc = (uint32_t)t[0] | ((uint32_t)t[1] << 8) | ((uint32_t)t[2] << 16) | ((uint32_t)t[3] << 24);
        uint32_t tis = (uint32_t)t[4] | ((uint32_t)t[5] << 8) | ((uint32_t)t[6] << 16) | ((uint32_t)t[7] << 24);
        if (orc_crc32(0, out + pos, made) != crc || (uint32_t)made != tis) {
            e = ORC_E_INVALID;
            break;
        }
        pos += made;
    }
    b->out[bi] = out;
    b->olen[bi] = pos;
    b->oerr[bi] = e;
}

/* pipeline.rs:1357-1636 (parallel_decode_bgzf); member decode by the textbook
 * inflate (the reference's default engine is libdeflate, whose arithmetic is
 * outside the reference tree; any failure maps to one Invalid error). */
static int decode_bgzf_path(const uint8_t* file, size_t len, size_t chunk_size, int nthreads, orc_sink_cb cb,
                            void* user, orc_stats* st, orc_error* err) {
    size_t capm = 1024, nm = 0;
    size_t* ms = (size_t*)malloc(capm * sizeof(size_t));
    size_t* me = (size_t*)malloc(capm * sizeof(size_t));
    size_t pos = 0;
    int walk_err = 0;
    while (pos < len) {
        uint32_t sz = orc_parse_bgzf_block_size(file + pos, len - pos);
        if (!sz) break;
        if (pos + sz > len) {
            walk_err = ORC_E_TRUNCATED;
            break;
        }
        if (nm == capm) {
            capm *= 2;
            ms = (size_t*)realloc(ms, capm * sizeof(size_t));
            me = (size_t*)realloc(me, capm * sizeof(size_t));
        }
        ms[nm] = pos;
        me[nm] = pos + sz;
        nm++;
        pos += sz;
    }
    /* batch ramp 64 KiB -> chunk target (pipeline.rs:1488-1551) */
    size_t target_max = chunk_size > 65536 ? chunk_size : 65536, target = 65536;
    size_t nbat = 0, capb = 64;
    size_t* bstart = (size_t*)malloc(capb * sizeof(size_t));
    size_t* bcount = (size_t*)malloc(capb * sizeof(size_t));
    size_t cur_first = 0, cur_bytes = 0;
    for (size_t m = 0; m < nm; m++) {
        cur_bytes += me[m] - ms[m];
        if (cur_bytes >= target || m + 1 == nm) {
            if (nbat == capb) {
                capb *= 2;
                bstart = (size_t*)realloc(bstart, capb * sizeof(size_t));
                bcount = (size_t*)realloc(bcount, capb * sizeof(size_t));
            }
            bstart[nbat] = cur_first;
            bcount[nbat] = m + 1 - cur_first;
            nbat++;
            cur_first = m + 1;
            if (cur_bytes >= target) target = target * 2 < target_max ? target * 2 : target_max;
            cur_bytes = 0;
        }
    }
    uint8_t** out = (uint8_t**)calloc(nbat + 1, sizeof(uint8_t*));
    size_t* olen = (size_t*)calloc(nbat + 1, sizeof(size_t));
    int* oerr = (int*)calloc(nbat + 1, sizeof(int));
    BgzfCtx bc = {file, ms, me, 0, out, olen, oerr, bstart, bcount};
    parallel_for(nbat, nthreads, bgzf_batch, &bc);
    int rc = ORC_OK, closed = 0;
    for (size_t b = 0; b < nbat; b++) {
        if (oerr[b]) {
            set_err(err, ORC_E_INVALID, "gzip member decode failed");
            rc = ORC_E_INVALID;
            break;
        }
        st->uncompressed_bytes += olen[b];
        if (cb && cb(user, out[b], olen[b])) {
            set_err(err, ORC_E_IO, "sink closed");
            rc = ORC_E_IO;
            closed = 1;
            break;
        }
        st->chunks_decoded++;
    }
    (void)closed;
    for (size_t b = 0; b < nbat; b++) free(out[b]);
    free(out);
    free(olen);
    free(oerr);
    free(bstart);
    free(bcount);
    free(ms);
    free(me);
    if (rc == ORC_OK && walk_err) {
        set_err(err, walk_err, "truncated bgzf member");
        rc = walk_err;
    }
    if (rc == ORC_OK && pos < len) {
        long hl = orc_parse_header(file + pos, len - pos);
        if (hl < 0) {
            set_err(err, (int)-hl, "bad gzip header");
            return (int)-hl;
        }
        if (nthreads == 1)
            rc = serial_member_path(file, len, pos + (size_t)hl, chunk_size, 0, cb, user, st, err);
        else
            rc = decode_member_path(file + pos + hl, len - pos - (size_t)hl, chunk_size, nthreads, cb, user, st, err);
    }
    return rc;
}

/* effective_threads (pipeline.rs:1187-1196): 0 = the machine's parallelism */
static int effective_threads(int nthreads) {
    if (nthreads > 0) return nthreads;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 1 ? (int)n : 1;
}

/* lib.rs:366-496 */
int orc_read_gz(const uint8_t* file, size_t len, size_t chunk_size, int nthreads, orc_sink_cb cb, void* user,
                orc_stats* st, orc_error* err) {
    memset(st, 0, sizeof *st);
    if (err) memset(err, 0, sizeof *err);
    if (len == 0) {
        set_err(err, ORC_E_IO, "empty input");
        return ORC_E_IO;
    }
    st->compressed_bytes = len;
    nthreads = effective_threads(nthreads);
    if (orc_parse_bgzf_block_size(file, len)) return decode_bgzf_path(file, len, chunk_size, nthreads, cb, user, st, err);
    long hl = orc_parse_header(file, len);
    if (hl < 0) {
        set_err(err, (int)-hl, "bad gzip header");
        return (int)-hl;
    }
    /* a single worker decodes serially (pipeline.rs:481-488) */
    if (nthreads == 1) return serial_member_path(file, len, (size_t)hl, chunk_size, 0, cb, user, st, err);
    return decode_member_path(file + hl, len - (size_t)hl, chunk_size, nthreads, cb, user, st, err);
}

/* lib.rs:378-408: a non-seekable (or empty) input with one worker takes the
 * streaming serial path, which reads every member header itself and knows no
 * BGZF; compressed_bytes stays 0 (lib.rs:402-407). */
int orc_read_gz_stream_serial(const uint8_t* file, size_t len, orc_sink_cb cb, void* user, orc_stats* st,
                              orc_error* err) {
    memset(st, 0, sizeof *st);
    if (err) memset(err, 0, sizeof *err);
    return serial_member_path(file, len, 0, 0, 1, cb, user, st, err);
}
