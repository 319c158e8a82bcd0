/*
 * rrg_synth.c -- seeded synthetic workloads and gzip/BGZF writers.
 *
 * Generators (all deterministic in the seed, independent of thread count):
 *   text   : order-2 Markov chain over 64 printable symbols ('!'..'`'), each of
 *            the 4096 context rows drawn from Dirichlet(0.3), newline after a
 *            geometric(mean 80) run of symbols.
 *   fastq  : 4-line records "@r<id>", 150 bp uniform ACGT with N at p=0.001,
 *            "+", and a 150-char Phred walk over '!'..'J'.
 *   mixed  : 1-8 MiB segments cycling through random bytes (stored), byte runs,
 *            period-3..300 repeats, a repeated 32 KiB random block, and Markov
 *            text at level 1 or 9.
 * Writers:
 *   gzip   : one gzip member, compressed by zlib in independent pieces that are
 *            primed with the previous 32 KiB and joined with sync flushes (the
 *            stream is a single valid member; back-references cross pieces).
 *   bgzf   : 65280-byte blocks, each a gzip member with the BC/BSIZE extra
 *            field, plus the 28-byte BGZF end-of-file block.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

/* ---- RNG: splitmix64 seeding + xoshiro256** -------------------------------- */
typedef struct {
    uint64_t s[4];
} Rng;
static uint64_t splitmix(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static void rng_seed(Rng* r, uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 4; i++) r->s[i] = splitmix(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_next(Rng* r) {
    uint64_t* s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}
static inline double rng_unif(Rng* r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static double rng_normal(Rng* r) {
    double u1 = rng_unif(r), u2 = rng_unif(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
/* Gamma(a) by Marsaglia-Tsang, with the U^(1/a) boost for a < 1 */
static double rng_gamma(Rng* r, double a) {
    if (a < 1.0) {
        double u = rng_unif(r);
        return rng_gamma(r, a + 1.0) * pow(u > 1e-300 ? u : 1e-300, 1.0 / a);
    }
    double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (;;) {
        double x = rng_normal(r), v = 1.0 + c * x;
        if (v <= 0) continue;
        v = v * v * v;
        double u = rng_unif(r);
        if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
        if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v;
    }
}

/* ---- parallel helper -------------------------------------------------------- */
typedef struct {
    void (*fn)(void*, size_t);
    void* ctx;
    size_t n;
    size_t next;
    pthread_mutex_t mu;
} Par;
static void* par_worker(void* a) {
    Par* p = (Par*)a;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        size_t i = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (i >= p->n) return 0;
        p->fn(p->ctx, i);
    }
}
static void par_for(size_t n, int nt, void (*fn)(void*, size_t), void* ctx) {
    Par p = {fn, ctx, n, 0, PTHREAD_MUTEX_INITIALIZER};
    if (nt < 1) nt = 1;
    if ((size_t)nt > n) nt = (int)n;
    if (nt <= 1) {
        par_worker(&p);
        return;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int i = 0; i < nt; i++) pthread_create(&th[i], 0, par_worker, &p);
    for (int i = 0; i < nt; i++) pthread_join(th[i], 0);
    free(th);
}

/* ---- Markov text ------------------------------------------------------------- */
typedef struct {
    uint16_t cdf[4096][64]; /* cumulative probabilities scaled to 65535 */
} TextModel;

static void text_model(TextModel* m, uint64_t seed) {
    Rng r;
    rng_seed(&r, seed, 0x7E57);
    for (int row = 0; row < 4096; row++) {
        double g[64], s = 0;
        for (int k = 0; k < 64; k++) s += (g[k] = rng_gamma(&r, 0.3));
        double acc = 0;
        for (int k = 0; k < 64; k++) {
            acc += g[k] / s;
            uint32_t v = (uint32_t)(acc * 65535.0 + 0.5);
            m->cdf[row][k] = (uint16_t)(v > 65535 ? 65535 : v);
        }
        m->cdf[row][63] = 65535;
    }
}

/* text of n bytes from stream `stream`; state starts at row 0 */
static void text_fill(const TextModel* m, uint64_t seed, uint64_t stream, uint8_t* out, size_t n) {
    Rng r;
    rng_seed(&r, seed, stream);
    unsigned a = 0, b = 0;
    size_t i = 0;
    while (i < n) {
        /* geometric run length with mean 80, then a newline */
        double u = rng_unif(&r);
        size_t run = 1 + (size_t)(log(1.0 - u) / log(1.0 - 1.0 / 80.0));
        for (size_t k = 0; k < run && i < n; k++) {
            const uint16_t* cdf = m->cdf[(a << 6) | b];
            unsigned x = (unsigned)(rng_next(&r) >> 48);
            unsigned s = 0;
            while (s < 63 && cdf[s] < x) s++;
            out[i++] = (uint8_t)(0x21 + s);
            a = b;
            b = s;
        }
        if (i < n) out[i++] = '\n';
    }
}

#define TEXT_BLOCK (4u << 20)
typedef struct {
    const TextModel* m;
    uint64_t seed;
    uint8_t* out;
    size_t n;
} TextCtx;
static void text_job(void* c, size_t i) {
    TextCtx* t = (TextCtx*)c;
    size_t off = i * (size_t)TEXT_BLOCK, len = t->n - off < TEXT_BLOCK ? t->n - off : TEXT_BLOCK;
    text_fill(t->m, t->seed, 1 + i, t->out + off, len);
}

int rrg_synth_text(uint64_t seed, uint8_t* out, size_t n, int nthreads) {
    TextModel* m = (TextModel*)malloc(sizeof(TextModel));
    if (!m) return -1;
    text_model(m, seed);
    TextCtx c = {m, seed, out, n};
    par_for((n + TEXT_BLOCK - 1) / TEXT_BLOCK, nthreads, text_job, &c);
    free(m);
    return 0;
}

/* ---- FASTQ-like -------------------------------------------------------------- */
#define FQ_RECS 65536u /* records per independent generation block */
typedef struct {
    uint64_t seed;
    uint8_t** blk;
    size_t* blen;
} FqCtx;

static size_t fq_block(uint64_t seed, size_t bi, uint8_t* o) {
    Rng r;
    rng_seed(&r, seed, 0xFA57000000ull + bi);
    static const char ACGT[4] = {'A', 'C', 'G', 'T'};
    size_t p = 0;
    for (size_t k = 0; k < FQ_RECS; k++) {
        uint64_t id = (uint64_t)bi * FQ_RECS + k;
        char num[24];
        int nd = 0;
        do {
            num[nd++] = (char)('0' + id % 10);
            id /= 10;
        } while (id);
        o[p++] = '@';
        o[p++] = 'r';
        while (nd) o[p++] = (uint8_t)num[--nd];
        o[p++] = '\n';
        for (int j = 0; j < 150; j++) {
            uint64_t x = rng_next(&r);
            o[p++] = (x >> 54) == 0 ? 'N' : (uint8_t)ACGT[x & 3]; /* p(N) = 2^-10 ~ 0.001 */
        }
        o[p++] = '\n';
        o[p++] = '+';
        o[p++] = '\n';
        /* Phred walk: sticky, drifting back toward Q36, occasional jumps */
        int q = 30 + (int)(rng_next(&r) % 10);
        for (int j = 0; j < 150; j++) {
            uint64_t x = rng_next(&r) % 100;
            if (x < 70) {
            } else if (x < 85) {
                q += q < 36 ? 1 : -1;
            } else if (x < 95) {
                q += (x & 1) ? 1 : -1;
            } else {
                q += (x & 1) ? 3 : -3;
            }
            if (q < 0) q = -q;
            if (q > 41) q = 82 - q;
            o[p++] = (uint8_t)('!' + q);
        }
        o[p++] = '\n';
    }
    return p;
}
static void fq_job(void* c, size_t i) {
    FqCtx* f = (FqCtx*)c;
    f->blk[i] = (uint8_t*)malloc(FQ_RECS * 340ull);
    f->blen[i] = fq_block(f->seed, i, f->blk[i]);
}

/* exactly n bytes (the last record is cut at n) */
int rrg_synth_fastq(uint64_t seed, uint8_t* out, size_t n, int nthreads) {
    size_t per = FQ_RECS * 300ull; /* lower bound of a block's size */
    size_t nb = n / per + 2;
    FqCtx c = {seed, (uint8_t**)calloc(nb, sizeof(uint8_t*)), (size_t*)calloc(nb, sizeof(size_t))};
    par_for(nb, nthreads, fq_job, &c);
    size_t p = 0;
    for (size_t i = 0; i < nb; i++) {
        size_t take = c.blen[i] < n - p ? c.blen[i] : n - p;
        memcpy(out + p, c.blk[i], take);
        p += take;
        free(c.blk[i]);
    }
    free(c.blk);
    free(c.blen);
    return p == n ? 0 : -1;
}

/* ---- gzip writer (pieces primed with the previous 32 KiB) --------------------- */
typedef struct {
    const uint8_t* in;
    size_t n;
    size_t piece;
    const int* seg_level; /* optional: level per segment */
    const size_t* seg_end;
    size_t nseg;
    int level;
    uint8_t** out;
    size_t* olen;
    int* err;
} GzCtx;

static int level_at(const GzCtx* g, size_t pos) {
    if (!g->seg_level) return g->level;
    size_t lo = 0, hi = g->nseg;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (g->seg_end[mid] <= pos) lo = mid + 1;
        else hi = mid;
    }
    return g->seg_level[lo < g->nseg ? lo : g->nseg - 1];
}

static void gz_piece(void* c, size_t i) {
    GzCtx* g = (GzCtx*)c;
    size_t off = i * g->piece, len = g->n - off < g->piece ? g->n - off : g->piece;
    int last = off + len == g->n;
    z_stream s;
    memset(&s, 0, sizeof s);
    int lvl = level_at(g, off);
    if (deflateInit2(&s, lvl, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
        g->err[i] = 1;
        return;
    }
    if (off) {
        size_t dl = off < 32768 ? off : 32768;
        deflateSetDictionary(&s, g->in + off - dl, (uInt)dl);
    }
    size_t cap = deflateBound(&s, (uLong)len) + 64 + len / 8;
    uint8_t* o = (uint8_t*)malloc(cap);
    s.next_out = o;
    s.avail_out = (uInt)cap;
    size_t p = off;
    int ok = 1;
    while (p < off + len && ok) {
        /* feed up to the next segment boundary so the level can change there */
        size_t end = off + len;
        if (g->seg_level) {
            int cur = level_at(g, p);
            size_t lo = 0, hi = g->nseg;
            while (lo < hi) {
                size_t mid = (lo + hi) / 2;
                if (g->seg_end[mid] <= p) lo = mid + 1;
                else hi = mid;
            }
            if (lo < g->nseg && g->seg_end[lo] < end) end = g->seg_end[lo];
            if (cur != lvl) {
                /* A new level: the stream so far ends in a sync flush and a new
                 * one starts at the new level, primed with the 32 KiB before
                 * (deflateParams of zlib 1.2.11 crashed inside deflate on some
                 * of these inputs). */
                s.next_in = 0;
                s.avail_in = 0;
                if (deflate(&s, Z_SYNC_FLUSH) == Z_STREAM_ERROR) ok = 0;
                Bytef* const no = s.next_out;
                const uInt na = s.avail_out;
                deflateEnd(&s);
                memset(&s, 0, sizeof s);
                if (deflateInit2(&s, cur, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
                    ok = 0;
                    break;
                }
                if (p) {
                    size_t dl = p < 32768 ? p : 32768;
                    deflateSetDictionary(&s, g->in + p - dl, (uInt)dl);
                }
                s.next_out = no;
                s.avail_out = na;
                lvl = cur;
            }
        }
        while (p < end && ok) {
            size_t take = end - p < (1u << 30) ? end - p : (1u << 30);
            s.next_in = (Bytef*)(g->in + p);
            s.avail_in = (uInt)take;
            int fl = (p + take == off + len) ? (last ? Z_FINISH : Z_SYNC_FLUSH) : Z_NO_FLUSH;
            int rc = deflate(&s, fl);
            if (rc == Z_STREAM_ERROR || s.avail_in) ok = 0;
            if (fl == Z_FINISH && rc != Z_STREAM_END) ok = 0;
            p += take;
        }
    }
    if (len == 0 && last) {
        s.avail_in = 0;
        if (deflate(&s, Z_FINISH) != Z_STREAM_END) ok = 0;
    }
    g->out[i] = o;
    g->olen[i] = cap - s.avail_out;
    g->err[i] = !ok;
    deflateEnd(&s);
}

/* Returns compressed size or 0 on failure.  out must hold rrg_gzip_bound(n). */
size_t rrg_gzip_bound(size_t n) { return n + n / 8 + (n >> 12) * 64 + 4096; }

static size_t gzip_common(const uint8_t* in, size_t n, int level, const int* seg_level, const size_t* seg_end,
                          size_t nseg, uint8_t* out, size_t cap, int nthreads, size_t piece) {
    if (piece < (1u << 20)) piece = 1u << 20;
    size_t np = n ? (n + piece - 1) / piece : 1;
    GzCtx g = {in, n, piece, seg_level, seg_end, nseg, level,
               (uint8_t**)calloc(np, sizeof(uint8_t*)), (size_t*)calloc(np, sizeof(size_t)), (int*)calloc(np, sizeof(int))};
    par_for(np, nthreads, gz_piece, &g);
    size_t p = 0;
    int bad = 0;
    static const uint8_t hdr[10] = {0x1f, 0x8b, 8, 0, 0, 0, 0, 0, 0, 3};
    if (cap < 18) bad = 1;
    else {
        memcpy(out, hdr, 10);
        p = 10;
    }
    for (size_t i = 0; i < np; i++) {
        if (g.err[i] || p + g.olen[i] + 8 > cap) bad = 1;
        if (!bad) {
            memcpy(out + p, g.out[i], g.olen[i]);
            p += g.olen[i];
        }
        free(g.out[i]);
    }
    free(g.out);
    free(g.olen);
    free(g.err);
    if (bad) return 0;
    uint32_t crc = 0;
    size_t q = 0;
    while (q < n) {
        size_t t = n - q < (1u << 30) ? n - q : (1u << 30);
        crc = (uint32_t)crc32(crc, in + q, (uInt)t);
        q += t;
    }
    uint32_t isz = (uint32_t)n;
    for (int k = 0; k < 4; k++) out[p++] = (uint8_t)(crc >> (8 * k));
    for (int k = 0; k < 4; k++) out[p++] = (uint8_t)(isz >> (8 * k));
    return p;
}

size_t rrg_gzip(const uint8_t* in, size_t n, int level, uint8_t* out, size_t cap, int nthreads, size_t piece) {
    return gzip_common(in, n, level, 0, 0, 0, out, cap, nthreads, piece);
}

/* ---- BGZF writer ------------------------------------------------------------ */
#define BGZF_BLOCK 65280u
typedef struct {
    const uint8_t* in;
    size_t n;
    int level;
    uint8_t* out; /* slot of 65536+64 per block */
    uint32_t* olen;
    int* err;
} BgCtx;
static void bg_job(void* c, size_t i) {
    BgCtx* b = (BgCtx*)c;
    size_t off = (size_t)i * BGZF_BLOCK, len = b->n - off < BGZF_BLOCK ? b->n - off : BGZF_BLOCK;
    uint8_t* o = b->out + (size_t)i * 65600u;
    z_stream s;
    memset(&s, 0, sizeof s);
    if (deflateInit2(&s, b->level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
        b->err[i] = 1;
        return;
    }
    s.next_in = (Bytef*)(b->in + off);
    s.avail_in = (uInt)len;
    s.next_out = o + 18;
    s.avail_out = 65536 - 26;
    int rc = deflate(&s, Z_FINISH);
    size_t clen = 65536 - 26 - s.avail_out;
    deflateEnd(&s);
    if (rc != Z_STREAM_END) {
        b->err[i] = 1;
        return;
    }
    static const uint8_t h[16] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 'B', 'C', 2, 0};
    memcpy(o, h, 16);
    size_t total = 18 + clen + 8;
    o[16] = (uint8_t)((total - 1) & 0xff);
    o[17] = (uint8_t)((total - 1) >> 8);
    uint32_t crc = (uint32_t)crc32(0, b->in + off, (uInt)len);
    uint8_t* t = o + 18 + clen;
    for (int k = 0; k < 4; k++) t[k] = (uint8_t)(crc >> (8 * k));
    for (int k = 0; k < 4; k++) t[4 + k] = (uint8_t)((uint32_t)len >> (8 * k));
    b->olen[i] = (uint32_t)total;
}

size_t rrg_bgzf_bound(size_t n) { return (n / BGZF_BLOCK + 2) * 65600u + 64; }

size_t rrg_bgzf(const uint8_t* in, size_t n, int level, uint8_t* out, size_t cap, int nthreads) {
    size_t nb = (n + BGZF_BLOCK - 1) / BGZF_BLOCK;
    uint8_t* tmp = (uint8_t*)malloc(nb * 65600u + 1);
    BgCtx b = {in, n, level, tmp, (uint32_t*)calloc(nb + 1, 4), (int*)calloc(nb + 1, sizeof(int))};
    par_for(nb, nthreads, bg_job, &b);
    size_t p = 0;
    int bad = 0;
    for (size_t i = 0; i < nb && !bad; i++) {
        if (b.err[i] || p + b.olen[i] > cap) {
            bad = 1;
            break;
        }
        memmove(out + p, tmp + i * 65600u, b.olen[i]);
        p += b.olen[i];
    }
    static const uint8_t eof[28] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 'B', 'C',
                                    2,    0,    0x1b, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!bad && p + 28 <= cap) {
        memcpy(out + p, eof, 28);
        p += 28;
    } else {
        bad = 1;
    }
    free(tmp);
    free(b.olen);
    free(b.err);
    return bad ? 0 : p;
}

/* ---- pigz-style members -------------------------------------------------------- */
/* Consecutive `member`-byte pieces, each an independent gzip member with a plain
 * 10-byte header (no extra field, so not BGZF), as parallel gzip writers emit. */
typedef struct {
    const uint8_t* in;
    size_t n, member, slot;
    int level;
    uint8_t* out;
    size_t* olen;
    int* err;
} MemCtx;
static void mem_job(void* c, size_t i) {
    MemCtx* b = (MemCtx*)c;
    size_t off = i * b->member, len = b->n - off < b->member ? b->n - off : b->member;
    uint8_t* o = b->out + i * b->slot;
    z_stream s;
    memset(&s, 0, sizeof s);
    if (deflateInit2(&s, b->level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
        b->err[i] = 1;
        return;
    }
    s.next_in = (Bytef*)(b->in + off);
    s.avail_in = (uInt)len;
    s.next_out = o + 10;
    s.avail_out = (uInt)(b->slot - 18);
    int rc = deflate(&s, Z_FINISH);
    size_t clen = b->slot - 18 - s.avail_out;
    deflateEnd(&s);
    if (rc != Z_STREAM_END) {
        b->err[i] = 1;
        return;
    }
    static const uint8_t h[10] = {0x1f, 0x8b, 8, 0, 0, 0, 0, 0, 0, 0xff};
    memcpy(o, h, 10);
    uint32_t crc = (uint32_t)crc32(0, b->in + off, (uInt)len);
    uint8_t* t = o + 10 + clen;
    for (int k = 0; k < 4; k++) t[k] = (uint8_t)(crc >> (8 * k));
    for (int k = 0; k < 4; k++) t[4 + k] = (uint8_t)((uint32_t)len >> (8 * k));
    b->olen[i] = 10 + clen + 8;
}

size_t rrg_members_bound(size_t n, size_t member) {
    const size_t nb = (n + member - 1) / member;
    return (nb ? nb : 1) * (compressBound((uLong)member) + 64) + 64;
}

size_t rrg_members(const uint8_t* in, size_t n, size_t member, int level, uint8_t* out, size_t cap, int nthreads) {
    if (member == 0 || member > (1u << 30)) return 0;
    const size_t nb = (n + member - 1) / member, slot = compressBound((uLong)member) + 64;
    uint8_t* tmp = (uint8_t*)malloc(nb * slot + 1);
    MemCtx b = {in, n, member, slot, level, tmp, (size_t*)calloc(nb + 1, sizeof(size_t)), (int*)calloc(nb + 1, sizeof(int))};
    par_for(nb, nthreads, mem_job, &b);
    size_t p = 0;
    int bad = 0;
    for (size_t i = 0; i < nb; i++) {
        if (b.err[i] || p + b.olen[i] > cap) {
            bad = 1;
            break;
        }
        memmove(out + p, tmp + i * slot, b.olen[i]);
        p += b.olen[i];
    }
    free(tmp);
    free(b.olen);
    free(b.err);
    return bad ? 0 : p;
}

/* ---- mixed-entropy single member --------------------------------------------- */
/* Fills raw[n]; writes per-segment levels into seg_level/seg_end (cap entries);
 * returns the number of segments. */
size_t rrg_synth_mixed(uint64_t seed, uint8_t* raw, size_t n, int* seg_level, size_t* seg_end, size_t cap,
                       int nthreads) {
    Rng r;
    rng_seed(&r, seed, 0x313D);
    TextModel* m = (TextModel*)malloc(sizeof(TextModel));
    text_model(m, seed ^ 0x55);
    size_t p = 0, ns = 0;
    unsigned kind = 0;
    (void)nthreads;
    while (p < n && ns < cap) {
        size_t len = (1u << 20) + (size_t)(rng_next(&r) % (7u << 20));
        if (len > n - p) len = n - p;
        uint8_t* o = raw + p;
        int lvl = 6;
        switch (kind % 5) {
            case 0: /* incompressible: stored blocks */
                for (size_t i = 0; i < len; i += 8) {
                    uint64_t x = rng_next(&r);
                    for (size_t k = 0; k < 8 && i + k < len; k++) o[i + k] = (uint8_t)(x >> (8 * k));
                }
                lvl = 0;
                break;
            case 1: { /* long byte runs */
                size_t i = 0;
                while (i < len) {
                    size_t run = 64 + (size_t)(rng_next(&r) % 65536);
                    uint8_t v = (uint8_t)rng_next(&r);
                    for (size_t k = 0; k < run && i < len; k++) o[i++] = v;
                }
                break;
            }
            case 2: { /* period-p repeats, p in [3,300] */
                size_t i = 0;
                while (i < len) {
                    size_t per = 3 + (size_t)(rng_next(&r) % 298), reps = 16 + (size_t)(rng_next(&r) % 4096);
                    uint8_t pat[300];
                    for (size_t k = 0; k < per; k++) pat[k] = (uint8_t)rng_next(&r);
                    for (size_t k = 0; k < per * reps && i < len; k++) o[i++] = pat[k % per];
                }
                break;
            }
            case 3: { /* a 32 KiB random block repeated: distance-32768 matches */
                uint8_t* blk = (uint8_t*)malloc(32768);
                for (size_t k = 0; k < 32768; k += 8) {
                    uint64_t x = rng_next(&r);
                    memcpy(blk + k, &x, 8);
                }
                for (size_t i = 0; i < len; i++) o[i] = blk[i & 32767];
                free(blk);
                break;
            }
            default: /* Markov text at level 1 or 9 */
                text_fill(m, seed, 0x1000 + ns, o, len);
                lvl = (rng_next(&r) & 1) ? 9 : 1;
                break;
        }
        seg_level[ns] = lvl;
        p += len;
        seg_end[ns] = p;
        ns++;
        kind++;
    }
    free(m);
    return ns;
}

size_t rrg_gzip_segments(const uint8_t* in, size_t n, const int* seg_level, const size_t* seg_end, size_t nseg,
                         uint8_t* out, size_t cap, int nthreads, size_t piece) {
    return gzip_common(in, n, 6, seg_level, seg_end, nseg, out, cap, nthreads, piece);
}

/*
 * A consumer for timing the host-output decode: an rrg_chunk_cb that copies
 * each chunk into one reused buffer (as a consumer that parses or writes the
 * bytes would touch them) and counts chunks and bytes.
 */
typedef struct {
    uint8_t* buf;
    size_t cap;
    uint64_t bytes, chunks;
} RrgCopySink;

int rrg_sink_copy(void* user, const uint8_t* data, size_t len) {
    RrgCopySink* s = (RrgCopySink*)user;
    size_t off = 0;
    while (off < len) {
        const size_t n = len - off < s->cap ? len - off : s->cap;
        memcpy(s->buf, data + off, n);
        off += n;
    }
    s->bytes += len;
    s->chunks += 1;
    return 0;
}

/* The same callback without touching the bytes: the decoder's own delivery rate. */
int rrg_sink_count(void* user, const uint8_t* data, size_t len) {
    RrgCopySink* s = (RrgCopySink*)user;
    (void)data;
    s->bytes += len;
    s->chunks += 1;
    return 0;
}
