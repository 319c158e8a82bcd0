// Host harness for the CPU tests: runs the finder's per-lane strict header
// check and its two precode prefilters (rusty_rapidgzip_amd/csrc/rrg_hdr.h)
// and the oracle's strict header parse on every bit position of a raw
// DEFLATE stream and reports where they disagree.  Usage: strict_header_check <file> <first bit> <end bit>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../oracle/rrg_oracle.h"
#include "../rusty_rapidgzip_amd/csrc/rrg_hdr.h"

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<uint8_t> data;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + got);
    std::fclose(f);
    const uint64_t nbits = (uint64_t)data.size() * 8;
    std::vector<uint64_t> words(data.size() / 8 + 4, 0);
    std::memcpy(words.data(), data.data(), data.size());
    const uint64_t nwords = words.size();
    auto fetch = [&](uint64_t q) -> uint64_t { return q < nwords ? words[q] : 0; };
    const uint64_t from = std::strtoull(argv[2], nullptr, 10), to = std::strtoull(argv[3], nullptr, 10);
    uint32_t tab[512];
    for (uint32_t k = 0; k < 512; ++k) tab[k] = rrg::precode3_entry(k);
    auto bit = [&](uint64_t q) -> uint32_t { return q < nbits ? (data[q >> 3] >> (q & 7)) & 1u : 0u; };
    uint64_t pass = 0, bad = 0;
    for (uint64_t p = from; p < to && p < nbits; ++p) {
        const bool a = rrg::strict_header_lane(fetch, nbits, p);
        const bool b = orc_strict_header_ok(data.data(), data.size(), p) != 0;
        pass += b;
        if (a != b) {
            if (bad < 10) std::printf("mismatch at bit %llu: lane %d oracle %d\n", (unsigned long long)p, a, b);
            ++bad;
        }
        // The finder's precode prefilters (bit by bit and by the 3-length
        // table) against the definition: HCLEN + 4 three-bit lengths after
        // the 17 header bits, a single used length or a complete code
        // (sum of 2^-len == 1), every bit inside the stream.
        const uint64_t q = p >> 6;
        const uint32_t sh = (uint32_t)(p & 63);
        const uint64_t w0 = fetch(q), w1 = fetch(q + 1), w2 = fetch(q + 2);
        const uint64_t lo = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;
        const uint64_t hi = sh ? (w1 >> sh) | (w2 << (64 - sh)) : w1;
        uint32_t hclen = 4;
        for (int i = 0; i < 4; ++i) hclen += bit(p + 13 + i) << i;
        bool def = p + 17 + 3ull * hclen <= nbits;
        if (def) {
            uint32_t used = 0, kraft = 0;  // kraft in units of 2^-7
            for (uint32_t i = 0; i < hclen; ++i) {
                const uint32_t len = bit(p + 17 + 3 * i) | (bit(p + 18 + 3 * i) << 1) | (bit(p + 19 + 3 * i) << 2);
                if (len) {
                    ++used;
                    kraft += 1u << (7 - len);
                }
            }
            def = used == 1 || (used >= 2 && kraft == 128);
        }
        const bool c = rrg::precode_ok(lo, hi, p, nbits), d = rrg::precode_ok_lut(tab, lo, hi, p, nbits);
        if (c != def || d != def || (b && !def)) {
            if (bad < 10)
                std::printf("precode mismatch at bit %llu: bitwise %d table %d definition %d strict %d\n",
                            (unsigned long long)p, c, d, def, b);
            ++bad;
        }
    }
    std::printf("positions %llu strict %llu mismatches %llu\n", (unsigned long long)(to - from),
                (unsigned long long)pass, (unsigned long long)bad);
    return bad ? 1 : 0;
}
