// Host harness for the CPU tests: runs the finder's per-lane strict header
// check (rusty_rapidgzip_amd/csrc/rrg_hdr.h) and the oracle's strict header
// parse on every bit position of a raw DEFLATE stream and reports where they
// disagree.  Usage: strict_header_check <file> <first bit> <end bit>
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
    uint64_t pass = 0, bad = 0;
    for (uint64_t p = from; p < to && p < nbits; ++p) {
        const bool a = rrg::strict_header_lane(fetch, nbits, p);
        const bool b = orc_strict_header_ok(data.data(), data.size(), p) != 0;
        pass += b;
        if (a != b) {
            if (bad < 10) std::printf("mismatch at bit %llu: lane %d oracle %d\n", (unsigned long long)p, a, b);
            ++bad;
        }
    }
    std::printf("positions %llu strict %llu mismatches %llu\n", (unsigned long long)(to - from),
                (unsigned long long)pass, (unsigned long long)bad);
    return bad ? 1 : 0;
}
