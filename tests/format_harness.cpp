// Compares blamm::put_g6 with std::to_chars(general, precision 6) on the
// double value of a float (the text the reference's occurrence writer gets
// from ostream's default float output).  Prints "mismatches N checked M" and
// the first few differing values.
//   format_harness random COUNT SEED      random bit patterns (finite floats)
//   format_harness range LO_BITS HI_BITS  every float bit pattern in [LO, HI)
//   format_harness ties                   m / 2^s for small m: exact decimal ties
#include <charconv>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "../blamm_amd/csrc/host/format_g6.hpp"

static uint64_t checked = 0, bad = 0;

static void check(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, sizeof f);
    if (!(f == f) || f - f != 0.f) return;   // nan / inf: only finite scores occur
    char a[32], b[32];
    char* ea = blamm::put_g6(a, f);
    char* eb = std::to_chars(b, b + sizeof b, (double)f, std::chars_format::general, 6).ptr;
    ++checked;
    if (ea - a != eb - b || std::memcmp(a, b, (size_t)(ea - a)) != 0) {
        if (bad < 10)
            std::printf("diff %08" PRIx32 " put_g6=%.*s to_chars=%.*s\n", u, (int)(ea - a), a, (int)(eb - b), b);
        ++bad;
    }
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "random" && argc == 4) {
        std::mt19937_64 g(std::strtoull(argv[3], nullptr, 10));
        const uint64_t n = std::strtoull(argv[2], nullptr, 10);
        for (uint64_t i = 0; i < n; ++i) check((uint32_t)g());
    } else if (mode == "range" && argc == 4) {
        const uint64_t lo = std::strtoull(argv[2], nullptr, 0), hi = std::strtoull(argv[3], nullptr, 0);
        for (uint64_t x = lo; x < hi; ++x) check((uint32_t)x);
    } else if (mode == "ties") {
        for (int s = 1; s <= 40; ++s)
            for (uint32_t m = 1; m < (1u << 16); ++m) {
                const float f = (float)m / (float)(1ull << s);
                uint32_t u;
                std::memcpy(&u, &f, sizeof u);
                check(u);
                check(u | 0x80000000u);
            }
    } else {
        return 2;
    }
    std::printf("mismatches %" PRIu64 " checked %" PRIu64 "\n", bad, checked);
    return 0;
}
