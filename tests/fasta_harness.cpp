// Test driver for the host FASTA reader (pack_fasta): prints the packed
// stream decoded back to bytes, its runs and its record names.
// Usage: fasta_harness maxLen file...
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../blamm_amd/csrc/host/host.hpp"

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const uint64_t maxLen = std::strtoull(argv[1], nullptr, 10);
    std::vector<std::string> files(argv + 2, argv + argc);
    try {
        const blamm::PackedStream s = blamm::pack_fasta(files, maxLen);
        std::printf("n %llu nseq %llu runs %zu names %zu\n", (unsigned long long)s.n, (unsigned long long)s.nseq,
                    s.runStart.size(), s.names.size());
        std::string text(s.n, '?');
        for (uint64_t i = 0; i < s.n; ++i) {
            const int b = (int)((s.codes[i >> 5] >> (2 * (i & 31))) & 3);
            const bool low = s.hasLower() && ((s.lower[i >> 6] >> (i & 63)) & 1);
            text[i] = (low ? "acgt" : "ACGT")[b];
        }
        std::printf("stream %s\n", text.c_str());
        for (size_t r = 0; r < s.runStart.size(); ++r)
            std::printf("run %llu %llu %llu\n", (unsigned long long)s.runStart[r], (unsigned long long)s.runSeq[r],
                        (unsigned long long)s.runPos[r]);
        for (const std::string& nm : s.names) std::printf("name %s\n", nm.c_str());
    } catch (const std::exception& e) {
        std::printf("error %s\n", e.what());
    }
    return 0;
}
