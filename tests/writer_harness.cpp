// Test driver for the occurrence writer (write_hits + OutFile): formats the
// hits of a binary file against the FASTA file's records and writes them with
// the given number of worker threads.
// Usage: writer_harness file.fa hits.bin payload threads out.txt
// hits.bin: records of {u64 pos, u32 col, f32 score}; columns c are named
// "M<c>", have length 5 + c % 7 and strand '+' for even c, '-' for odd c.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../blamm_amd/csrc/host/occurrences.hpp"

int main(int argc, char** argv)
{
    if (argc != 6) return 2;
    try {
        const blamm::PackedStream s = blamm::pack_fasta({argv[1]}, UINT64_MAX);
        std::vector<blamm_hit> hits;
        if (FILE* f = std::fopen(argv[2], "rb")) {
            blamm_hit h;
            while (std::fread(&h, sizeof(h), 1, f) == 1) hits.push_back(h);
            std::fclose(f);
        }
        std::vector<blamm::ColumnInfo> cols;
        for (int c = 0; c < 64; ++c) cols.push_back({"M" + std::to_string(c), (uint64_t)(5 + c % 7), c % 2 ? '-' : '+'});
        blamm::Group g;
        g.name = "g";
        g.seqNames = s.names;
        blamm::OutFile os(argv[5]);
        const uint64_t n = blamm::write_hits(os, hits.data(), hits.size(), std::strtoull(argv[3], nullptr, 10), s, 0,
                                             cols, g, std::atoi(argv[4]));
        os.close();
        std::printf("%llu\n", (unsigned long long)n);
    } catch (const std::exception& e) {
        std::printf("error %s\n", e.what());
        return 1;
    }
    return 0;
}
