// Occurrence formatting and writing (pwmscan.cpp:81-102).
#include "occurrences.hpp"

#include <algorithm>
#include <charconv>

#include "format_g6.hpp"

namespace blamm {

namespace {

// Append an unsigned integer.
inline void put_u64(std::string& s, uint64_t v)
{
    char b[24];
    auto r = std::to_chars(b, b + sizeof(b), v);
    s.append(b, r.ptr);
}

// Append a float as ostream's default formatting prints it (%g, 6 digits).
inline void put_score(std::string& s, float f)
{
    char b[32];
    s.append(b, put_g6(b, f));
}

}  // namespace

// One output line per hit whose position is below `payload`: seqName  blamm
// motif  pos  pos+L  score  strand  .  .  (pwmscan.cpp:88-95); formatted by
// `threads` workers, written in hit order.
uint64_t write_hits(OutFile& os, const blamm_hit* hits, size_t n, uint64_t payload, const PackedStream& s,
                    uint64_t base, const std::vector<ColumnInfo>& cols, const Group& g, int threads)
{
    if (n == 0) return 0;
    threads = (int)std::max<size_t>(1, std::min<size_t>((size_t)threads, n / 4096 + 1));
    std::vector<std::string> out((size_t)threads);
    std::vector<std::string> err((size_t)threads);
    std::vector<uint64_t> written((size_t)threads, 0);
    auto work = [&](int t) {
        const size_t b = n * (size_t)t / (size_t)threads, e = n * (size_t)(t + 1) / (size_t)threads;
        // a local string: the size fields of neighbouring strings in `out`
        // would share cache lines between workers
        std::string o;
        o.reserve((e - b) * 64);
        uint64_t kept = 0;
        for (size_t i = b; i < e; ++i) {
            const blamm_hit& h = hits[i];
            if (h.pos >= payload) continue;   // starts in the halo: the next segment reports it
            ++kept;
            uint64_t seq, pos;
            s.locate(base + h.pos, seq, pos);
            if (seq >= g.seqNames.size()) {
                err[(size_t)t] = "sequence index beyond the dictionary of group " + g.name +
                                 " (is " + g.name + "'s .dict stale?)";
                return;
            }
            const ColumnInfo& c = cols[h.col];
            o += g.seqNames[seq];
            o += "\tblamm\t";
            o += c.name;
            o += '\t';
            put_u64(o, pos);
            o += '\t';
            put_u64(o, pos + c.len);
            o += '\t';
            put_score(o, h.score);
            o += '\t';
            o += c.strand;
            o += "\t.\t.\n";
        }
        written[(size_t)t] = kept;
        out[(size_t)t] = std::move(o);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (const std::string& e : err)
        if (!e.empty()) throw std::runtime_error(e);
    os.write_parts(out);
    uint64_t total = 0;
    for (uint64_t w : written) total += w;
    return total;
}


}  // namespace blamm
