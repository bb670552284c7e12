// Occurrence formatting and writing (pwmscan.cpp:81-102).
#include "occurrences.hpp"

#include <algorithm>
#include <atomic>
#include <charconv>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>

#include "format_g6.hpp"

namespace blamm {

namespace {

inline char* put_text(char* p, const std::string& t)
{
    std::memcpy(p, t.data(), t.size());
    return p + t.size();
}

// Stream position -> (record, position in record) as PackedStream::locate
// computes it, with the binary search narrowed by a table of the last run
// starting at or before each 1024-base bucket (hits arrive in no particular
// order, and a search over all runs cost more than formatting the line).
class Locator {
public:
    explicit Locator(const PackedStream& s) : s_(s)
    {
        const size_t nb = (size_t)(s.n >> kShift) + 2;
        first_.assign(nb, 0);
        size_t r = 0;
        for (size_t b = 0; b < nb; ++b) {
            const uint64_t at = (uint64_t)b << kShift;
            while (r + 1 < s.runStart.size() && s.runStart[r + 1] <= at) ++r;
            first_[b] = (uint32_t)r;
        }
    }
    void locate(uint64_t q, uint64_t& seq, uint64_t& pos) const
    {
        const size_t b = (size_t)(q >> kShift);
        if (s_.runStart.empty() || b + 1 >= first_.size() || s_.runStart.size() > UINT32_MAX) {
            s_.locate(q, seq, pos);
            return;
        }
        // the run holding q lies in [first_[b], first_[b + 1]]
        size_t lo = first_[b], hi = (size_t)first_[b + 1] + 1;
        while (hi - lo > 1) {
            const size_t mid = (lo + hi) / 2;
            if (s_.runStart[mid] <= q) lo = mid; else hi = mid;
        }
        seq = s_.runSeq[lo];
        pos = s_.runPos[lo] + (q - s_.runStart[lo]);
    }

private:
    static constexpr int kShift = 10;
    const PackedStream& s_;
    std::vector<uint32_t> first_;
};

}  // namespace

// One output line per hit whose position is below `payload`: seqName  blamm
// motif  pos  pos+L  score  strand  .  .  (pwmscan.cpp:88-95), written in hit
// order.  The hits are cut into chunks that `threads` workers take in order;
// a chunk is written (pwrite) as soon as every chunk before it has been
// formatted, so writing overlaps the formatting of later chunks.
uint64_t write_hits(OutFile& os, const blamm_hit* hits, size_t n, uint64_t payload, const PackedStream& s,
                    uint64_t base, const std::vector<ColumnInfo>& cols, const Group& g, int threads)
{
    if (n == 0) return 0;
    threads = (int)std::max<size_t>(1, std::min<size_t>((size_t)threads, n / 4096 + 1));
    // the fixed text around each column's fields
    std::vector<std::string> pre(cols.size()), post(cols.size());
    size_t longest = 0;
    for (size_t c = 0; c < cols.size(); ++c) {
        pre[c] = "\tblamm\t" + cols[c].name + "\t";
        post[c] = std::string("\t") + cols[c].strand + "\t.\t.\n";
        longest = std::max(longest, pre[c].size() + post[c].size());
    }
    size_t name_max = 0;
    for (const std::string& nm : g.seqNames) name_max = std::max(name_max, nm.size());
    // two positions, two tabs and a score: at most 2 x 20 + 2 + 32 characters
    const size_t line_max = name_max + longest + 74;
    const Locator loc(s);

    struct Chunk {
        std::unique_ptr<char[]> buf;
        size_t len = 0;
        uint64_t at = 0;
        bool done = false;
    };
    const size_t nchunk = std::min<size_t>(n / 4096 + 1, (size_t)threads * 8);
    std::vector<Chunk> ch(nchunk);
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> kept{0};
    std::mutex mu;
    std::condition_variable cv;
    size_t placed = 0;                 // chunks [0, placed) have offsets
    uint64_t end = os.offset();
    std::string err;
    std::vector<std::string> werr((size_t)threads);
    const bool direct = os.seekable();

    auto format = [&](size_t c) {
        const size_t b = n * c / nchunk, e = n * (c + 1) / nchunk;
        ch[c].buf.reset(new char[(e - b) * line_max + 1]);
        char* const o0 = ch[c].buf.get();
        char* o = o0;
        uint64_t k = 0;
        for (size_t i = b; i < e; ++i) {
            const blamm_hit& h = hits[i];
            if (h.pos >= payload) continue;   // starts in the halo: the next segment reports it
            uint64_t seq, pos;
            loc.locate(base + h.pos, seq, pos);
            if (seq >= g.seqNames.size()) {
                std::lock_guard<std::mutex> lk(mu);
                err = "sequence index beyond the dictionary of group " + g.name + " (is " + g.name +
                      "'s .dict stale?)";
                break;
            }
            ++k;
            o = put_text(o, g.seqNames[seq]);
            o = put_text(o, pre[h.col]);
            o = std::to_chars(o, o + 20, pos).ptr;
            *o++ = '\t';
            o = std::to_chars(o, o + 20, pos + cols[h.col].len).ptr;
            *o++ = '\t';
            o = put_g6(o, h.score);
            o = put_text(o, post[h.col]);
        }
        kept += k;
        ch[c].len = (size_t)(o - o0);
    };
    auto work = [&](int t) {
        try {
            for (;;) {
                const size_t c = next++;
                if (c >= nchunk) break;
                try {
                    format(c);
                } catch (const std::exception& e) {   // still mark the chunk, or later ones would wait forever
                    werr[(size_t)t] = e.what();
                    ch[c].len = 0;
                }
                std::unique_lock<std::mutex> lk(mu);
                ch[c].done = true;
                while (placed < nchunk && ch[placed].done) {
                    ch[placed].at = end;
                    end += ch[placed].len;
                    ++placed;
                }
                cv.notify_all();
                if (!direct) continue;
                cv.wait(lk, [&] { return placed > c; });
                lk.unlock();
                os.write_at(ch[c].buf.get(), ch[c].len, ch[c].at);
                ch[c].buf.reset();
            }
        } catch (const std::exception& e) {
            werr[(size_t)t] = e.what();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    if (!err.empty()) throw std::runtime_error(err);
    for (const std::string& e : werr)
        if (!e.empty()) throw std::runtime_error(e);
    if (!direct)   // a pipe: in order, one chunk after the other
        for (const Chunk& c : ch) os.write_next(c.buf.get(), c.len);
    if (direct) os.advance(end - os.offset());
    return kept.load();
}

}  // namespace blamm
