// FASTA -> packed filtered stream.
//
// Semantics follow the reference reader (sequence.cpp:142-293):
//  - files are read in the given order as one batch; lines split on '\n'
//    (a '\r' stays in the line and is an invalid byte);
//  - empty lines are skipped; a line starting with '>' starts a new record
//    (position reset to 0);
//  - every other line advances the record position by its length; only
//    ACGTacgt are kept, any other byte ends a run;
//  - sequence data before the first record is an error;
//  - the stream stops at maxLen bases (the dict's TOT_SEQ_LENGTH).
// Instead of strings and per-block marker maps, the stream is packed to 2 bits
// per base, with a run table (start, record, position), by all host threads:
// see pack_fasta.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>

#include "host.hpp"

namespace blamm {

namespace {

// byte -> code | flags: bit 2 = valid base, bit 3 = lowercase
struct ByteTable {
    uint8_t t[256];
    ByteTable()
    {
        std::memset(t, 0, sizeof(t));
        const char* up = "ACGT";
        const char* lo = "acgt";
        for (int i = 0; i < 4; ++i) {
            t[(uint8_t)up[i]] = (uint8_t)(i | 4);
            t[(uint8_t)lo[i]] = (uint8_t)(i | 4 | 8);
        }
    }
};
const ByteTable kTable;

struct Packer {
    PackedStream s;
    uint64_t maxLen;
    bool full;
    bool havePrev = false;      // a base was appended before
    uint64_t prevSeq = 0, prevPos = 0;

    explicit Packer(uint64_t m) : maxLen(m), full(m == 0) {}

    void append(uint8_t code, uint64_t seq, uint64_t pos)
    {
        if (!(havePrev && prevSeq == seq && prevPos + 1 == pos)) {
            s.runStart.push_back(s.n);
            s.runSeq.push_back(seq);
            s.runPos.push_back(pos);
        }
        havePrev = true;
        prevSeq = seq;
        prevPos = pos;
        const uint64_t i = s.n;
        if ((i & 31) == 0) s.codes.push_back(0);
        s.codes.back() |= (uint64_t)(code & 3) << (2 * (i & 31));
        if (code & 8) {
            const size_t need = (size_t)(i >> 6) + 1;
            if (s.lower.size() < need) s.lower.resize(need, 0);
            s.lower[i >> 6] |= 1ull << (i & 63);
        }
        ++s.n;
        if (s.n >= maxLen) full = true;
    }

    void finish()
    {
        if (!s.lower.empty()) s.lower.resize((size_t)((s.n + 63) / 64), 0);
    }
};

}  // namespace

void PackedStream::locate(uint64_t q, uint64_t& seq, uint64_t& pos) const
{
    // last run start <= q
    size_t lo = 0, hi = runStart.size();
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (runStart[mid] <= q) lo = mid + 1; else hi = mid;
    }
    const size_t r = lo - 1;
    seq = runSeq[r];
    pos = runPos[r] + (q - runStart[r]);
}

uint64_t PackedStream::nextBreak(uint64_t q) const
{
    size_t lo = 0, hi = runStart.size();
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (runStart[mid] <= q) lo = mid + 1; else hi = mid;
    }
    return lo < runStart.size() ? runStart[lo] : n;
}

// Smallest chunk a file is cut into for the parallel passes (tests build the
// reader with tiny chunks to put chunk boundaries everywhere).
#ifndef BLAMM_FASTA_MIN_CHUNK
#define BLAMM_FASTA_MIN_CHUNK (1 << 20)
#endif
#ifndef BLAMM_FASTA_SEGMENT
#define BLAMM_FASTA_SEGMENT (256u << 20)
#endif

namespace {

// Bytes of a segment: a grow-only buffer that is not zero-filled (the reads
// write every byte before it is used), reused from segment to segment.
struct SegmentBytes {
    std::unique_ptr<char[]> p;
    size_t n = 0, cap = 0;
    const char* data() const { return p.get(); }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
    char operator[](size_t i) const { return p[i]; }
    // room for `want` bytes, keeping the first n
    void reserve(size_t want)
    {
        if (want <= cap) return;
        const size_t c = std::max(want, cap + cap / 2);
        std::unique_ptr<char[]> q(new char[c]);
        if (n) std::memcpy(q.get(), p.get(), n);
        p = std::move(q);
        cap = c;
    }
};

// A segment of a file (whole lines) cut into line-aligned chunks.
struct FileData {
    SegmentBytes bytes;
    std::vector<std::pair<size_t, size_t>> chunks;   // [begin, end) at line starts
};

// A file read in line-aligned segments of about BLAMM_FASTA_SEGMENT bytes
// (each read with parallel pread), so reading can stop once the stream is full
// and memory stays bounded on large genomes.
class SegmentReader {
public:
    SegmentReader(const std::string& fn, unsigned threads) : fn_(fn), threads_(threads)
    {
        fd_ = ::open(fn.c_str(), O_RDONLY);
        if (fd_ < 0) throw std::runtime_error("Could not open file: " + fn);
        struct stat st;
        if (::fstat(fd_, &st) != 0) throw std::runtime_error("Could not open file: " + fn);
        size_ = (size_t)st.st_size;
    }
    ~SegmentReader()
    {
        if (fd_ >= 0) ::close(fd_);
    }
    SegmentReader(const SegmentReader&) = delete;
    SegmentReader& operator=(const SegmentReader&) = delete;

    // The next segment: whole lines (the file's last line may lack its '\n');
    // false when the file is exhausted.
    bool next(FileData& d)
    {
        d.chunks.clear();
        // the previous segment's unfinished last line comes first
        d.bytes.n = 0;
        d.bytes.reserve(carry_.size());
        if (!carry_.empty()) std::memcpy(d.bytes.p.get(), carry_.data(), carry_.size());
        d.bytes.n = carry_.size();
        carry_.clear();
        if (d.bytes.empty() && off_ >= size_) return false;
        size_t cut = std::string::npos;
        while (off_ < size_) {
            const size_t want = std::min<size_t>(BLAMM_FASTA_SEGMENT, size_ - off_);
            const size_t base = d.bytes.n;
            d.bytes.reserve(base + want);
            read_at(d.bytes.p.get() + base, want, off_);
            d.bytes.n = base + want;
            off_ += want;
            // last line end among the new bytes
            const char* q = d.bytes.p.get();
            size_t nl = base + want;
            while (nl > base && q[nl - 1] != '\n') --nl;
            if (nl > base) {
                cut = nl;
                break;
            }
            // no line end in the new bytes: a line longer than a segment
        }
        if (off_ < size_ && cut != std::string::npos) {
            carry_.assign(d.bytes.data() + cut, d.bytes.n - cut);
            d.bytes.n = cut;
        }
        const size_t size = d.bytes.size();
        const size_t chunk = std::max<size_t>(size / (4 * threads_) + 1, BLAMM_FASTA_MIN_CHUNK);
        size_t b = 0;
        while (b < size) {
            size_t e = std::min(size, b + chunk);
            while (e < size && d.bytes[e - 1] != '\n') ++e;
            d.chunks.emplace_back(b, e);
            b = e;
        }
        return true;
    }

private:
    void read_at(char* dst, size_t n, size_t at)
    {
        const size_t piece = std::max<size_t>(n / threads_ + 1, 1 << 20);
        const size_t np = (n + piece - 1) / piece;
        std::vector<int> bad(np, 0);
        std::vector<std::thread> pool;
        for (size_t k = 0; k < np; ++k)
            pool.emplace_back([&, k]() {
                size_t o = k * piece;
                const size_t end = std::min(n, o + piece);
                while (o < end) {
                    const ssize_t r = ::pread(fd_, dst + o, end - o, (off_t)(at + o));
                    if (r <= 0) {
                        if (r < 0 && errno == EINTR) continue;
                        bad[k] = 1;
                        return;
                    }
                    o += (size_t)r;
                }
            });
        for (auto& t : pool) t.join();
        for (int x : bad)
            if (x) throw std::runtime_error("Could not read file: " + fn_);
    }

    std::string fn_;
    unsigned threads_;
    int fd_ = -1;
    size_t size_ = 0, off_ = 0;
    std::string carry_;
};

// What a chunk contributes, independent of what came before it.
struct ChunkSummary {
    uint64_t nrec = 0;            // '>' lines
    bool seqBeforeRec = false;    // a sequence line before its first '>' line
    uint64_t tailPos = 0;         // record positions advanced after its last '>' (all, if none)
    uint64_t nvalid = 0;          // ACGTacgt bytes
    bool lower = false;
    // the last valid base: record ('>' lines before it in the chunk) and
    // position (relative to the chunk's starting position when rec == 0)
    bool hasValid = false;
    uint64_t lastRec = 0, lastPos = 0;
};

ChunkSummary summarize(const char* p, size_t n)
{
    ChunkSummary c;
    uint64_t pos = 0;        // position in the current record (relative if no '>' yet)
    size_t i = 0;
    while (i < n) {
        // one line [i, e)
        const char* nl = (const char*)std::memchr(p + i, '\n', n - i);
        const size_t e = nl ? (size_t)(nl - p) : n;
        if (e > i) {
            if (p[i] == '>') {
                ++c.nrec;
                pos = 0;
            } else {
                if (c.nrec == 0) c.seqBeforeRec = true;
                for (size_t k = i; k < e; ++k) {
                    const uint8_t t = kTable.t[(uint8_t)p[k]];
                    if (t & 4) {
                        ++c.nvalid;
                        c.lower |= (t & 8) != 0;
                        c.hasValid = true;
                        c.lastRec = c.nrec;
                        c.lastPos = pos;
                    }
                    ++pos;
                }
            }
        }
        i = e + 1;
    }
    c.tailPos = pos;
    return c;
}

// State at a chunk start.
struct ChunkStart {
    uint64_t nseq = 0, recPos = 0, n0 = 0;
    bool havePrev = false;
    uint64_t prevSeq = 0, prevPos = 0;
};

// Output of one chunk's packing pass (runs and names in order).
struct ChunkOut {
    std::vector<uint64_t> runStart, runSeq, runPos;
    std::vector<std::string> names;
    uint64_t nrec = 0;      // '>' lines before the chunk's stop point
    uint64_t nbases = 0;    // bases appended
};

inline void or_word(uint64_t* w, uint64_t v)
{
    if (v) __atomic_fetch_or(w, v, __ATOMIC_RELAXED);
}

// Pack the chunk's valid bases at stream offsets st.n0 ...; stop after the
// base with stream index limit - 1.
ChunkOut pack_chunk(const char* p, size_t n, const ChunkStart& st, uint64_t limit, uint64_t* codes,
                    uint64_t* lower)
{
    ChunkOut o;
    uint64_t nseq = st.nseq, recPos = st.recPos, q = st.n0;
    bool havePrev = st.havePrev;
    uint64_t prevSeq = st.prevSeq, prevPos = st.prevPos;
    uint64_t word = 0, lword = 0;
    uint64_t wi = q >> 5, li = q >> 6;
    size_t i = 0;
    bool stop = q >= limit;
    while (i < n && !stop) {
        const char* nl = (const char*)std::memchr(p + i, '\n', n - i);
        const size_t e = nl ? (size_t)(nl - p) : n;
        if (e > i) {
            if (p[i] == '>') {
                ++nseq;
                ++o.nrec;
                recPos = 0;
                // record name: first whitespace-delimited token after '>'
                size_t k = i + 1;
                auto ws = [](char ch) { return ch == ' ' || ch == '\t' || ch == '\v' || ch == '\f' || ch == '\r'; };
                while (k < e && ws(p[k])) ++k;
                size_t m = k;
                while (m < e && !ws(p[m])) ++m;
                o.names.emplace_back(p + k, m - k);
            } else {
                for (size_t k = i; k < e; ++k) {
                    const uint8_t t = kTable.t[(uint8_t)p[k]];
                    if (t & 4) {
                        const uint64_t seq = nseq - 1;
                        if (!(havePrev && prevSeq == seq && prevPos + 1 == recPos)) {
                            o.runStart.push_back(q);
                            o.runSeq.push_back(seq);
                            o.runPos.push_back(recPos);
                        }
                        havePrev = true;
                        prevSeq = seq;
                        prevPos = recPos;
                        if ((q >> 5) != wi) {
                            or_word(&codes[wi], word);
                            word = 0;
                            wi = q >> 5;
                        }
                        word |= (uint64_t)(t & 3) << (2 * (q & 31));
                        if (lower) {
                            if ((q >> 6) != li) {
                                or_word(&lower[li], lword);
                                lword = 0;
                                li = q >> 6;
                            }
                            if (t & 8) lword |= 1ull << (q & 63);
                        }
                        ++q;
                        ++o.nbases;
                        if (q >= limit) {
                            stop = true;
                            break;
                        }
                    }
                    ++recPos;
                }
            }
        }
        i = e + 1;
    }
    if (o.nbases) {
        or_word(&codes[wi], word);
        if (lower) or_word(&lower[li], lword);
    }
    return o;
}

}  // namespace

// Files are read whole and packed in parallel: chunk summaries (records,
// positions, valid bases), a prefix over them, then every chunk packs its bases
// at their final stream offsets.  The result is the serial reader's, including
// where it stops: at maxLen bases, before any later record or file.
PackedStream pack_fasta(const std::vector<std::string>& files, uint64_t maxLen)
{
    const unsigned threads = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    PackedStream s;
    ChunkStart cur;           // reader state at the start of the next chunk
    std::vector<uint64_t> codes, lower;
    bool anyLower = false;
    auto parallel = [&](size_t count, const std::function<void(size_t)>& fn) {
        std::vector<std::thread> pool;
        std::atomic<size_t> next{0};
        for (unsigned t = 0; t < std::min<size_t>(threads, count); ++t)
            pool.emplace_back([&]() {
                for (size_t k; (k = next.fetch_add(1)) < count;) fn(k);
            });
        for (auto& t : pool) t.join();
    };
    bool done = false;
    for (const std::string& fn : files) {
        if (done || cur.n0 >= maxLen) break;   // the serial reader stops before opening the file
        SegmentReader rd(fn, threads);
        FileData d;
        while (!done && rd.next(d)) {
            const size_t nc = d.chunks.size();
            std::vector<ChunkSummary> sum(nc);
            parallel(nc, [&](size_t k) {
                sum[k] = summarize(d.bytes.data() + d.chunks[k].first, d.chunks[k].second - d.chunks[k].first);
            });
            // each chunk's starting state; `used` chunks are read, the last of
            // them only up to the maxLen-th base when the stream fills inside it
            std::vector<ChunkStart> start(nc);
            size_t used = 0;
            bool filled = false;
            ChunkStart c = cur;
            for (size_t k = 0; k < nc; ++k) {
                start[k] = c;
                used = k + 1;
                if (sum[k].seqBeforeRec && c.nseq == 0)
                    throw std::runtime_error("Input file does not appear to be in fasta format\n");
                const uint64_t take = std::min<uint64_t>(sum[k].nvalid, maxLen - c.n0);
                if (sum[k].lower && take > 0) anyLower = true;
                if (c.n0 + take >= maxLen) {
                    c.n0 += take;
                    filled = true;
                    break;
                }
                if (sum[k].hasValid) {
                    c.havePrev = true;
                    c.prevSeq = c.nseq + sum[k].lastRec - 1;
                    c.prevPos = sum[k].lastRec ? sum[k].lastPos : c.recPos + sum[k].lastPos;
                }
                c.recPos = sum[k].nrec ? sum[k].tailPos : c.recPos + sum[k].tailPos;
                c.nseq += sum[k].nrec;
                c.n0 += take;
            }
            codes.resize((size_t)((c.n0 + 31) / 32), 0);
            if (anyLower) lower.resize((size_t)((c.n0 + 63) / 64), 0);
            std::vector<ChunkOut> out(used);
            parallel(used, [&](size_t k) {
                out[k] = pack_chunk(d.bytes.data() + d.chunks[k].first, d.chunks[k].second - d.chunks[k].first,
                                    start[k], maxLen, codes.data(), anyLower ? lower.data() : nullptr);
            });
            for (size_t k = 0; k < used; ++k) {
                s.runStart.insert(s.runStart.end(), out[k].runStart.begin(), out[k].runStart.end());
                s.runSeq.insert(s.runSeq.end(), out[k].runSeq.begin(), out[k].runSeq.end());
                s.runPos.insert(s.runPos.end(), out[k].runPos.begin(), out[k].runPos.end());
                for (std::string& nm : out[k].names) s.names.push_back(std::move(nm));
            }
            if (filled) {
                // records are counted up to the stop point only
                cur = start[used - 1];
                cur.nseq += out[used - 1].nrec;
                cur.n0 += out[used - 1].nbases;
                done = true;
                break;
            }
            cur = c;
        }
    }
    s.n = cur.n0;
    s.nseq = cur.nseq;
    s.codes = std::move(codes);
    s.codes.resize((size_t)((s.n + 31) / 32), 0);
    if (anyLower) {
        s.lower = std::move(lower);
        s.lower.resize((size_t)((s.n + 63) / 64), 0);
    }
    return s;
}

PackedStream pack_bases(const char* bases, uint64_t n)
{
    Packer pk(n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t t = kTable.t[(uint8_t)bases[i]];
        if (!(t & 4)) throw std::runtime_error("pack_bases: byte outside ACGTacgt");
        pk.append(t, 0, i);
    }
    pk.finish();
    pk.s.nseq = n ? 1 : 0;
    return std::move(pk.s);
}

}  // namespace blamm
