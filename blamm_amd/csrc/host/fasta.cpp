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
// per base as it is read, with a run table (start, record, position).
#include <cstdio>
#include <cstring>
#include <memory>

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

PackedStream pack_fasta(const std::vector<std::string>& files, uint64_t maxLen)
{
    Packer pk(maxLen);
    uint64_t nseq = 0, recPos = 0;
    std::unique_ptr<char[]> buf(new char[1 << 22]);
    for (const std::string& fn : files) {
        if (pk.full) break;
        FILE* f = std::fopen(fn.c_str(), "rb");
        if (!f) throw std::runtime_error("Could not open file: " + fn);
        // line state: 0 = at line start, 1 = in header, 2 = in sequence line
        int state = 0, nameState = 0;
        size_t got;
        while (!pk.full && (got = std::fread(buf.get(), 1, 1 << 22, f)) > 0) {
            for (size_t i = 0; i < got && !pk.full; ++i) {
                const uint8_t ch = (uint8_t)buf[i];
                if (ch == '\n') {
                    state = 0;
                    continue;
                }
                if (state == 0) {
                    if (ch == '>') {
                        state = 1;
                        nameState = 0;
                        ++nseq;
                        recPos = 0;
                        pk.s.names.emplace_back();
                        continue;
                    }
                    if (nseq == 0) {
                        std::fclose(f);
                        throw std::runtime_error("Input file does not appear to be in fasta format\n");
                    }
                    state = 2;
                }
                if (state == 1) {
                    // record name: first whitespace-delimited token after '>'
                    const bool ws = ch == ' ' || ch == '\t' || ch == '\v' || ch == '\f' || ch == '\r';
                    if (nameState == 0 && !ws) nameState = 1;
                    else if (nameState == 1 && ws) nameState = 2;
                    if (nameState == 1) pk.s.names.back().push_back((char)ch);
                    continue;
                }
                const uint8_t t = kTable.t[ch];
                if (t & 4) pk.append(t, nseq - 1, recPos);
                ++recPos;
            }
        }
        std::fclose(f);
    }
    pk.finish();
    pk.s.nseq = nseq;
    return std::move(pk.s);
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
