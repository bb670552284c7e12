// Host side of `blamm` on MI355X: inputs (settings.cnf, .dict, motif files,
// hist_*.dat, FASTA) and outputs (occurrences, PWMthresholds.txt, .dict,
// hist_*.dat) with the reference's exact semantics, and the scan driver that
// feeds libblamm_hip.so.  Built with g++ and -ffp-contract=off so every float
// expression rounds like the reference's separate multiply/add.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace blamm {

// ---- settings.cnf (settings.cpp:34-74) ------------------------------------
struct Settings {
    size_t S_w = 250;
    size_t S_h = 1000;
    size_t minZeroArea = 64 * 64;
    size_t flushOutput = 100000;
    float pseudocount = 0.25f;
    bool fromFile = false;
    static Settings load(const std::string& path = "settings.cnf");
    void print() const;   // Settings::printSettings (settings.cpp:76-88)
};

// ---- .dict (species.cpp:84-143, 176-207) -----------------------------------
struct Group {
    std::string name;
    std::vector<std::string> files;          // std::set order: sorted, unique
    uint64_t totSeqLen = 0;
    std::array<uint64_t, 4> counts{{0, 0, 0, 0}};
    std::vector<std::string> seqNames;
    // Species::getNuclProbabilities (species.cpp:63-71)
    std::array<float, 4> probabilities(float pseudocount) const;
};
std::vector<Group> read_dict(const std::string& path);
void write_dict(const std::string& path, const std::vector<Group>& groups);
// `blamm dict` (dict.cpp:53-115): manifest -> groups with counts and names.
std::vector<Group> build_dict(const std::string& manifest);

// ---- motifs (motif.cpp:194-286, 323-449) -----------------------------------
struct Motif {
    std::string name;
    std::vector<std::array<uint64_t, 4>> pfm;
    bool rc = false;
    size_t size() const { return pfm.size(); }
    std::string baseName() const;      // motif.h:256-267
    bool isPermutation() const;        // motif.h:272-281
};
// MotifContainer::load: jaspar if the name ends in ".jaspar", else
// cluster-buster; sorted by length with std::sort (motif.cpp:365, 436).
std::vector<Motif> load_motifs(const std::string& path, bool withPermutations);
// addReverseComplements (motif.cpp:439-449): fwd at 2i, rc at 2i+1.
std::vector<Motif> add_reverse_complements(const std::vector<Motif>& m);
// Motif::PFM2PWM (motif.cpp:194-223); out has 4*L floats.
void pfm_to_pwm(const Motif& m, const std::array<uint64_t, 4>& bg, float pseudocount, float* out);
float max_score(const float* pwm, size_t L);   // motif.cpp:241-252
float min_score(const float* pwm, size_t L);   // motif.cpp:254-265

// ---- score histograms (motif.h:36-146, motif.cpp:47-132, 151-192) ----------
struct Histogram {
    size_t numBins = 0;
    float minScore = 0.f, maxScore = 0.f;
    std::vector<uint64_t> counts;
    float cutoff(float pvalue) const;                       // getScoreCutoff
    static Histogram load(const std::string& path);         // loadHistogram
    void write(const std::string& dir, const std::string& base,
               const std::string& label) const;             // writeGNUPlotFile
};
// computeTheoreticalSpectrum + generateTheoreticalHist (hist.cpp:162-175)
Histogram theoretical_histogram(const float* pwm, size_t L, const std::array<float, 4>& bg,
                                size_t numBins, size_t maxLength);
// The histogram as the scan reads it back from a .dat (min/max at 6
// significant digits, motif.cpp:77 then 120).
Histogram as_written(const Histogram& h);

// ---- packed sequence stream (sequence.cpp:142-293) -------------------------
// The filtered stream of one group: valid bases only, 2-bit codes, lowercase
// bits, and the runs that map stream positions back to (record, position).
struct PackedStream {
    uint64_t n = 0;
    std::vector<uint64_t> codes;       // 32 bases per word
    std::vector<uint64_t> lower;       // 64 bases per word; empty if no lowercase
    std::vector<uint64_t> runStart;    // stream position where each run starts
    std::vector<uint64_t> runSeq;      // record index of the run
    std::vector<uint64_t> runPos;      // position of its first base in the record
    uint64_t nseq = 0;                 // '>' records seen
    std::vector<std::string> names;    // their names (first token after '>')
    bool hasLower() const { return !lower.empty(); }
    // stream position -> (record index, position in record)
    void locate(uint64_t q, uint64_t& seq, uint64_t& pos) const;
    // first run start > q, or n (SeqBlock::getRemainingSeqLen)
    uint64_t nextBreak(uint64_t q) const;
};
PackedStream pack_fasta(const std::vector<std::string>& files, uint64_t maxLen);
// Pack an in-memory filtered stream (ACGTacgt only).
PackedStream pack_bases(const char* bases, uint64_t n);

// ---- the scan command (pwmscan.cpp:456-648) --------------------------------
int scan_main(int argc, char** argv);
int dict_main(int argc, char** argv);
int hist_main(int argc, char** argv);

}  // namespace blamm
