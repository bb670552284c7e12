// The `blamm dict`, `blamm hist` and `blamm scan` commands.
//
// Command lines, messages, side files and output formats follow the reference
// modules (dict.cpp:53-115, hist.cpp:177-266, pwmscan.cpp:456-648).  The scan
// body differs by design: each manifest group is packed once to 2 bits per
// base and handed to the MI355X engine (include/blamm_hip.h), which returns the
// compacted occurrence list; this file only formats it.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <thread>

#include "../../../include/blamm_hip.h"
#include "host.hpp"
#include "occurrences.hpp"

namespace blamm {

std::vector<Group> build_dict(const std::string& manifest)
{
    std::ifstream in(manifest);
    if (!in) throw std::runtime_error("Could not open file: " + manifest);
    std::vector<Group> groups;
    std::map<std::string, size_t> index;
    std::vector<std::set<std::string>> files;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::istringstream ls(line);
        std::string name, fasta;
        ls >> name >> fasta;
        if (name.empty() || fasta.empty())
            throw std::runtime_error("File " + manifest + " has incorrect format.\n"
                                     "Refer to the documention for more information.");
        if (!std::ifstream(fasta)) throw std::runtime_error("Could not open fasta file: " + fasta);
        auto it = index.find(name);
        if (it == index.end()) {
            it = index.emplace(name, groups.size()).first;
            groups.emplace_back();
            groups.back().name = name;
            files.emplace_back();
        }
        files[it->second].insert(fasta);
    }
    for (size_t g = 0; g < groups.size(); ++g) {
        std::cout << "Computing nucleotide composition for " << groups[g].name << " ...";
        std::cout.flush();
        groups[g].files.assign(files[g].begin(), files[g].end());
        const PackedStream s = pack_fasta(groups[g].files, UINT64_MAX);
        // per packed word (32 bases, A=00 C=01 G=10 T=11): the low and high
        // code bits at each base's even bit position, counted by popcount
        std::array<uint64_t, 4> cnt{{0, 0, 0, 0}};
        for (uint64_t w = 0; w < s.codes.size() && 32 * w < s.n; ++w) {
            const uint64_t nb = std::min<uint64_t>(32, s.n - 32 * w);
            const uint64_t m = nb == 32 ? 0x5555555555555555ull : 0x5555555555555555ull & ((1ull << (2 * nb)) - 1);
            const uint64_t lo = s.codes[w] & m, hi = (s.codes[w] >> 1) & m;
            cnt[0] += (uint64_t)__builtin_popcountll(m & ~(lo | hi));
            cnt[1] += (uint64_t)__builtin_popcountll(lo & ~hi);
            cnt[2] += (uint64_t)__builtin_popcountll(hi & ~lo);
            cnt[3] += (uint64_t)__builtin_popcountll(lo & hi);
        }
        groups[g].counts = cnt;
        groups[g].totSeqLen = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        groups[g].seqNames = s.names;
        std::cout << "\n";
    }
    return groups;
}

namespace {

void dict_usage()
{
    std::cout << "Usage: blamm dict [options] sequences.input\n"
                 "Goal: generate a dictionary for the input sequences\n\n"
                 " [options]\n  -h\t--help\t\tdisplay help message\n\n";
}

void hist_usage()
{
    std::cout << "Usage: blamm hist [options] motifs.input sequences.input\n"
                 "Goal: compute PWM score histograms\n\n"
                 " [options]\n"
                 "  -e\t--empirical\tcompute empirical histograms (on the GPU)\n"
                 "  -h\t--help\t\tdisplay help message\n\n"
                 " [options arg]\n"
                 "  -b\t--numbins\tnumber of bins per histogram [default = 250]\n"
                 "  -t\t--numthreads\tset the number of parallel threads [default = #cores]\n"
                 "  -l\t--length\tlength of sequence data to analyze [default = 10000000]\n"
                 "  -d\t--device\tHIP device index for -e [default = 0]\n\n"
                 " [file_options]\n"
                 "  -H\t--histdir\toutput directory for the histogram file(s) [default = .]\n\n";
}

void scan_usage()
{
    std::cout << "Usage: blamm scan [options] motifs.input sequences.input\n"
                 "Goal: find PWM matches in sequences\n\n"
                 " [options]\n"
                 "  -h\t--help\t\tdisplay help message\n"
                 "  -c\t--cuda\t\taccepted for compatibility; the MI355X engine is always used\n"
                 "  -rc\t--revcompl\talso search the reverse strand for occurrences\n\n"
                 " [options arg]\n"
                 "  -at\t--absthreshold\tset the minimal absolute score for a motif occurrence\n"
                 "  -rt\t--relthreshold\tset the minimal relative score [0..1] for a motif occurrence (default = 0.95)\n"
                 "  -pt\t--pthreshold\tcompute the motif score threshold from p-value [0..1] (default = 1E-4)\n"
                 "  -t\t--numthreads\tset the number of host threads [default = #cores]\n"
                 "  -p\t--precision\tbf16x2 (default), bf16x2lo (lo products on the MFMA too) or fp32\n"
                 "  -d\t--device\tHIP device index [default = 0]\n\n"
                 " [file_options]\n"
                 "  -H\t--histdir\tdirectory where the histogram file(s) are stored [default = .]\n"
                 "  -o\t--output\tfilename for the motif occurrences [default = occurences.txt]\n"
                 "  -T\t--tune\t\tpick the GPU tile shapes by timing a sample of each group;\n"
                 "    \t\t\tresults are cached in the given file and reused\n\n";
}

std::string dir_with_slash(std::string d)
{
    if (d.empty() || d.back() != '/') d.push_back('/');
    return d;
}

// Phase timings on stderr when BLAMM_TIMING is set (diagnostics only).
class PhaseClock {
public:
    PhaseClock() : on_(std::getenv("BLAMM_TIMING") != nullptr), t_(std::chrono::steady_clock::now()) {}
    void tick(const char* what)
    {
        const auto now = std::chrono::steady_clock::now();
        if (on_)
            std::cerr << "[blamm timing] " << what << " "
                      << std::chrono::duration<double, std::milli>(now - t_).count() << " ms" << std::endl;
        t_ = now;
    }

private:
    bool on_;
    std::chrono::steady_clock::time_point t_;
};

struct HipContext {
    blamm_hip_ctx* ctx = nullptr;
    ~HipContext()
    {
        if (ctx) blamm_hip_destroy(ctx);
    }
};

// Opens the device on a thread of its own (the HIP runtime's start-up and the
// library's code object load take ~50 ms) while the caller reads its inputs;
// take() waits for it.  Whatever it opened and nobody took is closed.
struct DeviceOpener {
    std::thread th;
    blamm_hip_ctx* ctx = nullptr;
    int ndev = 0, rc = BLAMM_OK;
    std::string err;
    void start(int device)
    {
        th = std::thread([this, device] {
            blamm_hip_device_count(&ndev);
            if (ndev == 0) return;
            rc = blamm_hip_create(device, &ctx);
            if (rc != BLAMM_OK) err = blamm_hip_last_error();   // the message is per thread
        });
    }
    // the opened context; throws the errors the opening met
    blamm_hip_ctx* take(const char* what)
    {
        if (th.joinable()) th.join();
        if (ndev == 0) throw std::runtime_error("HIP error: no devices found. Aborting...");
        if (rc != BLAMM_OK) throw std::runtime_error(std::string(what) + ": " + err);
        blamm_hip_ctx* c = ctx;
        ctx = nullptr;
        return c;
    }
    ~DeviceOpener()
    {
        if (th.joinable()) th.join();
        if (ctx) blamm_hip_destroy(ctx);
    }
};

// Bases per scan / histogram segment: 2^28, or BLAMM_SEGMENT_BASES rounded
// down to a multiple of 64 (at least 64).  Segments are cut at multiples of
// 64 bases so that their packed words and lowercase words start on word
// boundaries; each carries a halo of lmax - 1 bases.
uint64_t segment_bases()
{
    uint64_t seg = 1ull << 28;
    if (const char* e = std::getenv("BLAMM_SEGMENT_BASES")) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0) seg = std::max<uint64_t>(64, (uint64_t)v & ~63ull);
    }
    return std::min<uint64_t>(seg, 1ull << 30);
}

// Hits one segment may leave in the engine's pinned buffer before the next
// segment is cut shorter (64M hits = 1 GiB).
constexpr uint64_t kSegmentHits = 1ull << 26;

void hip_check(int rc, const char* what)
{
    if (rc != BLAMM_OK) throw std::runtime_error(std::string(what) + ": " + blamm_hip_last_error());
}

}  // namespace

int dict_main(int argc, char** argv)
{
    if (argc < 3) {
        dict_usage();
        return EXIT_FAILURE;
    }
    for (int i = 2; i < argc - 1; ++i) {
        const std::string a(argv[i]);
        dict_usage();
        return (a == "-h" || a == "--help") ? EXIT_SUCCESS : EXIT_FAILURE;
    }
    std::cout << "Welcome to blamm -- dictionary model" << std::endl;
    const std::string manifest(argv[argc - 1]);
    const std::vector<Group> groups = build_dict(manifest);
    write_dict(manifest + ".dict", groups);
    std::cout << "Done!  Dictionary written to " << manifest + ".dict" << std::endl;
    return EXIT_SUCCESS;
}

// Histogram::generateEmpiricalHist (hist.cpp:140-160) for one group: the
// scores of every valid window among the group's first maxLength bases,
// binned on the device (blamm_hip_hist), one histogram per motif written as
// writeGNUPlotFile does (hist.cpp:255-263).
static void empirical_histograms(HipContext& hc, DeviceOpener& opener, const Group& g, const std::vector<Motif>& motifs,
                          float pseudocount, size_t numBins, size_t maxLength, size_t lmax,
                          const std::string& histdir)
{
    const size_t ncols = motifs.size();
    if (ncols == 0) return;
    const size_t ldp = 4 * std::max<size_t>(lmax, 1);
    std::vector<float> P(ldp * ncols, 0.0f), thr(ncols, 0.0f), mins(ncols), maxs(ncols);
    std::vector<int> lengths(ncols);
    for (size_t c = 0; c < ncols; ++c) {
        pfm_to_pwm(motifs[c], g.counts, pseudocount, &P[c * ldp]);
        lengths[c] = (int)motifs[c].size();
        mins[c] = min_score(&P[c * ldp], motifs[c].size());
        maxs[c] = max_score(&P[c * ldp], motifs[c].size());
    }
    if (numBins > (1u << 24)) throw std::runtime_error("number of bins above 2^24");
    const PackedStream s = pack_fasta(g.files, maxLength);
    if (!hc.ctx) hc.ctx = opener.take("device setup");
    hip_check(blamm_hip_set_library(hc.ctx, P.data(), (int)ldp, (int)ncols, lengths.data(), thr.data(),
                                    BLAMM_PREC_BF16X2),
              "library upload");
    std::vector<uint64_t> counts(ncols * numBins, 0);
    // segments with a halo of lmax - 1; windows are counted in the segment
    // where they start
    const uint64_t seg = segment_bases(), halo = lmax - 1;
    for (uint64_t s0 = 0; s0 < s.n; s0 += seg) {
        const uint64_t payload = std::min(seg, s.n - s0);
        const uint64_t len = std::min(s.n - s0, payload + halo);
        std::vector<uint64_t> runs;
        for (size_t r = 0; r < s.runStart.size(); ++r)
            if (s.runStart[r] > s0 && s.runStart[r] < s0 + len) runs.push_back(s.runStart[r] - s0);
        std::vector<uint64_t> codes(blamm_hip_packed_words(len), 0), lower;
        std::memcpy(codes.data(), s.codes.data() + s0 / 32, ((len + 31) / 32) * 8);
        if (s.hasLower()) {
            lower.assign(blamm_hip_mask_words(len), 0);
            std::memcpy(lower.data(), s.lower.data() + s0 / 64, ((len + 63) / 64) * 8);
        }
        hip_check(blamm_hip_hist(hc.ctx, codes.data(), s.hasLower() ? lower.data() : nullptr, len, runs.data(),
                                 runs.size(), payload, mins.data(), maxs.data(), (uint32_t)numBins, counts.data()),
                  "histogram");
    }
    for (size_t c = 0; c < ncols; ++c) {
        Histogram h;
        h.numBins = numBins;
        h.minScore = mins[c];
        h.maxScore = maxs[c];
        h.counts.assign(counts.begin() + (ptrdiff_t)(c * numBins), counts.begin() + (ptrdiff_t)((c + 1) * numBins));
        h.write(histdir, "hist_" + g.name + "_" + motifs[c].name, motifs[c].name + " (" + g.name + ")");
    }
}

int hist_main(int argc, char** argv)
{
    if (argc < 4) {
        hist_usage();
        return EXIT_FAILURE;
    }
    size_t maxLength = 10000000, numBins = 250;
    size_t numThreads = std::max(1u, std::thread::hardware_concurrency());
    bool empirical = false;
    int device = 0;
    std::string histdir;
    for (int i = 2; i < argc - 2; ++i) {
        const std::string a(argv[i]);
        const bool more = i + 1 < argc - 2;
        if (a == "-h" || a == "--help") {
            hist_usage();
            return EXIT_SUCCESS;
        } else if ((a == "-l" || a == "--length") && more) {
            maxLength = (size_t)std::atoll(argv[++i]);
        } else if ((a == "-b" || a == "--numbins") && more) {
            numBins = (size_t)std::atoll(argv[++i]);
            if (numBins < 2) numBins = 2;
        } else if ((a == "-t" || a == "--numthreads") && more) {
            numThreads = (size_t)std::max(1, std::atoi(argv[++i]));
        } else if ((a == "-H" || a == "--histdir") && more) {
            histdir = dir_with_slash(argv[++i]);
        } else if (a == "-e" || a == "--empirical") {
            empirical = true;
        } else if ((a == "-d" || a == "--device") && more) {
            device = std::atoi(argv[++i]);
        } else {
            hist_usage();
            return EXIT_FAILURE;
        }
    }
    // -e: the device opens while the inputs and the first group's bases are
    // read; the histograms take it when they need it
    DeviceOpener opener;
    if (empirical) opener.start(device);
    std::cout << "Welcome to blamm -- histogram module" << std::endl;
    const Settings st = Settings::load();
    const std::vector<Group> groups = read_dict(std::string(argv[argc - 1]) + ".dict");
    const std::vector<Motif> motifs = load_motifs(argv[argc - 2], false);
    size_t lmax = 0;
    for (const Motif& m : motifs) lmax = std::max(lmax, m.size());
    std::cout << "Loaded " << motifs.size() << " motifs from disk";
    std::cout << "\nMaximum motif size: " << lmax << std::endl;
    if (empirical) {
        for (const Motif& m : motifs)
            if (m.size() < 1 || m.size() > BLAMM_MAX_MOTIF_LEN)
                throw std::runtime_error("motif " + m.name + " has length " + std::to_string(m.size()) +
                                         "; this build scores lengths 1.." + std::to_string(BLAMM_MAX_MOTIF_LEN));
        std::cout << "Using 1 GPU device (split-bf16 MFMA)" << std::endl;
    } else {
        std::cout << "Using " << numThreads << " thread(s)" << std::endl;
    }
    HipContext hc;
    for (const Group& g : groups) {
        const std::array<float, 4> bg = g.probabilities(st.pseudocount);
        {
            // Species::writeNuclProbabilities (species.cpp:73-82)
            const auto prec = std::cout.precision(3);
            std::cout << "Generating histograms for species: " << g.name << " [A: " << 100.0f * bg[0]
                      << "%, C: " << 100.0f * bg[1] << "%, G: " << 100.0f * bg[2] << "%, T: " << 100.0f * bg[3]
                      << "%]\n";
            std::cout.precision(prec);
        }
        if (empirical) {
            empirical_histograms(hc, opener, g, motifs, st.pseudocount, numBins, maxLength, lmax, histdir);
            continue;
        }
        // one motif per task on numThreads workers (each writes its own
        // files); the first failure in motif order is the one reported
        std::vector<std::string> errs(motifs.size());
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i; (i = next++) < motifs.size();) {
                const Motif& m = motifs[i];
                try {
                    std::vector<float> w(4 * m.size());
                    pfm_to_pwm(m, g.counts, st.pseudocount, w.data());
                    const Histogram h = theoretical_histogram(w.data(), m.size(), bg, numBins, maxLength);
                    h.write(histdir, "hist_" + g.name + "_" + m.name, m.name + " (" + g.name + ")");
                } catch (const std::exception& e) {
                    errs[i] = e.what();
                }
            }
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < std::min(numThreads, motifs.size()); ++t) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        for (const std::string& e : errs)
            if (!e.empty()) throw std::runtime_error(e);
    }
    return EXIT_SUCCESS;
}

int scan_main(int argc, char** argv)
{
    if (argc < 4) {
        scan_usage();
        return EXIT_FAILURE;
    }
    bool revCompl = false, absSpec = false, relSpec = false, pSpec = false;
    float absTh = 0.0f, relTh = 0.95f, pvalue = 0.0001f;
    size_t numThreads = std::max(1u, std::thread::hardware_concurrency());
    std::string histdir, outName = "occurrences.txt", tuneCache;
    int precision = BLAMM_PREC_BF16X2, device = 0;
    for (int i = 2; i < argc - 2; ++i) {
        const std::string a(argv[i]);
        const bool more = i + 1 < argc - 2;
        if (a == "-h" || a == "--help") {
            scan_usage();
            return EXIT_SUCCESS;
        } else if (a == "-rc" || a == "--revcompl") {
            revCompl = true;
        } else if (a == "-s" || a == "--simple") {
            throw std::runtime_error("the simple scan (-s) is not available in this build");
        } else if (a == "-c" || a == "--cuda") {
        } else if ((a == "-at" || a == "--absthreshold") && more) {
            absSpec = true;
            absTh = (float)std::atof(argv[++i]);
        } else if ((a == "-rt" || a == "--relthreshold") && more) {
            relSpec = true;
            relTh = (float)std::atof(argv[++i]);
            if (relTh < 0.0f || relTh > 1.0f)
                throw std::runtime_error("The relative threshold should be in range [0..1].");
        } else if ((a == "-pt" || a == "--pthreshold") && more) {
            pSpec = true;
            pvalue = (float)std::atof(argv[++i]);
            if (pvalue < 0.0f || pvalue > 1.0f) throw std::runtime_error("The p-value should be in range [0..1].");
        } else if ((a == "-t" || a == "--numthreads") && more) {
            const int t = std::atoi(argv[++i]);
            if (t < 1) throw std::runtime_error("Number of threads must be a non-zero positive number");
            numThreads = (size_t)t;
        } else if ((a == "-H" || a == "--histdir") && more) {
            histdir = dir_with_slash(argv[++i]);
        } else if ((a == "-o" || a == "--output") && more) {
            outName = argv[++i];
        } else if ((a == "-p" || a == "--precision") && more) {
            const std::string p(argv[++i]);
            if (p == "fp32") precision = BLAMM_PREC_FP32;
            else if (p == "bf16x2") precision = BLAMM_PREC_BF16X2;
            else if (p == "bf16x2lo") precision = BLAMM_PREC_BF16X2_LO;
            else throw std::runtime_error("precision must be fp32, bf16x2 or bf16x2lo");
        } else if ((a == "-d" || a == "--device") && more) {
            device = std::atoi(argv[++i]);
        } else if ((a == "-T" || a == "--tune") && more) {
            tuneCache = argv[++i];
        } else {
            scan_usage();
            return EXIT_FAILURE;
        }
    }
    if (!(absSpec || relSpec || pSpec)) relSpec = true;
    if (absSpec && relSpec) throw std::runtime_error("Specify either the absolute or relative threshold, not both.");
    if (absSpec && pSpec) throw std::runtime_error("Specify either the absolute or p-value threshold, not both.");
    if (relSpec && pSpec) throw std::runtime_error("Specify either the relative or p-value threshold, not both.");

    // the device opens while the inputs are read; the scan takes it when the
    // first group is ready, so input errors (dict, motifs, histograms) are
    // still reported first
    DeviceOpener opener;
    opener.start(device);

    std::cout << "Welcome to blamm -- PWM scan module" << std::endl;
    const Settings st = Settings::load();
    st.print();
    const std::string manifest(argv[argc - 1]);
    const std::vector<Group> groups = read_dict(manifest + ".dict");
    std::vector<Motif> motifs = load_motifs(argv[argc - 2], true);
    size_t lmax = 0;
    for (const Motif& m : motifs) lmax = std::max(lmax, m.size());
    std::cout << "Loaded " << motifs.size() << " motifs from disk\n";
    std::cout << "Maximum motif size: " << lmax << std::endl;
    if (revCompl) {
        motifs = add_reverse_complements(motifs);
        std::cout << "Scanning both forward and reverse strand of the input sequence(s)" << std::endl;
    } else {
        std::cout << "Scanning only the forward strand of the input sequence(s)" << std::endl;
    }
    for (const Motif& m : motifs)
        if (m.size() < 1 || m.size() > BLAMM_MAX_MOTIF_LEN)
            throw std::runtime_error("motif " + m.name + " has length " + std::to_string(m.size()) +
                                     "; this build scores lengths 1.." + std::to_string(BLAMM_MAX_MOTIF_LEN));
    if (absSpec) std::cout << "Absolute motif score threshold set to: " << absTh << std::endl;
    else if (relSpec) std::cout << "Relative motif score threshold set to: " << relTh << std::endl;
    else std::cout << "P-value motif score threshold set to: " << pvalue << std::endl;

    HipContext hc;
    auto open_device = [&]() {
        if (hc.ctx) return;
        hc.ctx = opener.take("device setup");
        std::cout << "Using 1 GPU device (" << (precision == BLAMM_PREC_FP32 ? "fp32" : precision == BLAMM_PREC_BF16X2_LO ? "split-bf16, lazy lo" : "split-bf16")
                  << " MFMA) and " << numThreads << " host thread(s)" << std::endl;
    };

    const size_t ncols = motifs.size();
    std::vector<ColumnInfo> cols(ncols);
    std::vector<int> lengths(ncols);
    for (size_t c = 0; c < ncols; ++c) {
        cols[c] = {motifs[c].name, motifs[c].size(), motifs[c].rc ? '-' : '+'};
        lengths[c] = (int)motifs[c].size();
    }
    const size_t ldp = 4 * std::max<size_t>(lmax, 1);

    std::ofstream cutoffs("PWMthresholds.txt");
    OutFile os(outName);
    uint64_t total = 0;
    PhaseClock clock;
    for (const Group& g : groups) {
        const std::array<float, 4> bgp = g.probabilities(st.pseudocount);
        {
            const auto prec = std::cout.precision(3);
            std::cout << "Scanning species: " << g.name << " [A: " << 100.0f * bgp[0] << "%, C: " << 100.0f * bgp[1]
                      << "%, G: " << 100.0f * bgp[2] << "%, T: " << 100.0f * bgp[3] << "%]\n";
            std::cout.precision(prec);
        }
        std::vector<float> P(ldp * ncols, 0.0f), thr(ncols);
        for (size_t c = 0; c < ncols; ++c) pfm_to_pwm(motifs[c], g.counts, st.pseudocount, &P[c * ldp]);
        std::map<std::string, float> pcut;   // one histogram per base name (fwd and rc share it)
        if (!absSpec && !relSpec) {
            // the .dat files load in parallel; the first one (in column order)
            // that fails is the one reported
            std::vector<std::string> names;
            std::set<std::string> seen;
            for (size_t c = 0; c < ncols; ++c)
                if (seen.insert(motifs[c].baseName()).second) names.push_back(motifs[c].baseName());
            std::vector<float> cut(names.size());
            std::vector<std::string> errs(names.size());
            std::atomic<size_t> next{0};
            auto load = [&]() {
                for (size_t i; (i = next++) < names.size();) {
                    try {
                        cut[i] = Histogram::load(histdir + "hist_" + g.name + "_" + names[i] + ".dat").cutoff(pvalue);
                    } catch (const std::exception& e) {
                        errs[i] = e.what();
                    }
                }
            };
            std::vector<std::thread> pool;
            const size_t nt = std::min<size_t>(numThreads, (names.size() + 15) / 16);
            for (size_t t = 1; t < nt; ++t) pool.emplace_back(load);
            load();
            for (auto& th : pool) th.join();
            for (size_t i = 0; i < names.size(); ++i) {
                if (!errs[i].empty()) throw std::runtime_error(errs[i]);
                pcut.emplace(names[i], cut[i]);
            }
        }
        for (size_t c = 0; c < ncols; ++c) {
            const float* w = &P[c * ldp];
            if (absSpec) {
                thr[c] = absTh;
            } else if (relSpec) {
                const float mx = max_score(w, motifs[c].size()), mn = min_score(w, motifs[c].size());
                thr[c] = relTh * (mx - mn) + mn;
            } else {
                thr[c] = pcut.at(motifs[c].baseName());
            }
        }
        for (size_t c = 0; c < ncols; ++c) {
            if (motifs[c].size() > 15) continue;
            const float* w = &P[c * ldp];
            cutoffs << g.name << "\t" << motifs[c].name << "\t" << min_score(w, motifs[c].size()) << "\t" << thr[c]
                    << "\t" << max_score(w, motifs[c].size()) << '\n';
        }
        clock.tick("thresholds");
        const PackedStream s = pack_fasta(g.files, g.totSeqLen);
        clock.tick("read and pack FASTA");
        open_device();
        clock.tick("device setup");
        // A NaN threshold (getScoreCutoff's 0/0 on an empty top bin, e.g. -pt 0)
        // passes no window in the reference (R >= NaN is false, kernel.cu:29;
        // pwmscan.cpp:117-119); +inf does the same on the device.
        std::vector<float> thr_dev(thr);
        for (float& t : thr_dev)
            if (std::isnan(t)) t = INFINITY;
        hip_check(blamm_hip_set_library(hc.ctx, P.data(), (int)ldp, (int)ncols, lengths.data(), thr_dev.data(),
                                        precision),
                  "library upload");
        clock.tick("library upload");
        // Segments of the stream (multiples of 64 bases) with a halo of
        // lmax - 1 bases; hits starting in a halo belong to the next segment.
        // A segment that leaves more than kSegmentHits hits makes the next
        // ones proportionally shorter, which bounds the pinned hit buffer.
        uint64_t seg = segment_bases();
        const uint64_t halo = lmax - 1;
        if (!absSpec && !relSpec && s.n > 0) {
            // -pt thresholds pass about p of a column's windows: room for three
            // times that many records and hits in a segment (at most what a
            // segment may leave), so the first scan usually runs once
            const double est = 3.0 * (double)pvalue * (double)std::min<uint64_t>(seg, s.n) * (double)ncols;
            const double cap = 1.25 * (double)kSegmentHits;
            hip_check(blamm_hip_reserve(hc.ctx, (uint64_t)std::min(est, cap) + (1u << 16)), "scratch reservation");
        }
        if (!tuneCache.empty() && s.n > 0) {
            // tile shapes from a sample of this group (the first 16 Mbp)
            const uint64_t m = std::min<uint64_t>(s.n, 1ull << 24);
            std::vector<uint64_t> runs;
            for (size_t r = 0; r < s.runStart.size(); ++r)
                if (s.runStart[r] > 0 && s.runStart[r] < m) runs.push_back(s.runStart[r]);
            std::vector<uint64_t> codes(blamm_hip_packed_words(m), 0), lower;
            std::memcpy(codes.data(), s.codes.data(), ((m + 31) / 32) * 8);
            if (s.hasLower()) {
                lower.assign(blamm_hip_mask_words(m), 0);
                std::memcpy(lower.data(), s.lower.data(), ((m + 63) / 64) * 8);
            }
            blamm_tuning t{};
            double ms = 0.0;
            hip_check(blamm_hip_autotune_host(hc.ctx, codes.data(), s.hasLower() ? lower.data() : nullptr, m,
                                              runs.data(), runs.size(), tuneCache.c_str(), &t, &ms),
                      "tile-shape tuning");
            std::cout << "GPU tile shapes: " << t.chunk_tiles << " tiles per item, " << t.blocks_per_wave
                      << " column blocks per wave, K padded to " << t.kpad_positions << " positions" << std::endl;
        }
        for (uint64_t s0 = 0; s0 < s.n;) {
            const uint64_t payload = std::min(seg, s.n - s0);
            const uint64_t len = std::min(s.n - s0, payload + halo);
            std::vector<uint64_t> runs;
            for (size_t r = 0; r < s.runStart.size(); ++r)
                if (s.runStart[r] > s0 && s.runStart[r] < s0 + len) runs.push_back(s.runStart[r] - s0);
            std::vector<uint64_t> codes(blamm_hip_packed_words(len), 0), lower;
            std::memcpy(codes.data(), s.codes.data() + s0 / 32, ((len + 31) / 32) * 8);
            if (s.hasLower()) {
                lower.assign(blamm_hip_mask_words(len), 0);
                std::memcpy(lower.data(), s.lower.data() + s0 / 64, ((len + 63) / 64) * 8);
            }
            // hits stay in the engine's pinned buffer until the next scan
            const blamm_hit* hits = nullptr;
            uint64_t nh = 0;
            hip_check(blamm_hip_scan_pinned(hc.ctx, codes.data(), s.hasLower() ? lower.data() : nullptr, len,
                                            runs.data(), runs.size(), &hits, &nh),
                      "scan");
            clock.tick("scan (upload, kernels, hit read-back)");
            total += write_hits(os, hits, nh, payload, s, s0, cols, g, (int)numThreads);
            clock.tick("format and write occurrences");
            s0 += payload;
            if (nh > kSegmentHits && payload == seg)
                seg = std::max<uint64_t>(64, (uint64_t)((double)seg * (double)kSegmentHits / (double)nh) & ~63ull);
        }
    }
    os.close();
    cutoffs.close();
    std::cout << "\nWrote " << total << " matches to " << outName << ".\n";
    return EXIT_SUCCESS;
}

}  // namespace blamm
