// Input/output formats of blamm with the reference's parsing and rounding.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <numeric>
#include <set>
#include <sstream>

#include "host.hpp"

namespace blamm {

// ---------------------------------------------------------------------------
// settings.cnf: "KEY value" lines, '#' comments, read from the working
// directory; unknown keys only warn (settings.cpp:34-74).
Settings Settings::load(const std::string& path)
{
    Settings s;
    std::ifstream in(path);
    if (!in) return s;
    s.fromFile = true;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "MATRIX_S_W") ls >> s.S_w;
        else if (key == "MATRIX_S_H") ls >> s.S_h;
        else if (key == "MATRIX_P_TILE_MIN_ZERO_AREA") ls >> s.minZeroArea;
        else if (key == "PSEUDOCOUNT") ls >> s.pseudocount;
        else if (key == "FLUSHOUTPUT") ls >> s.flushOutput;
        else std::cerr << "WARNING: settings.cnf contains unknown key: " << key << std::endl;
    }
    return s;
}

void Settings::print() const
{
    std::cout << (fromFile ? "Loaded configuration from file settings.cnf"
                           : "File settings.cnf not found, using default values")
              << std::endl;
    std::cout << "  MATRIX_S_W = " << S_w << "; MATRIX_S_H = " << S_h
              << "; MATRIX_P_TILE_MIN_ZERO_AREA = " << minZeroArea
              << "; PSEUDOCOUNT = " << pseudocount << "; FLUSHOUTPUT = " << flushOutput << "\n";
}

// ---------------------------------------------------------------------------
std::array<float, 4> Group::probabilities(float pc) const
{
    unsigned long long total = 0ull;
    for (uint64_t c : counts) total += c;
    float denom = (float)total;
    denom += 4.0f * pc;
    std::array<float, 4> p;
    for (int b = 0; b < 4; ++b) p[b] = ((float)counts[b] + pc) / denom;
    return p;
}

std::vector<Group> read_dict(const std::string& path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open file: " + path);
    std::string tag;
    size_t ngroups = 0;
    in >> tag >> ngroups;
    std::vector<Group> groups(ngroups);
    for (Group& g : groups) {
        size_t nfiles = 0, nseq = 0;
        in >> tag >> g.name >> tag >> nfiles;
        std::set<std::string> files;
        for (size_t i = 0; i < nfiles; ++i) {
            std::string f;
            in >> f;
            files.insert(f);
        }
        g.files.assign(files.begin(), files.end());
        in >> tag >> g.totSeqLen;
        in >> tag >> g.counts[0] >> g.counts[1] >> g.counts[2] >> g.counts[3];
        in >> tag >> nseq;
        g.seqNames.resize(nseq);
        for (std::string& s : g.seqNames) in >> s;
    }
    return groups;
}

void write_dict(const std::string& path, const std::vector<Group>& groups)
{
    std::ofstream out(path);
    if (!out) throw std::runtime_error("Cannot write to file: " + path);
    out << "NUM_SPECIES\t" << groups.size() << "\n";
    for (const Group& g : groups) {
        out << "SPECIES\t" << g.name << "\n";
        out << "NUM_FASTA_FILES\t" << g.files.size() << "\n";
        for (const std::string& f : g.files) out << f << "\n";
        out << "TOT_SEQ_LENGTH\t" << g.totSeqLen << "\n";
        out << "NUCL_COUNT_ACGT\t" << g.counts[0] << "\t" << g.counts[1] << "\t" << g.counts[2]
            << "\t" << g.counts[3] << "\n";
        out << "NUM_SEQUENCES\t" << g.seqNames.size() << "\n";
        for (const std::string& s : g.seqNames) out << s << "\n";
    }
}

// ---------------------------------------------------------------------------
std::string Motif::baseName() const
{
    long i = (long)name.size() - 1;
    while (i >= 0 && std::isdigit((unsigned char)name[i])) --i;
    if (i < 1 || name[i - 1] != '_' || name[i] != '_') return name;
    return name.substr(0, (size_t)(i - 1));
}

bool Motif::isPermutation() const
{
    long i = (long)name.size() - 1;
    while (i >= 0 && std::isdigit((unsigned char)name[i])) --i;
    return i >= 1 && name[i - 1] == '_' && name[i] == '_';
}

namespace {

// JASPAR blocks: ">ID rest-of-line" then four rows "X [ c c ... ]"; the first
// two whitespace tokens of a row are skipped and counts are read until the
// first token that is not an unsigned integer (motif.cpp:370-437).
void read_jaspar(const std::string& path, std::vector<Motif>& out)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Could not open file: " + path);
    while (in.good()) {
        std::string id, rest;
        in >> id;
        if (!id.empty()) id = id.substr(1);
        std::getline(in, rest);
        if (!in) break;
        std::array<std::vector<uint64_t>, 4> rows;
        for (auto& row : rows) {
            std::string line, skip;
            std::getline(in, line);
            std::istringstream ls(line);
            ls >> skip >> skip;
            size_t v;
            while (ls >> v) row.push_back(v);
        }
        Motif m;
        m.name = id;
        for (size_t p = 0; p < rows[0].size(); ++p) {
            if (p >= rows[1].size() || p >= rows[2].size() || p >= rows[3].size())
                throw std::runtime_error("Incorrect motif file format: " + path);
            m.pfm.push_back({{rows[0][p], rows[1][p], rows[2][p], rows[3][p]}});
        }
        out.push_back(std::move(m));
    }
}

// cluster-buster: ">name" lines, then one "A C G T" count line per position
// (motif.cpp:340-368).
void read_cluster_buster(const std::string& path, std::vector<Motif>& out)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Could not open file: " + path);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line[0] == '>') {
            Motif m;
            m.name = line.substr(1);
            out.push_back(std::move(m));
            continue;
        }
        if (out.empty()) throw std::runtime_error("Incorrect motif file format: " + path);
        std::istringstream ls(line);
        size_t a = 0, c = 0, g = 0, t = 0;
        ls >> a >> c >> g >> t;
        out.back().pfm.push_back({{a, c, g, t}});
    }
}

}  // namespace

std::vector<Motif> load_motifs(const std::string& path, bool withPermutations)
{
    std::vector<Motif> all;
    const std::string sfx = ".jaspar";
    if (path.size() > sfx.size() && path.compare(path.size() - sfx.size(), sfx.size(), sfx) == 0)
        read_jaspar(path, all);
    else
        read_cluster_buster(path, all);
    // Order by length with the same (unstable) std::sort and comparison as
    // the reference, so equal-length motifs land in the same columns.
    std::sort(all.begin(), all.end(), [](const Motif& x, const Motif& y) { return x.size() < y.size(); });
    std::vector<Motif> kept;
    for (Motif& m : all)
        if (withPermutations || !m.isPermutation()) kept.push_back(std::move(m));
    return kept;
}

std::vector<Motif> add_reverse_complements(const std::vector<Motif>& in)
{
    std::vector<Motif> out;
    out.reserve(2 * in.size());
    for (const Motif& m : in) {
        out.push_back(m);
        Motif r = m;
        std::reverse(r.pfm.begin(), r.pfm.end());
        for (auto& col : r.pfm) col = {{col[3], col[2], col[1], col[0]}};
        r.rc = !m.rc;
        out.push_back(std::move(r));
    }
    return out;
}

void pfm_to_pwm(const Motif& m, const std::array<uint64_t, 4>& bg, float pc, float* out)
{
    unsigned long long bgTotal = 0ull;
    for (uint64_t c : bg) bgTotal += c;
    float bgDenom = (float)bgTotal;
    bgDenom += 4.0f * pc;
    std::array<float, 4> prob;
    for (int b = 0; b < 4; ++b) prob[b] = ((float)bg[b] + pc) / bgDenom;
    if (m.rc) {   // the reverse strand sees the complemented background
        std::swap(prob[0], prob[3]);
        std::swap(prob[1], prob[2]);
    }
    for (size_t p = 0; p < m.size(); ++p) {
        unsigned long long colTotal = 0ull;
        for (uint64_t c : m.pfm[p]) colTotal += c;
        float denom = (float)colTotal;
        denom += 4.0f * pc;
        for (int b = 0; b < 4; ++b) {
            const float ppm = ((float)m.pfm[p][b] + pc) / denom;
            out[4 * p + b] = std::log2(ppm / prob[b]);   // log2f of the host libm
        }
    }
}

float max_score(const float* w, size_t L)
{
    float s = 0.0f;
    for (size_t p = 0; p < L; ++p)
        s += std::max(std::max(w[4 * p], w[4 * p + 1]), std::max(w[4 * p + 2], w[4 * p + 3]));
    return s;
}

float min_score(const float* w, size_t L)
{
    float s = 0.0f;
    for (size_t p = 0; p < L; ++p)
        s += std::min(std::min(w[4 * p], w[4 * p + 1]), std::min(w[4 * p + 2], w[4 * p + 3]));
    return s;
}

// ---------------------------------------------------------------------------
// p-value -> score: walk down from the top bin accumulating counts in double;
// interpolate inside the bin that reaches p * total; the bin coordinate is
// rounded to float before it is scaled (motif.cpp:47-69).
float Histogram::cutoff(float pvalue) const
{
    const float width = (maxScore - minScore) / (float)numBins;
    double total = 0.0;
    for (uint64_t c : counts) total += (double)c;
    double remaining = (double)pvalue * total;
    for (long i = (long)numBins - 1; i >= 0; --i) {
        const double c = (double)counts[(size_t)i];
        if (c < remaining) {
            remaining -= c;
            continue;
        }
        const double frac = remaining / c;
        const float binPos = (float)(frac * (double)i + (1.0 - frac) * (double)(i + 1));
        return binPos * width + minScore;
    }
    return maxScore;
}

Histogram Histogram::load(const std::string& path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Error: cannot read file " + path + ". Did you run the hist module?");
    Histogram h;
    in >> h.numBins >> h.minScore >> h.maxScore;
    h.counts.assign(h.numBins, 0);
    for (size_t i = 0; i < h.numBins; ++i) {
        float center;
        size_t v;
        in >> center >> v;
        h.counts[i] = v;
    }
    if (!in) throw std::runtime_error("Unexpected end-of-file reached");
    return h;
}

void Histogram::write(const std::string& dir, const std::string& base, const std::string& label) const
{
    const float width = (maxScore - minScore) / (float)numBins;
    {
        const std::string fn = dir + base + ".dat";
        std::ofstream out(fn);
        if (!out) throw std::runtime_error("Error: cannot write to file " + fn);
        out << numBins << "\t" << minScore << "\t" << maxScore << "\n";
        for (size_t i = 0; i < numBins; ++i) out << (0.5 + (double)i) * width + minScore << "\t" << counts[i] << "\n";
    }
    size_t ymax = 0;
    for (uint64_t c : counts) ymax = std::max<size_t>(ymax, c);
    ymax = (size_t)((double)ymax * 1.1);
    const std::string fn = dir + base + ".gnu";
    std::ofstream out(fn);
    if (!out) throw std::runtime_error("Error: cannot write to file " + fn);
    out << "set output \"" << base << ".ps\"\n"
        << "set key autotitle columnhead\n"
        << "set terminal postscript landscape\n"
        << "set terminal postscript noenhanced\n"
        << "set xrange [" << minScore << ":" << maxScore << "]\n"
        << "set yrange [" << 0 << ":" << ymax << "]\n"
        << "set xlabel 'PWM score'\n"
        << "set ylabel 'count'\n"
        << "plot \"" << base << ".dat\" using 1:2 title '" << label << "' with boxes\n";
}

Histogram as_written(const Histogram& h)
{
    Histogram r = h;
    std::ostringstream os;
    os << h.minScore << " " << h.maxScore;
    std::istringstream is(os.str());
    is >> r.minScore >> r.maxScore;
    return r;
}

// Score distribution of a PWM under an i.i.d. background, on the integer grid
// round(a*w + b) with a = numBins/(max-min), b = -a*min/L; probabilities are
// accumulated in float in ascending grid order, one position at a time
// (motif.cpp:151-192).  Each distinct grid score then *sets* its bin to
// (size_t)(maxLength * p) -- a later score in the same bin replaces an
// earlier one (motif.h:109-115, hist.cpp:162-175).
Histogram theoretical_histogram(const float* w, size_t L, const std::array<float, 4>& bg,
                                size_t numBins, size_t maxLength)
{
    Histogram h;
    h.numBins = numBins;
    h.minScore = min_score(w, L);
    h.maxScore = max_score(w, L);
    h.counts.assign(numBins, 0);
    const float range = h.maxScore - h.minScore;
    const float a = (float)numBins / range;
    const float b = -a * h.minScore / (float)L;
    std::vector<std::array<int, 4>> grid(L);
    for (size_t p = 0; p < L; ++p)
        for (int k = 0; k < 4; ++k) grid[p][k] = (int)std::round(a * w[4 * p + k] + b);
    std::map<int, float> dist, next;
    for (int k = 0; k < 4; ++k) dist[grid[0][k]] += bg[k];
    for (size_t p = 1; p < L; ++p) {
        next.clear();
        for (const auto& e : dist)
            for (int k = 0; k < 4; ++k) next[e.first + grid[p][k]] += e.second * bg[k];
        dist.swap(next);
    }
    std::map<float, float> spectrum;
    for (const auto& e : dist) spectrum.insert({((float)e.first - (float)L * b) / a, e.second});
    const float width = (h.maxScore - h.minScore) / (float)numBins;
    for (const auto& e : spectrum) {
        const size_t cnt = (size_t)((float)maxLength * e.second);
        int bin = (int)((e.first - h.minScore) / width);
        bin = std::max(0, std::min((int)numBins - 1, bin));
        h.counts[(size_t)bin] = cnt;
    }
    return h;
}

}  // namespace blamm
