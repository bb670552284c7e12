// libblamm_host.so: the C ABI declared in include/blamm_host.h.
#include <cstring>
#include <string>

#include "../../../include/blamm_host.h"
#include "host.hpp"

namespace {
thread_local std::string g_err;
int host_fail(const std::string& m)
{
    g_err = m;
    return -1;
}
}  // namespace

extern "C" {

const char* blamm_host_last_error(void) { return g_err.c_str(); }

int blamm_host_pwm(const uint64_t* pfm, int L, const uint64_t* bg, float pc, int rc, float* pwm)
{
    if (!pfm || !bg || !pwm || L < 1) return host_fail("pwm: bad arguments");
    blamm::Motif m;
    m.rc = rc != 0;
    for (int p = 0; p < L; ++p) m.pfm.push_back({{pfm[4 * p], pfm[4 * p + 1], pfm[4 * p + 2], pfm[4 * p + 3]}});
    blamm::pfm_to_pwm(m, {{bg[0], bg[1], bg[2], bg[3]}}, pc, pwm);
    return 0;
}

float blamm_host_max_score(const float* pwm, int L) { return blamm::max_score(pwm, (size_t)L); }
float blamm_host_min_score(const float* pwm, int L) { return blamm::min_score(pwm, (size_t)L); }

int blamm_host_theoretical_hist(const float* pwm, int L, const float* bg, uint64_t numBins, uint64_t maxLength,
                                uint64_t* counts, float* mn, float* mx)
{
    if (!pwm || !bg || !counts || L < 1 || numBins < 1) return host_fail("theoretical_hist: bad arguments");
    const blamm::Histogram h =
        blamm::theoretical_histogram(pwm, (size_t)L, {{bg[0], bg[1], bg[2], bg[3]}}, numBins, maxLength);
    std::memcpy(counts, h.counts.data(), numBins * sizeof(uint64_t));
    if (mn) *mn = h.minScore;
    if (mx) *mx = h.maxScore;
    return 0;
}

int blamm_host_pvalue_threshold(const float* pwm, int L, const uint64_t* bgc, float pc, uint64_t numBins,
                                uint64_t maxLength, float pvalue, float* thr)
{
    if (!pwm || !bgc || !thr || L < 1 || numBins < 1) return host_fail("pvalue_threshold: bad arguments");
    blamm::Group g;
    g.counts = {{bgc[0], bgc[1], bgc[2], bgc[3]}};
    const blamm::Histogram h = blamm::theoretical_histogram(pwm, (size_t)L, g.probabilities(pc), numBins, maxLength);
    *thr = blamm::as_written(h).cutoff(pvalue);
    return 0;
}

int blamm_host_pack(const char* bases, uint64_t n, uint64_t* codes, uint64_t* lower, int* has_lower)
{
    try {
        const blamm::PackedStream s = blamm::pack_bases(bases, n);
        std::memcpy(codes, s.codes.data(), s.codes.size() * 8);
        if (has_lower) *has_lower = s.hasLower() ? 1 : 0;
        if (lower) {
            std::memset(lower, 0, ((n + 63) / 64) * 8);
            if (s.hasLower()) std::memcpy(lower, s.lower.data(), s.lower.size() * 8);
        }
        return 0;
    } catch (const std::exception& e) {
        return host_fail(e.what());
    }
}

}  // extern "C"
