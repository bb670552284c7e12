/*
 * blamm_host.h -- C ABI of libblamm_host.so: the host-side pieces of the scan
 * path that callers other than the `blamm` executable need (bindings, tests,
 * benchmarks).  Each entry point restates a reference routine bit for bit;
 * nothing here touches the GPU.  Errors: int status (0 = ok, negative on
 * error) and blamm_host_last_error() on the calling thread.
 */
#ifndef BLAMM_HOST_H
#define BLAMM_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *blamm_host_last_error(void);

/* Motif::PFM2PWM (motif.cpp:194-223): pfm is L x 4 counts (ACGT), bg the
 * group's ACGT counts; rc != 0 marks a reverse-complement motif (its PFM must
 * already be reversed and complemented).  pwm receives L x 4 floats. */
int blamm_host_pwm(const uint64_t *pfm, int L, const uint64_t *bg, float pseudocount,
                   int rc, float *pwm);

/* getMaxScore / getMinScore (motif.cpp:241-265). */
float blamm_host_max_score(const float *pwm, int L);
float blamm_host_min_score(const float *pwm, int L);

/* Theoretical score histogram of `blamm hist` (motif.cpp:151-192,
 * hist.cpp:162-175) for a forward PWM under background probabilities bg:
 * counts[numBins], min/max score. */
int blamm_host_theoretical_hist(const float *pwm, int L, const float *bg, uint64_t numBins,
                                uint64_t maxLength, uint64_t *counts, float *minScore,
                                float *maxScore);

/* -pt threshold as `blamm scan` derives it from the .dat that `blamm hist`
 * writes for this motif (min/max at 6 significant digits, then
 * getScoreCutoff, motif.cpp:47-69). */
int blamm_host_pvalue_threshold(const float *pwm, int L, const uint64_t *bg_counts,
                                float pseudocount, uint64_t numBins, uint64_t maxLength,
                                float pvalue, float *threshold);

/* Pack an in-memory filtered stream (only ACGTacgt bytes) into 2-bit codes
 * (codes: (n+31)/32 words) and lowercase bits (lower: (n+63)/64 words, may be
 * NULL when the caller knows there is no lowercase).  *has_lower is set. */
int blamm_host_pack(const char *bases, uint64_t n, uint64_t *codes, uint64_t *lower,
                    int *has_lower);

#ifdef __cplusplus
}
#endif
#endif
