/*
 * blamm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference blamm `scan` path (and of the `dict` and
 * theoretical `hist` producers of its inputs).  It is the checker for the HIP
 * product: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and never as the thing measured or shipped.
 *
 * Every function cites the reference file:line it restates.  Pinning status of
 * each function is listed in the header comment of blamm_oracle.c.
 */
#ifndef BLAMM_ORACLE_H
#define BLAMM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One hit of the scan: stream position, column, FP64 score (and the float the
 * reference path would print). */
typedef struct {
    uint64_t pos;      /* position in the filtered stream */
    uint32_t col;      /* PWM column (reference column order) */
    uint32_t pad;
    double score64;    /* left-to-right FP64 sum */
    float score32;     /* (float)score64, or the sgemm FP32 value in sgemm mode */
    float pad2;
} oracle_hit;

/* Natural log2f of the host libm, exported so tests can pin PFM2PWM against
 * the same routine the reference calls (motif.cpp:219, std::log2(float)). */
float oracle_log2f(float x);

/* PFM -> PWM (motif.cpp:194-223).  pfm: L x 4 counts (ACGT), bg: 4 counts.
 * rc != 0 swaps the background (motif.cpp:204-207).  pwm: L x 4 floats. */
void oracle_pfm2pwm(const uint64_t *pfm, int L, const uint64_t *bg,
                    float pseudocount, int rc, float *pwm);

/* Motif::getMaxScore / getMinScore (motif.cpp:241-265). */
float oracle_max_score(const float *pwm, int L);
float oracle_min_score(const float *pwm, int L);

/* ScoreHistogram::getScoreCutoff (motif.cpp:47-69) over counts[numBins]. */
float oracle_score_cutoff(const uint64_t *counts, size_t numBins, float minScore,
                          float maxScore, float pvalue);

/* Reverse complement of a PFM/PWM in place (motif.cpp:267-286). */
void oracle_revcompl_pfm(uint64_t *pfm, int L);

/* FASTA filtering (sequence.cpp:142-293): files[nfiles] read in order; the
 * filtered stream (valid ACGTacgt bytes only) is truncated at maxLen.
 * Returns the number of bases; *stream (malloc'd), *runStart, *runSeq, *runPos
 * (malloc'd, *nruns entries) describe the maximal contiguous runs.
 * *nseq = number of '>' records.  Returns -1 on error (message in errbuf). */
int64_t oracle_filter_fasta(const char *const *files, int nfiles, uint64_t maxLen,
                            char **stream, uint64_t **runStart, uint64_t **runSeq,
                            uint64_t **runPos, uint64_t *nruns, uint64_t *nseq,
                            char *errbuf, size_t errlen);

/* Brute-force FP64 scan of a filtered stream (oracle predicate: a window at q
 * for column c is scored iff L_c <= remaining run length, score64 is the left
 * to right FP64 sum, lowercase bases contribute 0, hit iff score64 >= thr).
 * pwm: ncols x ldw x 4 floats (column c at pwm + c*ldw*4).  runStart holds the
 * stream positions where a run starts (sorted, first is 0).
 * Hits are appended to *hits (realloc'd); returns the number of hits. */
int64_t oracle_scan_fp64(const char *stream, uint64_t n, const uint64_t *runStart,
                         uint64_t nruns, const float *pwm, int ldw, const int *len,
                         const float *thr, int ncols, int nthreads, oracle_hit **hits);

/* The reference's sgemm scan (pwmscan.cpp:223-278 with matrix.h:292-306) over
 * the same stream, calling the host BLAS sgemm loaded by oracle_blas_open().
 * Columns must be sorted by non-decreasing length (motif.cpp:436). */
int64_t oracle_scan_sgemm(const char *stream, uint64_t n, const uint64_t *runStart,
                          uint64_t nruns, const float *pwm, int ldw, const int *len,
                          const float *thr, int ncols, int h, int w,
                          size_t minZeroArea, int nthreads, oracle_hit **hits);

/* Load an F77 sgemm from a shared library ("" = search the SciPy wheel). */
int oracle_blas_open(const char *path, char *errbuf, size_t errlen);

void oracle_free(void *p);

/* Theoretical spectrum (motif.cpp:151-192) and histogram fill
 * (hist.cpp:162-175): writes numBins counts.  Returns 0. */
int oracle_theoretical_hist(const float *pwm, int L, const float *bgProb,
                            size_t numBins, size_t maxLength, uint64_t *counts,
                            float *minScore, float *maxScore);

/* Empirical histograms (hist -e; hist.cpp:70-138, motif.h:58-66, 96-102):
 * for each column c and every window q < min(qend, n) that satisfies the run
 * rule, the FP32 left-to-right score (lowercase bases contribute 0) is binned
 * as clamp(int((score - minScore[c]) / width), 0, numBins - 1) with
 * width = (maxScore[c] - minScore[c]) / numBins in float and the x86 int
 * conversion; counts[c * numBins + bin] is incremented (added to). */
int oracle_empirical_hist(const char *stream, uint64_t n, const uint64_t *runStart,
                          uint64_t nruns, const float *pwm, int ldw, const int *len,
                          const float *minScore, const float *maxScore, int ncols,
                          size_t numBins, uint64_t qend, int nthreads, uint64_t *counts);

/* MotifContainer::load (motif.cpp:323-338): jaspar when the name ends in
 * ".jaspar" (loadJasparMotifs, motif.cpp:370-437), else cluster-buster
 * (loadCBMotifs, motif.cpp:340-368); sorted by length (stable here, the
 * reference's std::sort is not); permutation motifs (names ending in "__" and
 * digits, motif.h:272-281) dropped unless perms != 0.  On return *names is an
 * array of *n malloc'd strings, *lengths their lengths and *pfms the
 * concatenated L x 4 counts (all malloc'd; free with oracle_free).
 * Returns 0, or -1 when the file cannot be opened. */
int oracle_load_motifs(const char *fn, int perms, size_t *n, char ***names, int **lengths, uint64_t **pfms);

/* Motif::isPermutation / getBaseName (motif.h:256-281); base_name returns a
 * malloc'd string. */
int oracle_is_permutation(const char *name);
char *oracle_base_name(const char *name);

/* Command line entry (dict / hist / scan), same syntax as the reference. */
int oracle_main(int argc, char **argv);

#ifdef __cplusplus
}
#endif
#endif
