/*
 * blamm_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A plain-C restatement of the reference `blamm` scan path and of the two
 * producers of its inputs (`dict`, theoretical `hist`).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load or run it.
 *
 * Citations are to the reference tree (biointec/blamm) as file:line.
 *
 * Pinning status (what fixes each function besides this restatement; the
 * checks live in tests/pinning.py, run by tests/test_oracle_pinning.py):
 *   filter_fasta          hand-built FASTA cases with closed-form runs/positions
 *                         (N-runs, CRLF, lowercase, empty records, dict
 *                         truncation)                      check_fasta_cases
 *   pfm2pwm               closed form with the host libm log2f (bitwise), rc
 *                         identity W_rc[p][b] == W[L-1-p][3-b]
 *                                 check_pwm_values / _bitwise / _revcompl
 *   max/min score         brute force over all 4^L words  check_min_max_bruteforce
 *   score_cutoff          hand-written histograms with closed-form cutoffs,
 *                         float-coordinate cases           check_cutoff_*
 *   load_jaspar           the reference's example/motifs.jaspar, counts typed
 *                         from the file; hand-built layouts ("[1" swallowing a
 *                         count, "3]", blank lines, no final newline)
 *                                 check_load_jaspar_example / _layouts
 *   load_cb, is_permutation, base_name
 *                         hand-built cluster-buster file (names with blanks,
 *                         "__NN", "__", "X_1", "__5", "_5") with and without
 *                         permutations; a table of names
 *                                 check_load_cluster_buster / check_permutation_names
 *   dict writer           closed form on hand-built FASTA (case-folded counts,
 *                         group and file order, names, CRLF) check_dict_closed_form
 *   theoretical_hist      exact: designs where every float step is exact, bins
 *                         and counts fixed by hand (store-not-add collision,
 *                         round-half-away grid, float count product) and by an
 *                         exact-arithmetic brute force on 40 random dyadic
 *                         designs                          check_hist_dyadic_exact
 *                         plus support/count invariants on random PWMs
 *   scan_fp64             pure-Python brute force on tiny inputs; rc symmetry;
 *                         -at thresholds placed exactly on achievable sums (>=);
 *                         lowercase contributes 0         check_scan_*
 *   scan_sgemm            equals scan_fp64 except inside the FP32 error band;
 *                         independent of MATRIX_S_W/S_H and of the tiling
 *                                 tests/test_oracle_pinning.py
 *   empirical_hist        closed forms: L = 1 (bin counts = base counts),
 *                         dyadic PWMs whose sums and bin edges are exact
 *                         (pure-Python brute force), total = number of valid
 *                         windows, a window whose FP32 left-to-right sum and
 *                         exact sum fall on either side of a bin edge, width 0
 *                                 check_ehist_*
 * Planted mistakes: compiling with -DORACLE_MUTATE_<name> swaps one rule for a
 * plausible wrong one (GT, NO_BG_SWAP, LOWER_COUNTS, NO_RUN_CHECK,
 * CUTOFF_DOUBLE, HIST_ADD, RUNS_IGNORE_RECORD, HIST_ROUND, HIST_NO_RUN_CHECK,
 * HIST_FP64, HIST_RINT, HIST_COUNT_DOUBLE, JASPAR_ONE_SKIP, PERM_NEEDS_DIGIT,
 * DICT_UPPER_ONLY, DICT_FILE_ORDER); tests/test_oracle_mutations.py shows
 * each is caught by a pinning check.
 * The reference itself is not built here: its sources include the
 * CMake-generated config.h and need an external BLAS, so no run of the
 * reference binary pins any of the above.
 */
#define _GNU_SOURCE
#include "blamm_oracle.h"

#include <ctype.h>
#include <dlfcn.h>
#include <errno.h>
#include <glob.h>
#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

/* ------------------------------------------------------------------------ */
/* small utilities                                                           */
/* ------------------------------------------------------------------------ */

static void die(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
    exit(EXIT_FAILURE);
}

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) die("oracle: out of memory");
    return p;
}

static void *xrealloc(void *p, size_t n)
{
    p = realloc(p, n ? n : 1);
    if (!p) die("oracle: out of memory");
    return p;
}

static char *xstrdup(const char *s)
{
    char *r = strdup(s);
    if (!r) die("oracle: out of memory");
    return r;
}

void oracle_free(void *p) { free(p); }

float oracle_log2f(float x) { return log2f(x); }

/* whole-file read */
static char *read_file(const char *path, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    size_t cap = 1 << 16, n = 0;
    char *buf = xmalloc(cap + 1);
    for (;;) {
        if (n == cap) { cap *= 2; buf = xrealloc(buf, cap + 1); }
        size_t r = fread(buf + n, 1, cap - n, f);
        if (r == 0) break;
        n += r;
    }
    fclose(f);
    buf[n] = 0;
    *len = n;
    return buf;
}

/* A cursor emulating the std::istream operations the reference uses. */
typedef struct {
    const char *s;
    size_t n, i;
    int fail, eof;
} cursor;

static int is_ws(int c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

/* operator>>(std::string): skip whitespace, read a token (fails at EOF) */
static int cur_token(cursor *c, char **out)
{
    *out = NULL;
    if (c->fail) return 0;
    while (c->i < c->n && is_ws((unsigned char)c->s[c->i])) c->i++;
    if (c->i >= c->n) { c->fail = c->eof = 1; return 0; }
    size_t b = c->i;
    while (c->i < c->n && !is_ws((unsigned char)c->s[c->i])) c->i++;
    if (c->i >= c->n) c->eof = 1;
    *out = xmalloc(c->i - b + 1);
    memcpy(*out, c->s + b, c->i - b);
    (*out)[c->i - b] = 0;
    return 1;
}

/* std::getline: read up to '\n' (consumed); fails if nothing is extracted at EOF */
static int cur_getline(cursor *c, char **out)
{
    *out = NULL;
    if (c->fail) return 0;
    if (c->i >= c->n) { c->fail = c->eof = 1; return 0; }
    size_t b = c->i;
    while (c->i < c->n && c->s[c->i] != '\n') c->i++;
    size_t e = c->i;
    if (c->i < c->n) c->i++; else c->eof = 1;
    *out = xmalloc(e - b + 1);
    memcpy(*out, c->s + b, e - b);
    (*out)[e - b] = 0;
    return 1;
}

/* operator>>(size_t) on a token string: digits only (motif.cpp:396-403) */
static int parse_count(const char **p, uint64_t *v)
{
    const char *s = *p;
    while (*s && is_ws((unsigned char)*s)) s++;
    if (!isdigit((unsigned char)*s)) { *p = s; return 0; }
    uint64_t x = 0;
    while (isdigit((unsigned char)*s)) { x = x * 10 + (uint64_t)(*s - '0'); s++; }
    *v = x;
    *p = s;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* settings.cnf (settings.cpp:34-74)                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    size_t S_w, S_h, minZeroArea, flushOutput;
    float pseudocount;
} settings_t;

static void settings_load(settings_t *st)
{
    st->S_w = 250; st->S_h = 1000; st->minZeroArea = 64 * 64;
    st->flushOutput = 100000; st->pseudocount = 0.25f;
    size_t len;
    char *txt = read_file("settings.cnf", &len);
    if (!txt) return;
    cursor c = {txt, len, 0, 0, 0};
    char *line;
    while (cur_getline(&c, &line)) {
        if (line[0] == 0 || line[0] == '#') { free(line); continue; }
        char key[256] = {0};
        const char *s = line;
        while (*s && is_ws((unsigned char)*s)) s++;
        size_t k = 0;
        while (*s && !is_ws((unsigned char)*s) && k < sizeof(key) - 1) key[k++] = *s++;
        if (!strcmp(key, "MATRIX_S_W")) st->S_w = strtoull(s, NULL, 10);
        else if (!strcmp(key, "MATRIX_S_H")) st->S_h = strtoull(s, NULL, 10);
        else if (!strcmp(key, "MATRIX_P_TILE_MIN_ZERO_AREA")) st->minZeroArea = strtoull(s, NULL, 10);
        else if (!strcmp(key, "PSEUDOCOUNT")) st->pseudocount = strtof(s, NULL);
        else if (!strcmp(key, "FLUSHOUTPUT")) st->flushOutput = strtoull(s, NULL, 10);
        else fprintf(stderr, "WARNING: settings.cnf contains unknown key: %s\n", key);
        free(line);
    }
    free(txt);
}

/* ------------------------------------------------------------------------ */
/* motifs (motif.cpp:194-286, 323-449)                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    char *name;
    int L;
    uint64_t *pfm;   /* L x 4 */
    float *pwm;      /* L x 4 */
    int rc;
    float thr;
} motif_t;

void oracle_pfm2pwm(const uint64_t *pfm, int L, const uint64_t *bg,
                    float pc, int rc, float *pwm)
{
    /* motif.cpp:197-201: accumulate(..., 0ull) then float arithmetic */
    unsigned long long bsum = 0ull;
    for (int i = 0; i < 4; i++) bsum += bg[i];
    float bgTot = (float)bsum;
    bgTot += 4.0f * pc;
    float bgProb[4];
    for (int i = 0; i < 4; i++) bgProb[i] = ((float)bg[i] + pc) / bgTot;
#ifdef ORACLE_MUTATE_NO_BG_SWAP
    if (0) {
#else
    if (rc) {   /* motif.cpp:204-207 */
#endif
        float t = bgProb[0]; bgProb[0] = bgProb[3]; bgProb[3] = t;
        t = bgProb[1]; bgProb[1] = bgProb[2]; bgProb[2] = t;
    }
    for (int i = 0; i < L; i++) {   /* motif.cpp:212-222 */
        unsigned long long s = 0ull;
        for (int j = 0; j < 4; j++) s += pfm[4 * i + j];
        float tot = (float)s;
        tot += 4.0f * pc;
        for (int j = 0; j < 4; j++) {
            float ppm = ((float)pfm[4 * i + j] + pc) / tot;
            pwm[4 * i + j] = log2f(ppm / bgProb[j]);
        }
    }
}

float oracle_max_score(const float *pwm, int L)
{
    float s = 0.0f;   /* motif.cpp:241-252; std::max(a,b) returns a unless a < b */
    for (int p = 0; p < L; p++) {
        float ac = (pwm[4 * p] < pwm[4 * p + 1]) ? pwm[4 * p + 1] : pwm[4 * p];
        float gt = (pwm[4 * p + 2] < pwm[4 * p + 3]) ? pwm[4 * p + 3] : pwm[4 * p + 2];
        s += (ac < gt) ? gt : ac;
    }
    return s;
}

float oracle_min_score(const float *pwm, int L)
{
    float s = 0.0f;   /* motif.cpp:254-265; std::min returns a unless b < a */
    for (int p = 0; p < L; p++) {
        float ac = (pwm[4 * p + 1] < pwm[4 * p]) ? pwm[4 * p + 1] : pwm[4 * p];
        float gt = (pwm[4 * p + 3] < pwm[4 * p + 2]) ? pwm[4 * p + 3] : pwm[4 * p + 2];
        s += (gt < ac) ? gt : ac;
    }
    return s;
}

void oracle_revcompl_pfm(uint64_t *pfm, int L)
{
    /* motif.cpp:267-286: reverse positions, then complement (ACGT -> TGCA) */
    for (int i = 0; i < L / 2; i++)
        for (int j = 0; j < 4; j++) {
            uint64_t t = pfm[4 * i + j];
            pfm[4 * i + j] = pfm[4 * (L - 1 - i) + j];
            pfm[4 * (L - 1 - i) + j] = t;
        }
    for (int i = 0; i < L; i++) {
        uint64_t c[4] = {pfm[4 * i], pfm[4 * i + 1], pfm[4 * i + 2], pfm[4 * i + 3]};
        pfm[4 * i] = c[3]; pfm[4 * i + 1] = c[2]; pfm[4 * i + 2] = c[1]; pfm[4 * i + 3] = c[0];
    }
}

static void motif_push(motif_t **v, size_t *n, size_t *cap, const char *name)
{
    if (*n == *cap) { *cap = *cap ? *cap * 2 : 16; *v = xrealloc(*v, *cap * sizeof(motif_t)); }
    motif_t *m = &(*v)[(*n)++];
    memset(m, 0, sizeof(*m));
    m->name = xstrdup(name);
}

static void motif_add_char(motif_t *m, const uint64_t c[4])
{
    m->pfm = xrealloc(m->pfm, (size_t)(m->L + 1) * 4 * sizeof(uint64_t));
    memcpy(m->pfm + 4 * m->L, c, 4 * sizeof(uint64_t));
    m->L++;
}

/* loadJasparMotifs (motif.cpp:370-437) */
static void load_jaspar(const char *fn, motif_t **v, size_t *n, size_t *cap)
{
    size_t len;
    char *txt = read_file(fn, &len);
    if (!txt) die("Could not open file: %s", fn);
    cursor c = {txt, len, 0, 0, 0};
    while (!c.fail && !c.eof) {      /* while (ifs.good()) */
        char *name = NULL, *rest = NULL;
        cur_token(&c, &name);
        const char *nm = name ? name + 1 : "";
        if (name && name[0] == 0) nm = "";
        cur_getline(&c, &rest);
        if (c.fail) { free(name); free(rest); break; }
        uint64_t *freq[4] = {0};
        size_t nf[4] = {0};
        for (int r = 0; r < 4; r++) {
            char *line = NULL;
            cur_getline(&c, &line);
            if (!line) continue;
            const char *s = line;
            /* skip two tokens ("A" and "[") */
#ifdef ORACLE_MUTATE_JASPAR_ONE_SKIP
            for (int t = 0; t < 1; t++) {   /* wrong: leaves "[" in front of the counts */
#else
            for (int t = 0; t < 2; t++) {
#endif
                while (*s && is_ws((unsigned char)*s)) s++;
                while (*s && !is_ws((unsigned char)*s)) s++;
            }
            uint64_t v64;
            while (parse_count(&s, &v64)) {
                freq[r] = xrealloc(freq[r], (nf[r] + 1) * sizeof(uint64_t));
                freq[r][nf[r]++] = v64;
            }
            free(line);
        }
        motif_push(v, n, cap, nm);
        for (size_t i = 0; i < nf[0]; i++) {
            if (i >= nf[1] || i >= nf[2] || i >= nf[3])
                die("oracle: ragged count rows in motif %s", nm);
            uint64_t cc[4] = {freq[0][i], freq[1][i], freq[2][i], freq[3][i]};
            motif_add_char(&(*v)[*n - 1], cc);
        }
        for (int r = 0; r < 4; r++) free(freq[r]);
        free(name); free(rest);
    }
    free(txt);
}

/* loadCBMotifs (motif.cpp:340-368) */
static void load_cb(const char *fn, motif_t **v, size_t *n, size_t *cap)
{
    size_t len;
    char *txt = read_file(fn, &len);
    if (!txt) die("Could not open file: %s", fn);
    cursor c = {txt, len, 0, 0, 0};
    char *line;
    while (cur_getline(&c, &line)) {
        if (line[0] == 0) { free(line); continue; }
        if (line[0] == '>') { motif_push(v, n, cap, line + 1); free(line); continue; }
        if (*n == 0) die("Incorrect motif file format: %s", fn);
        const char *s = line;
        uint64_t cc[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; k++) if (!parse_count(&s, &cc[k])) break;
        motif_add_char(&(*v)[*n - 1], cc);
        free(line);
    }
    free(txt);
}

/* Motif::isPermutation (motif.h:272-281) */
static int is_permutation(const char *name)
{
    long i = (long)strlen(name) - 1;
    for (; i >= 0; i--) if (!isdigit((unsigned char)name[i])) break;
    if (i < 1) return 0;
#ifdef ORACLE_MUTATE_PERM_NEEDS_DIGIT
    if (i == (long)strlen(name) - 1) return 0;   /* wrong: "X__" is a permutation too */
#endif
    return name[i - 1] == '_' && name[i] == '_';
}

/* Motif::getBaseName (motif.h:256-267) */
static char *base_name(const char *name)
{
    long i = (long)strlen(name) - 1;
    for (; i >= 0; i--) if (!isdigit((unsigned char)name[i])) break;
    if (i < 1 || name[i - 1] != '_' || name[i] != '_') return xstrdup(name);
    char *r = xmalloc((size_t)i);
    memcpy(r, name, (size_t)(i - 1));
    r[i - 1] = 0;
    return r;
}

/* stable sort by length (the reference's std::sort is unstable; ties only
 * change the column order, which the normalised output does not see) */
static void sort_by_length(motif_t *v, size_t n)
{
    motif_t *tmp = xmalloc(n * sizeof(motif_t));
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * w) {
            size_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            size_t a = lo, b = mid, k = lo;
            while (a < mid && b < hi) tmp[k++] = (v[b].L < v[a].L) ? v[b++] : v[a++];
            while (a < mid) tmp[k++] = v[a++];
            while (b < hi) tmp[k++] = v[b++];
        }
        memcpy(v, tmp, n * sizeof(motif_t));
    }
    free(tmp);
}

/* MotifContainer::load (motif.cpp:323-338) + sort (motif.cpp:365, 436) */
static motif_t *load_motifs(const char *fn, int perms, size_t *nout)
{
    motif_t *all = NULL; size_t n = 0, cap = 0;
    size_t fl = strlen(fn);
    if (fl > 7 && !strcmp(fn + fl - 7, ".jaspar")) load_jaspar(fn, &all, &n, &cap);
    else load_cb(fn, &all, &n, &cap);
    sort_by_length(all, n);
    motif_t *out = xmalloc((n ? n : 1) * sizeof(motif_t));
    size_t k = 0;
    for (size_t i = 0; i < n; i++) {
        if (perms || !is_permutation(all[i].name)) out[k++] = all[i];
        else { free(all[i].name); free(all[i].pfm); }
    }
    free(all);
    *nout = k;
    return out;
}

/* addReverseComplements (motif.cpp:439-449): fwd at 2i, rc at 2i+1 */
static motif_t *add_rc(motif_t *v, size_t n)
{
    motif_t *r = xmalloc(2 * (n ? n : 1) * sizeof(motif_t));
    for (size_t i = 0; i < n; i++) {
        r[2 * i] = v[i];
        motif_t m = v[i];
        m.name = xstrdup(v[i].name);
        m.pfm = xmalloc((size_t)m.L * 4 * sizeof(uint64_t));
        memcpy(m.pfm, v[i].pfm, (size_t)m.L * 4 * sizeof(uint64_t));
        oracle_revcompl_pfm(m.pfm, m.L);
        m.rc = !v[i].rc;
        m.pwm = NULL;
        r[2 * i + 1] = m;
    }
    free(v);
    return r;
}

int oracle_load_motifs(const char *fn, int perms, size_t *n, char ***names, int **lengths, uint64_t **pfms)
{
    FILE *f = fopen(fn, "rb");
    if (!f) return -1;
    fclose(f);
    size_t k;
    motif_t *m = load_motifs(fn, perms, &k);
    size_t tot = 0;
    for (size_t i = 0; i < k; i++) tot += (size_t)m[i].L;
    *names = xmalloc((k ? k : 1) * sizeof(char *));
    *lengths = xmalloc((k ? k : 1) * sizeof(int));
    *pfms = xmalloc((tot ? tot : 1) * 4 * sizeof(uint64_t));
    size_t at = 0;
    for (size_t i = 0; i < k; i++) {
        (*names)[i] = m[i].name;
        (*lengths)[i] = m[i].L;
        if (m[i].L) memcpy(*pfms + 4 * at, m[i].pfm, (size_t)m[i].L * 4 * sizeof(uint64_t));
        at += (size_t)m[i].L;
        free(m[i].pfm);
    }
    free(m);
    *n = k;
    return 0;
}

int oracle_is_permutation(const char *name) { return is_permutation(name); }

char *oracle_base_name(const char *name) { return base_name(name); }

/* ------------------------------------------------------------------------ */
/* histograms (motif.cpp:47-132, 151-192; hist.cpp:162-175)                 */
/* ------------------------------------------------------------------------ */

float oracle_score_cutoff(const uint64_t *counts, size_t numBins, float minScore,
                          float maxScore, float pvalue)
{
    float width = (maxScore - minScore) / (float)numBins;
    double totObs = 0.0;
    for (size_t i = 0; i < numBins; i++) totObs += (double)counts[i];
    double bestObs = (double)pvalue * totObs;
    double curr = bestObs;
    for (long i = (long)numBins - 1; i >= 0; i--) {
        if ((double)counts[i] < curr) curr -= (double)counts[i];
        else {
            double frac = curr / (double)counts[i];
#ifdef ORACLE_MUTATE_CUTOFF_DOUBLE
            double cutoffi = frac * (double)i + (1.0 - frac) * (double)(i + 1);
            return (float)(cutoffi * (double)width + (double)minScore);
#else
            float cutoffi = (float)(frac * (double)i + (1.0 - frac) * (double)(i + 1));
            return cutoffi * width + minScore;
#endif
        }
    }
    return maxScore;
}

/* ScoreHistogram::loadHistogram (motif.cpp:109-132) */
static int load_hist(const char *path, uint64_t **counts, size_t *numBins,
                     float *minS, float *maxS)
{
    size_t len;
    char *txt = read_file(path, &len);
    if (!txt) return -1;
    char *s = txt, *e;
    errno = 0;
    unsigned long long nb = strtoull(s, &e, 10);
    if (e == s) { free(txt); return -2; }
    s = e;
    *minS = strtof(s, &e); if (e == s) { free(txt); return -2; } s = e;
    *maxS = strtof(s, &e); if (e == s) { free(txt); return -2; } s = e;
    *numBins = (size_t)nb;
    *counts = xmalloc((nb ? nb : 1) * sizeof(uint64_t));
    for (size_t i = 0; i < nb; i++) {
        (void)strtof(s, &e); if (e == s) { free(txt); return -2; } s = e;
        while (*s && is_ws((unsigned char)*s)) s++;
        if (!isdigit((unsigned char)*s)) { free(txt); return -2; }
        (*counts)[i] = strtoull(s, &e, 10); s = e;
    }
    free(txt);
    return 0;
}

/* Motif::computeTheoreticalSpectrum (motif.cpp:151-192) followed by the
 * setNumObservations fill of generateTheoreticalHist (hist.cpp:162-175).
 * std::map<int,float> is restated as a dense array with presence flags,
 * iterated in ascending key order exactly as the map is. */
int oracle_theoretical_hist(const float *pwm, int L, const float *bg,
                            size_t numBins, size_t maxLength, uint64_t *counts,
                            float *minScoreOut, float *maxScoreOut)
{
    float minS = oracle_min_score(pwm, L);
    float maxS = oracle_max_score(pwm, L);
    float range = maxS - minS;
    float a = (float)numBins / range;
    float b = -a * minS / (float)L;
    int *pi = xmalloc((size_t)L * 4 * sizeof(int));
    long lo = 0, hi = 0;
    for (int p = 0; p < L; p++) {
        int mn = 0, mx = 0;
        for (int k = 0; k < 4; k++) {
#ifdef ORACLE_MUTATE_HIST_RINT
            pi[4 * p + k] = (int)rintf(a * pwm[4 * p + k] + b);   /* wrong: ties to even */
#else
            pi[4 * p + k] = (int)roundf(a * pwm[4 * p + k] + b);
#endif
            if (k == 0 || pi[4 * p + k] < mn) mn = pi[4 * p + k];
            if (k == 0 || pi[4 * p + k] > mx) mx = pi[4 * p + k];
        }
        lo += mn; hi += mx;
    }
    /* dense range covering every partial sum */
    long plo = 0, phi = 0;
    for (int p = 0; p < L; p++) {
        int mn = pi[4 * p], mx = pi[4 * p];
        for (int k = 1; k < 4; k++) { if (pi[4 * p + k] < mn) mn = pi[4 * p + k]; if (pi[4 * p + k] > mx) mx = pi[4 * p + k]; }
        plo += mn < 0 ? mn : 0; phi += mx > 0 ? mx : 0;
    }
    (void)lo; (void)hi;
    long span = phi - plo + 1;
    float *cur = xmalloc((size_t)span * sizeof(float)), *nxt = xmalloc((size_t)span * sizeof(float));
    unsigned char *cp = xmalloc((size_t)span), *np = xmalloc((size_t)span);
    memset(cp, 0, (size_t)span);
    for (long i = 0; i < span; i++) cur[i] = 0.0f;
    for (int k = 0; k < 4; k++) {
        long key = pi[k] - plo;
        if (!cp[key]) { cp[key] = 1; cur[key] = 0.0f; }
        cur[key] += bg[k];
    }
    for (int pos = 1; pos < L; pos++) {
        memset(np, 0, (size_t)span);
        for (long i = 0; i < span; i++) nxt[i] = 0.0f;
        for (long i = 0; i < span; i++) {
            if (!cp[i]) continue;
            for (int k = 0; k < 4; k++) {
                long key = i + pi[4 * pos + k];
                if (!np[key]) { np[key] = 1; nxt[key] = 0.0f; }
                nxt[key] += cur[i] * bg[k];
            }
        }
        float *t = cur; cur = nxt; nxt = t;
        unsigned char *tp = cp; cp = np; np = tp;
    }
    /* histogram over [minS, maxS] with numBins bins (hist.cpp:250-253) */
    float width = (maxS - minS) / (float)numBins;
    for (size_t i = 0; i < numBins; i++) counts[i] = 0;
    float lastScore = -INFINITY; int have = 0;
    for (long i = 0; i < span; i++) {
        if (!cp[i]) continue;
        long key = i + plo;
        float score = ((float)key - (float)L * b) / a;
        /* map<float,float>::insert keeps the first entry for a repeated key;
         * keys arrive in ascending int order so a repeat is adjacent */
        if (have && score == lastScore) continue;
        have = 1; lastScore = score;
#ifdef ORACLE_MUTATE_HIST_COUNT_DOUBLE
        size_t cnt = (size_t)((double)maxLength * (double)cur[i]);   /* wrong: product in double */
#else
        float c = (float)maxLength * cur[i];
        size_t cnt = (size_t)c;
#endif
        int bin = (int)((score - minS) / width);
        if (bin < 0) bin = 0;
        if (bin > (int)numBins - 1) bin = (int)numBins - 1;
#ifdef ORACLE_MUTATE_HIST_ADD
        counts[bin] += cnt;
#else
        counts[bin] = cnt;   /* store, not add (motif.h:109-115) */
#endif
    }
    free(cur); free(nxt); free(cp); free(np); free(pi);
    *minScoreOut = minS; *maxScoreOut = maxS;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FASTA filtering (sequence.cpp:142-293)                                     */
/* ------------------------------------------------------------------------ */

static int valid_base(int c)
{
    return c == 'A' || c == 'a' || c == 'C' || c == 'c' || c == 'G' || c == 'g' || c == 'T' || c == 't';
}

int64_t oracle_filter_fasta(const char *const *files, int nfiles, uint64_t maxLen,
                            char **streamOut, uint64_t **runStartOut, uint64_t **runSeqOut,
                            uint64_t **runPosOut, uint64_t *nrunsOut, uint64_t *nseqOut,
                            char *errbuf, size_t errlen)
{
    size_t cap = 1 << 20, n = 0;
    char *stream = xmalloc(cap);
    size_t rcap = 64, nr = 0;
    uint64_t *rs = xmalloc(rcap * 8), *rq = xmalloc(rcap * 8), *rp = xmalloc(rcap * 8);
    uint64_t nseq = 0, curSeqLen = 0;
    /* last appended piece (for SeqBlock::isContiguous, sequence.cpp:35-50) */
    int havePiece = 0;
    uint64_t lastSeq = 0, lastEndPos = 0;
    int done = 0;
    for (int f = 0; f < nfiles && !done; f++) {
        FILE *fp = fopen(files[f], "rb");
        if (!fp) {
            snprintf(errbuf, errlen, "Could not open file: %s", files[f]);
            free(stream); free(rs); free(rq); free(rp);
            return -1;
        }
        char *line = NULL; size_t lcap = 0; ssize_t ll;
        while ((ll = getline(&line, &lcap, fp)) >= 0) {
            if (ll > 0 && line[ll - 1] == '\n') line[--ll] = 0;
            if (ll == 0) continue;                       /* sequence.cpp:152-153 */
            if (line[0] == '>') {                         /* sequence.cpp:156-163 */
                curSeqLen = 0;
                nseq++;
                continue;
            }
            if (nseq == 0) {                              /* sequence.cpp:169-170 */
                snprintf(errbuf, errlen, "Input file does not appear to be in fasta format\n");
                free(line); fclose(fp); free(stream); free(rs); free(rq); free(rp);
                return -1;
            }
            uint64_t seqIdx = nseq - 1, linePos = curSeqLen;
            curSeqLen += (uint64_t)ll;
            /* filterLine (sequence.cpp:181-214): split into valid runs */
            ssize_t i = 0;
            while (i < ll && !done) {
                while (i < ll && !valid_base((unsigned char)line[i])) i++;
                if (i >= ll) break;
                ssize_t b = i;
                while (i < ll && valid_base((unsigned char)line[i])) i++;
                uint64_t pieceLen = (uint64_t)(i - b), piecePos = linePos + (uint64_t)b;
                /* appendNextBlock truncation at maxFiltSeqLen (sequence.cpp:219-220) */
                if (n + pieceLen > maxLen) { pieceLen = maxLen - n; done = 1; }
                if (pieceLen == 0) break;
#ifdef ORACLE_MUTATE_RUNS_IGNORE_RECORD
                int contiguous = havePiece && lastEndPos == piecePos;
#else
                int contiguous = havePiece && lastSeq == seqIdx && lastEndPos == piecePos;
#endif
                if (!contiguous) {
                    if (nr == rcap) { rcap *= 2; rs = xrealloc(rs, rcap * 8); rq = xrealloc(rq, rcap * 8); rp = xrealloc(rp, rcap * 8); }
                    rs[nr] = n; rq[nr] = seqIdx; rp[nr] = piecePos; nr++;
                }
                while (n + pieceLen > cap) { cap *= 2; stream = xrealloc(stream, cap); }
                memcpy(stream + n, line + b, pieceLen);
                n += pieceLen;
                havePiece = 1; lastSeq = seqIdx; lastEndPos = piecePos + pieceLen;
                if (n >= maxLen) done = 1;
            }
        }
        free(line);
        fclose(fp);
    }
    *streamOut = stream; *runStartOut = rs; *runSeqOut = rq; *runPosOut = rp;
    *nrunsOut = nr; *nseqOut = nseq;
    return (int64_t)n;
}

/* ------------------------------------------------------------------------ */
/* scans                                                                     */
/* ------------------------------------------------------------------------ */

static int base_index(int ch)
{
    /* one-hot row of SeqMatrix::getNextSeqMatrix (sequence.cpp:312-319):
     * uppercase only; lowercase is a valid base with an all-zero row */
    switch (ch) { case 'A': return 0; case 'C': return 1; case 'G': return 2; case 'T': return 3; }
#ifdef ORACLE_MUTATE_LOWER_COUNTS
    switch (ch) { case 'a': return 0; case 'c': return 1; case 'g': return 2; case 't': return 3; }
#endif
    return 4;
}

typedef struct {
    const char *stream; uint64_t n;
    const uint64_t *runStart; uint64_t nruns;
    const float *pwm; int ldw; const int *len; const float *thr; int ncols;
    uint64_t q0, q1;
    oracle_hit *hits; size_t nh, cap;
} fp64_job;

static void push_hit(oracle_hit **h, size_t *n, size_t *cap, uint64_t pos, uint32_t col,
                     double s64, float s32)
{
    if (*n == *cap) { *cap = *cap ? *cap * 2 : 1024; *h = xrealloc(*h, *cap * sizeof(oracle_hit)); }
    oracle_hit *x = &(*h)[(*n)++];
    memset(x, 0, sizeof(*x));
    x->pos = pos; x->col = col; x->score64 = s64; x->score32 = s32;
}

/* first run start strictly greater than q (SeqBlock::getRemainingSeqLen,
 * sequence.cpp:68-79), or n */
static uint64_t next_break(const uint64_t *rs, uint64_t nr, uint64_t q, uint64_t n)
{
    uint64_t lo = 0, hi = nr;
    while (lo < hi) { uint64_t mid = (lo + hi) / 2; if (rs[mid] <= q) lo = mid + 1; else hi = mid; }
    return lo < nr ? rs[lo] : n;
}

static void *fp64_worker(void *arg)
{
    fp64_job *j = arg;
    for (uint64_t q = j->q0; q < j->q1; q++) {
        uint64_t rem = next_break(j->runStart, j->nruns, q, j->n) - q;
        for (int c = 0; c < j->ncols; c++) {
            int L = j->len[c];
#ifdef ORACLE_MUTATE_NO_RUN_CHECK
            if ((uint64_t)L > j->n - q) continue;
#else
            if ((uint64_t)L > rem) continue;          /* pwmscan.cpp:122-126 */
#endif
            const float *w = j->pwm + (size_t)c * (size_t)j->ldw * 4;
            double s = 0.0;
            for (int p = 0; p < L; p++) {
                int b = base_index((unsigned char)j->stream[q + (uint64_t)p]);
                if (b < 4) s += (double)w[4 * p + b];
            }
#ifdef ORACLE_MUTATE_GT
            if (s > (double)j->thr[c])
#else
            if (s >= (double)j->thr[c])                /* pwmscan.cpp:117-119 */
#endif
                push_hit(&j->hits, &j->nh, &j->cap, q, (uint32_t)c, s, (float)s);
        }
    }
    return NULL;
}

int64_t oracle_scan_fp64(const char *stream, uint64_t n, const uint64_t *runStart,
                         uint64_t nruns, const float *pwm, int ldw, const int *len,
                         const float *thr, int ncols, int nthreads, oracle_hit **hits)
{
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > n / 1024 + 1) nthreads = (int)(n / 1024 + 1);
    fp64_job *jobs = xmalloc((size_t)nthreads * sizeof(fp64_job));
    pthread_t *th = xmalloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; t++) {
        fp64_job *j = &jobs[t];
        memset(j, 0, sizeof(*j));
        j->stream = stream; j->n = n; j->runStart = runStart; j->nruns = nruns;
        j->pwm = pwm; j->ldw = ldw; j->len = len; j->thr = thr; j->ncols = ncols;
        j->q0 = n * (uint64_t)t / (uint64_t)nthreads;
        j->q1 = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, fp64_worker, j);
    }
    size_t total = 0;
    for (int t = 0; t < nthreads; t++) { pthread_join(th[t], NULL); total += jobs[t].nh; }
    oracle_hit *out = xmalloc((total ? total : 1) * sizeof(oracle_hit));
    size_t k = 0;
    for (int t = 0; t < nthreads; t++) {
        if (jobs[t].nh) memcpy(out + k, jobs[t].hits, jobs[t].nh * sizeof(oracle_hit));
        k += jobs[t].nh;
        free(jobs[t].hits);
    }
    free(jobs); free(th);
    *hits = out;
    return (int64_t)total;
}

/* ---- empirical histograms (hist -e) -------------------------------------- */

/* int(v) for a float as the reference's x86-64 hosts execute it (cvttss2si):
 * truncation toward zero, and INT_MIN for NaN or a value outside the int
 * range (ScoreHistogram::addObservation, motif.h:96-102). */
static int trunc_x86(float v)
{
    if (v >= -2147483648.0f && v < 2147483648.0f) return (int)v;
    return INT_MIN;
}

/* Bin of one observation (motif.h:96-102) with width = (max - min) / numBins
 * in float (motif.h:58-66). */
static int hist_bin(float score, float minScore, float width, size_t numBins)
{
#ifdef ORACLE_MUTATE_HIST_ROUND
    int b = (int)lrintf((score - minScore) / width);
#else
    int b = trunc_x86((score - minScore) / width);
#endif
    if (b < 0) b = 0;
    if (b > (int)numBins - 1) b = (int)numBins - 1;
    return b;
}

typedef struct {
    const char *stream; uint64_t n, qend;
    const uint64_t *runStart; uint64_t nruns;
    const float *pwm; int ldw; const int *len;
    const float *minScore, *maxScore;
    size_t numBins;
    uint64_t *counts;
    int c0, c1;
} ehist_job;

/* Histogram::extractObsScore (hist.cpp:70-91) over every offset of every
 * block (hist.cpp:131-136): each window that satisfies the run rule
 * (hist.cpp:83-86) adds its score to its motif's histogram.  The score is the
 * FP32 sum of the window's one-hot row times the PWM column, accumulated in
 * position order (the k order of an sgemm micro-kernel that keeps one
 * accumulator per output; the reference's BLAS summation order is not fixed
 * by the reference itself).  Lowercase bases contribute nothing. */
static void *ehist_worker(void *arg)
{
    ehist_job *j = arg;
    const uint64_t qe = j->qend < j->n ? j->qend : j->n;
    for (int c = j->c0; c < j->c1; c++) {
        const int L = j->len[c];
        const float *w = j->pwm + (size_t)c * (size_t)j->ldw * 4;
        const float width = (j->maxScore[c] - j->minScore[c]) / (float)j->numBins;
        uint64_t *cnt = j->counts + (size_t)c * j->numBins;
        uint64_t ri = 0;                                  /* first run start > q */
        for (uint64_t q = 0; q < qe; q++) {
            while (ri < j->nruns && j->runStart[ri] <= q) ri++;
            const uint64_t nb = ri < j->nruns ? j->runStart[ri] : j->n;
#ifdef ORACLE_MUTATE_HIST_NO_RUN_CHECK
            if ((uint64_t)L > j->n - q) continue;
#else
            if ((uint64_t)L > nb - q) continue;
#endif
#ifdef ORACLE_MUTATE_HIST_FP64
            double d = 0.0;
            for (int p = 0; p < L; p++) {
                int b = base_index((unsigned char)j->stream[q + (uint64_t)p]);
                if (b < 4) d += (double)w[4 * p + b];
            }
            const float s = (float)d;
#else
            float s = 0.0f;
            for (int p = 0; p < L; p++) {
                int b = base_index((unsigned char)j->stream[q + (uint64_t)p]);
                if (b < 4) s += w[4 * p + b];
            }
#endif
            cnt[hist_bin(s, j->minScore[c], width, j->numBins)]++;
        }
    }
    return NULL;
}

int oracle_empirical_hist(const char *stream, uint64_t n, const uint64_t *runStart,
                          uint64_t nruns, const float *pwm, int ldw, const int *len,
                          const float *minScore, const float *maxScore, int ncols,
                          size_t numBins, uint64_t qend, int nthreads, uint64_t *counts)
{
    if (numBins < 1 || ncols < 0) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ncols) nthreads = ncols > 0 ? ncols : 1;
    ehist_job *jobs = xmalloc((size_t)nthreads * sizeof(ehist_job));
    pthread_t *th = xmalloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; t++) {
        ehist_job *j = &jobs[t];
        j->stream = stream; j->n = n; j->qend = qend; j->runStart = runStart; j->nruns = nruns;
        j->pwm = pwm; j->ldw = ldw; j->len = len; j->minScore = minScore; j->maxScore = maxScore;
        j->numBins = numBins; j->counts = counts;
        j->c0 = (int)((int64_t)ncols * t / nthreads);
        j->c1 = (int)((int64_t)ncols * (t + 1) / nthreads);
        pthread_create(&th[t], NULL, ehist_worker, j);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(jobs); free(th);
    return 0;
}

/* ---- sgemm path ---------------------------------------------------------- */

typedef void (*sgemm_fn)(const char *, const char *, const int *, const int *, const int *,
                         const float *, const float *, const int *, const float *, const int *,
                         const float *, float *, const int *);
static sgemm_fn g_sgemm = NULL;

int oracle_blas_open(const char *path, char *errbuf, size_t errlen)
{
    if (g_sgemm) return 0;
    char found[4096] = {0};
    if (!path || !path[0]) {
        const char *env = getenv("BLAMM_ORACLE_BLAS");
        if (env && env[0]) snprintf(found, sizeof(found), "%s", env);
        else {
            glob_t g;
            const char *pats[] = {
                "/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
                "/usr/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
                "/usr/local/lib/python3*/site-packages/scipy.libs/libscipy_openblas*.so", NULL};
            for (int i = 0; pats[i] && !found[0]; i++) {
                if (glob(pats[i], 0, NULL, &g) == 0 && g.gl_pathc > 0)
                    snprintf(found, sizeof(found), "%s", g.gl_pathv[0]);
                globfree(&g);
            }
        }
    } else snprintf(found, sizeof(found), "%s", path);
    if (!found[0]) { snprintf(errbuf, errlen, "no BLAS library found"); return -1; }
    void *h = dlopen(found, RTLD_NOW | RTLD_LOCAL);
    if (!h) { snprintf(errbuf, errlen, "dlopen %s: %s", found, dlerror()); return -1; }
    const char *names[] = {"scipy_sgemm_", "sgemm_", NULL};
    for (int i = 0; names[i] && !g_sgemm; i++) g_sgemm = (sgemm_fn)dlsym(h, names[i]);
    if (!g_sgemm) { snprintf(errbuf, errlen, "no sgemm in %s", found); return -1; }
    /* BLAS threads = 1: blamm threads carry the parallelism (README.md:75-81) */
    void (*setnt)(int) = (void (*)(int))dlsym(h, "scipy_openblas_set_num_threads");
    if (!setnt) setnt = (void (*)(int))dlsym(h, "openblas_set_num_threads");
    if (setnt) setnt(1);
    return 0;
}

typedef struct { size_t colStart, colEnd, rowEnd; } tile_t;

/* generateMatrixTiles / findBestSplit / keepSplit (motif.cpp:451-540) */
static tile_t *make_tiles(const int *len, int ncols, size_t minZeroArea, size_t *ntOut)
{
    int Lmax = 0;
    for (int c = 0; c < ncols; c++) if (len[c] > Lmax) Lmax = len[c];
    size_t nt = 1;
    tile_t *t = xmalloc(sizeof(tile_t));
    t[0].colStart = 0; t[0].colEnd = (size_t)ncols; t[0].rowEnd = 4 * (size_t)Lmax;
    for (;;) {
        tile_t *nw = xmalloc(2 * nt * sizeof(tile_t));
        size_t nn = 0; int did = 0;
        for (size_t i = 0; i < nt; i++) {
            tile_t in = t[i];
            size_t bestZ = 0, bestJ = in.colStart + 1;
            for (size_t j = in.colStart + 1; j < in.colEnd; j++) {
                size_t numCols = j - in.colStart;
                size_t numRows = in.rowEnd - 4 * (size_t)len[j - 1];
                size_t z = numCols * numRows;
                if (z > bestZ) { bestZ = z; bestJ = j; }
            }
            tile_t left = {in.colStart, bestJ, 4 * (size_t)len[bestJ - 1]};
            tile_t right = {bestJ, in.colEnd, in.rowEnd};
            size_t zeroArea = (right.rowEnd - left.rowEnd) * (left.colEnd - left.colStart);
            if (zeroArea >= minZeroArea) { nw[nn++] = left; nw[nn++] = right; did = 1; }
            else nw[nn++] = in;
        }
        free(t); t = nw; nt = nn;
        if (!did) break;
    }
    *ntOut = nt;
    return t;
}

typedef struct {
    const char *stream; uint64_t n;
    const uint64_t *runStart; uint64_t nruns;
    const float *P; int ldp; const float *pwm; int ldw; const int *len; const float *thr; int ncols;
    const tile_t *tiles; size_t ntiles;
    int h, w, overlap;
    uint64_t nblocks;
    volatile uint64_t *nextBlock;
    pthread_mutex_t *mu;
    oracle_hit *hits; size_t nh, cap;
} sgemm_job;

static void *sgemm_worker(void *arg)
{
    sgemm_job *j = arg;
    const int h = j->h, w = j->w, ov = j->overlap;
    const size_t scols = 4 * (size_t)(w + ov);
    float *S = xmalloc((size_t)h * scols * sizeof(float));
    float *R = xmalloc((size_t)h * (size_t)j->ncols * sizeof(float));
    const uint64_t payload = (uint64_t)h * (uint64_t)w;
    for (;;) {
        pthread_mutex_lock(j->mu);
        uint64_t blk = *j->nextBlock;
        if (blk < j->nblocks) *j->nextBlock = blk + 1;
        pthread_mutex_unlock(j->mu);
        if (blk >= j->nblocks) break;
        uint64_t b0 = blk * payload;
        uint64_t bsz = j->n - b0;
        if (bsz > payload + (uint64_t)ov) bsz = payload + (uint64_t)ov;
        /* SeqMatrix::getNextSeqMatrix (sequence.cpp:299-340), column-major S */
        memset(S, 0, (size_t)h * scols * sizeof(float));
        size_t numOccRow = 0;
        for (int i = 0; i < h; i++) {
            int stop = 0;
            for (int c = 0; c < w; c++) {
                uint64_t bp = (uint64_t)i * (uint64_t)w + (uint64_t)c;
                if (bp >= bsz) { stop = 1; break; }
                int b = base_index((unsigned char)j->stream[b0 + bp]);
                if (b < 4) S[(size_t)(4 * c + b) * (size_t)h + (size_t)i] = 1.0f;
                numOccRow = (size_t)i + 1;
            }
            if (stop) break;
            for (int c = w; c < w + ov; c++) {
                uint64_t bp = (uint64_t)i * (uint64_t)w + (uint64_t)c;
                if (bp >= bsz) break;
                int b = base_index((unsigned char)j->stream[b0 + bp]);
                if (b < 4) S[(size_t)(4 * c + b) * (size_t)h + (size_t)i] = 1.0f;
                numOccRow = (size_t)i + 1;
            }
        }
        for (int off = 0; off < w; off++) {
            for (size_t t = 0; t < j->ntiles; t++) {   /* matrix.h:299-305 */
                int m = h, nn = (int)(j->tiles[t].colEnd - j->tiles[t].colStart);
                int k = (int)j->tiles[t].rowEnd, lda = h, ldb = j->ldp, ldc = h;
                float alpha = 1.0f, beta = 0.0f;
                g_sgemm("N", "N", &m, &nn, &k, &alpha, S + (size_t)4 * (size_t)off * (size_t)h, &lda,
                        j->P + j->tiles[t].colStart * (size_t)ldb, &ldb, &beta,
                        R + j->tiles[t].colStart * (size_t)ldc, &ldc);
            }
            /* extractOccurrences (pwmscan.cpp:104-133) */
            for (int c = 0; c < j->ncols; c++) {
                float th = j->thr[c];
                for (size_t i = 0; i < numOccRow; i++) {
                    float sc = R[(size_t)c * (size_t)h + i];
                    if (sc < th) continue;
                    uint64_t bp = (uint64_t)i * (uint64_t)w + (uint64_t)off;
                    uint64_t rem = 0;
                    if (bp < bsz) {
                        uint64_t nb = next_break(j->runStart, j->nruns, b0 + bp, j->n);
                        if (nb > b0 + bsz) nb = b0 + bsz;
                        rem = nb - (b0 + bp);
                    }
                    if ((uint64_t)j->len[c] > rem) continue;
                    const float *wt = j->pwm + (size_t)c * (size_t)j->ldw * 4;
                    double s = 0.0;
                    for (int p = 0; p < j->len[c]; p++) {
                        int b = base_index((unsigned char)j->stream[b0 + bp + (uint64_t)p]);
                        if (b < 4) s += (double)wt[4 * p + b];
                    }
                    push_hit(&j->hits, &j->nh, &j->cap, b0 + bp, (uint32_t)c, s, sc);
                }
            }
        }
    }
    free(S); free(R);
    return NULL;
}

int64_t oracle_scan_sgemm(const char *stream, uint64_t n, const uint64_t *runStart,
                          uint64_t nruns, const float *pwm, int ldw, const int *len,
                          const float *thr, int ncols, int h, int w,
                          size_t minZeroArea, int nthreads, oracle_hit **hits)
{
    if (!g_sgemm) return -1;
    for (int c = 1; c < ncols; c++) if (len[c] < len[c - 1]) return -2;
    int Lmax = 0;
    for (int c = 0; c < ncols; c++) if (len[c] > Lmax) Lmax = len[c];
    int ldp = 4 * Lmax;
    /* generateMatrix (motif.cpp:542-564): P is (4 Lmax) x ncols, column-major */
    float *P = xmalloc((size_t)ldp * (size_t)ncols * sizeof(float));
    memset(P, 0, (size_t)ldp * (size_t)ncols * sizeof(float));
    for (int c = 0; c < ncols; c++)
        for (int p = 0; p < len[c]; p++)
            for (int o = 0; o < 4; o++)
                P[(size_t)c * (size_t)ldp + (size_t)(4 * p + o)] = pwm[(size_t)c * (size_t)ldw * 4 + (size_t)(4 * p + o)];
    size_t ntiles;
    tile_t *tiles = make_tiles(len, ncols, minZeroArea, &ntiles);
    uint64_t payload = (uint64_t)h * (uint64_t)w;
    uint64_t nblocks = (n + payload - 1) / payload;
    if (nthreads < 1) nthreads = 1;
    volatile uint64_t next = 0;
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    sgemm_job *jobs = xmalloc((size_t)nthreads * sizeof(sgemm_job));
    pthread_t *th = xmalloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; t++) {
        sgemm_job *j = &jobs[t];
        memset(j, 0, sizeof(*j));
        j->stream = stream; j->n = n; j->runStart = runStart; j->nruns = nruns;
        j->P = P; j->ldp = ldp; j->pwm = pwm; j->ldw = ldw; j->len = len; j->thr = thr;
        j->ncols = ncols; j->tiles = tiles; j->ntiles = ntiles; j->h = h; j->w = w;
        j->overlap = Lmax - 1; j->nblocks = nblocks; j->nextBlock = &next; j->mu = &mu;
        pthread_create(&th[t], NULL, sgemm_worker, j);
    }
    size_t total = 0;
    for (int t = 0; t < nthreads; t++) { pthread_join(th[t], NULL); total += jobs[t].nh; }
    oracle_hit *out = xmalloc((total ? total : 1) * sizeof(oracle_hit));
    size_t k = 0;
    for (int t = 0; t < nthreads; t++) {
        if (jobs[t].nh) memcpy(out + k, jobs[t].hits, jobs[t].nh * sizeof(oracle_hit));
        k += jobs[t].nh;
        free(jobs[t].hits);
    }
    free(jobs); free(th); free(tiles); free(P);
    *hits = out;
    return (int64_t)total;
}

/* ------------------------------------------------------------------------ */
/* .dict (species.cpp:84-143, 176-207)                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    char *name;
    char **files; size_t nfiles;
    uint64_t totSeqLen;
    uint64_t counts[4];
    char **seqNames; size_t nseq;
} species_t;

static int cmp_str(const void *a, const void *b) { return strcmp(*(char *const *)a, *(char *const *)b); }

/* insert into a sorted unique set of strings (std::set<std::string>) */
static void set_insert(char ***v, size_t *n, const char *s)
{
    for (size_t i = 0; i < *n; i++) if (!strcmp((*v)[i], s)) return;
    *v = xrealloc(*v, (*n + 1) * sizeof(char *));
    (*v)[(*n)++] = xstrdup(s);
#ifndef ORACLE_MUTATE_DICT_FILE_ORDER   /* wrong when defined: files kept in manifest order */
    qsort(*v, *n, sizeof(char *), cmp_str);
#endif
}

static species_t *load_dict(const char *fn, size_t *nout)
{
    size_t len;
    char *txt = read_file(fn, &len);
    if (!txt) die("Cannot open file: %s", fn);
    cursor c = {txt, len, 0, 0, 0};
    char *tok;
#define NEXT() (cur_token(&c, &tok) ? tok : (die("oracle: truncated dict %s", fn), (char *)NULL))
    free(NEXT());
    size_t ns = strtoull(NEXT(), NULL, 10); free(tok);
    species_t *sp = xmalloc((ns ? ns : 1) * sizeof(species_t));
    for (size_t s = 0; s < ns; s++) {
        species_t *x = &sp[s];
        memset(x, 0, sizeof(*x));
        free(NEXT()); x->name = NEXT();
        free(NEXT()); size_t nf = strtoull(NEXT(), NULL, 10); free(tok);
        for (size_t f = 0; f < nf; f++) { set_insert(&x->files, &x->nfiles, NEXT()); free(tok); }
        free(NEXT()); x->totSeqLen = strtoull(NEXT(), NULL, 10); free(tok);
        free(NEXT());
        for (int k = 0; k < 4; k++) { x->counts[k] = strtoull(NEXT(), NULL, 10); free(tok); }
        free(NEXT()); x->nseq = strtoull(NEXT(), NULL, 10); free(tok);
        x->seqNames = xmalloc((x->nseq ? x->nseq : 1) * sizeof(char *));
        for (size_t q = 0; q < x->nseq; q++) x->seqNames[q] = NEXT();
    }
#undef NEXT
    free(txt);
    *nout = ns;
    return sp;
}

/* ------------------------------------------------------------------------ */
/* command line                                                              */
/* ------------------------------------------------------------------------ */

static int cmd_dict(int argc, char **argv)
{
    if (argc < 3) die("Usage: blamm_oracle dict sequences.mf");
    const char *mf = argv[argc - 1];
    char dictfn[4096];
    snprintf(dictfn, sizeof(dictfn), "%s.dict", mf);
    size_t len;
    char *txt = read_file(mf, &len);
    if (!txt) die("Could not open file: %s", mf);
    species_t *sp = NULL; size_t ns = 0;
    cursor c = {txt, len, 0, 0, 0};
    char *line;
    while (cur_getline(&c, &line)) {   /* dict.cpp:86-106 */
        if (!line[0]) { free(line); continue; }
        char name[4096] = {0}, file[4096] = {0};
        if (sscanf(line, "%4095s %4095s", name, file) < 2)
            die("File %s has incorrect format.", mf);
        FILE *t = fopen(file, "rb");
        if (!t) die("Could not open fasta file: %s", file);
        fclose(t);
        size_t i;
        for (i = 0; i < ns; i++) if (!strcmp(sp[i].name, name)) break;
        if (i == ns) {
            sp = xrealloc(sp, (ns + 1) * sizeof(species_t));
            memset(&sp[ns], 0, sizeof(species_t));
            sp[ns].name = xstrdup(name);
            ns++;
        }
        set_insert(&sp[i].files, &sp[i].nfiles, file);
        free(line);
    }
    free(txt);
    FILE *o = fopen(dictfn, "wb");
    if (!o) die("Cannot write to file: %s", dictfn);
    fprintf(o, "NUM_SPECIES\t%zu\n", ns);
    for (size_t s = 0; s < ns; s++) {
        /* computeNucleotideComposition (species.cpp:32-61) */
        char *stream; uint64_t *rs, *rq, *rp, nr, nseq; char err[512];
        int64_t n = oracle_filter_fasta((const char *const *)sp[s].files, (int)sp[s].nfiles,
                                        UINT64_MAX, &stream, &rs, &rq, &rp, &nr, &nseq, err, sizeof(err));
        if (n < 0) die("%s", err);
        uint64_t cnt[4] = {0, 0, 0, 0};
        for (int64_t i = 0; i < n; i++) {
#ifdef ORACLE_MUTATE_DICT_UPPER_ONLY
            int ch = (unsigned char)stream[i];   /* wrong: lowercase bases not counted */
            if (ch >= 'a') continue;
#else
            int ch = toupper((unsigned char)stream[i]);
#endif
            cnt[ch == 'A' ? 0 : ch == 'C' ? 1 : ch == 'G' ? 2 : 3]++;
        }
        /* sequence names: first token after '>' (sequence.cpp:155-161) */
        fprintf(o, "SPECIES\t%s\n", sp[s].name);
        fprintf(o, "NUM_FASTA_FILES\t%zu\n", sp[s].nfiles);
        for (size_t f = 0; f < sp[s].nfiles; f++) fprintf(o, "%s\n", sp[s].files[f]);
        fprintf(o, "TOT_SEQ_LENGTH\t%llu\n", (unsigned long long)(cnt[0] + cnt[1] + cnt[2] + cnt[3]));
        fprintf(o, "NUCL_COUNT_ACGT\t%llu\t%llu\t%llu\t%llu\n", (unsigned long long)cnt[0],
                (unsigned long long)cnt[1], (unsigned long long)cnt[2], (unsigned long long)cnt[3]);
        fprintf(o, "NUM_SEQUENCES\t%llu\n", (unsigned long long)nseq);
        for (size_t f = 0; f < sp[s].nfiles; f++) {
            FILE *fp = fopen(sp[s].files[f], "rb");
            char *ln = NULL; size_t lc = 0; ssize_t ll;
            while ((ll = getline(&ln, &lc, fp)) >= 0) {
                if (ll > 0 && ln[ll - 1] == '\n') ln[--ll] = 0;
                if (ll == 0 || ln[0] != '>') continue;
                char nm[4096] = {0};
                if (sscanf(ln + 1, "%4095s", nm) < 1) nm[0] = 0;
                fprintf(o, "%s\n", nm);
            }
            free(ln); fclose(fp);
        }
        free(stream); free(rs); free(rq); free(rp);
    }
    fclose(o);
    return 0;
}

static int cmd_hist(int argc, char **argv)
{
    if (argc < 4) die("Usage: blamm_oracle hist [options] motifs.input sequences.input");
    size_t maxLength = 10000000, numBins = 250;
    char histdir[4096] = "";
    int empirical = 0;
    for (int i = 2; i < argc - 2; i++) {   /* hist.cpp:187-215 */
        const char *a = argv[i];
        if ((!strcmp(a, "-l") || !strcmp(a, "--length")) && i + 1 < argc - 2) maxLength = strtoull(argv[++i], NULL, 10);
        else if ((!strcmp(a, "-b") || !strcmp(a, "--numbins")) && i + 1 < argc - 2) { numBins = strtoull(argv[++i], NULL, 10); if (numBins < 2) numBins = 2; }
        else if ((!strcmp(a, "-t") || !strcmp(a, "--numthreads")) && i + 1 < argc - 2) i++;
        else if ((!strcmp(a, "-H") || !strcmp(a, "--histdir")) && i + 1 < argc - 2) {
            snprintf(histdir, sizeof(histdir), "%s", argv[++i]);
            size_t l = strlen(histdir);
            if (l == 0 || histdir[l - 1] != '/') strcat(histdir, "/");
        } else if (!strcmp(a, "-e") || !strcmp(a, "--empirical")) empirical = 1;
        else die("Usage: blamm_oracle hist [options] motifs.input sequences.input");
    }
    settings_t st; settings_load(&st);
    char dictfn[4096];
    snprintf(dictfn, sizeof(dictfn), "%s.dict", argv[argc - 1]);
    size_t ns; species_t *sp = load_dict(dictfn, &ns);
    size_t nm; motif_t *m = load_motifs(argv[argc - 2], 0, &nm);
    uint64_t *counts = xmalloc(numBins * sizeof(uint64_t));
    int Lmax = 1;
    for (size_t i = 0; i < nm; i++) if (m[i].L > Lmax) Lmax = m[i].L;
    for (size_t s = 0; s < ns; s++) {
        unsigned long long bsum = sp[s].counts[0] + sp[s].counts[1] + sp[s].counts[2] + sp[s].counts[3];
        float bgTot = (float)bsum; bgTot += 4.0f * st.pseudocount;
        float bg[4];
        for (int k = 0; k < 4; k++) bg[k] = ((float)sp[s].counts[k] + st.pseudocount) / bgTot;
        /* empirical: every motif's histogram over the group's first maxLength
         * valid bases (hist.cpp:140-160) */
        uint64_t *ecounts = NULL;
        float *emin = NULL, *emax = NULL;
        if (empirical) {
            float *P = xmalloc((nm ? nm : 1) * (size_t)Lmax * 4 * sizeof(float));
            int *len = xmalloc((nm ? nm : 1) * sizeof(int));
            emin = xmalloc((nm ? nm : 1) * sizeof(float));
            emax = xmalloc((nm ? nm : 1) * sizeof(float));
            memset(P, 0, (nm ? nm : 1) * (size_t)Lmax * 4 * sizeof(float));
            for (size_t i = 0; i < nm; i++) {
                float *w = P + i * (size_t)Lmax * 4;
                oracle_pfm2pwm(m[i].pfm, m[i].L, sp[s].counts, st.pseudocount, 0, w);
                len[i] = m[i].L;
                emin[i] = oracle_min_score(w, m[i].L);
                emax[i] = oracle_max_score(w, m[i].L);
            }
            char *stream; uint64_t *rs, *rq, *rp, nr, nseq; char err[1024];
            int64_t n = oracle_filter_fasta((const char *const *)sp[s].files, (int)sp[s].nfiles, maxLength,
                                            &stream, &rs, &rq, &rp, &nr, &nseq, err, sizeof(err));
            if (n < 0) die("%s", err);
            ecounts = xmalloc((nm ? nm : 1) * numBins * sizeof(uint64_t));
            memset(ecounts, 0, (nm ? nm : 1) * numBins * sizeof(uint64_t));
            oracle_empirical_hist(stream, (uint64_t)n, rs, nr, P, Lmax, len, emin, emax, (int)nm, numBins,
                                  (uint64_t)n, (int)sysconf(_SC_NPROCESSORS_ONLN), ecounts);
            free(stream); free(rs); free(rq); free(rp); free(P); free(len);
        }
        for (size_t i = 0; i < nm; i++) {
            float *pwm = xmalloc((size_t)m[i].L * 4 * sizeof(float));
            oracle_pfm2pwm(m[i].pfm, m[i].L, sp[s].counts, st.pseudocount, 0, pwm);
            float mn, mx;
            if (empirical) {
                mn = emin[i]; mx = emax[i];
                memcpy(counts, ecounts + i * numBins, numBins * sizeof(uint64_t));
            } else {
                oracle_theoretical_hist(pwm, m[i].L, bg, numBins, maxLength, counts, &mn, &mx);
            }
            char fn[8192];
            snprintf(fn, sizeof(fn), "%shist_%s_%s.dat", histdir, sp[s].name, m[i].name);
            FILE *o = fopen(fn, "wb");
            if (!o) die("Error: cannot write to file %s", fn);
            float width = (mx - mn) / (float)numBins;   /* motif.cpp:71-84 */
            fprintf(o, "%zu\t%g\t%g\n", numBins, (double)mn, (double)mx);
            for (size_t b = 0; b < numBins; b++)
                fprintf(o, "%g\t%llu\n", (0.5 + (double)b) * (double)width + (double)mn, (unsigned long long)counts[b]);
            fclose(o);
            free(pwm);
        }
        free(ecounts); free(emin); free(emax);
    }
    free(counts);
    return 0;
}

static int cmp_hit(const void *a, const void *b)
{
    const oracle_hit *x = a, *y = b;
    if (x->pos != y->pos) return x->pos < y->pos ? -1 : 1;
    return (int)x->col - (int)y->col;
}

static int cmd_scan(int argc, char **argv)
{
    if (argc < 4) die("Usage: blamm_oracle scan [options] motifs.input sequences.input");
    int revCompl = 0, absSpec = 0, relSpec = 0, pSpec = 0, nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    float absTh = 0.0f, relTh = 0.95f, pvalue = 0.0001f;
    char histdir[4096] = "", outfn[4096] = "occurrences.txt", mode[32] = "fp64", blas[4096] = "";
    for (int i = 2; i < argc - 2; i++) {   /* pwmscan.cpp:469-514 */
        const char *a = argv[i];
        if (!strcmp(a, "-rc") || !strcmp(a, "--revcompl")) revCompl = 1;
        else if (!strcmp(a, "-c") || !strcmp(a, "--cuda")) ;
        else if ((!strcmp(a, "-at") || !strcmp(a, "--absthreshold")) && i + 1 < argc - 2) { absSpec = 1; absTh = (float)atof(argv[++i]); }
        else if ((!strcmp(a, "-rt") || !strcmp(a, "--relthreshold")) && i + 1 < argc - 2) {
            relSpec = 1; relTh = (float)atof(argv[++i]);
            if (relTh < 0.0f || relTh > 1.0f) die("The relative threshold should be in range [0..1].");
        } else if ((!strcmp(a, "-pt") || !strcmp(a, "--pthreshold")) && i + 1 < argc - 2) {
            pSpec = 1; pvalue = (float)atof(argv[++i]);
            if (pvalue < 0.0f || pvalue > 1.0f) die("The p-value should be in range [0..1].");
        } else if ((!strcmp(a, "-t") || !strcmp(a, "--numthreads")) && i + 1 < argc - 2) {
            nthreads = atoi(argv[++i]);
            if (nthreads < 1) die("Number of threads must be a non-zero positive number");
        } else if ((!strcmp(a, "-H") || !strcmp(a, "--histdir")) && i + 1 < argc - 2) {
            snprintf(histdir, sizeof(histdir), "%s", argv[++i]);
            size_t l = strlen(histdir);
            if (l == 0 || histdir[l - 1] != '/') strcat(histdir, "/");
        } else if ((!strcmp(a, "-o") || !strcmp(a, "--output")) && i + 1 < argc - 2) snprintf(outfn, sizeof(outfn), "%s", argv[++i]);
        else if (!strcmp(a, "--mode") && i + 1 < argc - 2) snprintf(mode, sizeof(mode), "%s", argv[++i]);
        else if (!strcmp(a, "--blas") && i + 1 < argc - 2) snprintf(blas, sizeof(blas), "%s", argv[++i]);
        else die("Usage: blamm_oracle scan [options] motifs.input sequences.input");
    }
    if (!(absSpec || relSpec || pSpec)) relSpec = 1;
    if (absSpec && relSpec) die("Specify either the absolute or relative threshold, not both.");
    if (absSpec && pSpec) die("Specify either the absolute or p-value threshold, not both.");
    if (relSpec && pSpec) die("Specify either the relative or p-value threshold, not both.");
    int sgemm = !strcmp(mode, "sgemm");
    if (sgemm) {
        char err[1024];
        if (oracle_blas_open(blas, err, sizeof(err))) die("oracle: %s", err);
    }
    settings_t st; settings_load(&st);
    char dictfn[4096];
    snprintf(dictfn, sizeof(dictfn), "%s.dict", argv[argc - 1]);
    size_t ns; species_t *sp = load_dict(dictfn, &ns);
    size_t nm; motif_t *m = load_motifs(argv[argc - 2], 1, &nm);
    if (revCompl) { m = add_rc(m, nm); nm *= 2; }
    int Lmax = 0;
    for (size_t i = 0; i < nm; i++) if (m[i].L > Lmax) Lmax = m[i].L;
    int ncols = (int)nm;
    float *pwm = xmalloc((size_t)ncols * (size_t)(Lmax ? Lmax : 1) * 4 * sizeof(float));
    int *len = xmalloc((size_t)(ncols ? ncols : 1) * sizeof(int));
    float *thr = xmalloc((size_t)(ncols ? ncols : 1) * sizeof(float));
    FILE *cut = fopen("PWMthresholds.txt", "wb");
    FILE *out = fopen(outfn, "wb");
    if (!cut || !out) die("oracle: cannot open output files");
    uint64_t total = 0;
    for (size_t s = 0; s < ns; s++) {
        memset(pwm, 0, (size_t)ncols * (size_t)(Lmax ? Lmax : 1) * 4 * sizeof(float));
        for (int c = 0; c < ncols; c++) {
            len[c] = m[c].L;
            oracle_pfm2pwm(m[c].pfm, m[c].L, sp[s].counts, st.pseudocount, m[c].rc,
                           pwm + (size_t)c * (size_t)Lmax * 4);
        }
        for (int c = 0; c < ncols; c++) {   /* pwmscan.cpp:599-616 */
            const float *w = pwm + (size_t)c * (size_t)Lmax * 4;
            if (absSpec) thr[c] = absTh;
            else if (relSpec) {
                float mx = oracle_max_score(w, m[c].L), mn = oracle_min_score(w, m[c].L);
                thr[c] = relTh * (mx - mn) + mn;
            } else {
                char *bn = base_name(m[c].name);
                char fn[8192];
                snprintf(fn, sizeof(fn), "%shist_%s_%s.dat", histdir, sp[s].name, bn);
                free(bn);
                uint64_t *counts; size_t nb; float hmn, hmx;
                int r = load_hist(fn, &counts, &nb, &hmn, &hmx);
                if (r == -1) die("Error: cannot read file %s. Did you run the hist module?", fn);
                if (r == -2) die("Unexpected end-of-file reached");
                thr[c] = oracle_score_cutoff(counts, nb, hmn, hmx, pvalue);
                free(counts);
            }
        }
        for (int c = 0; c < ncols; c++) {   /* pwmscan.cpp:619-625 */
            if (m[c].L > 15) continue;
            const float *w = pwm + (size_t)c * (size_t)Lmax * 4;
            fprintf(cut, "%s\t%s\t%g\t%g\t%g\n", sp[s].name, m[c].name,
                    (double)oracle_min_score(w, m[c].L), (double)thr[c], (double)oracle_max_score(w, m[c].L));
        }
        char *stream; uint64_t *rs, *rq, *rp, nr, nseq; char err[1024];
        int64_t n = oracle_filter_fasta((const char *const *)sp[s].files, (int)sp[s].nfiles,
                                        sp[s].totSeqLen, &stream, &rs, &rq, &rp, &nr, &nseq, err, sizeof(err));
        if (n < 0) die("%s", err);
        oracle_hit *hits = NULL;
        int64_t nh = sgemm
            ? oracle_scan_sgemm(stream, (uint64_t)n, rs, nr, pwm, Lmax, len, thr, ncols,
                                (int)st.S_h, (int)st.S_w, st.minZeroArea, nthreads, &hits)
            : oracle_scan_fp64(stream, (uint64_t)n, rs, nr, pwm, Lmax, len, thr, ncols, nthreads, &hits);
        if (nh < 0) die("oracle: scan failed (%lld)", (long long)nh);
        qsort(hits, (size_t)nh, sizeof(oracle_hit), cmp_hit);
        for (int64_t i = 0; i < nh; i++) {   /* writeOccToDisk (pwmscan.cpp:81-102) */
            uint64_t q = hits[i].pos;
            uint64_t lo = 0, hi = nr;
            while (lo < hi) { uint64_t mid = (lo + hi) / 2; if (rs[mid] <= q) lo = mid + 1; else hi = mid; }
            uint64_t r = lo - 1;
            uint64_t seqIdx = rq[r], seqPos = rp[r] + (q - rs[r]);
            if (seqIdx >= sp[s].nseq) die("oracle: sequence index beyond the dictionary");
            int c = (int)hits[i].col;
            fprintf(out, "%s\tblamm\t%s\t%llu\t%llu\t%g\t%c\t.\t.\n", sp[s].seqNames[seqIdx], m[c].name,
                    (unsigned long long)seqPos, (unsigned long long)(seqPos + (uint64_t)m[c].L),
                    (double)hits[i].score32, m[c].rc ? '-' : '+');
        }
        total += (uint64_t)nh;
        free(hits); free(stream); free(rs); free(rq); free(rp);
    }
    fclose(out); fclose(cut);
    printf("\nWrote %llu matches to %s.\n", (unsigned long long)total, outfn);
    return 0;
}

int oracle_main(int argc, char **argv)
{
    if (argc < 2) die("Usage: blamm_oracle dict|hist|scan [options] ...");
    if (!strcmp(argv[1], "dict")) return cmd_dict(argc, argv);
    if (!strcmp(argv[1], "hist")) return cmd_hist(argc, argv);
    if (!strcmp(argv[1], "scan")) return cmd_scan(argc, argv);
    die("Usage: blamm_oracle dict|hist|scan [options] ...");
    return 1;
}

#ifdef ORACLE_MAIN
int main(int argc, char **argv) { return oracle_main(argc, argv); }
#endif
