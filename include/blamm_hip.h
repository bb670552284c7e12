/*
 * blamm_hip.h -- C ABI of libblamm_hip.so, the MI355X scan engine behind
 * `blamm scan`.
 *
 * It replaces the reference's CUDA branch (compiled under HAVE_CUDA,
 * CMakeLists.txt:57-72, selected by -c/--cuda in pwmscan.cpp:479-480):
 *   - the per-thread device setup and uploads of scanThreadCUBLAS
 *     (pwmscan.cpp:297-357: cudaSetDevice, cudaMalloc of P/S/R/thresholds,
 *     cublasSetVector(P), cublasSetVector(threshold)),
 *   - the offset loop of cublasSgemm calls (matrix.h:314-323, pwmscan.cpp:385-389),
 *   - the threshold/compaction kernel kernel_wrapper/filterScore
 *     (kernel.cu:21-45, declared at pwmscan.cpp:41),
 *   - the blocking hit read-back (pwmscan.cpp:393-406).
 *
 * Plain pointers and sizes only.  Nothing throws across this boundary: every
 * entry point returns an int status (BLAMM_OK = 0, negative on error) and
 * blamm_hip_last_error() gives the calling thread's last message.
 * One context per host thread (thread-compatible, not thread-safe).
 *
 * Sequence layout ("packed stream"): the filtered stream of a manifest group
 * (only ACGTacgt bytes, runs concatenated, sequence.cpp:181-250) as 2-bit codes
 * A=0 C=1 G=2 T=3, base i in bits 2(i%32)..2(i%32)+1 of 64-bit word i/32.
 * Lowercase bases carry their code plus bit i%64 of lcmask word i/64; the
 * reference gives them an all-zero one-hot row (sequence.cpp:312-319), so they
 * contribute 0.  run_starts lists the stream positions where a new contiguous
 * run begins (a new record or a gap of invalid bytes, SeqBlock::append,
 * sequence.cpp:81-88); a window at q for a motif of length L is scored only if
 * q + L <= the first run start after q (or nbases) (pwmscan.cpp:122-126).
 */
#ifndef BLAMM_HIP_H
#define BLAMM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLAMM_HIP_ABI_VERSION 2

/* Longest motif the engine scores (positions); longer motifs are rejected by
 * blamm_hip_set_library with BLAMM_ERR_ARG. */
#define BLAMM_MAX_MOTIF_LEN 64

enum blamm_status {
    BLAMM_OK = 0,
    BLAMM_ERR_ARG = -1,       /* invalid argument */
    BLAMM_ERR_DEVICE = -2,    /* HIP runtime error / no device */
    BLAMM_ERR_NOMEM = -3,     /* device allocation failed */
    BLAMM_ERR_OVERFLOW = -4,  /* hit capacity too small: nothing was dropped,
                                 *nhits holds the required capacity */
    BLAMM_ERR_STATE = -5      /* call out of order (e.g. scan before library) */
};

enum blamm_precision {
    BLAMM_PREC_FP32 = 0,      /* FP32 MFMA, FP32 accumulate */
    BLAMM_PREC_BF16X2 = 1,    /* PWM split into bf16 hi+lo; the filter multiplies the hi
                                 terms on BF16 MFMA (FP32 accumulate) and keeps every
                                 window the proven hi-only band cannot rule out */
    BLAMM_PREC_BF16X2_LO = 2  /* PWM split into bf16 hi+lo; tile blocks the hi terms do
                                 not rule out also get their lo products on BF16 MFMA */
};

/* One occurrence.  pos: position in the packed stream of the scan call;
 * col: PWM column in the caller's order; score: the window's PWM score.
 * Replaces the (occIdx, occScore) pair of kernel.cu:30-31. */
typedef struct blamm_hit {
    uint64_t pos;
    uint32_t col;
    float score;
} blamm_hit;

/* Counters of the last completed scan on a context. */
typedef struct blamm_scan_stats {
    uint64_t candidates;      /* windows passing the MFMA filter band */
    uint64_t band_rescored;   /* windows rescored in FP64 */
    uint64_t hits;            /* occurrences reported */
    double kernel_ms;         /* device time of the scan kernels (HIP events) */
    double filter_ms;         /* of which: the MFMA filter kernels */
} blamm_scan_stats;

typedef struct blamm_hip_ctx blamm_hip_ctx;

int blamm_hip_abi_version(void);

/* Thread-local text of the last error on the calling thread ("" if none). */
const char *blamm_hip_last_error(void);

/* Number of visible HIP devices (replaces cudaGetDeviceCount, pwmscan.cpp:563). */
int blamm_hip_device_count(int *count);

/* Create / destroy a context on one device (replaces cudaSetDevice +
 * cublasCreate, pwmscan.cpp:309-313, and the cudaFree/cublasDestroy teardown,
 * pwmscan.cpp:428-436). */
int blamm_hip_create(int device, blamm_hip_ctx **ctx);
int blamm_hip_destroy(blamm_hip_ctx *ctx);

/* Upload one group's pattern matrix and thresholds; they stay resident until
 * replaced (BLAMM_ERR_STATE while a scan_device is in flight).  Replaces
 * cudaMalloc+cublasSetVector of P and of the per-column
 * thresholds (pwmscan.cpp:315-342).
 *   P:          column-major, ldp >= 4*max(lengths); P[c*ldp + 4*p + b] is the
 *               weight of base b (ACGT) at motif position p of column c
 *               (MotifContainer::generateMatrix, motif.cpp:542-564)
 *   lengths:    motif length of every column (1..BLAMM_MAX_MOTIF_LEN)
 *   thresholds: per-column score threshold; a window is an occurrence iff its
 *               score >= threshold (kernel.cu:29, pwmscan.cpp:117-119).  NaN
 *               is rejected: pass +INFINITY for a column that must report
 *               nothing (the reference's R >= NaN is never true)
 *   precision:  enum blamm_precision
 * Every weight must be finite.  The hit set is that of a left-to-right FP64
 * sum of the window's weights compared with (double)threshold in every
 * precision; the filter bands that make it so are per column and proven. */
int blamm_hip_set_library(blamm_hip_ctx *ctx, const float *P, int ldp, int ncols,
                          const int *lengths, const float *thresholds, int precision);

/* Largest filter band (MFMA score vs FP64) and largest reported-score error
 * bound over the resident library's columns. */
int blamm_hip_library_bounds(blamm_hip_ctx *ctx, double *max_filter_band,
                             double *max_score_error);

/* Words a packed-stream / lowercase-mask buffer must hold for nbases (the
 * engine reads a few words of zero padding past the end). */
uint64_t blamm_hip_packed_words(uint64_t nbases);
uint64_t blamm_hip_mask_words(uint64_t nbases);

/* Reserve room for n filter candidates (device scratch) and, for the
 * host-buffer scans, n hits in the context's device hit buffer (16 bytes
 * each, both).  Scans grow them on demand; reserving up front keeps device
 * allocation out of later calls and lets a first scan that fits run once.
 * BLAMM_ERR_STATE while a scan_device is in flight. */
int blamm_hip_reserve(blamm_hip_ctx *ctx, uint64_t n);

/* Scan host buffers.  packed has blamm_hip_packed_words(nbases) words; lcmask
 * is NULL when the stream has no lowercase base, else it has
 * blamm_hip_mask_words(nbases) words; run_starts are sorted.  Hits (unordered)
 * go to hits[0..cap); *nhits receives their number.  If cap is too small the
 * call returns BLAMM_ERR_OVERFLOW with *nhits = required capacity.
 * nbases < 2^32 - 512 per call. */
int blamm_hip_scan(blamm_hip_ctx *ctx, const uint64_t *packed, const uint64_t *lcmask,
                   uint64_t nbases, const uint64_t *run_starts, uint64_t nruns,
                   blamm_hit *hits, uint64_t cap, uint64_t *nhits);

/* As blamm_hip_scan, but the hits stay in a pinned buffer the context owns
 * (read back at full PCIe rate, no copy into caller memory): *hits points to
 * *nhits records, valid until the next call on this context. */
int blamm_hip_scan_pinned(blamm_hip_ctx *ctx, const uint64_t *packed, const uint64_t *lcmask,
                          uint64_t nbases, const uint64_t *run_starts, uint64_t nruns,
                          const blamm_hit **hits, uint64_t *nhits);

/* Enqueue a scan of device-resident buffers on `stream` (a hipStream_t, or
 * NULL for the context's own stream) without synchronising.  Buffer sizes are
 * passed so they can be checked.  The hit count is written to d_nhits
 * (device uint64).  Call blamm_hip_scan_finish() to synchronise and to learn
 * whether a capacity was exceeded. */
int blamm_hip_scan_device(blamm_hip_ctx *ctx, const uint64_t *d_packed, uint64_t packed_words,
                          const uint64_t *d_lcmask, uint64_t mask_words, uint64_t nbases,
                          const uint64_t *d_run_starts, uint64_t nruns,
                          blamm_hit *d_hits, uint64_t cap, uint64_t *d_nhits, void *stream);

/* Synchronise the last scan_device (or scan_device_noalloc).  On success
 * *nhits = hits produced.  BLAMM_ERR_OVERFLOW has two causes:
 *   - the hit buffer was too small: *nhits = the required capacity;
 *   - the filter records exceeded the candidate scratch: *nhits = 0 (the hit
 *     count is incomplete); after scan_device the scratch has been grown, so
 *     a retry succeeds (blamm_hip_last_stats().candidates holds the count).
 * Allocates only to grow the candidate scratch of scan_device. */
int blamm_hip_scan_finish(blamm_hip_ctx *ctx, uint64_t *nhits);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): blamm_hip_scan_device with its scratch -- the scan
 * counters and the filter records -- in a static device array of the library
 * (64 MiB per GPU, held until the process ends) instead of context buffers.
 * It calls no hipMalloc/hipFree, nor does blamm_hip_scan_finish after it.
 * Arguments, result and completion (blamm_hip_scan_finish) are those of
 * blamm_hip_scan_device.  A scan whose filter records exceed
 * BLAMM_NOALLOC_MAX_RECORDS is rejected by scan_finish with
 * BLAMM_ERR_OVERFLOW and *nhits = 0 (nothing is grown; split the stream).
 * Every context on one GPU shares that GPU's array, which keeps the last
 * call's contents: no-alloc scans that overlap on one GPU race. */
#define BLAMM_NOALLOC_MAX_RECORDS 4194304u
int blamm_hip_scan_device_noalloc(blamm_hip_ctx *ctx, const uint64_t *d_packed, uint64_t packed_words,
                                  const uint64_t *d_lcmask, uint64_t mask_words, uint64_t nbases,
                                  const uint64_t *d_run_starts, uint64_t nruns,
                                  blamm_hit *d_hits, uint64_t cap, uint64_t *d_nhits, void *stream);

/* Counters of the last finished scan. */
int blamm_hip_last_stats(blamm_hip_ctx *ctx, blamm_scan_stats *stats);

/* ------------------------------------------------------------------------
 * Tile shapes of the filter.  The reference picks column tiles of its
 * length-sorted pattern matrix by a recursive split that trades zero rows
 * against the number of sgemm calls (generateMatrixTiles / findBestSplit,
 * motif.cpp:451-540); here the same trade-off is three knobs of the MFMA
 * filter, set directly or by blamm_hip_autotune().
 * ------------------------------------------------------------------------ */
typedef struct blamm_tuning {
    uint32_t chunk_tiles;      /* window tile: 32-window tiles per work item */
    uint32_t blocks_per_wave;  /* motif tile: column blocks (32 bf16 / 16 fp32
                                  columns each) a wave scores at once: 1, 2
                                  (bf16 at most 2, fp32 at most 4) */
    uint32_t kpad_positions;   /* K-padding: a block's motif length is rounded
                                  up to a multiple of 4 or 8 positions */
} blamm_tuning;

int blamm_hip_get_tuning(blamm_hip_ctx *ctx, blamm_tuning *t);
/* Applies to later scans; re-lays the resident library out when the motif
 * tile or the K-padding changes. */
int blamm_hip_set_tuning(blamm_hip_ctx *ctx, const blamm_tuning *t);

/* Time full scans of a device-resident sample under candidate tunings and
 * keep the fastest.  cache_path (may be NULL): a text file of earlier
 * results, keyed by device, precision, library length histogram and sample
 * size; a matching entry is applied without timing, a new result is
 * appended.  *chosen / *chosen_ms (may be NULL) receive the setting and its
 * scan time. */
int blamm_hip_autotune(blamm_hip_ctx *ctx, const uint64_t *d_packed, uint64_t packed_words,
                       const uint64_t *d_lcmask, uint64_t mask_words, uint64_t nbases,
                       const uint64_t *d_run_starts, uint64_t nruns, const char *cache_path,
                       blamm_tuning *chosen, double *chosen_ms);
/* The same on a sample held in host memory (staged through the context's own
 * device buffers; buffer sizes as for blamm_hip_scan). */
int blamm_hip_autotune_host(blamm_hip_ctx *ctx, const uint64_t *packed, const uint64_t *lcmask,
                            uint64_t nbases, const uint64_t *run_starts, uint64_t nruns,
                            const char *cache_path, blamm_tuning *chosen, double *chosen_ms);

/* ------------------------------------------------------------------------
 * Batches: many requests in one persistent launch.  The reference scans one
 * manifest group's sequences as overlapping blocks handed to worker threads
 * (FastaBatch::getNextOverlappingBlock, sequence.cpp:274-293; scanPWMCUBLAS,
 * pwmscan.cpp:439-453); a batch instead packs several independent streams,
 * each with its own thresholds, into one work list scored by one launch
 * sequence that is captured in a hipGraph and replayed.
 * ------------------------------------------------------------------------ */
typedef struct blamm_request {
    const uint64_t *packed;      /* host, (nbases+31)/32 words, layout as above */
    const uint64_t *lcmask;      /* host, (nbases+63)/64 words, or NULL */
    uint64_t nbases;
    const uint64_t *run_starts;  /* sorted run starts of this stream */
    uint64_t nruns;
    const float *thresholds;     /* ncols thresholds in library column order;
                                    +INFINITY leaves a column out; NULL = the
                                    library's thresholds */
    uint64_t expected_hits;      /* admission estimate for the server; 0 = 1e-4
                                    per base and requested column */
} blamm_request;

typedef struct blamm_hip_batch blamm_hip_batch;

/* A batch slot on ctx's device with its own stream; device buffers start
 * sized for max_bases / max_requests and grow on demand.  It uses the
 * context's library and tuning: do not change them while a batch is in
 * flight. */
int blamm_hip_batch_create(blamm_hip_ctx *ctx, uint64_t max_bases, uint32_t max_requests,
                           blamm_hip_batch **batch);
int blamm_hip_batch_destroy(blamm_hip_batch *batch);
/* Device bytes the slot holds now. */
uint64_t blamm_hip_batch_device_bytes(const blamm_hip_batch *batch);
/* Stage nreq (< 65536) requests and launch them (asynchronous).  Request
 * buffers may be reused as soon as this returns.  Scratch and hit buffers are
 * grown up front to the requests' expected_hits (when given), so a batch is
 * rarely run twice. */
int blamm_hip_batch_submit(blamm_hip_batch *batch, const blamm_request *reqs, uint32_t nreq);
/* 0 when the batch has finished on the device (or none is in flight),
 * 1 while it runs. */
int blamm_hip_batch_query(blamm_hip_batch *batch);
/* Wait for the batch (replaying it after growing scratch or the hit buffer
 * when either overflowed); *nhits = hits of all requests. */
int blamm_hip_batch_wait(blamm_hip_batch *batch, uint64_t *nhits);
/* Hits of request `req` of the last finished batch, positions local to that
 * request's stream; valid until the next submit.  The device groups a batch's
 * hits by request and reads them back as 12-byte records (32-bit request-local
 * position, column, score); the first call after a wait rebuilds them as
 * blamm_hit. */
int blamm_hip_batch_hits(blamm_hip_batch *batch, uint32_t req, const blamm_hit **hits, uint64_t *n);
/* kernel_ms = the whole captured launch sequence of the last batch. */
int blamm_hip_batch_stats(blamm_hip_batch *batch, blamm_scan_stats *stats);

/* ------------------------------------------------------------------------
 * Server: continuous batching over a resident library.  No counterpart in
 * the reference (its scan is a one-shot command); built on the batch slots.
 * ------------------------------------------------------------------------ */
typedef struct blamm_server_config {
    uint64_t hbm_budget_bytes;    /* device bytes for batches in flight (0 = 8 GiB),
                                     split evenly over the slots */
    uint32_t slots;               /* batches in flight, 1..8 (0 = 2) */
    uint32_t max_batch_requests;  /* requests per batch (0 = 256) */
    uint64_t max_batch_bases;     /* bases per batch (0 = 2^24): smaller batches let
                                     one batch's hit read-back overlap the next
                                     batch's scan */
} blamm_server_config;

typedef struct blamm_request_info {
    int status;                   /* BLAMM_OK or the batch's error */
    uint32_t batch_requests;      /* requests in the batch it ran in */
    uint64_t nhits;
    uint64_t t_submit_ns, t_start_ns, t_done_ns;   /* steady clock */
} blamm_request_info;

typedef struct blamm_server_stats {
    uint64_t submitted, completed, batches, hits;
    uint64_t replays;             /* batches run again after growing scratch */
    double device_ms;             /* summed batch launch times */
    double host_submit_ms;        /* launcher thread: staging and launching batches */
    double host_wait_ms;          /* harvester and completer threads: waits for
                                     kernels and hit read-back */
} blamm_server_stats;

typedef struct blamm_hip_server blamm_hip_server;

/* Starts the server's three threads (a launcher that packs queued requests
 * into free batch slots, a harvester that waits for batches oldest first and
 * starts their hit read-back, a completer that waits for the read-backs and
 * finishes the requests), which own the batch slots.  While the server runs,
 * the context's library and tuning must not change. */
int blamm_hip_server_create(blamm_hip_ctx *ctx, const blamm_server_config *cfg,
                            blamm_hip_server **server);
/* Queue a request.  Its packed / lcmask / run_starts buffers must stay valid
 * until blamm_hip_server_wait() returns for it; thresholds are copied.
 * Requests larger than one slot's budget are rejected with BLAMM_ERR_ARG. */
int blamm_hip_server_submit(blamm_hip_server *server, const blamm_request *req, uint64_t *id);
/* BLAMM_OK when request id has finished, 1 on timeout (timeout_ms < 0 waits
 * without limit). */
int blamm_hip_server_wait(blamm_hip_server *server, uint64_t id, int64_t timeout_ms);
/* Copy the finished request's hits (request-local positions) and timing,
 * then forget the request (the 12-byte records read back from the device
 * become blamm_hit here, on the caller's thread).  BLAMM_ERR_OVERFLOW (request kept) when cap is
 * below info->nhits. */
int blamm_hip_server_result(blamm_hip_server *server, uint64_t id, blamm_hit *hits, uint64_t cap,
                            blamm_request_info *info);
/* Report a finished request's timing and hit count, drop its hits and forget
 * it (for callers that only need the count). */
int blamm_hip_server_release(blamm_hip_server *server, uint64_t id, blamm_request_info *info);
int blamm_hip_server_stats(blamm_hip_server *server, blamm_server_stats *stats);
/* Finishes queued and running requests, then stops the server's threads. */
int blamm_hip_server_destroy(blamm_hip_server *server);

/* ------------------------------------------------------------------------
 * Empirical score histograms (`blamm hist -e`).
 *
 * Replace Histogram::histThread (hist.cpp:93-138), whose sgemm over every
 * offset feeds Histogram::extractObsScore (hist.cpp:70-91): every window of
 * the stream that satisfies the run rule (hist.cpp:83-86) and starts before
 * qend adds one count to its column's histogram, in the bin of
 * ScoreHistogram::addObservation (motif.h:96-102),
 *   bin = clamp(int((score - min) / width), 0, nbins - 1),
 *   width = (max - min) / nbins (motif.h:58-66, float arithmetic),
 * where score is the FP32 sum of the window's weights in position order
 * (lowercase bases add nothing).  The library must have been uploaded with
 * BLAMM_PREC_BF16X2 (its thresholds are not used).  min_score/max_score hold
 * one float per library column; counts (ncols x nbins, column-major by
 * column: counts[col * nbins + bin]) is added to, so several streams (or
 * segments, with qend cutting the overlap) can be accumulated.  Synchronous.
 * ------------------------------------------------------------------------ */
int blamm_hip_hist(blamm_hip_ctx *ctx, const uint64_t *packed, const uint64_t *lcmask, uint64_t nbases,
                   const uint64_t *run_starts, uint64_t nruns, uint64_t qend, const float *min_score,
                   const float *max_score, uint32_t nbins, uint64_t *counts);
/* As blamm_hip_hist on device-resident buffers (sizes as for
 * blamm_hip_scan_device). */
int blamm_hip_hist_device(blamm_hip_ctx *ctx, const uint64_t *d_packed, uint64_t packed_words,
                          const uint64_t *d_lcmask, uint64_t mask_words, uint64_t nbases,
                          const uint64_t *d_run_starts, uint64_t nruns, uint64_t qend, const float *min_score,
                          const float *max_score, uint32_t nbins, uint64_t *counts);

#ifdef __cplusplus
}
#endif
#endif
