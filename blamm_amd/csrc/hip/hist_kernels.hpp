// Empirical score histograms (`blamm hist -e`) on the split-BF16 engine.
//
// The reference fills them with the scan's sgemm (Histogram::histThread,
// hist.cpp:93-138) and bins the FP32 score of every window that satisfies the
// run rule (Histogram::extractObsScore, hist.cpp:70-91) with
// ScoreHistogram::addObservation (motif.h:96-102):
//   bin = clamp(int((score - min) / width), 0, nbins - 1),
//   width = (max - min) / nbins (motif.h:58-66), all in float.
// The score binned here is the FP32 sum of the window's weights in position
// order (lowercase bases contribute nothing).
//
// hist_filter_kernel scores windows with the same hi+lo MFMA tiles as the
// scan and decides a window's bin from its MFMA score a when that is safe:
// x = fma(a, 1/width, -min/width) approximates the bin coordinate of the FP32
// score to within a per-column margin m (from the split and summation error
// bounds and the rounding of every float step, computed on the host), so when
// fract(x) lies in [m, 1 - m] the FP32 score's coordinate lies in the same
// unit interval and the bin is trunc(x), clamped.  Counts go to a
// workgroup-private LDS histogram that is added to HBM once per work item.
// Windows near a bin edge (about 2m of them) are staged per wave in LDS and
// binned exactly by the same wave (FP32 left-to-right sum of the window's
// weights, the reference's float arithmetic), so no record leaves the chip.
#pragma once

#include "scan_kernels.hpp"

namespace blamm {

struct HistSlot {
    float iw, c;       // x = fma(score, iw, c)
    float hm;          // bin decided from x iff |fract(x) - 1/2| <= hm
    float pad0;
    float mn, wd;      // histogram minimum and bin width, as the reference computes them
    uint32_t len;      // motif length; 0 = empty slot
    uint32_t pad;
};

struct HistArgs {
    const uint64_t* packed;
    const uint64_t* lcmask;        // MASK variants only
    const uint4* frag_hi;
    const uint4* frag_lo;
    const HistSlot* hs;
    const float* slot_w;           // [slot][kMaxLen][4] FP32 weights
    const uint64_t* run_starts;
    const uint32_t* groups;        // group_pack(blk0, nch, 1): one column block per item
    unsigned long long* counts;    // [slot][nbins]
    unsigned long long* counters;  // kCntBand: windows binned exactly; kCntQueue: work item head
    uint32_t nbases, nruns, qend, nbins;
    uint32_t ntiles, chunk_tiles, ngroups, nitems;
    uint32_t queue;                // work-item counter of this launch: counters[kCntQueue + queue]
};

// LDS histogram budget per workgroup (32 columns x up to 512 bins).
constexpr uint32_t kHistLdsBytes = 64 * 1024;
// Undecided windows staged per wave before exact binning (8 KiB per block
// next to the 64 KiB histogram).
constexpr int kHistStage = 128;

// ScoreHistogram::addObservation's bin (motif.h:96-102) of an FP32 score:
// float subtract and IEEE division, then int() as the reference's x86-64
// hosts execute it (truncation; NaN or out of range gives INT_MIN), clamped.
__device__ __forceinline__ uint32_t hist_bin_ref(float s, float mn, float wd, uint32_t nbins)
{
    const float v = (s - mn) / wd;
    int b = (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : (int)0x80000000u;
    b = max(b, 0);
    b = min(b, (int)nbins - 1);
    return (uint32_t)b;
}

// Exact bins of the staged windows (one record per lane: first window of
// the lane, slot, rows): FP32 left-to-right sum of the window's weights
// (lowercase bases skipped), the reference's binning, one count each.  The
// run rule was applied when the record was staged.  Returns the number of
// windows this lane binned.
template <bool LDS>
__device__ __forceinline__ uint32_t hist_exact_stage(const HistArgs& a, Stage& st, uint32_t* hl, uint32_t blk0,
                                                     uint32_t lane)
{
    uint32_t done = 0;
    for (uint32_t i = lane; i < st.n; i += 64) {
        const uint4 cd = st.buf[i];
        const uint32_t slot = cd.y;
        uint32_t rows = cd.z;
        const HistSlot h = a.hs[slot];
        const float* w = a.slot_w + (size_t)slot * kMaxLen * 4;
        while (rows) {
            const uint32_t r = (uint32_t)__builtin_ctz(rows);
            rows &= rows - 1u;
            const uint32_t q = cd.x + 8 * (r >> 2) + (r & 3);
            const uint64_t bits = load_bits64(a.packed, 2ull * q);
            const uint64_t bits2 = h.len > 32 ? load_bits64(a.packed, 2ull * q + 64) : 0ull;
            const uint64_t mbits = a.lcmask ? load_bits64(a.lcmask, q) : 0ull;
            float s = 0.f;
            for (uint32_t p = 0; p < h.len; ++p) {
                const float4 wp = *(const float4*)(w + 4 * p);
                const uint32_t b = (uint32_t)(p < 32 ? bits >> (2 * p) : bits2 >> (2 * (p - 32))) & 3u;
                const float x = b == 0 ? wp.x : b == 1 ? wp.y : b == 2 ? wp.z : wp.w;
                if (!((mbits >> p) & 1)) s += x;
            }
            const uint32_t bin = hist_bin_ref(s, h.mn, h.wd, a.nbins);
            if (LDS) atomicAdd(&hl[(slot - blk0 * kBf16Cols) * a.nbins + bin], 1u);
            else atomicAdd(&a.counts[(size_t)slot * a.nbins + bin], 1ull);
            ++done;
        }
    }
    st.n = 0;
    return done;
}

// Does the window at q of length len end at or before the first run start
// after q?  rnext is the first run start after the tile's first window and jr
// its index (nbases past the last run).
__device__ __forceinline__ bool hist_row_fits(const HistArgs& a, uint32_t q, uint32_t len, uint32_t jr,
                                              uint32_t rnext)
{
    uint32_t r = rnext;
    if (q >= rnext) {
        uint32_t k = jr + 1;
        while (k < a.nruns && (uint32_t)a.run_starts[k] <= q) ++k;
        r = k < a.nruns ? (uint32_t)a.run_starts[k] : a.nbases;
    }
    return q + len <= r;
}

// Bins of one finished tile block (16 windows per lane).  A window whose
// bin the MFMA score decides adds 1 to it; the others add 0 (no branch, one
// LDS atomic per window) and are returned as a row mask for exact binning.
// FAST: every window fits its run and starts before qend.
template <bool FAST, bool LDS>
__device__ __forceinline__ uint32_t hist_rows(const HistArgs& a, uint32_t* hl, const f32x16& acc, float iw, float cc,
                                              float hm, uint32_t len, uint32_t slot, uint32_t hoff, uint32_t qlane,
                                              uint32_t jr, uint32_t rnext)
{
    const uint32_t live = len != 0 ? 1u : 0u;
    const int last = (int)a.nbins - 1;
    uint32_t undecided = 0, valid = FAST ? (live ? 0xFFFFu : 0u) : 0u;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        uint32_t v = live;
        if (!FAST) {
            const uint32_t q = qlane + 8 * (r >> 2) + (r & 3);
            v = (live && q < a.qend && hist_row_fits(a, q, len, jr, rnext)) ? 1u : 0u;
            valid |= v << r;
        }
        const float x = __builtin_fmaf(acc[r], iw, cc);
        const bool ok = __builtin_fabsf(__builtin_amdgcn_fractf(x) - 0.5f) <= hm;
        const uint32_t bin = (uint32_t)min(max((int)x, 0), last);
        if (LDS) {
            atomicAdd(&hl[hoff + bin], ok ? v : 0u);
        } else if (ok && v) {
            atomicAdd(&a.counts[(size_t)slot * a.nbins + bin], 1ull);
        }
        undecided |= (ok ? 0u : 1u) << r;
    }
    return undecided & valid;
}

// Tiles t0, t0 + 4, ... < t1 of column blocks blk0 .. blk0+NB-1 (the wave's
// share of a work item).  hl: the workgroup's LDS histogram of the item's
// columns (LDS variant).
template <int NCH, int NB, bool MASK, bool LDS>
__device__ __forceinline__ uint32_t hist_range(const HistArgs& a, Stage& st, uint32_t* hl, uint32_t blk0,
                                               uint32_t t0, uint32_t t1, uint32_t lane)
{
    uint32_t exact = 0;
    bf16x8 bh[NB][NCH], bl[NB][NCH];
    float iw[NB], cc[NB], hm[NB];
    uint32_t len[NB], slot[NB], hoff[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
#pragma unroll
        for (int j = 0; j < NCH; ++j) {
            const size_t idx = ((size_t)(blk0 + b) * kMaxChunks + j) * 64 + lane;
            bh[b][j] = __builtin_bit_cast(bf16x8, a.frag_hi[idx]);
            bl[b][j] = __builtin_bit_cast(bf16x8, a.frag_lo[idx]);
        }
        slot[b] = (blk0 + b) * kBf16Cols + (lane & 31);
        const HistSlot h = a.hs[slot[b]];
        iw[b] = h.iw;
        cc[b] = h.c;
        hm[b] = h.hm;
        len[b] = h.len;
        hoff[b] = LDS ? (uint32_t)(b * kBf16Cols + (lane & 31)) * a.nbins : 0u;
    }
    const uint32_t off = (lane & 31) + 2 * (lane >> 5);   // first base of the lane's window, 0..33
    // pk[t]: dwords 2t + off/16 and the next, so the window of tile t starts
    // at bit sh < 32 of pk[t]
    const uint32_t sh = 2 * (off & 15);
    const u64_a4* __restrict__ pk = (const u64_a4*)((const uint32_t*)a.packed + (off >> 4));
    const uint32_t* __restrict__ mk = (const uint32_t*)a.lcmask + (off >> 5);

    // run cursor: first run start after the current tile's first window
    uint32_t jr, rnext;
    {
        const uint32_t q = t0 * kBf16Tile;
        uint32_t l = 0, h = a.nruns;
        while (l < h) {
            const uint32_t mid = (l + h) >> 1;
            if ((uint32_t)a.run_starts[mid] <= q) l = mid + 1; else h = mid;
        }
        jr = l;
        rnext = l < a.nruns ? (uint32_t)a.run_starts[l] : a.nbases;
    }
    for (uint32_t t = t0; t < t1; t += kWavesPerBlock) {
        const uint32_t q0 = t * kBf16Tile;
        const OneHotSrc bits = onehot_src<NCH>(pk[t], pk[t + 1], pk[t + 2], sh);
        const uint64_t mb = MASK ? window_mask<NCH>(mk[t], mk[t + 1], mk[t + 2], off & 31) : 0ull;
        f32x16 acc[NB];
        bf16_mfma_tile<NCH, NB, MASK>(acc, bh, bl, bits, mb);
        while (rnext <= q0) {
            ++jr;
            rnext = jr < a.nruns ? (uint32_t)a.run_starts[jr] : a.nbases;
        }
        // every window of the tile fits its run for every length <= 4 NCH
        const bool fast = q0 + (kBf16Tile - 1) + 4 * NCH <= rnext && q0 + kBf16Tile <= a.qend;
        const f32x16 (&av)[NB] = acc_read<AccRead::kGuard>(acc);
        const uint32_t qlane = q0 + 4 * (lane >> 5);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const uint32_t defer =
                fast ? hist_rows<true, LDS>(a, hl, av[b], iw[b], cc[b], hm[b], len[b], slot[b], hoff[b], qlane, jr,
                                            rnext)
                     : hist_rows<false, LDS>(a, hl, av[b], iw[b], cc[b], hm[b], len[b], slot[b], hoff[b], qlane, jr,
                                             rnext);
            const uint64_t m = __ballot(defer != 0u);
            if (m) {
                if (st.n > (uint32_t)(kHistStage - 64)) exact += hist_exact_stage<LDS>(a, st, hl, blk0, lane);
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (defer) st.buf[st.n + rank] = make_uint4(qlane, slot[b], defer, 0u);
                st.n += (uint32_t)__popcll(m);
            }
        }
    }
    exact += hist_exact_stage<LDS>(a, st, hl, blk0, lane);
    return exact;
}

// Persistent: each workgroup takes work items (chunk of chunk_tiles tiles,
// column block) from one queue; its four waves share the item's tiles and
// its LDS histogram.  Up to 128 VGPRs so that four waves per SIMD hide the
// latency of the per-window binning (measured: 8.3 ms against 12.3 ms with
// two blocks per wave at two waves per SIMD, 10 Mbp x 1,000 columns).
// Blocks of motifs longer than 32 positions (WIDE: up to 16 chunks of hi and
// lo fragments in registers) run in a launch of their own at two waves per
// SIMD.
template <bool MASK, bool LDS, bool WIDE>
__global__ __launch_bounds__(kThreads, WIDE ? 2 : 4) void hist_filter_kernel(HistArgs a)
{
    extern __shared__ uint32_t hist_lds[];
    __shared__ uint4 stage_lds[kWavesPerBlock][kHistStage];
    __shared__ uint32_t item_s;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Stage st{stage_lds[wave], 0u, nullptr};
    uint32_t exact = 0;
    for (;;) {
        if (threadIdx.x == 0) item_s = (uint32_t)atomicAdd(&a.counters[kCntQueue + a.queue], 1ull);
        __syncthreads();
        const uint32_t it = item_s;
        if (it >= a.nitems) break;
        const uint32_t k = it / a.ngroups, g = it - k * a.ngroups;
        const uint32_t gp = a.groups[g];
        const uint32_t blk0 = gp & 0xFFFFFu, nch = (gp >> 20) & 63u, nb = gp >> 26;
        const uint32_t t0 = k * a.chunk_tiles, t1 = min(t0 + a.chunk_tiles, a.ntiles);
        const uint32_t used = nb * kBf16Cols * a.nbins;
        if (LDS)
            for (uint32_t i = threadIdx.x; i < used; i += kThreads) hist_lds[i] = 0u;
        __syncthreads();
#define BLAMM_H(N, B) \
    case N * 8 + B: exact += hist_range<N, B, MASK, LDS>(a, st, hist_lds, blk0, t0 + wave, t1, lane); break;
        if constexpr (WIDE) {
            switch (nch * 8 + nb) {
                BLAMM_H(9, 1) BLAMM_H(10, 1) BLAMM_H(11, 1) BLAMM_H(12, 1)
                BLAMM_H(13, 1) BLAMM_H(14, 1) BLAMM_H(15, 1) BLAMM_H(16, 1)
                default: break;
            }
        } else {
            switch (nch * 8 + nb) {
                BLAMM_H(1, 1) BLAMM_H(2, 1) BLAMM_H(3, 1) BLAMM_H(4, 1)
                BLAMM_H(5, 1) BLAMM_H(6, 1) BLAMM_H(7, 1) BLAMM_H(8, 1)
                default: break;
            }
        }
#undef BLAMM_H
        if (LDS) {
            __syncthreads();
            unsigned long long* dst = a.counts + (size_t)blk0 * kBf16Cols * a.nbins;
            for (uint32_t i = threadIdx.x; i < used; i += kThreads) {
                const uint32_t v = hist_lds[i];
                if (v) atomicAdd(&dst[i], (unsigned long long)v);
            }
        }
        // the next item's first barrier orders these reads before its zeroing
    }
    // windows binned exactly (for statistics): one atomic per wave
    exact = __reduce_add_sync(~0ull, exact);
    if (lane == 0 && exact) atomicAdd(&a.counters[kCntBand], (unsigned long long)exact);
}

}  // namespace blamm
