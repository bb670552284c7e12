// Empirical score histograms (`blamm hist -e`) through the C ABI: replaces
// Histogram::histThread's sgemm and extractObsScore loop (hist.cpp:70-138)
// for one stream of a manifest group.  Included at the end of blamm_hip.hip.

namespace {

// Bin parameters of every library slot for histograms on [mins[col],
// maxs[col]] with nbins bins (kernels in hist_kernels.hpp).
void hist_slot_params(const blamm_hip_ctx* c, const float* mins, const float* maxs, uint32_t nbins,
                      std::vector<HistSlot>& hs)
{
    hs.assign((size_t)c->nslots, HistSlot{0.f, 0.f, -1.f, 0.f, 0.f, 0.f, 0u, 0u});
    std::vector<uint32_t> blk_nch((size_t)c->nslots / kBf16Cols + 1, kMaxChunks);
    for (const GroupInfo& g : c->groups)
        for (uint32_t i = 0; i < g.nb; ++i) blk_nch[g.blk0 + i] = g.nch;
    for (int s = 0; s < c->nslots; ++s) {
        const int col = c->slot_col_h[s];
        if (col < 0) continue;
        HistSlot h{0.f, 0.f, -1.f, 0.f, 0.f, 0.f, 0u, 0u};   // hm < 0: every window binned exactly
        const int L = c->lib_len[col];
        const float mn = mins[col], mx = maxs[col];
        const float wd = (mx - mn) / (float)nbins;          // motif.h:64
        h.mn = mn;
        h.wd = wd;
        h.len = (uint32_t)L;
        const float* w = c->lib_P.data() + (size_t)col * c->ldp;
        double smax = 0.0, ssplit = 0.0, rep = 0.0;   // sum_p max_b |w|, |hi| + |lo|, |w - hi - lo|
        for (int p = 0; p < L; ++p) {
            double m = 0.0, ms = 0.0, me = 0.0;
            for (int b = 0; b < 4; ++b) {
                const float x = w[4 * p + b];
                const float hf = bf16_to_float(bf16_rne(x));
                const float lf = bf16_to_float(bf16_rne(x - hf));
                m = std::max(m, (double)std::fabs(x));
                ms = std::max(ms, std::fabs((double)hf) + std::fabs((double)lf));
                me = std::max(me, std::fabs((double)x - (double)hf - (double)lf));
            }
            smax += m;
            ssplit += ms;
            rep += me;
        }
        // |MFMA score - FP32 left-to-right score|.  MFMA side: the split's
        // representation error plus the accumulation error.  A window's row
        // of the one-hot operand has at most 4 nonzero entries per
        // v_mfma_f32_32x32x16_bf16 (one base per position), so each of the
        // 2 nch MFMAs adds at most 4 nonzero exact products to its
        // accumulator: 5 operands, each contributing at most one (possibly
        // truncating, u = 2^-22) rounding relative to sum_p max(|hi| + |lo|),
        // i.e. gamma_{10 nch}.  FP32 side: L - 1 round-to-nearest adds.
        const double ds = rep + gamma_n(10.0 * blk_nch[(size_t)s / kBf16Cols], 0x1p-22) * ssplit +
                          gamma_n(L, 0x1p-24) * smax;
        if (std::isfinite(mn) && std::isfinite(mx) && wd == 0.f) {
            // (score - min) / 0 is +-inf or NaN for every score; int() of
            // those is INT_MIN, clamped to bin 0
            h.iw = 0.f;
            h.c = 0.5f;
            h.hm = 0.5f;
        } else if (std::isfinite(mn) && std::isfinite(wd) && wd > 0.f) {
            const float iw = 1.0f / wd;
            const float cf = -(mn * iw);
            const double inv = 1.0 / (double)wd;
            const double S = smax * (1.0 + 0x1p-20);
            // largest |bin coordinate| of any score, and the margin covering
            // |x - the reference's coordinate of the FP32 score|: the score
            // difference, the rounding of iw and c, of the fma, of the
            // reference's subtract and divide, and of fract
            const double qmax = (S + std::fabs((double)mn) + ds) * inv + 1.0;
            const double m = ds * inv * (1.0 + 0x1p-20) + (S + ds + 2.0 * std::fabs((double)mn)) * inv * 0x1p-22 +
                             (qmax + 1.0) * 0x1p-21 + 0x1p-22;
            if (std::isfinite(iw) && std::isfinite(cf) && qmax < 0x1p30 && m < 0.25) {
                // |fract(x) - 1/2| <= hm implies m <= fract(x) <= 1 - m; the
                // subtraction is exact for fract(x) >= 1/4 and rounds by at
                // most 2^-26 below, covered by the extra 2^-25
                const double hmd = 0.5 - m - 0x1p-25;
                float hmf = (float)hmd;
                if ((double)hmf > hmd) hmf = std::nextafter(hmf, 0.f);
                h.iw = iw;
                h.c = cf;
                h.hm = hmf;
            }
        }
        hs[(size_t)s] = h;
    }
}

template <bool MASK, bool LDS, bool WIDE>
const void* hist_kernel_ptr()
{
    return (const void*)hist_filter_kernel<MASK, LDS, WIDE>;
}

const void* hist_kernel(bool mask, bool lds, bool wide)
{
    if (wide)
        return mask ? (lds ? hist_kernel_ptr<true, true, true>() : hist_kernel_ptr<true, false, true>())
                    : (lds ? hist_kernel_ptr<false, true, true>() : hist_kernel_ptr<false, false, true>());
    return mask ? (lds ? hist_kernel_ptr<true, true, false>() : hist_kernel_ptr<true, false, false>())
                : (lds ? hist_kernel_ptr<false, true, false>() : hist_kernel_ptr<false, false, false>());
}

}  // namespace

extern "C" {

int blamm_hip_hist_device(blamm_hip_ctx* c, const uint64_t* d_packed, uint64_t packed_words,
                          const uint64_t* d_lcmask, uint64_t mask_words, uint64_t nbases,
                          const uint64_t* d_run_starts, uint64_t nruns, uint64_t qend, const float* min_score,
                          const float* max_score, uint32_t nbins, uint64_t* counts)
{
    if (!c) return fail(BLAMM_ERR_ARG, "hist: null context");
    if (!c->have_lib) return fail(BLAMM_ERR_STATE, "hist: no library uploaded");
    if (c->pending) return fail(BLAMM_ERR_STATE, "hist: a scan is in flight");
    if (!prec_bf(c->precision))
        return fail(BLAMM_ERR_ARG, "hist: the library must be uploaded in the split-BF16 form");
    if (!min_score || !max_score || !counts) return fail(BLAMM_ERR_ARG, "hist: null histogram range or counts");
    if (nbins < 1 || nbins > (1u << 24)) return fail(BLAMM_ERR_ARG, "hist: nbins outside 1..2^24");
    if (nbases >= (1ull << 32) - 512) return fail(BLAMM_ERR_ARG, "hist: nbases must be < 2^32 - 512 per call");
    if (nbases > 0 && !d_packed) return fail(BLAMM_ERR_ARG, "hist: null packed stream");
    if (nbases > 0 && packed_words < blamm_hip_packed_words(nbases))
        return fail(BLAMM_ERR_ARG, "hist: packed buffer shorter than blamm_hip_packed_words(nbases)");
    if (d_lcmask && mask_words < blamm_hip_mask_words(nbases))
        return fail(BLAMM_ERR_ARG, "hist: mask buffer shorter than blamm_hip_mask_words(nbases)");
    if (nruns > 0 && !d_run_starts) return fail(BLAMM_ERR_ARG, "hist: null run starts");
    if (nruns >= (1ull << 32)) return fail(BLAMM_ERR_ARG, "hist: too many runs");
    if ((uint64_t)c->nslots * nbins > (1ull << 34)) return fail(BLAMM_ERR_ARG, "hist: columns x bins too large");
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (qend > nbases) qend = nbases;

    std::vector<HistSlot> hs;
    hist_slot_params(c, min_score, max_score, nbins, hs);
    // one column block per work item (the kernel keeps one block's
    // fragments in registers at four waves per SIMD); its histograms live in
    // LDS when its 32 columns x nbins counters fit
    const bool lds = (uint64_t)nbins * kBf16Cols * 4 <= kHistLdsBytes;
    // blocks of motifs up to 32 positions first, then the wide ones (their
    // own launch at two waves per SIMD)
    std::vector<uint32_t> groups;
    for (int pass = 0; pass < 2; ++pass)
        for (const GroupInfo& g : c->groups)
            if ((g.nch > (uint32_t)kWideChunks) == (pass == 1))
                for (uint32_t i = 0; i < g.nb; ++i) groups.push_back(group_pack(g.blk0 + i, g.nch, 1));
    uint32_t nnarrow = 0;
    for (const GroupInfo& g : c->groups)
        if (g.nch <= (uint32_t)kWideChunks) nnarrow += g.nb;
    HIP_TRY(c->hist_slots.reserve(hs.size()));
    HIP_TRY(c->hist_groups.reserve(groups.size()));
    const size_t ncount = (size_t)c->nslots * nbins;
    HIP_TRY(c->hist_counts.reserve(ncount));
    HIP_TRY(hipMemcpyAsync(c->hist_slots.p, hs.data(), hs.size() * sizeof(HistSlot), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->hist_groups.p, groups.data(), groups.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           s));
    const bool mask = d_lcmask != nullptr;
    const uint32_t dyn = lds ? kBf16Cols * nbins * 4 : 0u;

    HistArgs a{};
    a.packed = d_packed;
    a.lcmask = d_lcmask;
    a.frag_hi = (const uint4*)c->frag_a.p;
    a.frag_lo = (const uint4*)c->frag_b.p;
    a.hs = c->hist_slots.p;
    a.slot_w = c->slot_w.p;
    a.run_starts = d_run_starts;
    a.counts = c->hist_counts.p;
    a.counters = c->counters.p;
    a.nbases = (uint32_t)nbases;
    a.nruns = (uint32_t)nruns;
    a.qend = (uint32_t)qend;
    a.nbins = nbins;
    a.ntiles = (uint32_t)((qend + kBf16Tile - 1) / kBf16Tile);   // windows start before qend

    HIP_TRY(hipMemsetAsync(c->hist_counts.p, 0, ncount * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, c->counters.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[0], s));
    for (int wide = 0; wide < 2; ++wide) {
        HistArgs w = a;
        w.groups = c->hist_groups.p + (wide ? nnarrow : 0);
        w.ngroups = wide ? (uint32_t)groups.size() - nnarrow : nnarrow;
        w.queue = (uint32_t)wide;
        if (w.ngroups == 0 || w.ntiles == 0) continue;
        const void* kp = hist_kernel(mask, lds, wide);
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, kThreads, dyn) != hipSuccess || per_cu < 1)
            per_cu = 1;
        const uint32_t grid = (uint32_t)c->num_cu * (uint32_t)per_cu;
        // about four items per workgroup, 16..1024 tiles per item
        const uint64_t want = ((uint64_t)w.ntiles * w.ngroups + 4ull * grid - 1) / (4ull * grid);
        w.chunk_tiles = (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(16, want));
        const uint64_t nchunks = (w.ntiles + w.chunk_tiles - 1) / w.chunk_tiles;
        if (nchunks * w.ngroups >= (1ull << 32)) return fail(BLAMM_ERR_ARG, "hist: work list too long");
        w.nitems = (uint32_t)(nchunks * w.ngroups);
#define BLAMM_HL(M, L, W) hipLaunchKernelGGL((hist_filter_kernel<M, L, W>), dim3(grid), dim3(kThreads), dyn, s, w)
        if (wide) {
            if (mask) { if (lds) BLAMM_HL(true, true, true); else BLAMM_HL(true, false, true); }
            else { if (lds) BLAMM_HL(false, true, true); else BLAMM_HL(false, false, true); }
        } else {
            if (mask) { if (lds) BLAMM_HL(true, true, false); else BLAMM_HL(true, false, false); }
            else { if (lds) BLAMM_HL(false, true, false); else BLAMM_HL(false, false, false); }
        }
#undef BLAMM_HL
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(c->ev[1], s));
    HIP_TRY(hipStreamSynchronize(s));
    unsigned long long cnt[kCntSlots];
    HIP_TRY(hipMemcpy(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    std::vector<unsigned long long> dev(ncount);
    HIP_TRY(hipMemcpy(dev.data(), c->hist_counts.p, ncount * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int sl = 0; sl < c->nslots; ++sl) {
        const int col = c->slot_col_h[sl];
        if (col < 0) continue;
        for (uint32_t b = 0; b < nbins; ++b)
            counts[(size_t)col * nbins + b] += dev[(size_t)sl * nbins + b];
    }
    float ms_f = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms_f, c->ev[0], c->ev[1]));
    c->stats = blamm_scan_stats{};
    c->stats.band_rescored = cnt[kCntBand];   // windows binned from their FP32 sum
    c->stats.filter_ms = ms_f;
    c->stats.kernel_ms = ms_f;
    return BLAMM_OK;
}

int blamm_hip_hist(blamm_hip_ctx* c, const uint64_t* packed, const uint64_t* lcmask, uint64_t nbases,
                   const uint64_t* run_starts, uint64_t nruns, uint64_t qend, const float* min_score,
                   const float* max_score, uint32_t nbins, uint64_t* counts)
{
    if (!c) return fail(BLAMM_ERR_ARG, "hist: null context");
    if (nbases > 0 && !packed) return fail(BLAMM_ERR_ARG, "hist: null packed stream");
    if (nruns > 0 && !run_starts) return fail(BLAMM_ERR_ARG, "hist: null run starts");
    if (nbases >= (1ull << 32) - 512) return fail(BLAMM_ERR_ARG, "hist: nbases must be < 2^32 - 512 per call");
    if (c->pending) return fail(BLAMM_ERR_STATE, "hist: a scan is in flight");
    HIP_TRY(hipSetDevice(c->device));
    BLAMM_TRY(stage_stream(c, packed, lcmask, nbases, run_starts, nruns));
    return blamm_hip_hist_device(c, c->packed.p, blamm_hip_packed_words(nbases), lcmask ? c->lcmask.p : nullptr,
                                 blamm_hip_mask_words(nbases), nbases, c->runs.p, nruns, qend, min_score,
                                 max_score, nbins, counts);
}

}  // extern "C"
