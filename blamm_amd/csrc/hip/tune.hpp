// Autotuner of the filter's tile shapes: motif tile (column blocks per wave),
// window tile (32-window tiles per work item) and K-padding (motif positions
// a block's length is rounded up to).  It times whole scans of a
// device-resident sample on the context's device, keeps the fastest setting
// and records it in a text cache keyed by device, precision, library shape
// and sample size, so later runs reuse it without timing.
//
// Included once, at the end of blamm_hip.hip.

namespace {

std::string tune_key(const blamm_hip_ctx* c, uint64_t nbases, bool mask)
{
    // FNV-1a over the column length histogram
    uint64_t h = 1469598103934665603ull;
    std::vector<uint32_t> hist(BLAMM_MAX_MOTIF_LEN + 1, 0);
    for (int L : c->lib_len) hist[L]++;
    for (uint32_t v : hist) {
        h ^= v;
        h *= 1099511628211ull;
    }
    int lg = 0;
    while ((1ull << (lg + 1)) <= std::max<uint64_t>(nbases, 1)) ++lg;
    std::ostringstream os;
    os << c->arch << "/cu" << c->num_cu << "/" << (c->precision == BLAMM_PREC_BF16X2 ? "bf16x2" : c->precision == BLAMM_PREC_BF16X2_LO ? "bf16x2lo" : "fp32") << "/cols"
       << c->ncols << "/len" << std::hex << h << std::dec << "/n2^" << lg << (mask ? "/mask" : "");
    return os.str();
}

bool tune_lookup(const char* path, const std::string& key, blamm_tuning* t, double* ms)
{
    if (!path || !*path) return false;
    std::ifstream in(path);
    std::string line;
    bool found = false;
    while (std::getline(in, line)) {
        std::istringstream is(line);
        std::string k;
        blamm_tuning v{};
        double m = 0;
        if (!(is >> k >> v.chunk_tiles >> v.blocks_per_wave >> v.kpad_positions >> m)) continue;
        if (k == key) {   // the last entry wins
            *t = v;
            *ms = m;
            found = true;
        }
    }
    return found;
}

// Time one full scan of the sample under the context's current tuning
// (best of `reps`, after one warm-up).
int time_scan(blamm_hip_ctx* c, const uint64_t* d_packed, uint64_t pw, const uint64_t* d_lcmask, uint64_t mw,
              uint64_t nbases, const uint64_t* d_runs, uint64_t nruns, int reps, double* best)
{
    *best = 1e300;
    for (int i = 0; i <= reps; ++i) {
        int rc;
        for (int attempt = 0;; ++attempt) {
            BLAMM_TRY(blamm_hip_scan_device(c, d_packed, pw, d_lcmask, mw, nbases, d_runs, nruns, nullptr, 0, nullptr,
                                            nullptr));
            uint64_t nh = 0;
            rc = blamm_hip_scan_finish(c, &nh);
            // cap 0: the hit count always overflows; only scratch overflow matters
            if (rc == BLAMM_OK || c->stats.candidates <= c->cand.n) break;
            if (attempt >= 4) return fail(BLAMM_ERR_DEVICE, "autotune: scratch retries exhausted");
        }
        if (i > 0) *best = std::min(*best, (double)c->stats.kernel_ms);
    }
    return BLAMM_OK;
}

}  // namespace

extern "C" {

int blamm_hip_autotune(blamm_hip_ctx* c, const uint64_t* d_packed, uint64_t packed_words, const uint64_t* d_lcmask,
                       uint64_t mask_words, uint64_t nbases, const uint64_t* d_run_starts, uint64_t nruns,
                       const char* cache_path, blamm_tuning* chosen, double* chosen_ms)
{
    if (!c) return fail(BLAMM_ERR_ARG, "autotune: null context");
    if (!c->have_lib) return fail(BLAMM_ERR_STATE, "autotune: no library uploaded");
    if (c->pending) return fail(BLAMM_ERR_STATE, "autotune: a scan is in flight");
    if (nbases == 0) return fail(BLAMM_ERR_ARG, "autotune: empty sample");
    const std::string key = tune_key(c, nbases, d_lcmask != nullptr);
    blamm_tuning best = c->tuning;
    double best_ms = 0.0;
    if (tune_lookup(cache_path, key, &best, &best_ms)) {
        BLAMM_TRY(blamm_hip_set_tuning(c, &best));
    } else {
        const uint32_t nbmax = prec_bf(c->precision) ? kNbMaxBf16 : kNbMaxFp32;
        const uint32_t ntiles = (uint32_t)((nbases + kBf16Tile - 1) / kBf16Tile);
        best_ms = 1e300;
        // layouts at the default window tile, then window tiles for the best layout
        for (uint32_t kp : {4u, 8u})
            for (uint32_t nb = nbmax; nb >= 1; nb /= 2) {
                blamm_tuning t{512, nb, kp};
                BLAMM_TRY(blamm_hip_set_tuning(c, &t));
                double ms;
                BLAMM_TRY(time_scan(c, d_packed, packed_words, d_lcmask, mask_words, nbases, d_run_starts, nruns, 2, &ms));
                if (ms < best_ms) {
                    best_ms = ms;
                    best = t;
                }
            }
        for (uint32_t ct : {64u, 128u, 256u, 1024u, 2048u}) {
            if (ct > 2 * ntiles) continue;
            blamm_tuning t{ct, best.blocks_per_wave, best.kpad_positions};
            BLAMM_TRY(blamm_hip_set_tuning(c, &t));
            double ms;
            BLAMM_TRY(time_scan(c, d_packed, packed_words, d_lcmask, mask_words, nbases, d_run_starts, nruns, 2, &ms));
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
        BLAMM_TRY(blamm_hip_set_tuning(c, &best));
        if (cache_path && *cache_path) {
            std::ofstream out(cache_path, std::ios::app);
            if (!out) return fail(BLAMM_ERR_ARG, std::string("autotune: cannot write cache ") + cache_path);
            out << key << " " << best.chunk_tiles << " " << best.blocks_per_wave << " " << best.kpad_positions << " "
                << best_ms << "\n";
        }
    }
    if (chosen) *chosen = best;
    if (chosen_ms) *chosen_ms = best_ms;
    return BLAMM_OK;
}

int blamm_hip_autotune_host(blamm_hip_ctx* c, const uint64_t* packed, const uint64_t* lcmask, uint64_t nbases,
                            const uint64_t* run_starts, uint64_t nruns, const char* cache_path,
                            blamm_tuning* chosen, double* chosen_ms)
{
    if (!c) return fail(BLAMM_ERR_ARG, "autotune: null context");
    if (nbases == 0 || !packed) return fail(BLAMM_ERR_ARG, "autotune: empty sample");
    if (nruns > 0 && !run_starts) return fail(BLAMM_ERR_ARG, "autotune: null run starts");
    if (nbases >= (1ull << 32) - 512) return fail(BLAMM_ERR_ARG, "autotune: sample must be < 2^32 - 512 bases");
    HIP_TRY(hipSetDevice(c->device));
    const uint64_t pw = blamm_hip_packed_words(nbases), mw = blamm_hip_mask_words(nbases);
    HIP_TRY(c->packed.reserve(pw));
    HIP_TRY(hipMemsetAsync(c->packed.p, 0, pw * 8, c->stream));
    HIP_TRY(hipMemcpyAsync(c->packed.p, packed, ((nbases + 31) / 32) * 8, hipMemcpyHostToDevice, c->stream));
    if (lcmask) {
        HIP_TRY(c->lcmask.reserve(mw));
        HIP_TRY(hipMemsetAsync(c->lcmask.p, 0, mw * 8, c->stream));
        HIP_TRY(hipMemcpyAsync(c->lcmask.p, lcmask, ((nbases + 63) / 64) * 8, hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(c->runs.reserve(nruns + 1));
    if (nruns > 0)
        HIP_TRY(hipMemcpyAsync(c->runs.p, run_starts, nruns * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return blamm_hip_autotune(c, c->packed.p, pw, lcmask ? c->lcmask.p : nullptr, mw, nbases, c->runs.p, nruns,
                              cache_path, chosen, chosen_ms);
}

}  // extern "C"
