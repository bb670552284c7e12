// Batch slots: many requests (streams of uneven length, each with its own
// thresholds and column subset) scored by one persistent filter launch over
// an explicit work list, then one finalize launch.  The launch sequence is
// captured once into a hipGraph per slot and replayed for every batch: the
// kernels read the batch's sizes from a device descriptor, so only buffer
// growth or a library relayout forces a new capture.
//
// Included once, at the end of blamm_hip.hip (same translation unit as the
// kernels and the library layout).

#include <memory>
#include <mutex>

namespace {

constexpr uint64_t kReqAlign = 64;          // bases; keeps both packed and mask words aligned
constexpr uint64_t kBatchMaxBases = (1ull << 32) - 4096;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Pinned host buffers for finished batches' hits.  A buffer stays with the
// requests that point into it (the server hands out slices without copying)
// and returns to the pool when the last of them lets go, so a busy server
// recycles pinned memory instead of allocating and freeing it.
struct HitBuf {
    blamm_hit* p = nullptr;
    size_t n = 0;
};

struct HitPool {
    std::mutex mu;
    std::vector<HitBuf> free;
    ~HitPool()
    {
        for (HitBuf& h : free) (void)hipHostFree(h.p);
    }
};

int acquire_hits(const std::shared_ptr<HitPool>& pool, size_t want, std::shared_ptr<HitBuf>* out)
{
    HitBuf got;
    {
        std::lock_guard<std::mutex> g(pool->mu);
        size_t best = pool->free.size();
        for (size_t i = 0; i < pool->free.size(); ++i)
            if (pool->free[i].n >= want && (best == pool->free.size() || pool->free[i].n < pool->free[best].n))
                best = i;
        if (best < pool->free.size()) {
            got = pool->free[best];
            pool->free.erase(pool->free.begin() + best);
        }
    }
    if (!got.p) {
        got.n = std::max<size_t>(want + want / 4, 1u << 16);
        hipError_t e = hipHostMalloc((void**)&got.p, got.n * sizeof(blamm_hit), hipHostMallocDefault);
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? BLAMM_ERR_NOMEM : BLAMM_ERR_DEVICE,
                        std::string("pinned hit buffer: ") + hipGetErrorString(e));
    }
    std::shared_ptr<HitPool> keep = pool;
    *out = std::shared_ptr<HitBuf>(new HitBuf(got), [keep](HitBuf* h) {
        {
            std::lock_guard<std::mutex> g(keep->mu);
            keep->free.push_back(*h);
        }
        delete h;
    });
    return BLAMM_OK;
}

}  // namespace

struct blamm_hip_batch {
    blamm_hip_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_done = nullptr;

    // device buffers
    DevBuf<uint64_t> packed, lcmask, runs;
    // the batch's own column blocks: each request's selected columns
    DevBuf<float> thr;              // per batch slot: the request's threshold
    DevBuf<uint32_t> slot_map;      // per batch slot: library slot (kNoSlot = empty)
    DevBuf<uint32_t> blk_req;       // per batch block: request
    DevBuf<uint32_t> groups;        // column groups (group_pack)
    DevBuf<uint4> frag_a, frag_b;   // gathered fragments (fp32 weights in frag_a)
    DevBuf<double> band, band_hi;   // per batch slot
    DevBuf<uint32_t> req_start, req_end, req_count;
    DevBuf<unsigned long long> req_off, cursor;
    DevBuf<uint2> items;
    DevBuf<uint4> cand;
    DevBuf<uint4> chunks;
    DevBuf<BatchDesc> desc;
    DevBuf<unsigned long long> counters;
    DevBuf<blamm_hit> hits, grouped;
    uint64_t buf_gen = 1;           // bumped whenever a device buffer moves

    // pinned staging
    HostBuf<uint8_t> stage;
    HostBuf<unsigned long long> cnt_h, off_h;
    std::shared_ptr<HitPool> pool = std::make_shared<HitPool>();
    std::shared_ptr<HitBuf> out;    // hits of the last finished batch, grouped by request

    // captured launch sequence per lowercase variant
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint64_t exec_buf_gen[2] = {0, 0}, exec_lib_gen[2] = {0, 0};

    // current batch
    bool pending = false;
    bool mask = false;
    bool done = false;
    uint32_t nreq = 0;
    std::vector<uint64_t> start_h, n_h;
    std::vector<uint64_t> out_off;
    blamm_scan_stats stats{};
    uint64_t replays = 0;
};

namespace {

void drop_graphs(blamm_hip_batch* b)
{
    for (int i = 0; i < 2; ++i) {
        if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
        b->exec[i] = nullptr;
    }
}

// Grow a device buffer; a move invalidates the captured graphs.
template <typename T>
int batch_grow(blamm_hip_batch* b, DevBuf<T>& buf, size_t want)
{
    if (want <= buf.n && buf.p) return BLAMM_OK;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(buf.reserve(want + want / 8));
    b->buf_gen++;
    return BLAMM_OK;
}

int batch_enqueue_kernels(blamm_hip_batch* b, bool mask, uint32_t grid)
{
    blamm_hip_ctx* c = b->ctx;
    const bool bf = prec_bf(c->precision);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, b->stream, b->counters.p);
    HIP_TRY(hipGetLastError());
    // the batch's column blocks from the library's fragments
    GatherArgs ga;
    ga.desc = b->desc.p;
    ga.slot_map = b->slot_map.p;
    ga.lib_a = (const uint4*)c->frag_a.p;
    ga.lib_b = (const uint4*)c->frag_b.p;
    ga.lib_w = (const float*)c->frag_a.p;
    ga.lib_band = c->slot_band_any.p;
    ga.lib_band_hi = c->slot_band_hi_any.p;
    ga.out_a = b->frag_a.p;
    ga.out_b = b->frag_b.p;
    ga.out_w = (float*)b->frag_a.p;
    ga.out_band = b->band.p;
    ga.out_band_hi = b->band_hi.p;
    if (bf) hipLaunchKernelGGL(gather_kernel<true>, dim3((uint32_t)c->num_cu * 4), dim3(256), 0, b->stream, ga);
    else hipLaunchKernelGGL(gather_kernel<false>, dim3((uint32_t)c->num_cu * 4), dim3(256), 0, b->stream, ga);
    HIP_TRY(hipGetLastError());
    FilterArgs a;
    a.packed = b->packed.p;
    a.lcmask = mask ? b->lcmask.p : nullptr;
    a.frag_a = b->frag_a.p;
    a.frag_b = b->frag_b.p;
    a.thr = b->thr.p;
    a.band = b->band.p;
    a.band_hi = b->band_hi.p;
    a.groups = b->groups.p;
    a.cand = b->cand.p;
    a.counters = b->counters.p;
    a.cand_cap = b->cand.n;
    const WorkQueue q{0, 0, 1, 1};
    const ListArgs l{b->desc.p, b->items.p, b->chunks.p};
    hipError_t e = launch_filter(bf, prec_lo(c->precision), mask, true, a, q, l, grid, b->stream);
    if (e != hipSuccess) return fail(BLAMM_ERR_DEVICE, std::string("batch filter launch: ") + hipGetErrorString(e));
    FinalizeArgs f = finalize_args(c, b->cand.p, b->cand.n, b->counters.p, b->packed.p, mask ? b->lcmask.p : nullptr,
                                   b->runs.p);
    f.slot_thr = b->thr.p;
    f.hits = b->hits.p;
    f.hit_cap = b->hits.n;
    f.desc = b->desc.p;
    f.req_end = b->req_end.p;
    f.slot_map = b->slot_map.p;
    f.blk_req = b->blk_req.p;
    e = launch_finalize(bf, true, f, (uint32_t)c->num_cu * 8, b->stream);
    if (e != hipSuccess) return fail(BLAMM_ERR_DEVICE, std::string("batch finalize launch: ") + hipGetErrorString(e));
    // group the hits by request on the device
    GroupArgs g;
    g.desc = b->desc.p;
    g.counters = b->counters.p;
    g.cap = std::min(b->hits.n, b->grouped.n);
    g.in = (const DevHit*)b->hits.p;
    g.out = (DevHit*)b->grouped.p;
    g.req_count = b->req_count.p;
    g.req_off = b->req_off.p;
    g.cursor = b->cursor.p;
    g.req_start = b->req_start.p;
    const uint32_t gg = (uint32_t)c->num_cu * 4;
    hipLaunchKernelGGL(group_clear_kernel, dim3(64), dim3(kGroupThreads), 0, b->stream, g);
    hipLaunchKernelGGL(group_count_kernel, dim3(gg), dim3(kGroupThreads), 0, b->stream, g);
    hipLaunchKernelGGL(group_scan_kernel, dim3(1), dim3(1024), 0, b->stream, g);
    hipLaunchKernelGGL(group_scatter_kernel, dim3(gg), dim3(kGroupThreads), 0, b->stream, g);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(BLAMM_ERR_DEVICE, std::string("batch grouping launch: ") + hipGetErrorString(e));
    return BLAMM_OK;
}

// Launch the captured sequence (capturing it first when buffers or the
// library moved since the last capture), then copy the counters back.
int batch_launch(blamm_hip_batch* b)
{
    blamm_hip_ctx* c = b->ctx;
    const int v = b->mask ? 1 : 0;
    if (!b->exec[v] || b->exec_buf_gen[v] != b->buf_gen || b->exec_lib_gen[v] != c->lib_gen) {
        if (b->exec[v]) (void)hipGraphExecDestroy(b->exec[v]);
        b->exec[v] = nullptr;
        const bool bf = prec_bf(c->precision);
        const uint32_t grid =
            (uint32_t)c->num_cu * (uint32_t)persistent_wgs_per_cu(bf, prec_lo(c->precision), b->mask, true);
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_enqueue_kernels(b, b->mask, grid);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != BLAMM_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(BLAMM_ERR_DEVICE, std::string("batch capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&b->exec[v], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            b->exec[v] = nullptr;
            return fail(BLAMM_ERR_DEVICE, std::string("batch graph instantiate: ") + hipGetErrorString(ei));
        }
        b->exec_buf_gen[v] = b->buf_gen;
        b->exec_lib_gen[v] = c->lib_gen;
    }
    HIP_TRY(hipEventRecord(b->ev_start, b->stream));
    HIP_TRY(hipGraphLaunch(b->exec[v], b->stream));
    HIP_TRY(hipEventRecord(b->ev_end, b->stream));
    HIP_TRY(hipMemcpyAsync(b->cnt_h.p, b->counters.p, kCntSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipMemcpyAsync(b->off_h.p, b->req_off.p, ((size_t)b->nreq + 1) * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipEventRecord(b->ev_done, b->stream));
    return BLAMM_OK;
}

}  // namespace

extern "C" {

int blamm_hip_batch_create(blamm_hip_ctx* c, uint64_t max_bases, uint32_t max_requests, blamm_hip_batch** out)
{
    if (!c || !out) return fail(BLAMM_ERR_ARG, "batch_create: null pointer");
    *out = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    blamm_hip_batch* b = new blamm_hip_batch();
    b->ctx = c;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&b->ev_end);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming);
    if (e == hipSuccess) e = b->counters.reserve(kCntSlots);
    if (e == hipSuccess) e = b->desc.reserve(1);
    if (e == hipSuccess) e = b->cnt_h.reserve(kCntSlots);
    const uint64_t nb = std::min<uint64_t>(max_bases, kBatchMaxBases);
    const uint64_t nr = std::max<uint32_t>(max_requests, 1);
    if (e == hipSuccess) e = b->packed.reserve(nb / 32 + nr * 2 + 16);
    if (e == hipSuccess) e = b->cand.reserve(std::max<uint64_t>(nb / 256, 1u << 16));
    if (e == hipSuccess) e = b->hits.reserve(std::max<uint64_t>(nb / 256, 1u << 16));
    if (e == hipSuccess) e = b->grouped.reserve(b->hits.n);
    if (e != hipSuccess) {
        blamm_hip_batch_destroy(b);
        return fail(e == hipErrorOutOfMemory ? BLAMM_ERR_NOMEM : BLAMM_ERR_DEVICE,
                    std::string("batch_create: ") + hipGetErrorString(e));
    }
    *out = b;
    return BLAMM_OK;
}

int blamm_hip_batch_destroy(blamm_hip_batch* b)
{
    if (!b) return BLAMM_OK;
    (void)hipSetDevice(b->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    b->packed.release(); b->lcmask.release(); b->runs.release(); b->thr.release(); b->slot_map.release();
    b->blk_req.release(); b->groups.release(); b->frag_a.release(); b->frag_b.release(); b->band.release();
    b->band_hi.release();
    b->req_start.release();
    b->req_end.release(); b->req_count.release(); b->req_off.release(); b->cursor.release();
    b->items.release(); b->cand.release(); b->chunks.release(); b->desc.release(); b->counters.release();
    b->hits.release(); b->grouped.release();
    b->stage.release(); b->cnt_h.release(); b->off_h.release();
    b->out.reset();
    if (b->ev_start) (void)hipEventDestroy(b->ev_start);
    if (b->ev_end) (void)hipEventDestroy(b->ev_end);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return BLAMM_OK;
}

uint64_t blamm_hip_batch_device_bytes(const blamm_hip_batch* b)
{
    if (!b) return 0;
    return b->packed.n * 8 + b->lcmask.n * 8 + b->runs.n * 8 + b->thr.n * 4 + b->slot_map.n * 4 + b->blk_req.n * 4 +
           b->groups.n * 4 + (b->frag_a.n + b->frag_b.n) * 16 + (b->band.n + b->band_hi.n) * 8 +
           (b->req_start.n + b->req_end.n + b->req_count.n) * 4 + (b->req_off.n + b->cursor.n) * 8 + b->items.n * 8 +
           b->cand.n * 16 + b->chunks.n * 16 + (b->hits.n + b->grouped.n) * sizeof(blamm_hit) + sizeof(BatchDesc) +
           kCntSlots * 8;
}

int blamm_hip_batch_submit(blamm_hip_batch* b, const blamm_request* reqs, uint32_t nreq)
{
    if (!b) return fail(BLAMM_ERR_ARG, "batch_submit: null batch");
    blamm_hip_ctx* c = b->ctx;
    if (!c->have_lib) return fail(BLAMM_ERR_STATE, "batch_submit: no library uploaded");
    if (b->pending) return fail(BLAMM_ERR_STATE, "batch_submit: previous batch not waited for");
    if (nreq == 0 || !reqs) return fail(BLAMM_ERR_ARG, "batch_submit: no requests");
    if (nreq >= (1u << 16)) return fail(BLAMM_ERR_ARG, "batch_submit: at most 65535 requests per batch");
    HIP_TRY(hipSetDevice(c->device));

    // Layout of the batch stream: request r occupies [start_r, start_r + n_r),
    // starts aligned to kReqAlign bases.
    b->start_h.assign(nreq, 0);
    b->n_h.assign(nreq, 0);
    bool mask = false;
    uint64_t total = 0, nruns = 0;
    for (uint32_t r = 0; r < nreq; ++r) {
        const blamm_request& q = reqs[r];
        if (q.nbases > 0 && !q.packed) return fail(BLAMM_ERR_ARG, "batch_submit: request " + std::to_string(r) + " has no packed stream");
        if (q.nruns > 0 && !q.run_starts) return fail(BLAMM_ERR_ARG, "batch_submit: request " + std::to_string(r) + " has null run starts");
        for (uint64_t i = 1; i < q.nruns; ++i)
            if (q.run_starts[i] < q.run_starts[i - 1])
                return fail(BLAMM_ERR_ARG, "batch_submit: run starts of request " + std::to_string(r) + " not sorted");
        if (q.thresholds)
            for (int j = 0; j < c->ncols; ++j)
                if (std::isnan(q.thresholds[j]))
                    return fail(BLAMM_ERR_ARG, "batch_submit: NaN threshold in request " + std::to_string(r));
        b->start_h[r] = total;
        b->n_h[r] = q.nbases;
        total = align_up(total + q.nbases, kReqAlign);
        if (total > kBatchMaxBases) return fail(BLAMM_ERR_ARG, "batch_submit: batch longer than 2^32 - 4096 bases");
        mask = mask || (q.lcmask != nullptr && q.nbases > 0);
        nruns += q.nruns + 1;
    }
    const uint64_t pw = total / 32 + 16, mw = mask ? total / 64 + 16 : 0;
    const size_t nslots = (size_t)c->nslots;
    const uint32_t cpb = (uint32_t)c->cols_per_blk;
    const uint32_t ct = std::max<uint32_t>(1, c->tuning.chunk_tiles);
    const bool bf = prec_bf(c->precision);
    const uint32_t nbmax = std::min<uint32_t>(c->tuning.blocks_per_wave, bf ? kNbMaxBf16 : kNbMaxFp32);
    const int kstep = (int)c->tuning.kpad_positions / 4;

    // Each request's selected columns (finite or -inf threshold), in library
    // slot order (non-decreasing length), packed into blocks of its own.
    std::vector<uint32_t> smap, breq, gpk;
    std::vector<float> sthr;
    std::vector<std::vector<uint32_t>> rgroups(nreq);
    for (uint32_t r = 0; r < nreq; ++r) {
        const float* thr = reqs[r].thresholds ? reqs[r].thresholds : c->lib_thr.data();
        std::vector<uint32_t> sel;
        for (size_t sl = 0; sl < nslots; ++sl) {
            const int col = c->slot_col_h[sl];
            if (col >= 0 && thr[col] != INFINITY) sel.push_back((uint32_t)sl);
        }
        if (sel.empty() || b->n_h[r] == 0) continue;
        const uint32_t blk_first = (uint32_t)breq.size();
        std::vector<int> nchs;
        for (size_t i = 0; i < sel.size(); i += cpb) {
            int lm = 0;
            for (uint32_t k = 0; k < cpb; ++k) {
                if (i + k < sel.size()) {
                    const uint32_t sl = sel[i + k];
                    const int col = c->slot_col_h[sl];
                    smap.push_back(sl);
                    sthr.push_back(thr[col]);
                    lm = std::max(lm, c->lib_len[col]);
                } else {
                    smap.push_back(kNoSlot);
                    sthr.push_back(INFINITY);
                }
            }
            int nch = (lm + 3) / 4;
            nch = std::min(kMaxChunks, (nch + kstep - 1) / kstep * kstep);
            nchs.push_back(nch);
            breq.push_back(r);
        }
        std::vector<GroupInfo> gr;
        for (size_t i = 0; i < nchs.size();) {
            size_t e = i;
            while (e < nchs.size() && nchs[e] == nchs[i]) ++e;
            size_t bb = i;
            const uint32_t nbg = nchs[i] > kWideChunks ? 1u : nbmax;
            for (uint32_t nb = nbg; nb >= 1; nb /= 2)
                while (e - bb >= nb) {
                    gr.push_back({blk_first + (uint32_t)bb, (uint32_t)nchs[i], nb});
                    bb += nb;
                }
            i = e;
        }
        std::stable_sort(gr.begin(), gr.end(), [](const GroupInfo& x, const GroupInfo& y) { return x.nch > y.nch; });
        for (const GroupInfo& g : gr) {
            rgroups[r].push_back((uint32_t)gpk.size());
            gpk.push_back(group_pack(g.blk0, g.nch, g.nb));
        }
    }
    const uint32_t nblk = (uint32_t)breq.size();
    if (nblk >= (1u << 20)) return fail(BLAMM_ERR_ARG, "batch_submit: too many column blocks in one batch");
    // Chunks and items, chunk k to sub-queue k % kXcds, chunk-major.
    std::vector<uint4> chunks;
    std::vector<std::vector<uint2>> sub(kXcds);
    for (uint32_t r = 0; r < nreq; ++r) {
        if (b->n_h[r] == 0 || rgroups[r].empty()) continue;
        const uint32_t tb = (uint32_t)(b->start_h[r] / kBf16Tile);
        const uint32_t te = (uint32_t)((b->start_h[r] + b->n_h[r] + kBf16Tile - 1) / kBf16Tile);
        for (uint32_t t = tb; t < te; t += ct) {
            const uint32_t k = (uint32_t)chunks.size();
            chunks.push_back(make_uint4(t, std::min(te, t + ct), r, 0));
            for (uint32_t g : rgroups[r]) sub[k % kXcds].push_back(make_uint2(k, g));
        }
    }
    BatchDesc d{};
    size_t nitems = 0;
    for (int x = 0; x < kXcds; ++x) {
        d.sub_begin[x] = (uint32_t)nitems;
        nitems += sub[x].size();
    }
    d.sub_begin[kXcds] = (uint32_t)nitems;
    if (nitems >= (1ull << 32)) return fail(BLAMM_ERR_ARG, "batch_submit: work list too long");
    d.nreq = nreq;
    d.nbases = (uint32_t)total;
    d.nruns = (uint32_t)nruns;
    d.nblk = nblk;

    // Device capacity (growth invalidates captured graphs).
    BLAMM_TRY(batch_grow(b, b->packed, pw));
    if (mask) BLAMM_TRY(batch_grow(b, b->lcmask, mw));
    BLAMM_TRY(batch_grow(b, b->runs, nruns + 1));
    const size_t nbs = std::max<size_t>(smap.size(), 1);
    BLAMM_TRY(batch_grow(b, b->thr, nbs));
    BLAMM_TRY(batch_grow(b, b->slot_map, nbs));
    BLAMM_TRY(batch_grow(b, b->band, nbs));
    BLAMM_TRY(batch_grow(b, b->band_hi, nbs));
    BLAMM_TRY(batch_grow(b, b->blk_req, std::max<size_t>(nblk, 1)));
    BLAMM_TRY(batch_grow(b, b->groups, std::max<size_t>(gpk.size(), 1)));
    const size_t fr = (size_t)std::max<uint32_t>(nblk, 1) * (bf ? (size_t)kMaxChunks * 64 : (size_t)kMaxLen * 16);
    BLAMM_TRY(batch_grow(b, b->frag_a, fr));
    if (bf) BLAMM_TRY(batch_grow(b, b->frag_b, fr));
    else BLAMM_TRY(batch_grow(b, b->frag_b, 1));
    BLAMM_TRY(batch_grow(b, b->req_start, nreq));
    BLAMM_TRY(batch_grow(b, b->req_end, nreq));
    BLAMM_TRY(batch_grow(b, b->req_count, nreq));
    BLAMM_TRY(batch_grow(b, b->req_off, (size_t)nreq + 1));
    BLAMM_TRY(batch_grow(b, b->cursor, nreq));
    if (b->off_h.n < (size_t)nreq + 1) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        HIP_TRY(b->off_h.reserve((size_t)nreq + 1 + nreq / 8));
    }
    BLAMM_TRY(batch_grow(b, b->items, std::max<size_t>(nitems, 1)));
    // room for the expected candidates and hits (records <= hits + rescored
    // windows; 1.5x margin), so the batch rarely needs a second run
    uint64_t expect = 0;
    for (uint32_t r = 0; r < nreq; ++r) expect += reqs[r].expected_hits;
    if (expect) {
        expect += expect / 2 + 4096;
        BLAMM_TRY(batch_grow(b, b->cand, expect));
        BLAMM_TRY(batch_grow(b, b->hits, expect));
        BLAMM_TRY(batch_grow(b, b->grouped, b->hits.n));
    }
    BLAMM_TRY(batch_grow(b, b->chunks, std::max<size_t>(chunks.size(), 1)));

    // Stage everything in pinned memory, then one copy per buffer.
    const size_t o_pk = 0;
    const size_t o_mk = align_up(o_pk + pw * 8, 256);
    const size_t o_rn = align_up(o_mk + mw * 8, 256);
    const size_t o_tb = align_up(o_rn + (nruns + 1) * 8, 256);        // thresholds
    const size_t o_sm = align_up(o_tb + sthr.size() * 4, 256);        // slot map
    const size_t o_br = align_up(o_sm + smap.size() * 4, 256);        // block requests
    const size_t o_gp = align_up(o_br + (size_t)nblk * 4, 256);       // groups
    const size_t o_re = align_up(o_gp + gpk.size() * 4, 256);
    const size_t o_rs = align_up(o_re + (size_t)nreq * 4, 256);
    const size_t o_ch = align_up(o_rs + (size_t)nreq * 4, 256);
    const size_t o_it = align_up(o_ch + chunks.size() * 16, 256);
    const size_t o_ds = align_up(o_it + nitems * 8, 256);
    const size_t o_end = o_ds + sizeof(BatchDesc);
    if (o_end > b->stage.n) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        HIP_TRY(b->stage.reserve(o_end + o_end / 8));
    }
    uint8_t* S = b->stage.p;
    uint64_t* pk = (uint64_t*)(S + o_pk);
    uint64_t* mk = (uint64_t*)(S + o_mk);
    uint64_t* rn = (uint64_t*)(S + o_rn);
    uint32_t* re = (uint32_t*)(S + o_re);
    uint32_t* rs = (uint32_t*)(S + o_rs);
    std::memset(pk, 0, pw * 8);
    if (mask) std::memset(mk, 0, mw * 8);
    size_t ir = 0;
    for (uint32_t r = 0; r < nreq; ++r) {
        const blamm_request& q = reqs[r];
        const uint64_t s0 = b->start_h[r];
        if (q.nbases > 0) {
            std::memcpy(pk + s0 / 32, q.packed, ((q.nbases + 31) / 32) * 8);
            if (q.lcmask) std::memcpy(mk + s0 / 64, q.lcmask, ((q.nbases + 63) / 64) * 8);
        }
        rn[ir++] = s0;
        for (uint64_t i = 0; i < q.nruns; ++i)
            if (q.run_starts[i] > 0 && q.run_starts[i] < q.nbases) rn[ir++] = s0 + q.run_starts[i];
        re[r] = (uint32_t)(s0 + q.nbases);
        rs[r] = (uint32_t)s0;
    }
    d.nruns = (uint32_t)ir;
    std::memcpy(S + o_tb, sthr.data(), sthr.size() * 4);
    std::memcpy(S + o_sm, smap.data(), smap.size() * 4);
    std::memcpy(S + o_br, breq.data(), (size_t)nblk * 4);
    std::memcpy(S + o_gp, gpk.data(), gpk.size() * 4);
    std::memcpy(S + o_ch, chunks.data(), chunks.size() * 16);
    size_t at = 0;
    for (int x = 0; x < kXcds; ++x) {
        std::memcpy(S + o_it + at * 8, sub[x].data(), sub[x].size() * 8);
        at += sub[x].size();
    }
    std::memcpy(S + o_ds, &d, sizeof(d));

    hipStream_t s = b->stream;
    HIP_TRY(hipMemcpyAsync(b->packed.p, S + o_pk, pw * 8, hipMemcpyHostToDevice, s));
    if (mask) HIP_TRY(hipMemcpyAsync(b->lcmask.p, S + o_mk, mw * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->runs.p, S + o_rn, std::max<size_t>(ir, 1) * 8, hipMemcpyHostToDevice, s));
    if (!smap.empty()) {
        HIP_TRY(hipMemcpyAsync(b->thr.p, S + o_tb, sthr.size() * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(b->slot_map.p, S + o_sm, smap.size() * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(b->blk_req.p, S + o_br, (size_t)nblk * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(b->groups.p, S + o_gp, gpk.size() * 4, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipMemcpyAsync(b->req_end.p, S + o_re, (size_t)nreq * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->req_start.p, S + o_rs, (size_t)nreq * 4, hipMemcpyHostToDevice, s));
    if (!chunks.empty()) HIP_TRY(hipMemcpyAsync(b->chunks.p, S + o_ch, chunks.size() * 16, hipMemcpyHostToDevice, s));
    if (nitems) HIP_TRY(hipMemcpyAsync(b->items.p, S + o_it, nitems * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->desc.p, S + o_ds, sizeof(BatchDesc), hipMemcpyHostToDevice, s));

    b->mask = mask;
    b->nreq = nreq;
    b->done = false;
    BLAMM_TRY(batch_launch(b));
    b->pending = true;
    return BLAMM_OK;
}

int blamm_hip_batch_query(blamm_hip_batch* b)
{
    if (!b) return fail(BLAMM_ERR_ARG, "batch_query: null batch");
    if (!b->pending) return BLAMM_OK;
    const hipError_t e = hipEventQuery(b->ev_done);
    if (e == hipSuccess) return BLAMM_OK;
    if (e == hipErrorNotReady) return 1;
    return fail(BLAMM_ERR_DEVICE, std::string("batch_query: ") + hipGetErrorString(e));
}

int blamm_hip_batch_wait(blamm_hip_batch* b, uint64_t* nhits)
{
    if (!b) return fail(BLAMM_ERR_ARG, "batch_wait: null batch");
    if (!b->pending) return fail(BLAMM_ERR_STATE, "batch_wait: no batch in flight");
    blamm_hip_ctx* c = b->ctx;
    HIP_TRY(hipSetDevice(c->device));
    unsigned long long cand = 0, nh = 0;
    for (int attempt = 0;; ++attempt) {
        hipError_t e = hipEventSynchronize(b->ev_done);
        if (e != hipSuccess) {
            b->pending = false;
            return fail(BLAMM_ERR_DEVICE, std::string("batch_wait: ") + hipGetErrorString(e));
        }
        cand = b->cnt_h.p[kCntCand];
        nh = b->cnt_h.p[kCntHits];
        if (cand <= b->cand.n && nh <= b->hits.n) break;
        if (attempt >= 4) {
            b->pending = false;
            return fail(BLAMM_ERR_DEVICE, "batch_wait: capacity retries exhausted");
        }
        // Grow what overflowed and replay the batch (its inputs are still on
        // the device).
        b->replays++;
        if (cand > b->cand.n) BLAMM_TRY(batch_grow(b, b->cand, cand + cand / 4 + 1024));
        if (nh > b->hits.n) {
            BLAMM_TRY(batch_grow(b, b->hits, nh + nh / 4 + 1024));
            BLAMM_TRY(batch_grow(b, b->grouped, b->hits.n));
        }
        BLAMM_TRY(batch_launch(b));
    }
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev_start, b->ev_end));
    b->stats.candidates = cand;
    b->stats.band_rescored = b->cnt_h.p[kCntBand];
    b->stats.hits = nh;
    b->stats.kernel_ms = ms;
    b->stats.filter_ms = 0.f;
    // hits arrive grouped by request with request-local positions
    std::shared_ptr<HitBuf> buf;
    BLAMM_TRY(acquire_hits(b->pool, std::max<uint64_t>(nh, 1), &buf));
    if (nh > 0) {
        HIP_TRY(hipMemcpyAsync(buf->p, b->grouped.p, nh * sizeof(blamm_hit), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
    }
    b->out = buf;
    b->out_off.assign(b->off_h.p, b->off_h.p + b->nreq + 1);
    b->pending = false;
    b->done = true;
    if (nhits) *nhits = nh;
    return BLAMM_OK;
}

int blamm_hip_batch_hits(blamm_hip_batch* b, uint32_t req, const blamm_hit** hits, uint64_t* n)
{
    if (!b || !hits || !n) return fail(BLAMM_ERR_ARG, "batch_hits: null pointer");
    if (!b->done) return fail(BLAMM_ERR_STATE, "batch_hits: no finished batch");
    if (req >= b->nreq) return fail(BLAMM_ERR_ARG, "batch_hits: request index out of range");
    *hits = b->out->p + b->out_off[req];
    *n = b->out_off[req + 1] - b->out_off[req];
    return BLAMM_OK;
}

int blamm_hip_batch_stats(blamm_hip_batch* b, blamm_scan_stats* st)
{
    if (!b || !st) return fail(BLAMM_ERR_ARG, "batch_stats: null pointer");
    *st = b->stats;
    return BLAMM_OK;
}

}  // extern "C"
