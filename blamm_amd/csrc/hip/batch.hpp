// Batch slots: many requests (streams of uneven length, each with its own
// thresholds and column subset) scored by one persistent filter launch over
// an explicit work list, then one finalize launch.  The launch sequence is
// captured once into a hipGraph per slot and replayed for every batch: the
// kernels read the batch's sizes from a device descriptor, so only buffer
// growth or a library relayout forces a new capture.
//
// Included once, at the end of blamm_hip.hip (same translation unit as the
// kernels and the library layout).

#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kReqAlign = 64;          // bases; keeps both packed and mask words aligned
constexpr uint64_t kBatchMaxBases = (1ull << 32) - 4096;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Pinned host buffers for finished batches' hits.  A buffer stays with the
// requests that point into it (the server hands out slices without copying)
// and returns to the pool when the last of them lets go, so a busy server
// recycles pinned memory instead of allocating and freeing it.
struct HitBuf {
    GroupedHit* p = nullptr;
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

// GroupedHit (request-local 32-bit position) -> blamm_hit.
void expand_hits(const GroupedHit* in, uint64_t n, blamm_hit* out)
{
    for (uint64_t i = 0; i < n; ++i) {
        out[i].pos = in[i].pos;
        out[i].col = in[i].col;
        out[i].score = in[i].score;
    }
}

// The same for a whole batch: large ones are split over up to 16 threads (a
// single thread moves ~10 GB/s, a batch of 20M hits is 0.6 GB).
void expand_hits_parallel(const GroupedHit* in, uint64_t n, blamm_hit* out)
{
    const uint64_t per = 1ull << 21;
    const unsigned nt = (unsigned)std::min<uint64_t>({(n + per - 1) / per, 16ull,
                                                      std::max(1u, std::thread::hardware_concurrency())});
    if (nt <= 1) {
        expand_hits(in, n, out);
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) {
        const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
        pool.emplace_back([=] { expand_hits(in + b, e - b, out + b); });
    }
    expand_hits(in, n / nt, out);
    for (auto& th : pool) th.join();
}

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
        hipError_t e = hipHostMalloc((void**)&got.p, got.n * sizeof(GroupedHit), hipHostMallocDefault);
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

// Hit read-back on the GPU's DMA engine.  hipMemcpyAsync performs large
// device-to-host copies with blit kernels on the compute units, which then
// compete with the other slots' kernels: skipping the read-back altogether
// raised serving throughput by a third, although PCIe carried only ~20 GB/s.
// An HSA async copy runs on an SDMA engine instead.  Coherence: the copy
// starts after the batch's completion event (a default HIP event, whose
// completion releases the kernels' writes to system scope) has been waited
// for, and the completer waits for the copy's completion signal with acquire
// semantics.  When the agents cannot be resolved the batch keeps
// hipMemcpyAsync.
//
// Engine choice: MI355X has 16 SDMA engines; the agent reports the first
// HSA_AMD_AGENT_INFO_NUM_SDMA_ENG (2) as the engines for host traffic and the
// rest as xGMI engines.  Measured alone, device-to-host: engines 0-3 ~56 GB/s,
// 4-7 12.7 GB/s, 8-15 7-10 GB/s.  Measured under a serving run (3 slots,
// p = 1e-3): engines 0, 2 and 3 56.8 GB/s, engine 1 (the runtime's preferred
// device-to-host engine, where its own queues sit) 30 GB/s.  Left to the
// runtime, the read-backs moved between such engines and ran at ~30 GB/s for
// long stretches.  They are therefore issued on the first host engine;
// without an engine count the runtime picks.
struct DmaReadback {
    bool ready = false;
    hsa_agent_t gpu{}, cpu{};
    hsa_signal_t sig{};
    uint64_t ticks_per_s = 0;
    uint32_t engine = 0;   // hsa_amd_sdma_engine_id_t bit, 0 = the runtime's choice
};

// Start copying `bytes` from device memory to pinned host memory; completion
// brings d.sig from 1 to 0.
hsa_status_t dma_start(DmaReadback& d, void* dst, const void* src, size_t bytes)
{
    hsa_signal_store_screlease(d.sig, 1);
    if (d.engine)
        return hsa_amd_memory_async_copy_on_engine(dst, d.cpu, src, d.gpu, bytes, 0, nullptr, d.sig,
                                                   (hsa_amd_sdma_engine_id_t)d.engine, false);
    return hsa_amd_memory_async_copy(dst, d.cpu, src, d.gpu, bytes, 0, nullptr, d.sig);
}

hsa_status_t find_cpu_agent(hsa_agent_t a, void* out)
{
    hsa_device_type_t t;
    if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) == HSA_STATUS_SUCCESS && t == HSA_DEVICE_TYPE_CPU) {
        *(hsa_agent_t*)out = a;
        return HSA_STATUS_INFO_BREAK;
    }
    return HSA_STATUS_SUCCESS;
}

// The GPU agent that owns `device_ptr`, a CPU agent and a completion signal.
bool dma_setup(DmaReadback& d, const void* device_ptr)
{
    hsa_amd_pointer_info_t info{};
    info.size = sizeof(info);
    if (hsa_amd_pointer_info(device_ptr, &info, nullptr, nullptr, nullptr) != HSA_STATUS_SUCCESS) return false;
    if (info.type != HSA_EXT_POINTER_TYPE_HSA) return false;
    hsa_device_type_t t;
    if (hsa_agent_get_info(info.agentOwner, HSA_AGENT_INFO_DEVICE, &t) != HSA_STATUS_SUCCESS ||
        t != HSA_DEVICE_TYPE_GPU)
        return false;
    d.gpu = info.agentOwner;
    d.cpu.handle = 0;
    (void)hsa_iterate_agents(find_cpu_agent, &d.cpu);
    if (d.cpu.handle == 0) return false;
    if (hsa_system_get_info(HSA_SYSTEM_INFO_TIMESTAMP_FREQUENCY, &d.ticks_per_s) != HSA_STATUS_SUCCESS ||
        d.ticks_per_s == 0)
        return false;
    if (hsa_signal_create(0, 0, nullptr, &d.sig) != HSA_STATUS_SUCCESS) return false;
    d.engine = 0;
    uint32_t nhost = 0, avail = 0;
    if (hsa_agent_get_info(d.gpu, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_NUM_SDMA_ENG, &nhost) == HSA_STATUS_SUCCESS &&
        nhost > 0 && nhost <= 16 &&
        hsa_amd_memory_copy_engine_status(d.cpu, d.gpu, &avail) == HSA_STATUS_SUCCESS) {
        const uint32_t host = ((1u << nhost) - 1u) & avail;
        d.engine = host & (~host + 1u);   // first available host engine (0 if none)
    }
    d.ready = true;
    return true;
}

// Gathered fragments per column block of a batch (uint4 elements of frag_a,
// and of frag_b in the split-BF16 forms).
size_t batch_frag_elems(bool bf)
{
    return bf ? (size_t)kMaxChunks * 64 : (size_t)kMaxLen * 16;
}

// Device bytes of one column block of a batch: its fragments and the
// per-slot threshold, slot map and band tables.
uint64_t batch_block_bytes(const blamm_hip_ctx* c)
{
    const bool bf = prec_bf(c->precision);
    return (bf ? 2 : 1) * batch_frag_elems(bf) * sizeof(uint4) + (uint64_t)c->cols_per_blk * (4 + 4 + 8 + 8);
}

// Host-written tables of a batch live in one device arena laid out like their
// pinned staging image, so that one copy uploads them all.  Each table has a
// capacity; a batch that needs more grows the table geometrically (the arena
// moves and the captured graphs are dropped).  The work items come last, so
// the copy stops at the end of the batch's items.
enum MetaTable {
    kMetaDesc, kMetaReqStart, kMetaReqEnd, kMetaThr, kMetaSlotMap, kMetaBlkReq, kMetaGroups, kMetaChunks,
    kMetaRuns, kMetaItems, kMetaTables
};
constexpr size_t kMetaElem[kMetaTables] = {sizeof(BatchDesc), 4, 4, 4, 4, 4, 4, 16, 8, 8};

template <typename T>
struct MetaView {
    T* p = nullptr;
};

}  // namespace

struct blamm_hip_batch {
    blamm_hip_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_done = nullptr, ev_copy = nullptr;

    // device buffers
    DevBuf<uint64_t> packed, lcmask, run_bits;
    // host-written tables (views into `meta`)
    DevBuf<uint8_t> meta;
    size_t meta_cap[kMetaTables] = {};
    size_t meta_off[kMetaTables + 1] = {};
    MetaView<BatchDesc> desc;
    MetaView<uint32_t> req_start, req_end;
    // the batch's own column blocks: each request's selected columns
    MetaView<float> thr;            // per batch slot: the request's threshold
    MetaView<uint32_t> slot_map;    // per batch slot: library slot (kNoSlot = empty)
    MetaView<uint32_t> blk_req;     // per batch block: request
    MetaView<uint32_t> groups;      // column groups (group_pack)
    MetaView<uint4> chunks;         // (t0, t1, request, 0)
    MetaView<uint64_t> runs;        // run starts of the batch stream
    MetaView<uint2> items;          // (chunk, group), per XCD sub-queue
    // device-written
    DevBuf<uint4> frag_a, frag_b;   // gathered fragments (fp32 weights in frag_a)
    DevBuf<double> band, band_hi;   // per batch slot
    DevBuf<uint32_t> req_count;
    DevBuf<unsigned long long> req_off, cursor;
    DevBuf<uint4> cand;
    DevBuf<unsigned long long> counters;
    DevBuf<DevHit> hits;            // finalize output
    DevBuf<GroupedHit> grouped;     // grouped by request, read back
    uint64_t buf_gen = 1;           // bumped whenever a device buffer moves

    // pinned staging
    HostBuf<uint8_t> stage;
    HostBuf<unsigned long long> cnt_h, off_h;
    std::shared_ptr<HitPool> pool = std::make_shared<HitPool>();
    std::shared_ptr<HitBuf> copying;    // the batch's hits while they are copied back
    uint64_t copy_nh = 0;
    DmaReadback dma;
    bool copy_dma = false;          // the copy in flight runs on the DMA engine
    std::shared_ptr<HitBuf> out;    // hits of the last finished batch, grouped by request
    // out as blamm_hit, built by the first batch_hits call; grow-only and not
    // zero-filled (every element read is written first)
    std::unique_ptr<blamm_hit[]> expanded;
    uint64_t expanded_cap = 0;
    bool have_expanded = false;

    // captured launch sequence per lowercase variant
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint64_t exec_buf_gen[2] = {0, 0}, exec_lib_gen[2] = {0, 0};

    // A server with several slots sets this: the filter then leaves one
    // workgroup slot per CU free, so the other slots' small kernels and
    // finalize run beside it instead of waiting for its persistent workgroups
    // (serving +5-6% at p = 1e-5; a lone batch is 3% slower with it).
    bool leave_room = false;

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

// Grow a device buffer (by at least half, so a slot serving batches of
// varying size settles after a few growths); a move invalidates the captured
// graphs.
template <typename T>
int batch_grow(blamm_hip_batch* b, DevBuf<T>& buf, size_t want)
{
    if (want <= buf.n && buf.p) return BLAMM_OK;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(buf.reserve(std::max(want + want / 4, buf.n + buf.n / 2)));
    b->buf_gen++;
    return BLAMM_OK;
}

// Make room for `need` elements of every host-written table.
int meta_fit(blamm_hip_batch* b, const size_t (&need)[kMetaTables])
{
    bool grow = !b->meta.p;
    for (int x = 0; x < kMetaTables; ++x) grow = grow || need[x] > b->meta_cap[x];
    if (!grow) return BLAMM_OK;
    for (int x = 0; x < kMetaTables; ++x)
        if (need[x] > b->meta_cap[x] || b->meta_cap[x] == 0)
            b->meta_cap[x] = std::max<size_t>({need[x] + need[x] / 4, b->meta_cap[x] + b->meta_cap[x] / 2, 1});
    size_t off = 0;
    for (int x = 0; x < kMetaTables; ++x) {
        b->meta_off[x] = off;
        off = align_up(off + b->meta_cap[x] * kMetaElem[x], 256);
    }
    b->meta_off[kMetaTables] = off;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(b->meta.reserve(off));
    uint8_t* m = b->meta.p;
    b->desc.p = (BatchDesc*)(m + b->meta_off[kMetaDesc]);
    b->req_start.p = (uint32_t*)(m + b->meta_off[kMetaReqStart]);
    b->req_end.p = (uint32_t*)(m + b->meta_off[kMetaReqEnd]);
    b->thr.p = (float*)(m + b->meta_off[kMetaThr]);
    b->slot_map.p = (uint32_t*)(m + b->meta_off[kMetaSlotMap]);
    b->blk_req.p = (uint32_t*)(m + b->meta_off[kMetaBlkReq]);
    b->groups.p = (uint32_t*)(m + b->meta_off[kMetaGroups]);
    b->chunks.p = (uint4*)(m + b->meta_off[kMetaChunks]);
    b->runs.p = (uint64_t*)(m + b->meta_off[kMetaRuns]);
    b->items.p = (uint2*)(m + b->meta_off[kMetaItems]);
    b->buf_gen++;
    return BLAMM_OK;
}

int batch_enqueue_kernels(blamm_hip_batch* b, bool mask, uint32_t grid)
{
    blamm_hip_ctx* c = b->ctx;
    const bool bf = prec_bf(c->precision);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, b->stream, b->counters.p);
    GroupArgs g;
    g.desc = b->desc.p;
    g.counters = b->counters.p;
    g.cap = b->hits.n;
    g.out_cap = b->grouped.n;
    g.in = b->hits.p;
    g.out = b->grouped.p;
    g.req_count = b->req_count.p;
    g.req_off = b->req_off.p;
    g.cursor = b->cursor.p;
    g.req_start = b->req_start.p;
    // finalize counts hits per request into req_count
    hipLaunchKernelGGL(group_clear_kernel, dim3(64), dim3(kGroupThreads), 0, b->stream, g);
    // run-start bits of the batch stream (request starts included)
    hipLaunchKernelGGL(run_bits_clear_kernel, dim3((uint32_t)c->num_cu * 4), dim3(256), 0, b->stream, b->run_bits.p,
                       0u, (const BatchDesc*)b->desc.p);
    hipLaunchKernelGGL(run_bits_set_kernel, dim3((uint32_t)c->num_cu), dim3(256), 0, b->stream, b->run_bits.p,
                       (const uint64_t*)b->runs.p, 0u, 0u, (const BatchDesc*)b->desc.p);
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
    f.req_count = b->req_count.p;
    f.run_bits = b->run_bits.p;
    e = launch_finalize(bf, true, f, (uint32_t)c->num_cu * 8, b->stream);
    if (e != hipSuccess) return fail(BLAMM_ERR_DEVICE, std::string("batch finalize launch: ") + hipGetErrorString(e));
    // group the hits by request on the device
    const uint32_t gg = (uint32_t)c->num_cu * 4;
    hipLaunchKernelGGL(group_scan_kernel, dim3(1), dim3(1024), 0, b->stream, g);
    hipLaunchKernelGGL(group_scatter_kernel, dim3(gg), dim3(kGroupThreads), 0, b->stream, g);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(BLAMM_ERR_DEVICE, std::string("batch grouping launch: ") + hipGetErrorString(e));
    return BLAMM_OK;
}

// The batch's counters and per-request hit offsets, stored by one workgroup
// straight into pinned host memory once the batch's kernels are done (the
// stream orders it after them; the completion event after it releases the
// stores to the host).  Two small hipMemcpyAsync read-backs would instead sit
// in one of the runtime's DMA queues until the kernels end; this way the hit
// read-back is the batch's only work on a DMA engine.
__global__ __launch_bounds__(256) void export_counts_kernel(const unsigned long long* counters,
                                                            const unsigned long long* req_off, uint32_t nreq,
                                                            unsigned long long* cnt_out, unsigned long long* off_out)
{
    for (uint32_t i = threadIdx.x; i < (uint32_t)kCntSlots; i += blockDim.x) cnt_out[i] = counters[i];
    for (uint32_t i = threadIdx.x; i <= nreq; i += blockDim.x) off_out[i] = req_off[i];
}

// Launch the captured sequence (capturing it first when buffers or the
// library moved since the last capture), then export the counters.
int batch_launch(blamm_hip_batch* b)
{
    blamm_hip_ctx* c = b->ctx;
    const int v = b->mask ? 1 : 0;
    if (!b->exec[v] || b->exec_buf_gen[v] != b->buf_gen || b->exec_lib_gen[v] != c->lib_gen) {
        if (b->exec[v]) (void)hipGraphExecDestroy(b->exec[v]);
        b->exec[v] = nullptr;
        const bool bf = prec_bf(c->precision);
        const int wgs = persistent_wgs_per_cu(bf, prec_lo(c->precision), b->mask, true);
        const uint32_t grid = (uint32_t)c->num_cu * (uint32_t)std::max(1, wgs - (b->leave_room ? 1 : 0));
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
    if (b->off_h.n < (size_t)b->nreq + 1 || b->req_off.n < (size_t)b->nreq + 1 || b->cnt_h.n < (size_t)kCntSlots)
        return fail(BLAMM_ERR_STATE, "batch launch: counter buffers smaller than the batch");
    hipLaunchKernelGGL(export_counts_kernel, dim3(1), dim3(256), 0, b->stream, b->counters.p, b->req_off.p, b->nreq,
                       b->cnt_h.p, b->off_h.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(b->ev_done, b->stream));
    return BLAMM_OK;
}

// Waiting for a batch, in two steps so that a server can overlap them:
// batch_collect waits for the kernels (replaying the batch after growing the
// candidate or hit buffers when either overflowed) and starts the copy of the
// grouped hits into a pinned buffer; batch_land waits for that copy.
int batch_collect(blamm_hip_batch* b, uint64_t* nhits)
{
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
        if (cand <= b->cand.n && nh <= b->hits.n && nh <= b->grouped.n) break;
        if (attempt >= 4) {
            b->pending = false;
            return fail(BLAMM_ERR_DEVICE, "batch_wait: capacity retries exhausted");
        }
        // Grow what overflowed and replay the batch (its inputs are still on
        // the device).
        b->replays++;
        int rc = BLAMM_OK;
        if (cand > b->cand.n) rc = batch_grow(b, b->cand, cand + cand / 4 + 1024);
        if (rc == BLAMM_OK && nh > b->hits.n) rc = batch_grow(b, b->hits, nh + nh / 4 + 1024);
        if (rc == BLAMM_OK) rc = batch_grow(b, b->grouped, b->hits.n);
        if (rc == BLAMM_OK) rc = batch_launch(b);
        if (rc != BLAMM_OK) {
            b->pending = false;
            return rc;
        }
    }
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev_start, b->ev_end));
    b->stats.candidates = cand;
    b->stats.band_rescored = b->cnt_h.p[kCntBand];
    b->stats.hits = nh;
    b->stats.kernel_ms = ms;
    b->stats.filter_ms = 0.f;
    // hits arrive grouped by request with request-local positions
    b->copying.reset();
    b->copy_dma = false;
    int rc = acquire_hits(b->pool, std::max<uint64_t>(nh, 1), &b->copying);
    if (rc == BLAMM_OK && nh > 0 && b->dma.ready) {
        if (dma_start(b->dma, b->copying->p, b->grouped.p, nh * sizeof(GroupedHit)) == HSA_STATUS_SUCCESS) {
            b->copy_dma = true;
        } else {
            hsa_signal_store_screlease(b->dma.sig, 0);
            b->dma.ready = false;   // this batch reads back through HIP from now on
        }
    }
    if (rc == BLAMM_OK && nh > 0 && !b->copy_dma) {
        hipError_t e = hipMemcpyAsync(b->copying->p, b->grouped.p, nh * sizeof(GroupedHit), hipMemcpyDeviceToHost,
                                      b->stream);
        if (e == hipSuccess) e = hipEventRecord(b->ev_copy, b->stream);
        if (e != hipSuccess) rc = fail(BLAMM_ERR_DEVICE, std::string("batch hit copy: ") + hipGetErrorString(e));
    }
    if (rc != BLAMM_OK) {
        b->pending = false;
        return rc;
    }
    b->copy_nh = nh;
    if (nhits) *nhits = nh;
    return BLAMM_OK;
}

uint64_t now_ns_batch()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Wait for the DMA read-back (completion signal back to 0; negative on a
// copy error), giving up after 120 s.
int dma_wait(blamm_hip_batch* b)
{
    const uint64_t t0 = now_ns_batch();
    for (;;) {
        const hsa_signal_value_t v = hsa_signal_wait_scacquire(b->dma.sig, HSA_SIGNAL_CONDITION_LT, 1,
                                                               b->dma.ticks_per_s, HSA_WAIT_STATE_BLOCKED);
        if (v == 0) return BLAMM_OK;
        if (v < 0) return fail(BLAMM_ERR_DEVICE, "batch hit copy: the DMA engine reported an error");
        if (now_ns_batch() - t0 > 120ull * 1000000000ull)
            return fail(BLAMM_ERR_DEVICE, "batch hit copy: no completion after 120 s");
    }
}

int batch_land(blamm_hip_batch* b)
{
    if (b->copy_nh > 0 && b->copy_dma) {
        const int rc = dma_wait(b);
        b->copy_dma = false;
        if (rc != BLAMM_OK) {
            b->pending = false;
            return rc;
        }
    } else if (b->copy_nh > 0) {
        const hipError_t e = hipEventSynchronize(b->ev_copy);
        if (e != hipSuccess) {
            b->pending = false;
            return fail(BLAMM_ERR_DEVICE, std::string("batch hit copy: ") + hipGetErrorString(e));
        }
    }
    b->out = std::move(b->copying);
    b->have_expanded = false;
    b->out_off.assign(b->off_h.p, b->off_h.p + b->nreq + 1);
    b->pending = false;
    b->done = true;
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
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_copy, hipEventDisableTiming);
    if (e == hipSuccess) e = b->counters.reserve(kCntSlots);
    if (e == hipSuccess) e = b->cnt_h.reserve(kCntSlots);
    const uint64_t nb = std::min<uint64_t>(max_bases, kBatchMaxBases);
    const uint64_t nr = std::max<uint32_t>(max_requests, 1);
    if (e == hipSuccess) e = b->packed.reserve(nb / 32 + nr * 2 + 16);
    if (e == hipSuccess) e = b->cand.reserve(std::max<uint64_t>(nb / 256, 1u << 16));
    if (e == hipSuccess) e = b->hits.reserve(std::max<uint64_t>(nb / 256, 1u << 16));
    if (e == hipSuccess) e = b->grouped.reserve(b->hits.n);
    if (e == hipSuccess) (void)dma_setup(b->dma, b->grouped.p);
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
    if (b->copy_dma) (void)dma_wait(b);
    if (b->dma.ready || b->dma.sig.handle) (void)hsa_signal_destroy(b->dma.sig);
    drop_graphs(b);
    b->packed.release(); b->lcmask.release(); b->run_bits.release(); b->meta.release();
    b->frag_a.release(); b->frag_b.release(); b->band.release();
    b->band_hi.release();
    b->req_count.release(); b->req_off.release(); b->cursor.release();
    b->cand.release(); b->counters.release();
    b->hits.release(); b->grouped.release();
    b->stage.release(); b->cnt_h.release(); b->off_h.release();
    b->out.reset();
    b->copying.reset();
    if (b->ev_start) (void)hipEventDestroy(b->ev_start);
    if (b->ev_end) (void)hipEventDestroy(b->ev_end);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    if (b->ev_copy) (void)hipEventDestroy(b->ev_copy);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return BLAMM_OK;
}

uint64_t blamm_hip_batch_device_bytes(const blamm_hip_batch* b)
{
    if (!b) return 0;
    return (b->packed.n + b->lcmask.n + b->run_bits.n) * 8 + b->meta.n + (b->frag_a.n + b->frag_b.n) * 16 +
           (b->band.n + b->band_hi.n) * 8 + b->req_count.n * 4 + (b->req_off.n + b->cursor.n) * 8 + b->cand.n * 16 +
           b->hits.n * sizeof(DevHit) + b->grouped.n * sizeof(GroupedHit) + kCntSlots * 8;
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
    BLAMM_TRY(batch_grow(b, b->run_bits, total / 64 + 2));
    const size_t nbs = std::max<size_t>(smap.size(), 1);
    size_t need[kMetaTables];
    need[kMetaDesc] = 1;
    need[kMetaReqStart] = need[kMetaReqEnd] = nreq;
    need[kMetaThr] = need[kMetaSlotMap] = nbs;
    need[kMetaBlkReq] = std::max<size_t>(nblk, 1);
    need[kMetaGroups] = std::max<size_t>(gpk.size(), 1);
    need[kMetaChunks] = std::max<size_t>(chunks.size(), 1);
    need[kMetaRuns] = nruns + 1;
    need[kMetaItems] = std::max<size_t>(nitems, 1);
    BLAMM_TRY(meta_fit(b, need));
    BLAMM_TRY(batch_grow(b, b->band, nbs));
    BLAMM_TRY(batch_grow(b, b->band_hi, nbs));
    const size_t fr = (size_t)std::max<uint32_t>(nblk, 1) * batch_frag_elems(bf);
    BLAMM_TRY(batch_grow(b, b->frag_a, fr));
    if (bf) BLAMM_TRY(batch_grow(b, b->frag_b, fr));
    else BLAMM_TRY(batch_grow(b, b->frag_b, 1));
    BLAMM_TRY(batch_grow(b, b->req_count, nreq));
    BLAMM_TRY(batch_grow(b, b->req_off, (size_t)nreq + 1));
    BLAMM_TRY(batch_grow(b, b->cursor, nreq));
    if (b->off_h.n < (size_t)nreq + 1) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        HIP_TRY(b->off_h.reserve((size_t)nreq + 1 + nreq / 8));
    }
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

    // Stage everything in pinned memory: the packed stream, the lowercase
    // mask, then the tables in the arena's layout; three copies in all.
    const size_t o_pk = 0;
    const size_t o_mk = align_up(o_pk + pw * 8, 256);
    const size_t o_mt = align_up(o_mk + mw * 8, 256);
    const size_t o_end = o_mt + b->meta_off[kMetaTables];
    if (o_end > b->stage.n) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        HIP_TRY(b->stage.reserve(o_end + o_end / 2));
    }
    uint8_t* S = b->stage.p;
    uint8_t* M = S + o_mt;
    auto table = [&](int x) { return M + b->meta_off[x]; };
    uint64_t* pk = (uint64_t*)(S + o_pk);
    uint64_t* mk = (uint64_t*)(S + o_mk);
    uint64_t* rn = (uint64_t*)table(kMetaRuns);
    uint32_t* re = (uint32_t*)table(kMetaReqEnd);
    uint32_t* rs = (uint32_t*)table(kMetaReqStart);
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
    std::memcpy(table(kMetaThr), sthr.data(), sthr.size() * 4);
    std::memcpy(table(kMetaSlotMap), smap.data(), smap.size() * 4);
    std::memcpy(table(kMetaBlkReq), breq.data(), (size_t)nblk * 4);
    std::memcpy(table(kMetaGroups), gpk.data(), gpk.size() * 4);
    std::memcpy(table(kMetaChunks), chunks.data(), chunks.size() * 16);
    size_t at = 0;
    for (int x = 0; x < kXcds; ++x) {
        std::memcpy(table(kMetaItems) + at * 8, sub[x].data(), sub[x].size() * 8);
        at += sub[x].size();
    }
    std::memcpy(table(kMetaDesc), &d, sizeof(d));

    hipStream_t s = b->stream;
    HIP_TRY(hipMemcpyAsync(b->packed.p, S + o_pk, pw * 8, hipMemcpyHostToDevice, s));
    if (mask) HIP_TRY(hipMemcpyAsync(b->lcmask.p, S + o_mk, mw * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->meta.p, M, b->meta_off[kMetaItems] + std::max<size_t>(nitems, 1) * 8,
                           hipMemcpyHostToDevice, s));

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
    BLAMM_TRY(batch_collect(b, nhits));
    return batch_land(b);
}

int blamm_hip_batch_hits(blamm_hip_batch* b, uint32_t req, const blamm_hit** hits, uint64_t* n)
{
    if (!b || !hits || !n) return fail(BLAMM_ERR_ARG, "batch_hits: null pointer");
    if (!b->done) return fail(BLAMM_ERR_STATE, "batch_hits: no finished batch");
    if (req >= b->nreq) return fail(BLAMM_ERR_ARG, "batch_hits: request index out of range");
    if (!b->have_expanded) {
        const uint64_t nh = b->out_off[b->nreq];
        if (nh > b->expanded_cap || !b->expanded) {
            const uint64_t cap = std::max<uint64_t>(nh + nh / 4, 1);
            b->expanded.reset(new (std::nothrow) blamm_hit[cap]);
            b->expanded_cap = b->expanded ? cap : 0;
            if (!b->expanded) return fail(BLAMM_ERR_NOMEM, "batch_hits: no host memory for the hits");
        }
        expand_hits_parallel(b->out->p, nh, b->expanded.get());
        b->have_expanded = true;
    }
    *hits = b->expanded.get() + b->out_off[req];
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
