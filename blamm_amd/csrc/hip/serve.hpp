// Serving loop: a resident library, requests admitted by an HBM budget and
// batched as they arrive (continuous batching).  The server owns `slots`
// batch slots and three threads, one per stage of a slot:
//  - the launcher takes a free slot whenever one exists, packs the queued
//    requests (first come, first served) that fit the slot's share of the
//    budget into one batch and launches it while other slots' batches are
//    still on the device;
//  - the harvester waits for the oldest batch in flight and starts the copy
//    of its hits (grouped by request) into pinned memory;
//  - the completer waits for the oldest copy, hands the hits to the
//    requests and frees the slot.
// So a batch's hit read-back overlaps the next batch's kernels and never
// holds up a launch.
// Every request's result is the same as if it had been scanned alone:
// thresholds, run rule and scores are per request and per window.
//
// Included once, at the end of blamm_hip.hip.

#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

namespace {

uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct ServeRequest {
    blamm_request req{};
    std::vector<float> thr;          // private copy of the request's thresholds
    uint64_t bytes = 0;              // admission estimate
    int status = 1;                  // 1 = queued/running, else final status
    std::string error;
    std::shared_ptr<HitBuf> buf;                   // the batch's hits, grouped by request
    uint64_t off = 0, nhits = 0;                   // this request's slice of buf
    uint64_t t_submit = 0, t_start = 0, t_done = 0;
    uint32_t batch_requests = 0;
};

struct ServeSlot {
    blamm_hip_batch* batch = nullptr;
    std::vector<uint64_t> ids;       // requests of the batch in flight
    // free -> launching (launcher) -> in flight -> harvesting (harvester) ->
    // copying -> completing (completer) -> free
    enum State { kFree, kLaunching, kInFlight, kHarvesting, kCopying, kCompleting } state = kFree;
    uint64_t launched = 0;           // launch order, for harvesting and completing oldest first
};

}  // namespace

struct blamm_hip_server {
    blamm_hip_ctx* ctx = nullptr;
    blamm_server_config cfg{};
    uint64_t slot_budget = 0;
    std::vector<ServeSlot> slots;
    std::mutex mu;
    // cv_work: a request was queued or a slot was freed (launcher);
    // cv_flight: a batch was launched (harvester); cv_copy: a hit copy was
    // started (completer); cv_done: requests finished
    std::condition_variable cv_work, cv_flight, cv_copy, cv_done;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, ServeRequest> reqs;
    uint64_t next_id = 1, launches = 0;
    bool stop = false;
    std::thread launcher, harvester, completer;
    blamm_server_stats stats{};
};

namespace {

// Device bytes a request occupies inside a batch: its stream, mask, run
// starts and run-start bits; its column blocks (fragments and per-slot
// tables); its work items; and per expected hit a filter record (16 B, about
// one per hit), a finalize output (16 B) and a grouped read-back record.
uint64_t request_bytes(const blamm_hip_ctx* c, const blamm_request& q)
{
    const uint64_t n = q.nbases + kReqAlign;
    uint64_t active = 0;
    for (int j = 0; j < c->ncols; ++j)
        if (!q.thresholds || q.thresholds[j] != INFINITY) ++active;
    const uint64_t hits = q.expected_hits ? q.expected_hits : (uint64_t)(1e-4 * (double)q.nbases * (double)active) + 1;
    const uint64_t tiles = n / kBf16Tile + 1;
    const uint64_t cpb = (uint64_t)std::max(1, c->cols_per_blk);
    const uint64_t blocks = (active + cpb - 1) / cpb;
    return n / 4 + (q.lcmask ? n / 8 : 0) + 8 * (q.nruns + 1) + n / 8 + blocks * batch_block_bytes(c) +
           8 * (tiles / std::max<uint32_t>(1, c->tuning.chunk_tiles) + 1) * c->groups.size() +
           (32 + sizeof(GroupedHit)) * hits;
}

void serve_finish(blamm_hip_server* S, ServeSlot& sl, int rc, std::unique_lock<std::mutex>& lk)
{
    // called with the lock held; the batch has been waited for (rc)
    const uint64_t t = now_ns();
    // the requests share the batch's grouped hits (no per-request copy)
    std::shared_ptr<HitBuf> buf;
    if (rc == BLAMM_OK) buf = sl.batch->out;
    for (uint32_t r = 0; r < (uint32_t)sl.ids.size(); ++r) {
        auto it = S->reqs.find(sl.ids[r]);
        if (it == S->reqs.end()) continue;
        ServeRequest& q = it->second;
        if (rc == BLAMM_OK) {
            q.buf = buf;
            q.off = sl.batch->out_off[r];
            q.nhits = sl.batch->out_off[r + 1] - sl.batch->out_off[r];
            q.status = BLAMM_OK;
        } else {
            q.status = rc;
            q.error = blamm_hip_last_error();
        }
        q.t_done = t;
        S->stats.completed++;
        S->stats.hits += q.nhits;
    }
    sl.ids.clear();
    sl.state = ServeSlot::kFree;
    (void)lk;
    S->cv_done.notify_all();
    S->cv_work.notify_all();
}

// Launcher thread: fill free slots with batches of queued requests.
void serve_launch_loop(blamm_hip_server* S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        ServeSlot* sl = nullptr;
        S->cv_work.wait(lk, [&] {
            if (S->stop && S->queue.empty()) return true;
            if (S->queue.empty()) return false;
            for (ServeSlot& x : S->slots)
                if (x.state == ServeSlot::kFree) {
                    sl = &x;
                    return true;
                }
            return false;
        });
        if (!sl) break;   // stopping with an empty queue
        std::vector<uint64_t> ids;
        std::vector<blamm_request> batch;
        uint64_t bytes = 0, bases = 0;
        while (!S->queue.empty() && ids.size() < S->cfg.max_batch_requests) {
            const ServeRequest& q = S->reqs[S->queue.front()];
            if (!ids.empty() && (bytes + q.bytes > S->slot_budget ||
                                 bases + q.req.nbases + kReqAlign > S->cfg.max_batch_bases))
                break;
            bytes += q.bytes;
            bases += q.req.nbases + kReqAlign;
            ids.push_back(S->queue.front());
            batch.push_back(q.req);
            S->queue.pop_front();
        }
        const uint64_t t = now_ns();
        for (uint64_t id : ids) {
            S->reqs[id].t_start = t;
            S->reqs[id].batch_requests = (uint32_t)ids.size();
        }
        sl->ids = ids;
        sl->state = ServeSlot::kLaunching;
        S->stats.batches++;
        lk.unlock();
        const int rc = blamm_hip_batch_submit(sl->batch, batch.data(), (uint32_t)batch.size());
        const uint64_t t1 = now_ns();
        lk.lock();
        S->stats.host_submit_ms += (t1 - t) / 1e6;
        if (rc != BLAMM_OK) {
            serve_finish(S, *sl, rc, lk);
            S->cv_flight.notify_all();
        } else {
            sl->launched = ++S->launches;
            sl->state = ServeSlot::kInFlight;
            S->cv_flight.notify_all();
        }
    }
    S->cv_flight.notify_all();
}

// Harvester thread: wait for batches oldest first and start their hit
// copies.
void serve_harvest_loop(blamm_hip_server* S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        ServeSlot* oldest = nullptr;
        S->cv_flight.wait(lk, [&] {
            oldest = nullptr;
            bool launching = false;
            for (ServeSlot& x : S->slots) {
                if (x.state == ServeSlot::kInFlight && (!oldest || x.launched < oldest->launched)) oldest = &x;
                launching = launching || x.state == ServeSlot::kLaunching;
            }
            return oldest != nullptr || (S->stop && S->queue.empty() && !launching);
        });
        if (!oldest) break;
        oldest->state = ServeSlot::kHarvesting;
        lk.unlock();
        uint64_t nh = 0;
        const uint64_t t0 = now_ns();
        const uint64_t rp0 = oldest->batch->replays;
        const int rc = batch_collect(oldest->batch, &nh);
        blamm_scan_stats bs{};
        blamm_hip_batch_stats(oldest->batch, &bs);
        const uint64_t t1 = now_ns();
        lk.lock();
        S->stats.device_ms += bs.kernel_ms;
        S->stats.host_wait_ms += (t1 - t0) / 1e6;
        S->stats.replays += oldest->batch->replays - rp0;
        if (rc != BLAMM_OK) {
            serve_finish(S, *oldest, rc, lk);
        } else {
            oldest->state = ServeSlot::kCopying;
            S->cv_copy.notify_all();
        }
    }
    S->cv_copy.notify_all();
}

// Completer thread: wait for hit copies oldest first, hand the hits to the
// requests and free the slots.
void serve_complete_loop(blamm_hip_server* S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        ServeSlot* oldest = nullptr;
        S->cv_copy.wait(lk, [&] {
            oldest = nullptr;
            bool busy = false;
            for (ServeSlot& x : S->slots) {
                if (x.state == ServeSlot::kCopying && (!oldest || x.launched < oldest->launched)) oldest = &x;
                busy = busy || x.state == ServeSlot::kLaunching || x.state == ServeSlot::kInFlight ||
                       x.state == ServeSlot::kHarvesting;
            }
            return oldest != nullptr || (S->stop && S->queue.empty() && !busy);
        });
        if (!oldest) break;
        oldest->state = ServeSlot::kCompleting;
        lk.unlock();
        const uint64_t t0 = now_ns();
        const int rc = batch_land(oldest->batch);
        const uint64_t t1 = now_ns();
        lk.lock();
        S->stats.host_wait_ms += (t1 - t0) / 1e6;
        serve_finish(S, *oldest, rc, lk);
    }
}

}  // namespace

extern "C" {

int blamm_hip_server_create(blamm_hip_ctx* c, const blamm_server_config* cfg, blamm_hip_server** out)
{
    if (!c || !out) return fail(BLAMM_ERR_ARG, "server_create: null pointer");
    *out = nullptr;
    if (!c->have_lib) return fail(BLAMM_ERR_STATE, "server_create: no library uploaded");
    blamm_server_config k{};
    if (cfg) k = *cfg;
    if (k.hbm_budget_bytes == 0) k.hbm_budget_bytes = 8ull << 30;
    if (k.slots == 0) k.slots = 2;
    if (k.slots > 8) return fail(BLAMM_ERR_ARG, "server_create: at most 8 slots");
    if (k.max_batch_requests == 0) k.max_batch_requests = 256;
    if (k.max_batch_bases == 0) k.max_batch_bases = 1ull << 24;
    k.max_batch_bases = std::min<uint64_t>(k.max_batch_bases, kBatchMaxBases);
    if (k.max_batch_requests >= (1u << 16)) return fail(BLAMM_ERR_ARG, "server_create: max_batch_requests too large");
    blamm_hip_server* S = new blamm_hip_server();
    S->ctx = c;
    S->cfg = k;
    S->slot_budget = k.hbm_budget_bytes / k.slots;
    S->slots.resize(k.slots);
    for (ServeSlot& sl : S->slots) {
        const int rc = blamm_hip_batch_create(c, std::min<uint64_t>(S->slot_budget, 1ull << 30), k.max_batch_requests,
                                              &sl.batch);
        if (rc != BLAMM_OK) {
            for (ServeSlot& s2 : S->slots) blamm_hip_batch_destroy(s2.batch);
            delete S;
            return rc;
        }
        sl.batch->leave_room = k.slots > 1;
    }
    S->launcher = std::thread(serve_launch_loop, S);
    S->harvester = std::thread(serve_harvest_loop, S);
    S->completer = std::thread(serve_complete_loop, S);
    *out = S;
    return BLAMM_OK;
}

int blamm_hip_server_submit(blamm_hip_server* S, const blamm_request* q, uint64_t* id)
{
    if (!S || !q || !id) return fail(BLAMM_ERR_ARG, "server_submit: null pointer");
    const blamm_hip_ctx* c = S->ctx;
    if (q->nbases > 0 && !q->packed) return fail(BLAMM_ERR_ARG, "server_submit: no packed stream");
    if (q->nruns > 0 && !q->run_starts) return fail(BLAMM_ERR_ARG, "server_submit: null run starts");
    if (q->nbases + kReqAlign > kBatchMaxBases) return fail(BLAMM_ERR_ARG, "server_submit: request too long");
    for (uint64_t i = 1; i < q->nruns; ++i)
        if (q->run_starts[i] < q->run_starts[i - 1]) return fail(BLAMM_ERR_ARG, "server_submit: run starts not sorted");
    ServeRequest r;
    r.req = *q;
    if (q->thresholds) {
        for (int j = 0; j < c->ncols; ++j)
            if (std::isnan(q->thresholds[j])) return fail(BLAMM_ERR_ARG, "server_submit: NaN threshold");
        r.thr.assign(q->thresholds, q->thresholds + c->ncols);
    }
    r.bytes = request_bytes(c, *q);
    if (r.bytes > S->slot_budget)
        return fail(BLAMM_ERR_ARG, "server_submit: request needs ~" + std::to_string(r.bytes) +
                                       " device bytes; the per-batch budget is " + std::to_string(S->slot_budget));
    r.t_submit = now_ns();
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->stop) return fail(BLAMM_ERR_STATE, "server_submit: server is shutting down");
        *id = S->next_id++;
        ServeRequest& slot = S->reqs[*id];
        slot = std::move(r);
        if (!slot.thr.empty()) slot.req.thresholds = slot.thr.data();
        S->queue.push_back(*id);
        S->stats.submitted++;
    }
    S->cv_work.notify_one();
    return BLAMM_OK;
}

int blamm_hip_server_wait(blamm_hip_server* S, uint64_t id, int64_t timeout_ms)
{
    if (!S) return fail(BLAMM_ERR_ARG, "server_wait: null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(BLAMM_ERR_ARG, "server_wait: unknown request");
    auto ready = [&] { return S->reqs[id].status != 1; };
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return 1;
    return BLAMM_OK;
}

int blamm_hip_server_result(blamm_hip_server* S, uint64_t id, blamm_hit* hits, uint64_t cap,
                            blamm_request_info* info)
{
    if (!S) return fail(BLAMM_ERR_ARG, "server_result: null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(BLAMM_ERR_ARG, "server_result: unknown request");
    ServeRequest& q = it->second;
    if (q.status == 1) return fail(BLAMM_ERR_STATE, "server_result: request not finished");
    if (info) {
        info->status = q.status;
        info->nhits = q.nhits;
        info->t_submit_ns = q.t_submit;
        info->t_start_ns = q.t_start;
        info->t_done_ns = q.t_done;
        info->batch_requests = q.batch_requests;
    }
    if (q.status != BLAMM_OK) {
        const int st = q.status;
        const std::string msg = q.error;
        S->reqs.erase(it);
        return fail(st, "request failed: " + msg);
    }
    if (q.nhits > cap) return fail(BLAMM_ERR_OVERFLOW, "server_result: hit capacity too small");
    if (q.nhits && !hits) return fail(BLAMM_ERR_ARG, "server_result: null hit buffer");
    // the request's slice of the batch's read-back stays alive through buf
    const std::shared_ptr<HitBuf> buf = q.buf;
    const uint64_t off = q.off, n = q.nhits;
    S->reqs.erase(it);
    lk.unlock();
    if (n) expand_hits(buf->p + off, n, hits);
    return BLAMM_OK;
}

int blamm_hip_server_release(blamm_hip_server* S, uint64_t id, blamm_request_info* info)
{
    if (!S) return fail(BLAMM_ERR_ARG, "server_release: null server");
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(BLAMM_ERR_ARG, "server_release: unknown request");
    ServeRequest& q = it->second;
    if (q.status == 1) return fail(BLAMM_ERR_STATE, "server_release: request not finished");
    if (info) {
        info->status = q.status;
        info->nhits = q.nhits;
        info->t_submit_ns = q.t_submit;
        info->t_start_ns = q.t_start;
        info->t_done_ns = q.t_done;
        info->batch_requests = q.batch_requests;
    }
    const int st = q.status;
    S->reqs.erase(it);
    return st == BLAMM_OK ? BLAMM_OK : fail(st, "request failed");
}

int blamm_hip_server_stats(blamm_hip_server* S, blamm_server_stats* st)
{
    if (!S || !st) return fail(BLAMM_ERR_ARG, "server_stats: null pointer");
    std::lock_guard<std::mutex> g(S->mu);
    *st = S->stats;
    return BLAMM_OK;
}

int blamm_hip_server_destroy(blamm_hip_server* S)
{
    if (!S) return BLAMM_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_flight.notify_all();
    S->cv_copy.notify_all();
    if (S->launcher.joinable()) S->launcher.join();
    if (S->harvester.joinable()) S->harvester.join();
    if (S->completer.joinable()) S->completer.join();
    for (ServeSlot& sl : S->slots) blamm_hip_batch_destroy(sl.batch);
    delete S;
    return BLAMM_OK;
}

}  // extern "C"
