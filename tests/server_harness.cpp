// CPU harness for the serving loop's threads and locking (serve.hpp), built
// with ThreadSanitizer or AddressSanitizer + UBSan.  The device side is
// replaced by a stand-in batch that "scans" on the CPU: each request of a
// batch gets a deterministic list of hits computed from its own stream (so a
// result is the same alone and in any mix), collect and land sleep for a
// moment as kernels and read-backs would.  Everything above the batch API
// (admission, the launcher / harvester / completer threads, the slot state
// machine, result hand-out and shutdown) is the library's own code.
//
// Usage: server_harness clients requests_per_client slots
// Prints "ok <requests> <hits>" or "error <message>".
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../include/blamm_hip.h"

// --- what serve.hpp takes from the library's main unit ---------------------
typedef int hipError_t;
static hipError_t hipSetDevice(int) { return 0; }

namespace {
thread_local std::string g_last_error;
int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}
constexpr uint64_t kReqAlign = 64;
constexpr uint64_t kBatchMaxBases = (1ull << 32) - 4096;
constexpr uint32_t kBf16Tile = 32;

struct GroupedHit {
    uint32_t pos;
    uint32_t col;
    float score;
};
struct HitBuf {
    GroupedHit* p = nullptr;
    size_t n = 0;
};
void expand_hits(const GroupedHit* in, uint64_t n, blamm_hit* out)
{
    for (uint64_t i = 0; i < n; ++i) out[i] = {in[i].pos, in[i].col, in[i].score};
}

// The hits the stand-in reports for one request: every 97th position of the
// stream whose packed 2-bit code is odd, in the columns the request keeps.
std::vector<GroupedHit> fake_hits(const blamm_request& q, int ncols)
{
    std::vector<GroupedHit> h;
    for (uint64_t i = 0; i < q.nbases; i += 97) {
        const uint32_t code = (uint32_t)(q.packed[i / 32] >> (2 * (i % 32))) & 3u;
        if (!(code & 1)) continue;
        for (int c = 0; c < ncols; ++c)
            if (!q.thresholds || q.thresholds[c] != INFINITY) h.push_back({(uint32_t)i, (uint32_t)c, (float)(i % 13)});
    }
    return h;
}
}  // namespace

struct blamm_hip_ctx {
    int device = 0;
    bool have_lib = true;
    int ncols = 6, nslots = 6, cols_per_blk = 32;
    blamm_tuning tuning{512, 2, 4};
    std::vector<int> groups = std::vector<int>(3);
};

struct blamm_hip_batch {
    blamm_hip_ctx* ctx = nullptr;
    std::vector<std::vector<GroupedHit>> req_hits;   // the batch "on the device"
    std::shared_ptr<HitBuf> copying, out;
    std::vector<uint64_t> out_off;
    uint64_t replays = 0;
    bool pending = false;
    bool leave_room = false;
};

namespace {
uint64_t batch_block_bytes(const blamm_hip_ctx*) { return 32768; }

std::shared_ptr<HitBuf> new_hitbuf(size_t n)
{
    HitBuf* h = new HitBuf;
    h->n = std::max<size_t>(n, 1);
    h->p = new GroupedHit[h->n];
    return std::shared_ptr<HitBuf>(h, [](HitBuf* x) {
        delete[] x->p;
        delete x;
    });
}

int batch_collect(blamm_hip_batch* b, uint64_t* nhits)
{
    std::this_thread::sleep_for(std::chrono::microseconds(200));   // kernels
    uint64_t n = 0;
    for (const auto& v : b->req_hits) n += v.size();
    b->copying = new_hitbuf(n);
    b->out_off.assign(1, 0);
    uint64_t at = 0;
    for (const auto& v : b->req_hits) {
        std::copy(v.begin(), v.end(), b->copying->p + at);
        at += v.size();
        b->out_off.push_back(at);
    }
    if (nhits) *nhits = n;
    return BLAMM_OK;
}

int batch_land(blamm_hip_batch* b)
{
    std::this_thread::sleep_for(std::chrono::microseconds(100));   // read-back
    b->out = std::move(b->copying);
    b->pending = false;
    return BLAMM_OK;
}
}  // namespace

extern "C" {
const char* blamm_hip_last_error(void) { return g_last_error.c_str(); }
int blamm_hip_batch_create(blamm_hip_ctx* c, uint64_t, uint32_t, blamm_hip_batch** out)
{
    *out = new blamm_hip_batch;
    (*out)->ctx = c;
    return BLAMM_OK;
}
int blamm_hip_batch_destroy(blamm_hip_batch* b)
{
    delete b;
    return BLAMM_OK;
}
int blamm_hip_batch_submit(blamm_hip_batch* b, const blamm_request* reqs, uint32_t nreq)
{
    if (b->pending) return fail(BLAMM_ERR_STATE, "batch_submit: previous batch not waited for");
    b->req_hits.clear();
    for (uint32_t r = 0; r < nreq; ++r) b->req_hits.push_back(fake_hits(reqs[r], b->ctx->ncols));
    b->pending = true;
    return BLAMM_OK;
}
int blamm_hip_batch_stats(blamm_hip_batch*, blamm_scan_stats* st)
{
    *st = blamm_scan_stats{};
    st->kernel_ms = 0.2;
    return BLAMM_OK;
}
}

#include "../blamm_amd/csrc/hip/serve.hpp"

// --- clients ----------------------------------------------------------------
int main(int argc, char** argv)
{
    if (argc != 4) return 2;
    const int clients = std::atoi(argv[1]), per = std::atoi(argv[2]);
    blamm_hip_ctx ctx;
    blamm_server_config cfg{};
    cfg.hbm_budget_bytes = 64ull << 20;
    cfg.slots = (uint32_t)std::atoi(argv[3]);
    cfg.max_batch_requests = 8;
    cfg.max_batch_bases = 200000;
    blamm_hip_server* srv = nullptr;
    if (blamm_hip_server_create(&ctx, &cfg, &srv) != BLAMM_OK) {
        std::printf("error create: %s\n", blamm_hip_last_error());
        return 1;
    }
    std::atomic<uint64_t> done{0}, hits{0};
    std::atomic<bool> bad{false};
    std::mutex err_mu;
    std::string err;
    auto client = [&](int id) {
        std::vector<std::vector<uint64_t>> streams(per);
        std::vector<std::vector<float>> thr(per);
        std::vector<uint64_t> ids(per);
        std::vector<blamm_request> reqs(per);
        for (int k = 0; k < per; ++k) {
            const uint64_t n = 1 + (uint64_t)((id * 7919 + k * 104729) % 40000);
            streams[k].resize((n + 31) / 32 + 1);
            for (size_t w = 0; w < streams[k].size(); ++w)
                streams[k][w] = 0x9E3779B97F4A7C15ull * (w + 1 + 1000ull * id + 7ull * k);
            thr[k].assign(ctx.ncols, 1.0f);
            thr[k][(id + k) % ctx.ncols] = INFINITY;   // each request leaves one column out
            reqs[k] = blamm_request{streams[k].data(), nullptr, n, nullptr, 0, thr[k].data(), 0};
            if (blamm_hip_server_submit(srv, &reqs[k], &ids[k]) != BLAMM_OK) {
                std::lock_guard<std::mutex> g(err_mu);
                err = std::string("submit: ") + blamm_hip_last_error();
                bad = true;
                return;
            }
        }
        for (int k = 0; k < per; ++k) {
            // (no timeout: gcc 11's ThreadSanitizer does not see the
            // pthread_cond_clockwait behind a timed condition-variable wait)
            if (blamm_hip_server_wait(srv, ids[k], -1) != BLAMM_OK) {
                bad = true;
                return;
            }
            const std::vector<GroupedHit> want = fake_hits(reqs[k], ctx.ncols);
            blamm_request_info info{};
            if (k % 3 == 2) {   // some callers only want the count
                if (blamm_hip_server_release(srv, ids[k], &info) != BLAMM_OK || info.nhits != want.size()) bad = true;
            } else {
                std::vector<blamm_hit> got(want.size() + 1);
                if (blamm_hip_server_result(srv, ids[k], got.data(), got.size(), &info) != BLAMM_OK ||
                    info.nhits != want.size())
                    bad = true;
                for (size_t i = 0; i < want.size() && !bad; ++i)
                    if (got[i].pos != want[i].pos || got[i].col != want[i].col || got[i].score != want[i].score)
                        bad = true;
            }
            hits += info.nhits;
            done++;
        }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < clients; ++i) th.emplace_back(client, i);
    for (auto& t : th) t.join();
    blamm_server_stats st{};
    blamm_hip_server_stats(srv, &st);
    // a last request left queued: destroy must finish it
    std::vector<uint64_t> tail(64, 0x5555555555555555ull);
    blamm_request last{tail.data(), nullptr, 2000, nullptr, 0, nullptr, 0};
    uint64_t last_id = 0;
    blamm_hip_server_submit(srv, &last, &last_id);
    blamm_hip_server_destroy(srv);
    if (bad || st.completed != (uint64_t)clients * per) {
        std::printf("error %s completed %llu\n", err.c_str(), (unsigned long long)st.completed);
        return 1;
    }
    std::printf("ok %llu %llu\n", (unsigned long long)done.load(), (unsigned long long)hits.load());
    return 0;
}
