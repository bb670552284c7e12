// libblamm_hip.so: the C ABI declared in include/blamm_hip.h.
//
// Replaces the reference's HAVE_CUDA scan worker (scanThreadCUBLAS,
// pwmscan.cpp:297-437) and kernel.cu.  Host work here is limited to laying the
// library out for the MFMA kernels (column blocks, fragment swizzle, bf16
// split, proven filter bands), to building work lists and to moving buffers;
// scoring happens on the device in scan_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/blamm_hip.h"
#include "instances.hpp"

using namespace blamm;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(e_ == hipErrorOutOfMemory ? BLAMM_ERR_NOMEM : BLAMM_ERR_DEVICE,       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

#define BLAMM_TRY(expr)                \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != BLAMM_OK) return rc_; \
    } while (0)

// Column blocks one wave holds in registers at most.
constexpr int kNbMaxBf16 = 2;
constexpr int kNbMaxFp32 = 4;

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Grow to at least `want` elements (contents are not preserved).
    hipError_t reserve(size_t want)
    {
        if (want <= n && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = std::max<size_t>(want, 1);
        return e;
    }
};

// Pinned host staging (hipHostMalloc; memory is never registered after the fact).
template <typename T>
struct HostBuf {
    T* p = nullptr;
    size_t n = 0;
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t reserve(size_t want)
    {
        if (want <= n && p) return hipSuccess;
        release();
        hipError_t e = hipHostMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T), hipHostMallocDefault);
        if (e == hipSuccess) n = std::max<size_t>(want, 1);
        return e;
    }
};

uint16_t bf16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return (uint16_t)(u >> 16);
}

float bf16_to_float(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// gamma_n = n u / (1 - n u): bound factor for any summation order of n terms.
double gamma_n(double n, double u) { return n * u / (1.0 - n * u); }

struct GroupInfo {
    uint32_t blk0, nch, nb;
};

}  // namespace

// Scratch of blamm_hip_scan_device_noalloc: the scan counters (first 256
// bytes) and BLAMM_NOALLOC_MAX_RECORDS filter records of 16 bytes, 64 MiB in
// all.  Every GPU the library runs on holds its own copy until the process
// ends; all contexts on one GPU share it, so no-alloc scans that overlap on
// one GPU race.  Host code never uses the array's name as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device in blamm_hip_create with that device current.
constexpr uint64_t kNoAllocHeader = 256;
static_assert(kCntSlots * sizeof(unsigned long long) <= kNoAllocHeader, "counters exceed the scratch header");
static __device__ __align__(256) uint8_t scratch[kNoAllocHeader + (uint64_t)BLAMM_NOALLOC_MAX_RECORDS * 16];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
std::string g_scratch_err[kMaxDevices];

// Device address of the current device's scratch (taken once per device).
void noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) {
        g_scratch_ptr[device] = p;
        g_scratch_err[device].clear();
    } else {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    }
}

}  // namespace

struct blamm_hip_ctx {
    int device = 0;
    int num_cu = 256;
    std::string arch;
    int filter_rounds = 1;      // resident-wave multiples of the persistent grid
    blamm_tuning tuning{512, 2, 4};
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

    // library as given (kept for relayout under another tuning)
    bool have_lib = false;
    int precision = 0;
    int ncols = 0;
    int ldp = 0;
    std::vector<float> lib_P, lib_thr;
    std::vector<int> lib_len;
    uint64_t lib_gen = 0;       // bumped by every relayout

    // device layout of the library
    int nslots = 0;
    int cols_per_blk = 32;
    std::vector<GroupInfo> groups;             // longest motifs first
    std::vector<int> slot_col_h;               // -1 = empty slot
    std::vector<double> band_h, seq_h;         // per slot
    double max_filter_band = 0.0, max_score_error = 0.0;
    DevBuf<uint8_t> frag_a, frag_b;
    DevBuf<float> slot_thr, slot_w, slot_w2;   // w2: FP32 sums of position pairs
    DevBuf<double> slot_band, slot_band_any, slot_seq;   // band_any: valid in a block of any length
    DevBuf<double> slot_band_hi, slot_band_hi_any;      // hi-only stage of the bf16 filter
    DevBuf<int> slot_len;
    DevBuf<uint32_t> slot_col, groups_d;

    // empirical histograms
    DevBuf<HistSlot> hist_slots;
    DevBuf<uint32_t> hist_groups;
    DevBuf<unsigned long long> hist_counts;

    // scratch
    DevBuf<uint4> cand;
    DevBuf<unsigned long long> counters;
    DevBuf<uint64_t> run_bits;        // run-start bits of the stream being scanned

    // host-buffer staging
    DevBuf<uint64_t> packed, lcmask, runs, nhits_dev;
    DevBuf<blamm_hit> hits;
    HostBuf<blamm_hit> hits_pinned;   // blamm_hip_scan_pinned's result

    // last scan
    bool pending = false;
    bool pending_noalloc = false;             // records and counters in the no-alloc scratch
    uint64_t pending_hit_cap = 0;
    uint64_t pending_cand_cap = 0;            // record capacity the filter was launched with
    unsigned long long* pending_counters = nullptr;
    hipStream_t pending_stream = nullptr;
    blamm_scan_stats stats{};
};

namespace {

// Precision modes: FP32, or bf16 with the filter in its hi-only or lazy-lo
// form (both keep the PWM split into hi and lo terms on the device).
inline bool prec_bf(int p) { return p == BLAMM_PREC_BF16X2 || p == BLAMM_PREC_BF16X2_LO; }
inline bool prec_lo(int p) { return p == BLAMM_PREC_BF16X2_LO; }

template <bool BF, bool MASK, bool LIST>
const void* filter_kernel_ptr()
{
    return (const void*)filter_persistent_kernel<BF, MASK, LIST>;
}

const void* filter_kernel(bool bf, bool lo, bool mask, bool list)
{
    if (bf && lo) {
        if (mask) return list ? (const void*)filter_lo_kernel<true, true> : (const void*)filter_lo_kernel<true, false>;
        return list ? (const void*)filter_lo_kernel<false, true> : (const void*)filter_lo_kernel<false, false>;
    }
    if (bf) {
        if (mask) return list ? filter_kernel_ptr<true, true, true>() : filter_kernel_ptr<true, true, false>();
        return list ? filter_kernel_ptr<true, false, true>() : filter_kernel_ptr<true, false, false>();
    }
    if (mask) return list ? filter_kernel_ptr<false, true, true>() : filter_kernel_ptr<false, true, false>();
    return list ? filter_kernel_ptr<false, false, true>() : filter_kernel_ptr<false, false, false>();
}

// Workgroups of the persistent filter resident per CU (register/LDS limits).
int persistent_wgs_per_cu(bool bf, bool lo, bool mask, bool list)
{
    int n = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, filter_kernel(bf, lo, mask, list), kThreads, 0);
    return (e == hipSuccess && n > 0) ? n : 1;
}

hipError_t launch_filter(bool bf, bool lo, bool mask, bool list, const FilterArgs& a, const WorkQueue& q,
                         const ListArgs& l, uint32_t grid, hipStream_t s)
{
#define BLAMM_LAUNCH(B, M, L) \
    hipLaunchKernelGGL((filter_persistent_kernel<B, M, L>), dim3(grid), dim3(kThreads), 0, s, a, q, l)
#define BLAMM_LAUNCH_LO(M, L) hipLaunchKernelGGL((filter_lo_kernel<M, L>), dim3(grid), dim3(kThreads), 0, s, a, q, l)
    if (bf && lo) {
        if (mask) { if (list) BLAMM_LAUNCH_LO(true, true); else BLAMM_LAUNCH_LO(true, false); }
        else { if (list) BLAMM_LAUNCH_LO(false, true); else BLAMM_LAUNCH_LO(false, false); }
    } else if (bf) {
        if (mask) { if (list) BLAMM_LAUNCH(true, true, true); else BLAMM_LAUNCH(true, true, false); }
        else { if (list) BLAMM_LAUNCH(true, false, true); else BLAMM_LAUNCH(true, false, false); }
    } else {
        if (mask) { if (list) BLAMM_LAUNCH(false, true, true); else BLAMM_LAUNCH(false, true, false); }
        else { if (list) BLAMM_LAUNCH(false, false, true); else BLAMM_LAUNCH(false, false, false); }
    }
#undef BLAMM_LAUNCH
#undef BLAMM_LAUNCH_LO
    return hipGetLastError();
}

hipError_t launch_finalize(bool bf, bool list, const FinalizeArgs& f, uint32_t grid, hipStream_t s)
{
    if (bf) {
        if (list) hipLaunchKernelGGL((finalize_kernel<16, true>), dim3(grid), dim3(kThreads), 0, s, f);
        else hipLaunchKernelGGL((finalize_kernel<16, false>), dim3(grid), dim3(kThreads), 0, s, f);
    } else {
        if (list) hipLaunchKernelGGL((finalize_kernel<4, true>), dim3(grid), dim3(kThreads), 0, s, f);
        else hipLaunchKernelGGL((finalize_kernel<4, false>), dim3(grid), dim3(kThreads), 0, s, f);
    }
    return hipGetLastError();
}

__global__ void store_count_kernel(const unsigned long long* counters, uint64_t* out)
{
    if (threadIdx.x == 0) *out = counters[kCntHits];
}

int check_tuning(const blamm_tuning& t, int precision)
{
    if (t.chunk_tiles < 1 || t.chunk_tiles > (1u << 20)) return fail(BLAMM_ERR_ARG, "tuning: chunk_tiles outside 1..2^20");
    const uint32_t nbmax = prec_bf(precision) ? kNbMaxBf16 : kNbMaxFp32;
    if (t.blocks_per_wave != 1 && t.blocks_per_wave != 2 && t.blocks_per_wave != 4)
        return fail(BLAMM_ERR_ARG, "tuning: blocks_per_wave must be 1, 2 or 4");
    if (t.blocks_per_wave > nbmax)
        return fail(BLAMM_ERR_ARG, "tuning: blocks_per_wave above " + std::to_string(nbmax) + " for this precision");
    if (t.kpad_positions != 4 && t.kpad_positions != 8)
        return fail(BLAMM_ERR_ARG, "tuning: kpad_positions must be 4 or 8");
    return BLAMM_OK;
}

// Lay the stored library out for the kernels under the context's tuning and
// upload it.
int relayout(blamm_hip_ctx* c)
{
    const bool bf = prec_bf(c->precision);
    const int ncols = c->ncols, ldp = c->ldp;
    const int* lengths = c->lib_len.data();
    const float* P = c->lib_P.data();
    const int cols_per_blk = bf ? kBf16Cols : kFp32Cols;
    const uint32_t nbmax = std::min<uint32_t>(c->tuning.blocks_per_wave, bf ? kNbMaxBf16 : kNbMaxFp32);
    const int kstep = (int)c->tuning.kpad_positions / 4;   // chunk granularity

    // Columns by non-decreasing length (stable); consecutive runs of
    // cols_per_blk form a block whose chunk count is set by its longest
    // column, rounded up to the K-padding granularity.
    std::vector<int> order(ncols);
    for (int j = 0; j < ncols; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return lengths[x] < lengths[y]; });
    std::vector<std::vector<int>> blocks;   // each: cols_per_blk entries, -1 = empty
    std::vector<int> blk_nch;
    for (int s = 0; s < ncols; s += cols_per_blk) {
        std::vector<int> b(cols_per_blk, -1);
        int lm = 0;
        for (int k = 0; k < cols_per_blk && s + k < ncols; ++k) {
            b[k] = order[s + k];
            lm = std::max(lm, lengths[order[s + k]]);
        }
        blocks.push_back(b);
        int nch = (lm + 3) / 4;
        nch = std::min(kMaxChunks, (nch + kstep - 1) / kstep * kstep);
        blk_nch.push_back(nch);
    }
    // Groups: runs of equal chunk count cut into nbmax blocks (one block for
    // motifs longer than 4 kWideChunks positions), the remainder into
    // smaller powers of two, so no padding block is scored.
    std::vector<GroupInfo> groups;
    int max_nch = 1;
    for (size_t i = 0; i < blocks.size();) {
        size_t e = i;
        while (e < blocks.size() && blk_nch[e] == blk_nch[i]) ++e;
        size_t b = i;
        max_nch = std::max(max_nch, blk_nch[i]);
        const uint32_t nbg = blk_nch[i] > kWideChunks ? 1u : nbmax;
        for (uint32_t nb = nbg; nb >= 1; nb /= 2)
            while (e - b >= nb) {
                groups.push_back({(uint32_t)b, (uint32_t)blk_nch[i], nb});
                b += nb;
            }
        i = e;
    }
    std::stable_sort(groups.begin(), groups.end(), [](const GroupInfo& x, const GroupInfo& y) { return x.nch > y.nch; });
    const int nblk = (int)blocks.size();
    const int nslots = nblk * cols_per_blk;
    if (nblk >= (1 << 20)) return fail(BLAMM_ERR_ARG, "set_library: too many columns");

    std::vector<float> s_thr(nslots, INFINITY), s_w((size_t)nslots * kMaxLen * 4, 0.f);
    std::vector<int> s_len(nslots, 1), s_colh(nslots, -1);
    std::vector<uint32_t> s_col(nslots, 0);
    std::vector<double> s_band(nslots, 0.0), s_band_any(nslots, 0.0), s_seq(nslots, 0.0);
    std::vector<double> s_band_hi(nslots, 0.0), s_band_hi_any(nslots, 0.0);
    std::vector<uint16_t> hi16((size_t)nslots * kMaxLen * 4, 0), lo16((size_t)nslots * kMaxLen * 4, 0);
    double max_band = 0.0, max_err = 0.0;
    const double u24 = std::ldexp(1.0, -24), u53 = std::ldexp(1.0, -53);
    for (int bi = 0; bi < nblk; ++bi) {
        const int nch = blk_nch[bi];
        for (int k = 0; k < cols_per_blk; ++k) {
            const int col = blocks[bi][k];
            if (col < 0) continue;
            const int slot = bi * cols_per_blk + k;
            const int L = lengths[col];
            const float* w = P + (size_t)col * ldp;
            double sum_max = 0.0, sum_split = 0.0, rep_err = 0.0, sum_lo = 0.0;
            for (int p = 0; p < L; ++p) {
                double mx = 0.0, mxs = 0.0, me = 0.0, ml = 0.0;
                for (int b = 0; b < 4; ++b) {
                    const float x = w[4 * p + b];
                    s_w[((size_t)slot * kMaxLen + p) * 4 + b] = x;
                    mx = std::max(mx, (double)std::fabs(x));
                    if (bf) {
                        const uint16_t h = bf16_rne(x);
                        const float hf = bf16_to_float(h);
                        const uint16_t l = bf16_rne(x - hf);   // x - hf is exact in FP32
                        const float lf = bf16_to_float(l);
                        hi16[((size_t)slot * kMaxLen + p) * 4 + b] = h;
                        lo16[((size_t)slot * kMaxLen + p) * 4 + b] = l;
                        me = std::max(me, std::fabs((double)x - (double)hf - (double)lf));
                        mxs = std::max(mxs, std::fabs((double)hf) + std::fabs((double)lf));
                        ml = std::max(ml, std::fabs((double)lf));
                    } else {
                        mxs = std::max(mxs, (double)std::fabs(x));
                    }
                }
                sum_max += mx;
                sum_split += mxs;
                rep_err += me;
                sum_lo += ml;
            }
            // |FP64 left-to-right sum - exact sum| (binary64, L-1 additions)
            const double fp64_err = gamma_n(L, u53) * sum_max;
            // MFMA filter band: representation error of the split plus any
            // summation order of every product the window's dot product sees
            // (32 per chunk for hi+lo, 16 per chunk for FP32), with unit
            // roundoff 2^-22 to allow for truncating accumulation.
            const double nterms = bf ? 32.0 * nch : 16.0 * nch;
            const double band = 2.0 * (rep_err + gamma_n(nterms, std::ldexp(1.0, -22)) * sum_split + fp64_err) + 1e-30;
            // a batch may place the column in a block of any chunk count the
            // library has
            const double nterms_any = bf ? 32.0 * max_nch : 16.0 * max_nch;
            s_band_any[slot] =
                2.0 * (rep_err + gamma_n(nterms_any, std::ldexp(1.0, -22)) * sum_split + fp64_err) + 1e-30;
            // hi-only score H': the lo terms it leaves out are bounded by
            // sum_p max_b |lo|; rounding as for the full sum
            s_band_hi[slot] =
                2.0 * (sum_lo + rep_err + gamma_n(nterms, std::ldexp(1.0, -22)) * sum_split + fp64_err) + 1e-30;
            s_band_hi_any[slot] =
                2.0 * (sum_lo + rep_err + gamma_n(nterms_any, std::ldexp(1.0, -22)) * sum_split + fp64_err) + 1e-30;
            // Sequential FP32 band used by finalize (RNE adds, L-1 of them).
            const double seq = 2.0 * (gamma_n(L, u24) * sum_max + fp64_err) + 1e-30;
            s_thr[slot] = c->lib_thr[col];
            s_len[slot] = L;
            s_col[slot] = (uint32_t)col;
            s_colh[slot] = col;
            s_band[slot] = band;
            s_seq[slot] = seq;
            max_band = std::max(max_band, band);
            max_err = std::max(max_err, seq);
        }
    }

    // Fragment swizzle.
    std::vector<uint8_t> fa, fb;
    if (bf) {
        // [block][chunk][lane] uint4: lane l holds column l%32, positions
        // 4j + 2(l/32) + {0,1}, 4 bases each, as bf16 pairs.
        fa.assign((size_t)nblk * kMaxChunks * 64 * 16, 0);
        fb.assign(fa.size(), 0);
        for (int bi = 0; bi < nblk; ++bi)
            for (int j = 0; j < kMaxChunks; ++j)
                for (int l = 0; l < 64; ++l) {
                    const int slot = bi * kBf16Cols + (l & 31);
                    uint16_t vh[8], vl[8];
                    for (int i = 0; i < 8; ++i) {
                        const int p = 4 * j + 2 * (l >> 5) + i / 4, b = i % 4;
                        vh[i] = hi16[((size_t)slot * kMaxLen + p) * 4 + b];
                        vl[i] = lo16[((size_t)slot * kMaxLen + p) * 4 + b];
                    }
                    const size_t off = (((size_t)bi * kMaxChunks + j) * 64 + l) * 16;
                    std::memcpy(&fa[off], vh, 16);
                    std::memcpy(&fb[off], vl, 16);
                }
    } else {
        // [block][position][lane] float: lane l holds W_{col l%16}[p][l/16].
        fa.assign((size_t)nblk * kMaxLen * 64 * 4, 0);
        for (int bi = 0; bi < nblk; ++bi)
            for (int p = 0; p < kMaxLen; ++p)
                for (int l = 0; l < 64; ++l) {
                    const int slot = bi * kFp32Cols + (l & 15);
                    const float x = s_w[((size_t)slot * kMaxLen + p) * 4 + (l >> 4)];
                    std::memcpy(&fa[(((size_t)bi * kMaxLen + p) * 64 + l) * 4], &x, 4);
                }
        fb.assign(16, 0);
    }
    std::vector<uint32_t> gpk(groups.size());
    for (size_t g = 0; g < groups.size(); ++g) gpk[g] = group_pack(groups[g].blk0, groups[g].nch, groups[g].nb);

    c->have_lib = false;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(c->frag_a.reserve(fa.size()));
    HIP_TRY(c->frag_b.reserve(fb.size()));
    HIP_TRY(c->slot_thr.reserve(nslots));
    HIP_TRY(c->slot_band.reserve(nslots));
    HIP_TRY(c->slot_band_any.reserve(nslots));
    HIP_TRY(c->slot_band_hi.reserve(nslots));
    HIP_TRY(c->slot_band_hi_any.reserve(nslots));
    HIP_TRY(c->slot_seq.reserve(nslots));
    HIP_TRY(c->slot_w.reserve(s_w.size()));
    // finalize's pair table: entry b1 + 4 b2 of pair q is w[2q][b1] + w[2q+1][b2]
    // rounded to FP32 (rows past a motif's length are zero)
    std::vector<float> s_w2((size_t)nslots * (kMaxLen / 2) * 16, 0.f);
    for (size_t sl = 0; sl < (size_t)nslots; ++sl)
        for (int q = 0; q < kMaxLen / 2; ++q)
            for (int b1 = 0; b1 < 4; ++b1)
                for (int b2 = 0; b2 < 4; ++b2)
                    s_w2[(sl * (kMaxLen / 2) + q) * 16 + b1 + 4 * b2] =
                        s_w[(sl * kMaxLen + 2 * q) * 4 + b1] + s_w[(sl * kMaxLen + 2 * q + 1) * 4 + b2];
    HIP_TRY(c->slot_w2.reserve(s_w2.size()));
    HIP_TRY(hipMemcpy(c->slot_w2.p, s_w2.data(), s_w2.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c->slot_len.reserve(nslots));
    HIP_TRY(c->slot_col.reserve(nslots));
    HIP_TRY(c->groups_d.reserve(gpk.size()));
    HIP_TRY(hipMemcpy(c->frag_a.p, fa.data(), fa.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->frag_b.p, fb.data(), fb.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_thr.p, s_thr.data(), nslots * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_band.p, s_band.data(), nslots * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_band_any.p, s_band_any.data(), nslots * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_band_hi.p, s_band_hi.data(), nslots * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_band_hi_any.p, s_band_hi_any.data(), nslots * sizeof(double),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_seq.p, s_seq.data(), nslots * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_w.p, s_w.data(), s_w.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_len.p, s_len.data(), nslots * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->slot_col.p, s_col.data(), nslots * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->groups_d.p, gpk.data(), gpk.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->nslots = nslots;
    c->cols_per_blk = cols_per_blk;
    c->groups = groups;
    c->slot_col_h = s_colh;
    c->band_h = s_band;
    c->seq_h = s_seq;
    c->max_filter_band = max_band;
    c->max_score_error = max_err;
    c->lib_gen++;
    c->have_lib = true;
    return BLAMM_OK;
}

FilterArgs filter_args(const blamm_hip_ctx* c, const uint64_t* packed, const uint64_t* lcmask, const float* thr,
                       uint4* cand, uint64_t cand_cap, unsigned long long* counters)
{
    FilterArgs a;
    a.packed = packed;
    a.lcmask = lcmask;
    a.frag_a = c->frag_a.p;
    a.frag_b = c->frag_b.p;
    a.thr = thr;
    a.band = c->slot_band.p;
    a.band_hi = c->slot_band_hi.p;
    a.groups = c->groups_d.p;
    a.cand = cand;
    a.counters = counters;
    a.cand_cap = cand_cap;
    return a;
}

FinalizeArgs finalize_args(const blamm_hip_ctx* c, const uint4* cand, uint64_t cand_cap, unsigned long long* counters,
                           const uint64_t* packed, const uint64_t* lcmask, const uint64_t* runs)
{
    FinalizeArgs f;
    f.cand = cand;
    f.counters = counters;
    f.counters_rw = counters;
    f.cand_cap = cand_cap;
    f.packed = packed;
    f.lcmask = lcmask;
    f.nbases = 0;
    f.run_starts = runs;
    f.nruns = 0;
    f.run_bits = nullptr;
    f.slot_len = c->slot_len.p;
    f.slot_w = c->slot_w.p;
    f.slot_w2 = c->slot_w2.p;
    f.slot_thr = c->slot_thr.p;
    f.slot_seq = c->slot_seq.p;
    f.slot_col = c->slot_col.p;
    f.hits = nullptr;
    f.hit_cap = 0;
    f.desc = nullptr;
    f.req_end = nullptr;
    f.slot_map = nullptr;
    f.blk_req = nullptr;
    f.req_count = nullptr;
    return f;
}

}  // namespace

extern "C" {

int blamm_hip_abi_version(void) { return BLAMM_HIP_ABI_VERSION; }

const char* blamm_hip_last_error(void) { return g_last_error.c_str(); }

uint64_t blamm_hip_packed_words(uint64_t nbases) { return (nbases + 31) / 32 + 16; }

uint64_t blamm_hip_mask_words(uint64_t nbases) { return (nbases + 63) / 64 + 16; }

int blamm_hip_device_count(int* count)
{
    if (!count) return fail(BLAMM_ERR_ARG, "device_count: null pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return BLAMM_OK;
}

int blamm_hip_create(int device, blamm_hip_ctx** out)
{
    if (!out) return fail(BLAMM_ERR_ARG, "create: null output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(BLAMM_ERR_DEVICE, "no HIP device found");
    if (device < 0 || device >= n) return fail(BLAMM_ERR_ARG, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0)
        return fail(BLAMM_ERR_DEVICE, "libblamm_hip.so is built for gfx950, device is " + arch);
    blamm_hip_ctx* c = new blamm_hip_ctx();
    c->device = device;
    c->arch = arch.substr(0, arch.find(':'));
    c->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (const char* e = std::getenv("BLAMM_CHUNK_TILES")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= (1 << 20)) c->tuning.chunk_tiles = (uint32_t)v;
    }
    if (const char* e = std::getenv("BLAMM_FILTER_ROUNDS")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= 64) c->filter_rounds = v;
    }
    noalloc_scratch_init(device);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
    if (e == hipSuccess) e = c->counters.reserve(kCntSlots);
    if (e != hipSuccess) {
        blamm_hip_destroy(c);
        return fail(BLAMM_ERR_DEVICE, std::string("create: ") + hipGetErrorString(e));
    }
    *out = c;
    return BLAMM_OK;
}

int blamm_hip_destroy(blamm_hip_ctx* c)
{
    if (!c) return BLAMM_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->frag_a.release(); c->frag_b.release(); c->slot_thr.release(); c->slot_band.release();
    c->slot_band_any.release(); c->slot_band_hi.release(); c->slot_band_hi_any.release();
    c->slot_seq.release(); c->slot_w.release(); c->slot_w2.release();
    c->slot_len.release(); c->slot_col.release(); c->groups_d.release(); c->cand.release();
    c->counters.release(); c->run_bits.release();
    c->packed.release(); c->lcmask.release(); c->runs.release(); c->nhits_dev.release();
    c->hits.release();
    c->hits_pinned.release();
    c->hist_slots.release(); c->hist_groups.release(); c->hist_counts.release();
    for (int i = 0; i < 4; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BLAMM_OK;
}

int blamm_hip_set_library(blamm_hip_ctx* c, const float* P, int ldp, int ncols,
                          const int* lengths, const float* thresholds, int precision)
{
    if (!c) return fail(BLAMM_ERR_ARG, "set_library: null context");
    if (!P || !lengths || !thresholds || ncols <= 0)
        return fail(BLAMM_ERR_ARG, "set_library: empty library");
    if (precision != BLAMM_PREC_FP32 && !prec_bf(precision))
        return fail(BLAMM_ERR_ARG, "set_library: unknown precision");
    int lmax = 0;
    for (int j = 0; j < ncols; ++j) {
        if (lengths[j] < 1 || lengths[j] > BLAMM_MAX_MOTIF_LEN)
            return fail(BLAMM_ERR_ARG, "set_library: motif length of column " + std::to_string(j) +
                                           " outside 1.." + std::to_string(BLAMM_MAX_MOTIF_LEN));
        lmax = std::max(lmax, lengths[j]);
        if (std::isnan(thresholds[j]))
            return fail(BLAMM_ERR_ARG, "set_library: NaN threshold in column " + std::to_string(j));
    }
    if (ldp < 4 * lmax) return fail(BLAMM_ERR_ARG, "set_library: ldp < 4 * max motif length");
    if (c->pending) return fail(BLAMM_ERR_STATE, "set_library: a scan is in flight (call scan_finish first)");
    for (int j = 0; j < ncols; ++j)
        for (int k = 0; k < 4 * lengths[j]; ++k)
            if (!std::isfinite(P[(size_t)j * ldp + k]))
                return fail(BLAMM_ERR_ARG, "set_library: non-finite weight in column " + std::to_string(j) +
                                               " (PSEUDOCOUNT must keep every PWM entry finite)");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_lib = false;
    c->precision = precision;
    c->ncols = ncols;
    c->ldp = ldp;
    c->lib_P.assign(P, P + (size_t)ncols * ldp);
    c->lib_len.assign(lengths, lengths + ncols);
    c->lib_thr.assign(thresholds, thresholds + ncols);
    const uint32_t nbmax = prec_bf(precision) ? kNbMaxBf16 : kNbMaxFp32;
    if (c->tuning.blocks_per_wave > nbmax) c->tuning.blocks_per_wave = nbmax;
    return relayout(c);
}

int blamm_hip_library_bounds(blamm_hip_ctx* c, double* band, double* err)
{
    if (!c || !c->have_lib) return fail(BLAMM_ERR_STATE, "library_bounds: no library");
    if (band) *band = c->max_filter_band;
    if (err) *err = c->max_score_error;
    return BLAMM_OK;
}

int blamm_hip_get_tuning(blamm_hip_ctx* c, blamm_tuning* t)
{
    if (!c || !t) return fail(BLAMM_ERR_ARG, "get_tuning: null pointer");
    *t = c->tuning;
    return BLAMM_OK;
}

int blamm_hip_set_tuning(blamm_hip_ctx* c, const blamm_tuning* t)
{
    if (!c || !t) return fail(BLAMM_ERR_ARG, "set_tuning: null pointer");
    if (c->pending) return fail(BLAMM_ERR_STATE, "set_tuning: a scan is in flight");
    BLAMM_TRY(check_tuning(*t, c->have_lib ? c->precision : BLAMM_PREC_FP32));
    const bool relay = c->have_lib && (t->blocks_per_wave != c->tuning.blocks_per_wave ||
                                       t->kpad_positions != c->tuning.kpad_positions);
    c->tuning = *t;
    if (relay) {
        HIP_TRY(hipSetDevice(c->device));
        return relayout(c);
    }
    return BLAMM_OK;
}

int blamm_hip_reserve(blamm_hip_ctx* c, uint64_t n)
{
    if (!c) return fail(BLAMM_ERR_ARG, "reserve: null context");
    if (c->pending) return fail(BLAMM_ERR_STATE, "reserve: a scan is in flight (call scan_finish first)");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(c->cand.reserve(n));
    HIP_TRY(c->hits.reserve(n));
    return BLAMM_OK;
}

}  // extern "C"

namespace {

// Argument checks shared by the device-buffer scan entry points.
int check_scan_args(const blamm_hip_ctx* c, const uint64_t* d_packed, uint64_t packed_words, const uint64_t* d_lcmask,
                    uint64_t mask_words, uint64_t nbases, const uint64_t* d_run_starts, uint64_t nruns,
                    const blamm_hit* d_hits, uint64_t cap)
{
    if (!c) return fail(BLAMM_ERR_ARG, "scan: null context");
    if (!c->have_lib) return fail(BLAMM_ERR_STATE, "scan: no library uploaded");
    if (c->pending) return fail(BLAMM_ERR_STATE, "scan: previous scan not finished");
    if (nbases >= (1ull << 32) - 512) return fail(BLAMM_ERR_ARG, "scan: nbases must be < 2^32 - 512 per call");
    if (nbases > 0 && !d_packed) return fail(BLAMM_ERR_ARG, "scan: null packed stream");
    if (nbases > 0 && packed_words < blamm_hip_packed_words(nbases))
        return fail(BLAMM_ERR_ARG, "scan: packed buffer shorter than blamm_hip_packed_words(nbases)");
    if (d_lcmask && mask_words < blamm_hip_mask_words(nbases))
        return fail(BLAMM_ERR_ARG, "scan: mask buffer shorter than blamm_hip_mask_words(nbases)");
    if (nruns > 0 && !d_run_starts) return fail(BLAMM_ERR_ARG, "scan: null run starts");
    if (nruns >= (1ull << 32)) return fail(BLAMM_ERR_ARG, "scan: too many runs");
    if (cap > 0 && !d_hits) return fail(BLAMM_ERR_ARG, "scan: null hit buffer");
    return BLAMM_OK;
}

// Enqueue counter clear, filter, finalize (and the hit-count copy) on s, with
// the filter records in cand[0..cand_cap) and the counters in `counters`.
// With run_bits (nbases / 64 + 2 words) finalize checks the run rule from
// run-start bits built here; without, it searches the run starts.
int enqueue_scan(blamm_hip_ctx* c, const uint64_t* d_packed, const uint64_t* d_lcmask, uint64_t nbases,
                 const uint64_t* d_run_starts, uint64_t nruns, blamm_hit* d_hits, uint64_t cap, uint64_t* d_nhits,
                 uint4* cand, uint64_t cand_cap, unsigned long long* counters, uint64_t* run_bits, hipStream_t s)
{
    const bool bf = prec_bf(c->precision);
    const bool mask = d_lcmask != nullptr;
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counters);
    HIP_TRY(hipGetLastError());
    if (nruns == 0) run_bits = nullptr;   // nothing to search
    if (run_bits) {
        const uint32_t words = (uint32_t)(nbases / 64 + 2);
        hipLaunchKernelGGL(run_bits_clear_kernel, dim3(std::min<uint32_t>((words + 255) / 256, 4096)), dim3(256), 0, s,
                           run_bits, (uint32_t)nbases, (const BatchDesc*)nullptr);
        hipLaunchKernelGGL(run_bits_set_kernel, dim3(std::min<uint32_t>((uint32_t)((nruns + 255) / 256), 4096)),
                           dim3(256), 0, s, run_bits, d_run_starts, (uint32_t)nruns, (uint32_t)nbases,
                           (const BatchDesc*)nullptr);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(c->ev[0], s));
    if (nbases > 0) {
        const uint32_t ntiles = (uint32_t)((nbases + kBf16Tile - 1) / kBf16Tile);   // 32-window tiles
        const FilterArgs a = filter_args(c, d_packed, d_lcmask, c->slot_thr.p, cand, cand_cap, counters);
        WorkQueue q{};
        q.ntiles = ntiles;
        q.chunk_tiles = std::max<uint32_t>(1, std::min<uint32_t>(c->tuning.chunk_tiles, ntiles));
        q.nchunks = (ntiles + q.chunk_tiles - 1) / q.chunk_tiles;
        q.ngroups = (uint32_t)c->groups.size();
        if ((uint64_t)((q.nchunks + kXcds - 1) / kXcds) * q.ngroups >= (1ull << 32))
            return fail(BLAMM_ERR_ARG, "scan: work queue too long");
        const ListArgs l{nullptr, nullptr, nullptr};
        const uint32_t grid =
            (uint32_t)c->num_cu * (uint32_t)persistent_wgs_per_cu(bf, prec_lo(c->precision), mask, false) *
            (uint32_t)c->filter_rounds;
        hipError_t e = launch_filter(bf, prec_lo(c->precision), mask, false, a, q, l, grid, s);
        if (e != hipSuccess) return fail(BLAMM_ERR_DEVICE, std::string("filter launch: ") + hipGetErrorString(e));
    }
    HIP_TRY(hipEventRecord(c->ev[1], s));
    {
        FinalizeArgs f = finalize_args(c, cand, cand_cap, counters, d_packed, d_lcmask, d_run_starts);
        f.nbases = (uint32_t)nbases;
        f.nruns = (uint32_t)nruns;
        f.run_bits = run_bits;
        f.hits = d_hits;
        f.hit_cap = cap;
        hipError_t e = launch_finalize(bf, false, f, (uint32_t)c->num_cu * 8, s);
        if (e != hipSuccess) return fail(BLAMM_ERR_DEVICE, std::string("finalize launch: ") + hipGetErrorString(e));
    }
    HIP_TRY(hipEventRecord(c->ev[2], s));
    if (d_nhits) {
        hipLaunchKernelGGL(store_count_kernel, dim3(1), dim3(64), 0, s, counters, d_nhits);
        HIP_TRY(hipGetLastError());
    }
    c->pending = true;
    c->pending_hit_cap = cap;
    c->pending_cand_cap = cand_cap;
    c->pending_counters = counters;
    c->pending_stream = s;
    return BLAMM_OK;
}

}  // namespace

extern "C" {

int blamm_hip_scan_device(blamm_hip_ctx* c, const uint64_t* d_packed, uint64_t packed_words,
                          const uint64_t* d_lcmask, uint64_t mask_words, uint64_t nbases,
                          const uint64_t* d_run_starts, uint64_t nruns, blamm_hit* d_hits,
                          uint64_t cap, uint64_t* d_nhits, void* stream)
{
    BLAMM_TRY(check_scan_args(c, d_packed, packed_words, d_lcmask, mask_words, nbases, d_run_starts, nruns, d_hits,
                              cap));
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (c->cand.n == 0) HIP_TRY(c->cand.reserve(1u << 20));
    if (nruns > 0 && c->run_bits.n < nbases / 64 + 2) HIP_TRY(c->run_bits.reserve(nbases / 64 + 2));
    c->pending_noalloc = false;
    return enqueue_scan(c, d_packed, d_lcmask, nbases, d_run_starts, nruns, d_hits, cap, d_nhits, c->cand.p,
                        c->cand.n, c->counters.p, c->run_bits.p, s);
}

int blamm_hip_scan_device_noalloc(blamm_hip_ctx* c, const uint64_t* d_packed, uint64_t packed_words,
                                  const uint64_t* d_lcmask, uint64_t mask_words, uint64_t nbases,
                                  const uint64_t* d_run_starts, uint64_t nruns, blamm_hit* d_hits,
                                  uint64_t cap, uint64_t* d_nhits, void* stream)
{
    BLAMM_TRY(check_scan_args(c, d_packed, packed_words, d_lcmask, mask_words, nbases, d_run_starts, nruns, d_hits,
                              cap));
    if (c->device < 0 || c->device >= kMaxDevices)
        return fail(BLAMM_ERR_DEVICE, "scan_noalloc: device index beyond the scratch table");
    uint8_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = (uint8_t*)g_scratch_ptr[c->device];
        if (!base) return fail(BLAMM_ERR_DEVICE, "scan_noalloc: no scratch address (" + g_scratch_err[c->device] + ")");
    }
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    c->pending_noalloc = true;
    return enqueue_scan(c, d_packed, d_lcmask, nbases, d_run_starts, nruns, d_hits, cap, d_nhits,
                        (uint4*)(base + kNoAllocHeader), (uint64_t)BLAMM_NOALLOC_MAX_RECORDS,
                        (unsigned long long*)base, nullptr, s);
}

int blamm_hip_scan_finish(blamm_hip_ctx* c, uint64_t* nhits)
{
    if (!c) return fail(BLAMM_ERR_ARG, "scan_finish: null context");
    if (!c->pending) return fail(BLAMM_ERR_STATE, "scan_finish: no scan in flight");
    c->pending = false;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->pending_stream));
    unsigned long long cnt[kCntSlots];
    HIP_TRY(hipMemcpy(cnt, c->pending_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    float ms_f = 0.f, ms_all = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms_f, c->ev[0], c->ev[1]));
    HIP_TRY(hipEventElapsedTime(&ms_all, c->ev[0], c->ev[2]));
    c->stats.candidates = cnt[kCntCand];
    c->stats.band_rescored = cnt[kCntBand];
    c->stats.hits = cnt[kCntHits];
    c->stats.kernel_ms = ms_all;
    c->stats.filter_ms = ms_f;
    if (nhits) *nhits = cnt[kCntHits];
    if (cnt[kCntCand] > c->pending_cand_cap) {
        // Filter records were dropped, so the hit count is incomplete: report
        // 0 hits and the overflow.
        if (nhits) *nhits = 0;
        if (c->pending_noalloc)
            return fail(BLAMM_ERR_OVERFLOW, "scan_noalloc: " + std::to_string(cnt[kCntCand]) +
                                                " filter records, more than the no-alloc scratch holds (" +
                                                std::to_string(c->pending_cand_cap) +
                                                "); scan a shorter stream or fewer columns per call");
        // Grow the scratch so that a retry holds every candidate.
        HIP_TRY(c->cand.reserve((size_t)(cnt[kCntCand] + cnt[kCntCand] / 4 + 1024)));
        return fail(BLAMM_ERR_OVERFLOW, "scan: candidate scratch overflowed (grown; retry the scan)");
    }
    if (cnt[kCntHits] > c->pending_hit_cap)
        return fail(BLAMM_ERR_OVERFLOW, "scan: hit capacity too small");
    return BLAMM_OK;
}

int blamm_hip_last_stats(blamm_hip_ctx* c, blamm_scan_stats* st)
{
    if (!c || !st) return fail(BLAMM_ERR_ARG, "last_stats: null pointer");
    *st = c->stats;
    return BLAMM_OK;
}

}  // extern "C"

namespace {

// Upload a host stream into the context's staging buffers (c->packed,
// c->lcmask, c->runs), on the context's stream.
int stage_stream(blamm_hip_ctx* c, const uint64_t* packed, const uint64_t* lcmask, uint64_t nbases,
                 const uint64_t* run_starts, uint64_t nruns)
{
    const uint64_t pw = blamm_hip_packed_words(nbases), mw = blamm_hip_mask_words(nbases);
    HIP_TRY(c->packed.reserve(pw));
    HIP_TRY(hipMemsetAsync(c->packed.p, 0, pw * 8, c->stream));
    if (nbases > 0)
        HIP_TRY(hipMemcpyAsync(c->packed.p, packed, ((nbases + 31) / 32) * 8, hipMemcpyHostToDevice, c->stream));
    if (lcmask) {
        HIP_TRY(c->lcmask.reserve(mw));
        HIP_TRY(hipMemsetAsync(c->lcmask.p, 0, mw * 8, c->stream));
        HIP_TRY(hipMemcpyAsync(c->lcmask.p, lcmask, ((nbases + 63) / 64) * 8, hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(c->runs.reserve(nruns + 1));
    if (nruns > 0)
        HIP_TRY(hipMemcpyAsync(c->runs.p, run_starts, nruns * 8, hipMemcpyHostToDevice, c->stream));
    return BLAMM_OK;
}

// Upload a host stream and run the scan into c->hits (grown until every hit
// fits); *nhits = hits on the device.
int scan_staged(blamm_hip_ctx* c, const uint64_t* packed, const uint64_t* lcmask, uint64_t nbases,
                const uint64_t* run_starts, uint64_t nruns, uint64_t cap_limit, uint64_t* nhits)
{
    BLAMM_TRY(stage_stream(c, packed, lcmask, nbases, run_starts, nruns));
    const uint64_t pw = blamm_hip_packed_words(nbases), mw = blamm_hip_mask_words(nbases);
    uint64_t dcap = std::min<uint64_t>(cap_limit, std::max<uint64_t>(c->hits.n, 1u << 16));
    for (int attempt = 0; attempt < 8; ++attempt) {
        HIP_TRY(c->hits.reserve(std::max<uint64_t>(dcap, 1)));
        int rc = blamm_hip_scan_device(c, c->packed.p, pw, lcmask ? c->lcmask.p : nullptr, mw, nbases,
                                       c->runs.p, nruns, c->hits.p, dcap, nullptr, c->stream);
        if (rc != BLAMM_OK) return rc;
        uint64_t n = 0;
        rc = blamm_hip_scan_finish(c, &n);
        if (rc == BLAMM_OK) {
            *nhits = n;
            return BLAMM_OK;
        }
        if (rc != BLAMM_ERR_OVERFLOW) return rc;
        if (n == 0 && c->stats.candidates > dcap) {
            // the candidate scratch overflowed (and was grown): size the hits
            // for as many as there are records, which usually holds them all,
            // so the retry is the last run
            dcap = std::min<uint64_t>(cap_limit, c->stats.candidates);
        }
        if (n > dcap) {   // hit buffer too small
            if (n > cap_limit) {
                *nhits = n;
                return fail(BLAMM_ERR_OVERFLOW, "scan: hit capacity too small");
            }
            dcap = n;
        }
    }
    return fail(BLAMM_ERR_DEVICE, "scan: capacity retries exhausted");
}

}  // namespace

extern "C" {

int blamm_hip_scan_pinned(blamm_hip_ctx* c, const uint64_t* packed, const uint64_t* lcmask, uint64_t nbases,
                          const uint64_t* run_starts, uint64_t nruns, const blamm_hit** hits, uint64_t* nhits)
{
    if (!c) return fail(BLAMM_ERR_ARG, "scan: null context");
    if (!nhits || !hits) return fail(BLAMM_ERR_ARG, "scan: null output");
    if (!c->have_lib) return fail(BLAMM_ERR_STATE, "scan: no library uploaded");
    if (nbases > 0 && !packed) return fail(BLAMM_ERR_ARG, "scan: null packed stream");
    if (nbases >= (1ull << 32) - 512) return fail(BLAMM_ERR_ARG, "scan: nbases must be < 2^32 - 512 per call");
    HIP_TRY(hipSetDevice(c->device));
    uint64_t n = 0;
    BLAMM_TRY(scan_staged(c, packed, lcmask, nbases, run_starts, nruns, UINT64_MAX, &n));
    if (n > c->hits_pinned.n) HIP_TRY(c->hits_pinned.reserve(n + n / 8));
    if (n > 0) HIP_TRY(hipMemcpy(c->hits_pinned.p, c->hits.p, n * sizeof(blamm_hit), hipMemcpyDeviceToHost));
    *hits = c->hits_pinned.p;
    *nhits = n;
    return BLAMM_OK;
}

}  // extern "C"

extern "C" {

int blamm_hip_scan(blamm_hip_ctx* c, const uint64_t* packed, const uint64_t* lcmask, uint64_t nbases,
                   const uint64_t* run_starts, uint64_t nruns, blamm_hit* hits, uint64_t cap,
                   uint64_t* nhits)
{
    if (!c) return fail(BLAMM_ERR_ARG, "scan: null context");
    if (!nhits) return fail(BLAMM_ERR_ARG, "scan: null nhits");
    if (!c->have_lib) return fail(BLAMM_ERR_STATE, "scan: no library uploaded");
    if (nbases > 0 && !packed) return fail(BLAMM_ERR_ARG, "scan: null packed stream");
    if (cap > 0 && !hits) return fail(BLAMM_ERR_ARG, "scan: null hit buffer");
    if (nbases >= (1ull << 32) - 512) return fail(BLAMM_ERR_ARG, "scan: nbases must be < 2^32 - 512 per call");
    HIP_TRY(hipSetDevice(c->device));
    uint64_t n = 0;
    const int rc = scan_staged(c, packed, lcmask, nbases, run_starts, nruns, cap, &n);
    *nhits = n;
    if (rc != BLAMM_OK) return rc;
    if (n > 0) HIP_TRY(hipMemcpy(hits, c->hits.p, n * sizeof(blamm_hit), hipMemcpyDeviceToHost));
    return BLAMM_OK;
}

}  // extern "C"

#include "batch.hpp"
#include "tune.hpp"
#include "serve.hpp"
#include "hist.hpp"
