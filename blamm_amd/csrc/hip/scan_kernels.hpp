// Device kernels of the MI355X scan engine (gfx950, wave64).
//
// The scan computes, for every window q of the packed stream and every PWM
// column c, score(q,c) = sum_p W_c[p][base(q+p)].  The reference expresses this
// as R = S * P with S the one-hot strip matrix (SeqMatrix, sequence.cpp:299-340)
// and P the pattern matrix (motif.cpp:542-564), one sgemm per column tile and
// per offset (pwmscan.cpp:385-389), followed by a threshold/compaction pass over
// R (kernel.cu:21-45).  Here the one-hot operand is built in registers straight
// from 2-bit packed bases, the product runs on MFMA with FP32 accumulation, and
// the threshold test plus hit compaction run in the epilogue, so R never
// reaches HBM.  The MFMA score only filters: a window becomes a candidate when
// it clears thr - band (band = proven bound on |MFMA score - FP64 score|).
// finalize_kernel then applies the run-boundary rule and decides every
// candidate exactly against the FP64 left-to-right sum.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace blamm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaveSize = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kThreads = kWaveSize * kWavesPerBlock;
constexpr int kMaxLen = 64;          // positions per column slot
constexpr int kMaxChunks = kMaxLen / 4;
constexpr int kWideChunks = 8;       // above this many chunks a wave holds one column block
constexpr int kLoChunks = 16;        // lazy-lo form: chunk fragments per wave in LDS (NB x NCH)
constexpr int kBf16Cols = 32;        // columns per bf16 column block (MFMA N)
constexpr int kBf16Tile = 32;        // windows per bf16 tile (MFMA M)
constexpr int kFp32Cols = 16;        // columns per fp32 column block
constexpr int kFp32Tile = 16;        // windows per fp32 tile

// Batch slot without a column.
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// Counter slots in device memory.
enum { kCntCand = 0, kCntHits = 1, kCntBand = 2, kCntQueue = 4, kCntSlots = 12 };

// 64 consecutive bits of a little-endian bit stream starting at bit `bit`.
__device__ __forceinline__ uint64_t load_bits64(const uint64_t* __restrict__ words, uint64_t bit)
{
    const uint64_t wi = bit >> 6;
    const uint32_t sh = (uint32_t)(bit & 63);
    const uint64_t w0 = words[wi];
    const uint64_t w1 = words[wi + 1];
    return sh ? ((w0 >> sh) | (w1 << (64 - sh))) : w0;
}

// Reserve `cnt` consecutive slots of a device counter for the set lanes of
// `mask` and return this lane's slot (only meaningful for set lanes).  One
// atomic per wave; the counter is 64-bit so it never wraps.
__device__ __forceinline__ uint64_t wave_reserve(unsigned long long* counter, uint64_t mask)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t cnt = (uint32_t)__popcll(mask);
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(counter, (unsigned long long)cnt);
    base = __shfl(base, 0);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    return (uint64_t)base + rank;
}

struct FilterArgs {
    const uint64_t* packed;      // 2-bit codes
    const uint64_t* lcmask;      // lowercase bits (MASK variants only)
    const void* frag_a;          // bf16: hi fragments; fp32: weights
    const void* frag_b;          // bf16: lo fragments
    const float* thr;            // per slot: threshold
    const double* band;          // per slot: filter band (MFMA score vs FP64 score)
    const double* band_hi;       // per slot: band of the hi-only score (bf16 lazy-lo stage)
    const uint32_t* groups;      // column groups, see group_pack()
    uint4* cand;                 // (first window of a lane, slot, rows above the threshold, 0)
    unsigned long long* counters;
    uint64_t cand_cap;
};

// Largest float <= t - b and smallest float >= t + b (b >= 0), for the
// filter and rescoring thresholds.  The double result is pushed a few of its
// own ulps outward before the directed conversion, which covers the rounding
// of the double subtraction itself.  Infinite thresholds stay as they are.
__device__ __forceinline__ float thr_minus(float t, double b)
{
    if (!__builtin_isfinite(t)) return t;
    double d = (double)t - b;
    d -= __builtin_fabs(d) * 0x1p-50 + 0x1p-1000;
    return __double2float_rd(d);
}

__device__ __forceinline__ float thr_plus(float t, double b)
{
    if (!__builtin_isfinite(t)) return t;
    double d = (double)t + b;
    d += __builtin_fabs(d) * 0x1p-50 + 0x1p-1000;
    return __double2float_ru(d);
}

// A column group is nb consecutive blocks of one length class (nch chunks of
// 4 positions), scored together by one wave.
__host__ __device__ constexpr uint32_t group_pack(uint32_t blk0, uint32_t nch, uint32_t nb)
{
    return blk0 | (nch << 20) | (nb << 26);
}

// Per-wave staging of filter candidates in LDS.  A single global counter hit
// by every candidate serialises on one L2 atomic unit (tens of millions of
// candidates per scan), so each wave collects up to kStage candidates in its
// own LDS slice and reserves global space with one atomic per flush.  Only
// the owning wave touches its slice; LDS operations of one wave complete in
// program order, so no barrier is needed.  256 entries (16 KiB per block)
// need half the reservations of 128 (measured: -0.5% per scan of 100 Mbp x
// 2,000 columns).
constexpr int kStage = 256;

struct Stage {
    uint4* buf;        // this wave's LDS slice
    uint32_t n;        // staged entries (wave-uniform)
    uint4* lo;         // bf16: this wave's lo fragments (LDS), [block][chunk][lane]
};

template <class Args>   // FilterArgs or HistArgs: cand, counters, cand_cap
__device__ __forceinline__ void stage_flush(const Args& a, Stage& st)
{
    if (st.n == 0) return;
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&a.counters[kCntCand], (unsigned long long)st.n);
    base = __shfl(base, 0);
    for (uint32_t i = lane; i < st.n; i += 64) {
        const uint64_t at = (uint64_t)base + i;
        if (at < a.cand_cap) a.cand[at] = st.buf[i];
    }
    st.n = 0;
}

// Record the lanes whose accumulator maximum clears the filter threshold.
// A record (first window of the lane, slot, row mask) stands for the lane's
// windows (16 rows of the bf16 tile, 4 of the fp32 sub-tile) whose MFMA score
// clears the filter threshold; finalize_kernel decides them exactly.  The row
// mask is built only when some lane of the wave passes (rare), so the common
// path stays a max, a compare and a ballot.
template <int NR>
__device__ __forceinline__ void emit_rows(const FilterArgs& a, Stage& st, const float (&acc)[NR], float th,
                                          uint32_t qlane, uint32_t slot)
{
    float mx = acc[0];
#pragma unroll
    for (int r = 1; r < NR; ++r) mx = fmaxf(mx, acc[r]);
    const bool pred = mx >= th;
    const uint64_t m = __ballot(pred);
    if (m) {
        uint32_t rows = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) rows |= (acc[r] >= th ? 1u : 0u) << r;
        if (st.n > (uint32_t)(kStage - 64)) stage_flush(a, st);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (pred) st.buf[st.n + rank] = make_uint4(qlane, slot, rows, 0u);
        st.n += (uint32_t)__popcll(m);
    }
}

// Wait states between matrix-core results and the VALU code that reads them.
// The compiler counts the wait states a v_mfma result needs before a VALU
// read; for v_mfma_f32_16x16x4_f32 on gfx950 its count proved too small now
// and then (the FP32 filter lost a candidate in about one scan in ten).  Where
// a read can follow the last MFMA of an accumulator closely, 16 wait states
// are placed here and pinned by scheduling barriers.
__device__ __forceinline__ void mfma_result_guard()
{
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_nop 7\n\ts_nop 7" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// Every VALU read of an MFMA accumulator goes through acc_read: kGuard places
// the guard's wait states first; kCovered states, at the call, that the
// MFMAs of a whole further tile were issued after the last MFMA that wrote
// the accumulator (the software-pipelined loops), which covers the result
// latency many times over.
enum class AccRead { kGuard, kCovered };

template <AccRead R, typename V>
__device__ __forceinline__ const V& acc_read(const V& acc)
{
    if constexpr (R == AccRead::kGuard) mfma_result_guard();
    return acc;
}

// 64 bits of the 128-bit value (hi:lo) starting at bit s (0 <= s < 64),
// without a branch for s == 0.
__device__ __forceinline__ uint64_t funnel64(uint64_t lo, uint64_t hi, uint32_t s, uint32_t s_inv)
{
    return (lo >> s) | ((hi << 1) << s_inv);   // s_inv = 63 - s
}

__device__ __forceinline__ uint32_t funnel32(uint32_t lo, uint32_t hi, uint32_t s)
{
    return __builtin_amdgcn_alignbit(hi, lo, s);
}

// Tiles are 32 consecutive windows; a wave walks tiles [t0, t1) of its
// range.  The FP32 loop keeps a register window of kGroup+1 packed words and
// prefetches the next group's words one group ahead.
constexpr int kGroup = 4;

// Two forms of the bf16 filter (template flag LO), both software-pipelined
// over two accumulator sets and both exact after finalize:
//  - hi only (LO = false, BLAMM_PREC_BF16X2): a tile block whose hi-only
//    score can reach the threshold records its windows at or above
//    thr - band_hi for finalize to decide; lo products are never issued.
//  - lazy lo (LO = true, BLAMM_PREC_BF16X2_LO): such a block gets its lo
//    products (lo fragments kept in a per-wave LDS slice) and records its
//    windows at or above thr - band.
// At 100 Mbp x 2,000 columns the hi-only form issues 11% fewer MFMAs and no
// lo-fragment LDS traffic for 14% more records (filter 21.9 -> 19.2 ms at two
// waves per SIMD).

// ---------------------------------------------------------------------------
// Split-BF16 filter: W = hi + lo (bf16 each); one-hot operand is exact in bf16.
// MFMA v_mfma_f32_32x32x16_bf16: A = one-hot (32 windows x 16 = 4 positions x
// 4 bases), B = PWM chunk (16 x 32 columns), D = 32x32 FP32 scores.
// Lane l: A rows/ B cols index l%32, K half h = l/32 holds positions 2h, 2h+1
// of the chunk; D lane l holds column l%32, rows 8(r/4) + 4h + r%4.
// Each wave keeps the hi terms of NB column blocks (NCH chunks) in registers
// and streams window tiles through them (the lazy-lo form keeps the lo terms
// in its LDS slice; the histogram kernel keeps both in registers).
// ---------------------------------------------------------------------------
// The one-hot operand of chunk j for the lane's window: bf16 1.0 (0x3F80) in
// lane b of four per base, built in registers (no memory latency; the VALU
// work overlaps earlier MFMAs); a lowercase base gets an all-zero row.
// Its source, the lane's bases as shift counts: byte j of dword 0 / dword 1
// of p[h] holds 16 x the base at chunk 4h+j's first / second position of the
// lane (bits 4-5 of the byte, the rest zero), so the position's one-hot pair
// of dwords is 0x3F80 shifted by that byte.  v_lshlrev_b64 reads only the low
// six bits of its count, so a chunk costs one 64-bit shift of p[h] (none for
// chunks 0 and 4) and one 64-bit shift per position.
struct OneHotSrc {
    uint64_t p[4];
};

// Packed words as loaded by a lane: the lane's pointer is offset to the
// dword holding its window's first base, which is 4-byte aligned only.
typedef uint64_t u64_a4 __attribute__((aligned(4)));

// Shift counts of the lane's window whose first base is at bit s (< 32) of
// the dword pair w0; the window's bases continue through w1 (bases 16..47)
// and w2 (bases 48..63).  Every 4 chunks need 32 more bits.
__device__ __forceinline__ uint64_t onehot_counts(uint32_t b)
{
    constexpr uint32_t M = 0x30303030u;
    uint64_t p = ((uint64_t)((b << 2) & M) << 32) | ((b << 4) & M);
    // an empty asm (no instruction) keeps the compiler from folding the
    // masks back into every per-chunk extraction
    asm("" : "+v"(p));
    return p;
}

template <int NCH>
__device__ __forceinline__ OneHotSrc onehot_src(uint64_t w0, uint64_t w1, uint64_t w2, uint32_t s)
{
    OneHotSrc r;
    r.p[0] = onehot_counts(funnel32((uint32_t)w0, (uint32_t)(w0 >> 32), s));
    r.p[1] = NCH > 4 ? onehot_counts(funnel32((uint32_t)(w0 >> 32), (uint32_t)w1, s)) : 0;
    r.p[2] = NCH > 8 ? onehot_counts(funnel32((uint32_t)w1, (uint32_t)(w1 >> 32), s)) : 0;
    r.p[3] = NCH > 12 ? onehot_counts(funnel32((uint32_t)(w1 >> 32), (uint32_t)w2, s)) : 0;
    return r;
}

// Lowercase bits of the lane's window (bit p = position p), from consecutive
// mask dwords m0, m1, m2 starting at bit s (< 32) of m0.
template <int NCH>
__device__ __forceinline__ uint64_t window_mask(uint32_t m0, uint32_t m1, uint32_t m2, uint32_t s)
{
    uint64_t m = funnel32(m0, m1, s);
    if constexpr (NCH > kWideChunks) m |= (uint64_t)funnel32(m1, m2, s) << 32;
    return m;
}

template <bool MASK>
__device__ __forceinline__ bf16x8 onehot_chunk(const OneHotSrc& src, uint64_t mb, int j)
{
    uint64_t c = src.p[j >> 2];
    if (j & 3) {
        // the shift amount goes through an empty asm so that the compiler,
        // not knowing it is below 32, keeps one v_lshrrev_b64 instead of
        // splitting it into two dword shifts
        uint32_t sa = 8 * (j & 3);
        asm("" : "+s"(sa));
        c >>= sa;
    }
    uint64_t o0 = 0x3F80ull << ((uint32_t)c & 63u);
    uint64_t o1 = 0x3F80ull << ((uint32_t)(c >> 32) & 63u);
    if (MASK) {
        if ((mb >> (4 * j)) & 1) o0 = 0;
        if ((mb >> (4 * j + 1)) & 1) o1 = 0;
    }
    const u32x4 av = {(uint32_t)o0, (uint32_t)(o0 >> 32), (uint32_t)o1, (uint32_t)(o1 >> 32)};
    return __builtin_bit_cast(bf16x8, av);
}

template <int NCH, int NB, bool MASK>
__device__ __forceinline__ void bf16_mfma_tile(f32x16 (&acc)[NB], const bf16x8 (&bh)[NB][NCH],
                                               const bf16x8 (&bl)[NB][NCH], const OneHotSrc& bits, uint64_t mb)
{
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = (f32x16){0.f};
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
        const bf16x8 A = onehot_chunk<MASK>(bits, mb, j);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, bh[b][j], acc[b], 0, 0, 0);
            acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, bl[b][j], acc[b], 0, 0, 0);
        }
    }
}

// ---------------------------------------------------------------------------
// Lazy lo.  The hi products alone give a score H'; a window's FP64 score can
// reach thr only if H' >= thr - band_hi, where band_hi bounds the lo terms,
// the split's representation error and every rounding (proved per column on
// the host).  So a tile block whose hi-only maximum stays below that bound in
// every lane cannot hold a candidate and its lo products are skipped; blocks
// that pass get their lo products and the exact-band test of the plain form
// (modes 1, 2), or record their windows with H' >= thr - band_hi as they are
// (mode 3; the record's row mask then marks those windows).
// ---------------------------------------------------------------------------
template <int NCH, int NB, bool MASK>
__device__ __forceinline__ void bf16_hi_tile(f32x16 (&acc)[NB], const bf16x8 (&bh)[NB][NCH],
                                             const OneHotSrc& bits, uint64_t mb)
{
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = (f32x16){0.f};
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
        const bf16x8 A = onehot_chunk<MASK>(bits, mb, j);
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, bh[b][j], acc[b], 0, 0, 0);
    }
}

template <int NB>
__device__ __forceinline__ float tile_max(const f32x16& v)
{
    float mx = v[0];
#pragma unroll
    for (int r = 1; r < 16; ++r) mx = fmaxf(mx, v[r]);
    return mx;
}

// Finish a tile whose hi products are in acc (R: how the first read is kept
// clear of the matrix cores).  Lazy-lo form: the lo fragments come from the
// wave's LDS slice.
template <int NCH, int NB, bool MASK, bool LO, AccRead R>
__device__ __forceinline__ void bf16_finish_tile(const FilterArgs& a, Stage& st, f32x16 (&acc)[NB],
                                                 const float (&th)[NB], const float (&thh)[NB], const OneHotSrc& bits,
                                                 uint64_t mb, uint32_t q0, uint32_t blk0, uint32_t lane)
{
    uint32_t need = 0;
    const f32x16 (&hv)[NB] = acc_read<R>(acc);
#pragma unroll
    for (int b = 0; b < NB; ++b)
        if (__ballot(tile_max<NB>(hv[b]) >= thh[b])) need |= 1u << b;
    if (need == 0) return;
    const uint32_t qlane = q0 + 4 * (lane >> 5);
    if constexpr (!LO) {
        (void)th; (void)bits; (void)mb;
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (need & (1u << b)) {
                float v[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) v[r] = hv[b][r];
                emit_rows<16>(a, st, v, thh[b], qlane, (blk0 + b) * kBf16Cols + (lane & 31));
            }
    } else {
        // one uniform branch per block around its whole chunk loop (a branch
        // per MFMA would copy the accumulator at every join)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (!(need & (1u << b))) continue;
            bf16x8 B[NCH];
#pragma unroll
            for (int j = 0; j < NCH; ++j) B[j] = __builtin_bit_cast(bf16x8, st.lo[(b * NCH + j) * 64 + lane]);
#pragma unroll
            for (int j = 0; j < NCH; ++j)
                acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(onehot_chunk<MASK>(bits, mb, j), B[j], acc[b], 0, 0, 0);
        }
        const f32x16 (&fv)[NB] = acc_read<AccRead::kGuard>(acc);
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (need & (1u << b)) {
                float v[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) v[r] = fv[b][r];
                emit_rows<16>(a, st, v, th[b], qlane, (blk0 + b) * kBf16Cols + (lane & 31));
            }
    }
}

// Tiles [t0, t1) of column blocks blk0 .. blk0+NB-1.  Lane-private bases
// of tile t come from packed words t and t+1 of the lane's pointer (a sliding
// pair).  Software-pipelined: the hi products of tile t are issued before
// tile t-1 is finished.
template <int NCH, int NB, bool MASK, bool LO>
__device__ __forceinline__ void bf16_range(const FilterArgs& a, Stage& st, uint32_t blk0, uint32_t t0, uint32_t t1,
                                           uint32_t lane)
{
    static_assert(!LO || NB * NCH <= kLoChunks, "lo fragments exceed the wave's LDS slice");
    static_assert(NCH <= kWideChunks || NB == 1, "wide blocks are scored one per wave");
    const uint4* __restrict__ fh = (const uint4*)a.frag_a;
    const uint4* __restrict__ fl = (const uint4*)a.frag_b;
    bf16x8 bh[NB][NCH];
    float th[NB], thh[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
#pragma unroll
        for (int j = 0; j < NCH; ++j) {
            const size_t idx = ((size_t)(blk0 + b) * kMaxChunks + j) * 64 + lane;
            bh[b][j] = __builtin_bit_cast(bf16x8, fh[idx]);
            if (LO) st.lo[(b * NCH + j) * 64 + lane] = fl[idx];   // each lane reads back only its own
        }
        const uint32_t slot = (blk0 + b) * kBf16Cols + (lane & 31);
        th[b] = thr_minus(a.thr[slot], a.band[slot]);
        thh[b] = thr_minus(a.thr[slot], a.band_hi[slot]);
    }

    const uint32_t off = (lane & 31) + 2 * (lane >> 5);   // first base of the lane's window, 0..33
    // pk[t]: dwords 2t + off/16 and the next of the packed stream, so the
    // window of tile t starts at bit sh < 32 of pk[t] and continues through
    // pk[t + 1] (and pk[t + 2] for motifs longer than 48 positions)
    const uint32_t sh = 2 * (off & 15);
    const u64_a4* __restrict__ pk = (const u64_a4*)((const uint32_t*)a.packed + (off >> 4));
    const uint32_t* __restrict__ mk = (const uint32_t*)a.lcmask + (off >> 5);
    f32x16 accA[NB], accB[NB];
    OneHotSrc bA, bB{};
    uint64_t mA = 0, mB = 0;
    uint32_t m0 = 0, m1 = 0, m2 = 0;
    bA = onehot_src<NCH>(pk[t0], pk[t0 + 1], pk[t0 + 2], sh);
    if (MASK) mA = window_mask<NCH>(mk[t0], mk[t0 + 1], mk[t0 + 2], off & 31);
    bf16_hi_tile<NCH, NB, MASK>(accA, bh, bA, mA);
    uint64_t w0 = pk[t0 + 1], w1 = pk[t0 + 2], w2 = pk[t0 + 3];
    if (MASK) { m0 = mk[t0 + 1]; m1 = mk[t0 + 2]; m2 = mk[t0 + 3]; }
    uint32_t t = t0 + 1;
    bool pendingA = true;           // accA holds tile t-1, not yet finished
    for (; t + 1 < t1; t += 2) {
        const uint64_t w3 = pk[t + 3], w4 = pk[t + 4];
        uint32_t m3 = 0, m4 = 0;
        if (MASK) { m3 = mk[t + 3]; m4 = mk[t + 4]; }
        bB = onehot_src<NCH>(w0, w1, w2, sh);
        if (MASK) mB = window_mask<NCH>(m0, m1, m2, off & 31);
        bf16_hi_tile<NCH, NB, MASK>(accB, bh, bB, mB);
        bf16_finish_tile<NCH, NB, MASK, LO, AccRead::kCovered>(a, st, accA, th, thh, bA, mA, (t - 1) * kBf16Tile, blk0,
                                                              lane);
        bA = onehot_src<NCH>(w1, w2, w3, sh);
        if (MASK) mA = window_mask<NCH>(m1, m2, m3, off & 31);
        bf16_hi_tile<NCH, NB, MASK>(accA, bh, bA, mA);
        bf16_finish_tile<NCH, NB, MASK, LO, AccRead::kCovered>(a, st, accB, th, thh, bB, mB, t * kBf16Tile, blk0, lane);
        w0 = w2; w1 = w3; w2 = w4;
        if (MASK) { m0 = m2; m1 = m3; m2 = m4; }
    }
    if (t < t1) {                   // one tile left: t
        bB = onehot_src<NCH>(w0, w1, w2, sh);
        if (MASK) mB = window_mask<NCH>(m0, m1, m2, off & 31);
        bf16_hi_tile<NCH, NB, MASK>(accB, bh, bB, mB);
        bf16_finish_tile<NCH, NB, MASK, LO, AccRead::kCovered>(a, st, accA, th, thh, bA, mA, (t - 1) * kBf16Tile, blk0,
                                                              lane);
        pendingA = false;
        bf16_finish_tile<NCH, NB, MASK, LO, AccRead::kGuard>(a, st, accB, th, thh, bB, mB, t * kBf16Tile, blk0, lane);
    }
    if (pendingA)
        bf16_finish_tile<NCH, NB, MASK, LO, AccRead::kGuard>(a, st, accA, th, thh, bA, mA, (t - 1) * kBf16Tile, blk0,
                                                            lane);
}

// ---------------------------------------------------------------------------
// FP32 filter: v_mfma_f32_16x16x4_f32, one motif position per MFMA.
// A = one-hot (16 windows x 4 bases), B = W_c[p][.] (4 x 16 columns).
// Lane l: A[l%16][l/16], B[l/16][l%16]; D lane l holds column l%16,
// rows 4(l/16) + r.  A 32-window tile is two 16-window sub-tiles.
// ---------------------------------------------------------------------------
template <int NCH, int NB, bool MASK>
__device__ __forceinline__ void fp32_subtile(const FilterArgs& a, Stage& st, const float (&bw)[NB][4 * NCH],
                                             const float (&th)[NB], uint32_t q0,
                                             uint64_t bits, uint64_t bits2, uint64_t mb, uint32_t k, uint32_t blk0,
                                             uint32_t lane)
{
    constexpr int NP = 4 * NCH;
    f32x4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const uint32_t base = (uint32_t)((p < 32 ? bits >> (2 * p) : bits2 >> (2 * (p - 32)))) & 3u;
        bool on = (base == k);
        if (MASK) on = on && !((mb >> p) & 1);
        const float av = on ? 1.0f : 0.0f;
#pragma unroll
        for (int b = 0; b < NB; ++b)
            acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bw[b][p], acc[b], 0, 0, 0);
    }
    const uint32_t qlane = q0 + 4 * k;     // windows qlane + 0..3
    const f32x4 (&av)[NB] = acc_read<AccRead::kGuard>(acc);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const float v[4] = {av[b][0], av[b][1], av[b][2], av[b][3]};
        emit_rows<4>(a, st, v, th[b], qlane, (blk0 + b) * kFp32Cols + (lane & 15));
    }
}

// Tiles [t0, t1) of column blocks blk0 .. blk0+NB-1 (FP32 path).  A register
// window of kGroup + 2 packed words covers a group of tiles and the 64 bases
// past its last window; the next group's words are loaded a group ahead.
template <int NCH, int NB, bool MASK>
__device__ __forceinline__ void fp32_range(const FilterArgs& a, Stage& st, uint32_t blk0, uint32_t t0, uint32_t t1,
                                           uint32_t lane)
{
    static_assert(NCH <= kWideChunks || NB == 1, "wide blocks are scored one per wave");
    const float* __restrict__ fw = (const float*)a.frag_a;
    constexpr int NP = 4 * NCH;
    constexpr bool WIDE = NCH > kWideChunks;
    float bw[NB][NP];
    float th[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
#pragma unroll
        for (int p = 0; p < NP; ++p) bw[b][p] = fw[((size_t)(blk0 + b) * kMaxLen + p) * 64 + lane];
        const uint32_t slot = (blk0 + b) * kFp32Cols + (lane & 15);
        th[b] = thr_minus(a.thr[slot], a.band[slot]);
    }
    const uint32_t i = lane & 15, k = lane >> 4;
    const uint32_t s0 = 2 * i, s1 = 2 * (i + 16);    // bit offsets of the two sub-tiles
    const uint64_t* __restrict__ pk = a.packed;
    const uint32_t* __restrict__ mk = (const uint32_t*)a.lcmask;
    uint64_t w[kGroup + 2];
    uint32_t mw[kGroup + 2];
#pragma unroll
    for (int u = 0; u <= kGroup + 1; ++u) {
        w[u] = pk[t0 + u];
        mw[u] = MASK ? mk[t0 + u] : 0u;
    }
    for (uint32_t t = t0; t < t1; t += kGroup) {
        uint64_t nx[kGroup];
        uint32_t nm[kGroup];
        if (t + kGroup < t1) {
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                nx[u] = pk[t + kGroup + 2 + u];
                nm[u] = MASK ? mk[t + kGroup + 2 + u] : 0u;
            }
        }
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
            if (t + u < t1) {
                const uint32_t q0 = (t + u) * kBf16Tile;
                const uint64_t b0 = funnel64(w[u], w[u + 1], s0, 63 - s0);
                const uint64_t b1 = funnel64(w[u], w[u + 1], s1, 63 - s1);
                const uint64_t c0 = WIDE ? funnel64(w[u + 1], w[u + 2], s0, 63 - s0) : 0ull;
                const uint64_t c1 = WIDE ? funnel64(w[u + 1], w[u + 2], s1, 63 - s1) : 0ull;
                uint64_t m0 = 0, m1 = 0;
                if (MASK) {
                    m0 = funnel32(mw[u], mw[u + 1], i);
                    m1 = funnel32(mw[u], mw[u + 1], i + 16);
                    if (WIDE) {
                        m0 |= (uint64_t)funnel32(mw[u + 1], mw[u + 2], i) << 32;
                        m1 |= (uint64_t)funnel32(mw[u + 1], mw[u + 2], i + 16) << 32;
                    }
                }
                fp32_subtile<NCH, NB, MASK>(a, st, bw, th, q0, b0, c0, m0, k, blk0, lane);
                fp32_subtile<NCH, NB, MASK>(a, st, bw, th, q0 + kFp32Tile, b1, c1, m1, k, blk0, lane);
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            w[u] = w[kGroup + u];
            mw[u] = mw[kGroup + u];
        }
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
            w[u + 2] = nx[u];
            mw[u + 2] = nm[u];
        }
    }
}

// Work of the persistent filter is a set of items (tile chunk, column group).
// Chunks are dealt to kXcds sub-queues (chunk k to sub-queue k % kXcds); a
// workgroup starts on the sub-queue of its XCD (workgroups are dispatched
// round-robin, so block b runs on XCD b % kXcds) and steals from the others
// when it drains.  Inside a sub-queue items run chunk-major: all column groups
// of a chunk follow each other, so the chunk's packed words are fetched into
// that XCD's L2 once and reused by every group.  Groups are numbered longest
// motifs first.
//   implicit form (one stream, whole library): item j of sub-queue x is chunk
//     (j / ngroups) * kXcds + x, group j % ngroups;
//   list form (a batch of requests): sub-queue x holds
//     items[sub_begin[x] .. sub_begin[x+1]), each (chunk, group), and a chunk
//     is (first tile, end tile, request); groups, fragments and per-slot
//     tables then describe the batch's own column blocks (each request's
//     selected columns packed into blocks of their own).
constexpr int kXcds = 8;
struct WorkQueue {
    uint32_t ntiles;
    uint32_t nchunks;
    uint32_t chunk_tiles;
    uint32_t ngroups;
};

// Batch descriptor in device memory (list form).  The host writes it before
// each launch, so one captured graph serves every batch that fits the
// buffers it was captured with.
// A batch hit as the host receives it: position within its request, the
// caller's column and the score, 12 bytes (a quarter less read-back than
// the 16-byte blamm_hit, which the host rebuilds on demand).
struct GroupedHit {
    uint32_t pos;
    uint32_t col;
    float score;
};

struct BatchDesc {
    uint32_t sub_begin[kXcds + 1];
    uint32_t nreq;
    uint32_t nbases;
    uint32_t nruns;
    uint32_t nblk;          // column blocks of the batch
    uint32_t pad[3];
};

struct ListArgs {
    const BatchDesc* desc;
    const uint2* items;     // (chunk, group)
    const uint4* chunks;    // (t0, t1, request, 0)
};

// One item: tiles [t0, t1) against one column group.
template <bool BF, bool MASK, bool LO>
__device__ __forceinline__ void filter_item(const FilterArgs& a, Stage& st, uint32_t grp,
                                            uint32_t t0, uint32_t t1, uint32_t lane)
{
    const uint32_t blk0 = grp & 0xFFFFFu, nch = (grp >> 20) & 63u, nb = grp >> 26;
    if constexpr (BF) {
        switch (nch * 8 + nb) {
#define BLAMM_BF(N, B) case N * 8 + B: bf16_range<N, B, MASK, LO>(a, st, blk0, t0, t1, lane); break;
        BLAMM_BF(1, 1) BLAMM_BF(2, 1) BLAMM_BF(3, 1) BLAMM_BF(4, 1)
        BLAMM_BF(5, 1) BLAMM_BF(6, 1) BLAMM_BF(7, 1) BLAMM_BF(8, 1)
        BLAMM_BF(1, 2) BLAMM_BF(2, 2) BLAMM_BF(3, 2) BLAMM_BF(4, 2)
        BLAMM_BF(5, 2) BLAMM_BF(6, 2) BLAMM_BF(7, 2) BLAMM_BF(8, 2)
        BLAMM_BF(9, 1) BLAMM_BF(10, 1) BLAMM_BF(11, 1) BLAMM_BF(12, 1)
        BLAMM_BF(13, 1) BLAMM_BF(14, 1) BLAMM_BF(15, 1) BLAMM_BF(16, 1)
#undef BLAMM_BF
        default: break;
        }
    } else {
        switch (nch * 8 + nb) {
#define BLAMM_F32(N, B) case N * 8 + B: fp32_range<N, B, MASK>(a, st, blk0, t0, t1, lane); break;
        BLAMM_F32(1, 1) BLAMM_F32(2, 1) BLAMM_F32(3, 1) BLAMM_F32(4, 1)
        BLAMM_F32(5, 1) BLAMM_F32(6, 1) BLAMM_F32(7, 1) BLAMM_F32(8, 1)
        BLAMM_F32(1, 2) BLAMM_F32(2, 2) BLAMM_F32(3, 2) BLAMM_F32(4, 2)
        BLAMM_F32(5, 2) BLAMM_F32(6, 2) BLAMM_F32(7, 2) BLAMM_F32(8, 2)
        BLAMM_F32(1, 4) BLAMM_F32(2, 4) BLAMM_F32(3, 4) BLAMM_F32(4, 4)
        BLAMM_F32(5, 4) BLAMM_F32(6, 4) BLAMM_F32(7, 4) BLAMM_F32(8, 4)
        BLAMM_F32(9, 1) BLAMM_F32(10, 1) BLAMM_F32(11, 1) BLAMM_F32(12, 1)
        BLAMM_F32(13, 1) BLAMM_F32(14, 1) BLAMM_F32(15, 1) BLAMM_F32(16, 1)
#undef BLAMM_F32
        default: break;
        }
    }
}

// One launch for the whole work set: every resident wave pulls items until
// all sub-queues are empty (one atomic per item).
template <bool BF, bool MASK, bool LIST, bool LO>
__device__ __forceinline__ void filter_body(const FilterArgs& a, const WorkQueue& q, const ListArgs& l)
{
    const uint32_t lane = threadIdx.x & 63;
    __shared__ uint4 stage_lds[kWavesPerBlock][kStage];
    // lazy-lo form: per-wave lo fragments (kLoChunks chunk fragments x 64 lanes)
    __shared__ uint4 lo_lds[LO ? kWavesPerBlock * kLoChunks * 64 : 1];
    Stage st{stage_lds[threadIdx.x >> 6], 0u, LO ? lo_lds + (threadIdx.x >> 6) * kLoChunks * 64 : lo_lds};
    const uint32_t home = blockIdx.x % kXcds;
    for (uint32_t hop = 0; hop < (uint32_t)kXcds; ++hop) {
        const uint32_t x = (home + hop) % kXcds;
        uint32_t items = 0, base = 0;
        if constexpr (LIST) {
            base = __builtin_amdgcn_readfirstlane(l.desc->sub_begin[x]);
            items = __builtin_amdgcn_readfirstlane(l.desc->sub_begin[x + 1]) - base;
        } else {
            if (x >= q.nchunks) continue;
            items = ((q.nchunks - x + kXcds - 1) / kXcds) * q.ngroups;
        }
        for (;;) {
            uint32_t j = 0;
            if (lane == 0) j = atomicAdd((unsigned int*)&a.counters[kCntQueue + x], 1u);
            j = __builtin_amdgcn_readfirstlane(__shfl(j, 0));
            if (j >= items) break;
            uint32_t grp, t0, t1;
            if constexpr (LIST) {
                const uint2 it = l.items[base + j];
                const uint4 ch = l.chunks[__builtin_amdgcn_readfirstlane(it.x)];
                grp = __builtin_amdgcn_readfirstlane(a.groups[__builtin_amdgcn_readfirstlane(it.y)]);
                t0 = __builtin_amdgcn_readfirstlane(ch.x);
                t1 = __builtin_amdgcn_readfirstlane(ch.y);
            } else {
                const uint32_t k = (j / q.ngroups) * kXcds + x;
                grp = __builtin_amdgcn_readfirstlane(a.groups[j % q.ngroups]);
                t0 = k * q.chunk_tiles;
                t1 = min(q.ntiles, t0 + q.chunk_tiles);
            }
            filter_item<BF, MASK, LO>(a, st, grp, t0, t1, lane);
        }
    }
    stage_flush(a, st);
}

// The bf16 filter in its hi-only form (BF) and the FP32 filter.  The bf16
// form fits in 168 VGPRs for three waves per SIMD (a few values spilled,
// reloaded only where staged records are flushed and between work items;
// 1.6% faster than two waves at 180 VGPRs); the FP32 form keeps its 210-220.
template <bool BF, bool MASK, bool LIST>
__global__ __launch_bounds__(kThreads, BF ? 3 : 1) void filter_persistent_kernel(FilterArgs a, WorkQueue q, ListArgs l)
{
    filter_body<BF, MASK, LIST, false>(a, q, l);
}

// The bf16 filter in its lazy-lo form (80 KiB of LDS per block: two blocks
// per CU, so two waves per SIMD and at most 256 registers).
template <bool MASK, bool LIST>
__global__ __launch_bounds__(kThreads, 2) void filter_lo_kernel(FilterArgs a, WorkQueue q, ListArgs l)
{
    filter_body<true, MASK, LIST, true>(a, q, l);
}

struct FinalizeArgs {
    const uint4* cand;
    const unsigned long long* counters;   // kCntCand read, kCntHits/kCntBand written
    unsigned long long* counters_rw;
    uint64_t cand_cap;
    const uint64_t* packed;
    const uint64_t* lcmask;               // may be null
    uint32_t nbases;
    const uint64_t* run_starts;
    uint32_t nruns;
    const uint64_t* run_bits;             // bit i set: a run starts at base i (null: search run_starts)
    const int* slot_len;
    const float* slot_w;                  // [slot][kMaxLen][4] FP32 weights
    const float* slot_w2;                 // [slot][kMaxLen / 2][16] FP32 sums of position pairs
    const float* slot_thr;                // the threshold per slot (list form: per batch slot)
    const double* slot_seq;               // sequential-FP32 band per slot
    const uint32_t* slot_col;             // caller's column index
    void* hits;                           // blamm_hit[]
    uint64_t hit_cap;
    const BatchDesc* desc;                // list form: nbases, nruns from here
    const uint32_t* req_end;              // list form: end of each request's bases
    const uint32_t* slot_map;             // list form: batch slot -> library slot
    const uint32_t* blk_req;              // list form: request of each batch block
    uint32_t* req_count;                  // list form: hits per request (cleared before the launch)
};

struct DevHit {
    uint64_t pos;
    uint32_t col;
    float score;
};

// Requests counted in LDS by finalize and the grouping kernels; batches with
// more requests count with global atomics.
constexpr int kGroupLds = 1024;
constexpr int kGroupThreads = 256;

// For the lanes with `on` set, add one per lane to counter[key] with one
// atomic per distinct key in the wave, and return the counter's value before
// the wave's addition plus the lane's rank among the lanes sharing its key
// (so lanes of one key get consecutive values).  The hits of one wave come
// from the records of one filter item, so there is usually a single key.
template <typename T>
__device__ __forceinline__ T wave_add_by_key(T* counter, uint32_t key, bool on)
{
    const uint32_t lane = threadIdx.x & 63;
    uint64_t todo = __ballot(on);
    T mine = 0;
    while (todo) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(todo);
        const uint32_t k = (uint32_t)__shfl((int)key, (int)leader);
        const bool mem = on && key == k;
        const uint64_t same = __ballot(mem);
        T base = 0;
        if (lane == leader) base = atomicAdd(&counter[k], (T)__popcll(same));
        base = __shfl(base, (int)leader);
        if (mem)
            mine = base + (T)__builtin_amdgcn_mbcnt_hi((uint32_t)(same >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
        todo &= ~same;
    }
    return mine;
}

// Hits staged per wave before one atomic reserves their range: every flush
// is an atomic on one counter, and at ~30M hits per scan those atomics bound
// the kernel (512: 1.37 ms, 256: 1.8 ms at 100 Mbp x 2,000 columns; 1,024
// halves the resident waves and is slower).
constexpr int kHitStage = 512;

// Windows a filter record stands for: record window q covers q + off(r).
template <int NR>
__device__ __forceinline__ uint32_t record_offset(int r)
{
    return NR == 16 ? (uint32_t)(8 * (r >> 2) + (r & 3)) : (uint32_t)r;
}

// Exact decision for every window of every filter record: the run-boundary
// rule (a window must end at or before the next run start, pwmscan.cpp:122-126),
// then the left-to-right FP32 sum; a window whose FP32 sum lies within its
// proven band of the threshold is rescored in FP64 (the only FP64 step).
// One thread per record; its NR windows share the column's weights and a
// 128-bit window of packed bases that is shifted one position per step.
// List form: records carry slot | request << 16, thresholds are per request,
// a window must also end inside its request, and hits report
// pos = request << 32 | batch position.
template <int NR, bool LIST>
__global__ __launch_bounds__(kThreads) void finalize_kernel(FinalizeArgs a)
{
    uint32_t nbases = a.nbases, nruns = a.nruns;
    if constexpr (LIST) {
        nbases = __builtin_amdgcn_readfirstlane(a.desc->nbases);
        nruns = __builtin_amdgcn_readfirstlane(a.desc->nruns);
    }
    const uint64_t ncand_raw = a.counters[kCntCand];
    const uint64_t ncand = ncand_raw < a.cand_cap ? ncand_raw : a.cand_cap;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = (uint64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    DevHit* hits = (DevHit*)a.hits;
    // hits are staged per wave in LDS and reserved with one atomic per flush
    __shared__ DevHit hit_lds[kWavesPerBlock][kHitStage];
    DevHit* sbuf = hit_lds[threadIdx.x >> 6];
    uint32_t sn = 0;
    // list form: hits per request, for the grouping that follows
    __shared__ uint32_t req_lds[LIST ? kGroupLds : 1];
    uint32_t nreq = 0;
    bool req_in_lds = false;
    if constexpr (LIST) {
        nreq = __builtin_amdgcn_readfirstlane(a.desc->nreq);
        req_in_lds = nreq <= (uint32_t)kGroupLds;
        if (req_in_lds)
            for (uint32_t i = threadIdx.x; i < nreq; i += kThreads) req_lds[i] = 0u;
        __syncthreads();
    }
    auto flush = [&]() {
        if (sn == 0) return;
        unsigned long long at0 = 0;
        if (lane == 0) at0 = atomicAdd(&a.counters_rw[kCntHits], (unsigned long long)sn);
        at0 = __shfl(at0, 0);
        for (uint32_t i = lane; i < sn; i += 64)
            if ((uint64_t)at0 + i < a.hit_cap) hits[at0 + i] = sbuf[i];
        sn = 0;
    };
    for (uint64_t base = wave * 64; base < ncand; base += nw * 64) {
        const uint64_t idx = base + lane;
        bool have = idx < ncand;
        // slot: threshold slot of the record; lib: library slot holding the
        // column's weights (the same slot outside the list form)
        uint32_t q0 = 0, slot = 0, lib = 0, req = 0, qend = nbases, rows = 0;
        if (have) {
            const uint4 cd = a.cand[idx];
            q0 = cd.x;
            slot = lib = cd.y;
            rows = cd.z;
            if constexpr (LIST) {
                constexpr uint32_t cpb = NR == 16 ? kBf16Cols : kFp32Cols;
                req = a.blk_req[slot / cpb];
                lib = a.slot_map[slot];
                qend = min(nbases, a.req_end[req]);
                if (lib == kNoSlot) have = false, lib = 0;   // empty slots never pass the filter
            }
        }
        if (!have) rows = 0;
        const int L = have ? a.slot_len[lib] : 0;
        const float* w = a.slot_w + (size_t)lib * kMaxLen * 4;
        const float* w2 = a.slot_w2 + (size_t)lib * (kMaxLen / 2) * 16;
        uint32_t nb1 = 0;
        if (have && !a.run_bits) {
            // first run start after q0
            uint32_t l = 0, h = nruns;
            while (l < h) {
                const uint32_t mid = (l + h) >> 1;
                if (a.run_starts[mid] <= q0) l = mid + 1; else h = mid;
            }
            nb1 = l < nruns ? (uint32_t)a.run_starts[l] : nbases;
        }
        const float thr = have ? a.slot_thr[slot] : 0.f;
        const double seq = have ? a.slot_seq[lib] : 0.0;
        const float shi = thr_plus(thr, seq);
        const float slo = thr_minus(thr, seq);
        const uint32_t col = have ? a.slot_col[lib] : 0u;
        // every lane takes its next marked window until no lane has one left
        // (uniform loop: the hit staging below stays wave-wide)
        while (__ballot(rows != 0u)) {
            const bool act = rows != 0u;
            const uint32_t r = act ? (uint32_t)__builtin_ctz(rows) : 0u;
            rows &= rows - 1u;
            const uint32_t q = q0 + record_offset<NR>(r);
            // bases 0..31 of the window in bits, 32..63 in bits2 (loaded for
            // motifs longer than 32 positions only)
            uint64_t bits = 0, bits2 = 0, mbits = 0;
            if (act) {
                bits = load_bits64(a.packed, 2ull * q);
                if (L > 32) bits2 = load_bits64(a.packed, 2ull * q + 64);
                if (a.lcmask) mbits = load_bits64(a.lcmask, q);
            }
            // FP32 sum of the window over its positions in order.  Without a
            // lowercase base in the window it adds one table entry per pair of
            // positions (the FP32 sum of the pair's two weights, indexed by the
            // pair's 4 bits of packed bases; a zero weight row completes an odd
            // length), so each lane makes half the dependent weight loads; with
            // one it adds the weights of the uppercase positions one by one.
            // Either way |sum - exact| <= gamma_(L-1) sum_p max_b |w|, which
            // the sequential band (slot_seq) covers.  Four loads at a time so
            // they overlap (rows up to kMaxLen exist for every slot).
            float s = 0.f;
            const int Lw = act ? L : 0;
            const uint64_t lenmask = L >= 64 ? ~0ull : (1ull << L) - 1ull;
            if (!(mbits & lenmask)) {
                const int np = (Lw + 1) >> 1;
                for (int q0 = 0; q0 < np; q0 += 4) {
                    const uint64_t src = q0 < 16 ? bits >> (4 * q0) : bits2 >> (4 * (q0 - 16));
                    float x[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) x[u] = w2[16 * (q0 + u) + (uint32_t)((src >> (4 * u)) & 15u)];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (q0 + u < np) s += x[u];
                }
            } else
            for (int p0 = 0; p0 < Lw; p0 += 4) {
                const uint64_t src = p0 < 32 ? bits >> (2 * p0) : bits2 >> (2 * (p0 - 32));
                float x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = w[4 * (p0 + u) + (uint32_t)((src >> (2 * u)) & 3u)];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (p0 + u < Lw && !((mbits >> (p0 + u)) & 1)) s += x[u];
            }
            bool hit = false;
            float score = s;
            if (act && q < qend && s >= slo) {
                // run-boundary rule: no run starts at q + 1 .. q + L - 1 and
                // the window ends inside the stream (and its request)
                bool inside = (uint64_t)q + (uint64_t)L <= (uint64_t)min(nbases, qend);
                if (a.run_bits) {
                    const uint64_t starts = load_bits64(a.run_bits, (uint64_t)q + 1);
                    inside = inside && (starts & ((1ull << (L - 1)) - 1ull)) == 0;
                } else {
                    // next run start after q
                    uint32_t nb = nb1;
                    if (q >= nb1) {
                        uint32_t l = 0, h = nruns;
                        while (l < h) {
                            const uint32_t mid = (l + h) >> 1;
                            if (a.run_starts[mid] <= q) l = mid + 1; else h = mid;
                        }
                        nb = l < nruns ? (uint32_t)a.run_starts[l] : nbases;
                    }
                    inside = inside && (uint64_t)q + (uint64_t)L <= (uint64_t)nb;
                }
                if (inside) {
                    if (s >= shi) {
                        hit = true;
                    } else {
                        // within the FP32 band: FP64 left-to-right sum
                        double d = 0.0;
                        for (int p = 0; p < L; ++p) {
                            if ((mbits >> p) & 1) continue;
                            const uint64_t src = p < 32 ? bits >> (2 * p) : bits2 >> (2 * (p - 32));
                            d += (double)w[4 * p + (src & 3)];
                        }
                        atomicAdd(&a.counters_rw[kCntBand], 1ull);
                        if (d >= (double)thr) {
                            hit = true;
                            score = (float)d;
                        }
                    }
                }
            }
            const uint64_t m = __ballot(hit);
            if (m) {
                if constexpr (LIST) {
                    if (req_in_lds) (void)wave_add_by_key(req_lds, req, hit);
                    else (void)wave_add_by_key(a.req_count, req, hit);
                }
                if (sn > (uint32_t)(kHitStage - 64)) flush();
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (hit) {
                    DevHit hh;
                    hh.pos = LIST ? ((uint64_t)req << 32 | q) : (uint64_t)q;
                    hh.col = col;
                    hh.score = score;
                    sbuf[sn + rank] = hh;
                }
                sn += (uint32_t)__popcll(m);
            }
        }
    }
    flush();
    if constexpr (LIST) {
        if (req_in_lds) {
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < nreq; i += kThreads)
                if (req_lds[i]) atomicAdd(&a.req_count[i], req_lds[i]);
        }
    }
}

#ifndef BLAMM_INSTANTIATE   // small kernels: launched from the main unit only
static __global__ void zero_counters_kernel(unsigned long long* c)
{
    if (threadIdx.x < kCntSlots) c[threadIdx.x] = 0ull;
}

// Run-start bits for finalize_kernel: bit i is set when a run starts at base
// i (0 < i < nbases).  The clear kernel zeroes words [0, nbases / 64 + 2)
// (what load_bits64 may read below q + 1 <= nbases); the set kernel then
// sets one bit per run start.  List form: sizes from the batch descriptor.
static __global__ __launch_bounds__(256) void run_bits_clear_kernel(uint64_t* bits, uint32_t nbases,
                                                                    const BatchDesc* desc)
{
    const uint32_t n = desc ? desc->nbases : nbases;
    const uint64_t words = (uint64_t)n / 64 + 2;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (uint64_t)gridDim.x * 256)
        bits[i] = 0ull;
}

static __global__ __launch_bounds__(256) void run_bits_set_kernel(uint64_t* bits, const uint64_t* starts,
                                                                  uint32_t nruns, uint32_t nbases,
                                                                  const BatchDesc* desc)
{
    const uint32_t n = desc ? desc->nbases : nbases;
    const uint32_t nr = desc ? desc->nruns : nruns;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < nr; i += gridDim.x * 256) {
        const uint64_t st = starts[i];
        if (st > 0 && st < n) atomicOr((unsigned long long*)&bits[st >> 6], 1ull << (st & 63));
    }
}
#endif

// ---------------------------------------------------------------------------
// Hits of a batch grouped by request on the device, so the host receives each
// request's hits contiguously with request-local positions: finalize_kernel
// counts the hits per request, group_scan_kernel turns the counts into
// offsets, group_scatter_kernel places the hits (12-byte GroupedHit).  Range
// reservations are taken per block and request in LDS when the batch has at
// most kGroupLds requests (one global atomic per block and request).
// ---------------------------------------------------------------------------

struct GroupArgs {
    const BatchDesc* desc;
    const unsigned long long* counters;   // kCntHits
    uint64_t cap;                         // finalize output capacity
    uint64_t out_cap;                     // grouped output capacity
    const DevHit* in;                     // finalize output, pos = request << 32 | batch position
    GroupedHit* out;                      // grouped, request-local positions
    uint32_t* req_count;                  // [nreq] hits per request, cleared by group_clear_kernel
    unsigned long long* req_off;          // [nreq + 1] exclusive offsets
    unsigned long long* cursor;           // [nreq] next free slot per request
    const uint32_t* req_start;            // [nreq] first batch position of each request
};

#ifndef BLAMM_INSTANTIATE
static __global__ __launch_bounds__(kGroupThreads) void group_clear_kernel(GroupArgs g)
{
    const uint32_t nreq = g.desc->nreq;
    for (uint32_t i = blockIdx.x * kGroupThreads + threadIdx.x; i < nreq; i += gridDim.x * kGroupThreads)
        g.req_count[i] = 0u;
}

__device__ __forceinline__ uint64_t group_nhits(const GroupArgs& g)
{
    const uint64_t n = g.counters[kCntHits];
    return n < g.cap ? n : g.cap;
}

// One block: exclusive prefix over the request counts.
static __global__ __launch_bounds__(1024) void group_scan_kernel(GroupArgs g)
{
    __shared__ unsigned long long part[1024];
    const uint32_t nreq = g.desc->nreq;
    const uint32_t per = (nreq + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = min(nreq, b + per);
    unsigned long long s = 0;
    for (uint32_t i = b; i < e; ++i) s += g.req_count[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const unsigned long long v = threadIdx.x >= d ? part[threadIdx.x - d] : 0ull;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long off = part[threadIdx.x] - s;   // exclusive
    for (uint32_t i = b; i < e; ++i) {
        g.req_off[i] = off;
        g.cursor[i] = off;
        off += g.req_count[i];
    }
    if (threadIdx.x == 1023) g.req_off[nreq] = part[1023];
}

// Each block places a contiguous range of hits: counts per request in LDS,
// one global range per request and block, then ranks within the range.  The
// per-request counts include hits that finalize could not store (capacity
// overflow: the batch is then run again), so every write is bounds-checked.
// (Writing the grouped hits straight into pinned host memory instead of
// device memory measured slower: the kernel then waits on PCIe.)
static __global__ __launch_bounds__(kGroupThreads) void group_scatter_kernel(GroupArgs g)
{
    __shared__ uint32_t cnt[kGroupLds];
    __shared__ unsigned long long base[kGroupLds];
    const uint32_t nreq = g.desc->nreq;
    const uint64_t n = group_nhits(g);
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t b0 = (uint64_t)blockIdx.x * per;
    const uint64_t b1 = b0 + per < n ? b0 + per : n;
    // the loops run while any lane of the wave has a hit left (the key
    // aggregation is wave-wide)
    const uint64_t wb = b0 + (threadIdx.x & ~63u);
    auto place = [&](const DevHit& h, uint32_t r, uint64_t at) {
        if (at < g.out_cap) {
            GroupedHit o;
            o.pos = (uint32_t)(h.pos & 0xFFFFFFFFull) - g.req_start[r];
            o.col = h.col;
            o.score = h.score;
            g.out[at] = o;
        }
    };
    if (nreq > (uint32_t)kGroupLds) {
        for (uint64_t w = wb; w < b1; w += kGroupThreads) {
            const uint64_t i = w + (threadIdx.x & 63);
            const bool on = i < b1;
            DevHit h{};
            if (on) h = g.in[i];
            const uint32_t r = on ? (uint32_t)(h.pos >> 32) : 0u;
            const unsigned long long at = wave_add_by_key(g.cursor, r, on);
            if (on) place(h, r, at);
        }
        return;
    }
    for (uint32_t i = threadIdx.x; i < nreq; i += kGroupThreads) cnt[i] = 0u;
    __syncthreads();
    for (uint64_t w = wb; w < b1; w += kGroupThreads) {
        const uint64_t i = w + (threadIdx.x & 63);
        const bool on = i < b1;
        const uint32_t r = on ? (uint32_t)(g.in[i].pos >> 32) : 0u;
        (void)wave_add_by_key(cnt, r, on);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nreq; i += kGroupThreads) {
        base[i] = cnt[i] ? atomicAdd(&g.cursor[i], (unsigned long long)cnt[i]) : 0ull;
        cnt[i] = 0u;
    }
    __syncthreads();
    for (uint64_t w = wb; w < b1; w += kGroupThreads) {
        const uint64_t i = w + (threadIdx.x & 63);
        const bool on = i < b1;
        DevHit h{};
        if (on) h = g.in[i];
        const uint32_t r = on ? (uint32_t)(h.pos >> 32) : 0u;
        const uint32_t k = wave_add_by_key(cnt, r, on);
        if (on) place(h, r, base[r] + k);
    }
}
#endif  // BLAMM_INSTANTIATE
static_assert(kCntQueue + kXcds <= kCntSlots, "queue counters");

// Column blocks of a batch: batch slot s takes the fragments of library slot
// slot_map[s] (an empty slot gets zero weights).  Bands are the library's
// bands for the longest block (valid for a column in any block).
struct GatherArgs {
    const BatchDesc* desc;                // nblk
    const uint32_t* slot_map;
    const uint4* lib_a;                   // bf16: [block][chunk][lane] hi / fp32: weights as uint4 view unused
    const uint4* lib_b;
    const float* lib_w;                   // fp32: [block][position][lane]
    const double* lib_band;               // per library slot, any block
    const double* lib_band_hi;            // per library slot, any block (hi-only stage)
    uint4* out_a;
    uint4* out_b;
    float* out_w;
    double* out_band;
    double* out_band_hi;
};

template <bool BF>
__global__ __launch_bounds__(256) void gather_kernel(GatherArgs g)
{
    const uint32_t nblk = g.desc->nblk;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    const uint64_t t0 = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    constexpr uint32_t cpb = BF ? kBf16Cols : kFp32Cols;
    if constexpr (BF) {
        const uint64_t total = (uint64_t)nblk * kMaxChunks * 64;
        for (uint64_t i = t0; i < total; i += stride) {
            const uint32_t l = (uint32_t)(i & 63), j = (uint32_t)((i >> 6) % kMaxChunks);
            const uint32_t b = (uint32_t)(i / (64 * kMaxChunks));
            const uint32_t lib = g.slot_map[b * cpb + (l & 31)];
            uint4 va = make_uint4(0, 0, 0, 0), vb = va;
            if (lib != kNoSlot) {
                const size_t src = ((size_t)(lib / cpb) * kMaxChunks + j) * 64 + (lib % cpb) + 32 * (l >> 5);
                va = g.lib_a[src];
                vb = g.lib_b[src];
            }
            g.out_a[i] = va;
            g.out_b[i] = vb;
        }
    } else {
        const uint64_t total = (uint64_t)nblk * kMaxLen * 64;
        for (uint64_t i = t0; i < total; i += stride) {
            const uint32_t l = (uint32_t)(i & 63), p = (uint32_t)((i >> 6) % kMaxLen);
            const uint32_t b = (uint32_t)(i / (64 * kMaxLen));
            const uint32_t lib = g.slot_map[b * cpb + (l & 15)];
            float v = 0.f;
            if (lib != kNoSlot) v = g.lib_w[((size_t)(lib / cpb) * kMaxLen + p) * 64 + (lib % cpb) + 16 * (l >> 4)];
            g.out_w[i] = v;
        }
    }
    for (uint64_t s = t0; s < (uint64_t)nblk * cpb; s += stride) {
        const uint32_t lib = g.slot_map[s];
        g.out_band[s] = lib != kNoSlot ? g.lib_band[lib] : 0.0;
        g.out_band_hi[s] = lib != kNoSlot ? g.lib_band_hi[lib] : 0.0;
    }
}

}  // namespace blamm
