// 2D multi-spin-coded Metropolis half sweep (64 spins per uint64 word) and its exact
// observables, for gfx950.
//
// Layout (src/pyising/lattice.py:12-13 of the reference): colour c is a row-major (rows, nw)
// uint64 array, nw = Lx / 128; bit b of word x of row y is colour site j = 64 x + b, 1 = +1.
// Row y's colour-c sites sit at x_full = 2 j + p with p = (y + c) & 1, so the horizontal
// neighbour in the other colour is j + 1 when p = 1 and j - 1 when p = 0
// (src/pyising/backends/numpy_backend.py:103-110).
//
// Decision (src/pyising/backends/numpy_backend.py:54-83): with x_k = own ^ nb_k, lanes with
// >= 2 anti-aligned neighbours flip (a2); lanes with exactly one (a1) or none (a0) flip iff
// u < thr4 (thr8), where u's bit 31 - k is bit b of random level k, levels 2l and 2l + 1 coming
// from Philox4x32-10((u32)w, t, colour, 1 + l; k0, k1) as r0 | r1 << 32 and r2 | r3 << 32
// (src/pyising/philox.py:67-78). The comparison is decided MSB-first and stops once no lane of
// the word is undecided, so the number of Philox calls per word is data dependent (about 3.8
// on average at T_c, up to 16). Any lane may evaluate any (word, level) call: the stream is
// counter-based, which is what lets the kernels below regroup work.
//
// Cost model (measured with SQ counters on MI355X): Philox4x32-10 is 36 VALU ops (philox_lane:
// the wave-uniform halves of rounds 0-2 run on the scalar unit), every op including
// v_mad_u64_u32 at the full wave64 rate, so the kernel is VALU-issue bound and its time is
// (lane-slots per word) x (ops per slot) plus a per-launch cost (dispatch, ramp, tail) of a few
// microseconds. What keeps both factors low:
//  * the level index lp is wave-uniform everywhere, so the threshold bits of a level are
//    known to the whole wave: the host packs the 4-bit case code of every level pair into one
//    64-bit argument and each call costs one scalar switch into straight-line v_bitop3 code
//    (0-3 ops per 32-bit half and level) instead of a masked general update;
//  * loops are wave-uniform (no per-lane exit), so there is no exec-mask bookkeeping; a
//    decided lane's updates are no-ops;
//  * classification is six v_bitop3/XOR ops per 32-bit half; when a wave's 64 words lie in
//    one row (Lx % 8192 == 0) its row pointers are scalar and the horizontal boundary bit
//    moves between lanes by a DPP wavefront shift;
//  * queue compaction (mode 0): `cap` Philox calls per word unrolled (computed before the spin
//    loads are consumed, hiding their latency), then the ~19% of words still undecided run
//    one call per round packed into full waves, so lockstep no longer pays the slowest of 64
//    words; streaming stores keep the new spins out of the kernel-boundary L2 write-back.
#include "msc_core.hpp"

#include <atomic>
#include <mutex>
#include <type_traits>

// ---------------------------------------------------------------------------- kernels ----
// stats (optional, diagnostics only): [0] += Philox calls the words needed, [1] += lane-slots
// the SIMDs executed (64 x wave iterations), [2] += words.

// Lockstep: one word per lane, the wave pays the maximum call count of its 64 words.
// LAZY = false is the eager ablation (all 16 calls, src/pyising/backends/cuda/ising.cu:135).
template <bool LAZY, bool STATS, int LAY>
__global__ __launch_bounds__(256) void msc_half_lockstep(MscArgs a, u64 codes,
                                                        unsigned long long* stats)
{
    const u32 w = xcd_band_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool valid = w < a.n_words;  // wave-uniform unless LAY == MSC_GENERAL
    MscState s = {};
    if (valid) s = msc_classify_wave<LAY>(a, w);
    u32 calls = 0, slots = 0;
    msc_run<LAZY, STATS>(a, codes, w, 0u, 16u, s, calls, slots);
    if (valid) a.own[w] = (u64)s.pwl | ((u64)s.pwh << 32);
    if (STATS) msc_stats(stats, calls, slots, valid ? 1u : 0u);
}

// Queue compaction (the default): one workgroup per W = 256 * R words, CAP unrolled lockstep
// calls, then rounds of one call per packed survivor (msc_queue_tile); streaming stores.
// R = 2 uses 14 KB of LDS and at most 64 VGPRs: 8 workgroups (32 waves) per CU. R = 2 tiles
// run phase A's level-pair switches once for both rounds (JOINT), which the R = 4 tiles'
// register budget does not allow.
static_assert(sizeof(MscQueueLds<1>) <= sizeof(MscQueueLds<2>), "graded tiles share the LDS");

// Graded tail: workgroups [0, n_big) take 256 R-word tiles (XCD band mapped); the ones after,
// dispatched last, take 256-word tiles (R = 2 only), so the final generation of workgroups
// finishes sooner and the launch's tail is shorter.
template <int R, int CAP, int LAY, bool STATS>
__global__ __launch_bounds__(256) void msc_half_queue(MscArgs a, u64 codes, u32 n_big,
                                                     unsigned long long* stats)
{
    __shared__ MscQueueLds<R> lds;
    using Leave = std::conditional_t<STATS, MscStay, MscLeave>;  // the statistics need every wave
    u32 calls = 0, slots = 0, words = 0;
    if (R == 1 || blockIdx.x < n_big) {
        const u32 base = xcd_band_block(blockIdx.x, n_big) * (256u * R);
        msc_queue_tile<R, CAP, LAY, STATS, false, true, false, Leave, R == 2>(
            a, codes, base, lds, calls, slots, words);
    } else if constexpr (R == 2) {
        const u32 base = n_big * 512u + (blockIdx.x - n_big) * 256u;
        msc_queue_tile<1, CAP, LAY, STATS, false, true, false, Leave>(
            a, codes, base, *reinterpret_cast<MscQueueLds<1>*>(&lds), calls, slots, words);
    }
    if (STATS) msc_stats(stats, calls, slots, words);
}

// ------------------------------------------------------------------------ observables ----
// Colour-0 word s with other-colour neighbours (c, hz, up, dn): ones += popc(s) + popc(c),
// anti += popc(s ^ nb) over the four; M = 2 ones - N, E = 2 anti - 2 N
// (src/pyising/backends/observables_torch.py:85-111).
__global__ __launch_bounds__(256) void msc_obs_kernel(const u64* __restrict__ w0,
                                                      const u64* __restrict__ w1, u32 rows, u32 nw,
                                                      u32 n_words, FastDiv div_nw,
                                                      unsigned long long* partials)
{
    u64 ones = 0, anti = 0;
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 w = blockIdx.x * blockDim.x + threadIdx.x; w < n_words; w += stride) {
        const u32 y = fast_div(w, div_nw);
        const u32 x = w - y * nw;
        const u64 rbase = (u64)y * nw;
        const u32 p = y & 1u;
        const u64 s = w0[w], c = w1[w];
        const u32 xn = p ? ((x + 1 == nw) ? 0u : x + 1) : ((x == 0) ? nw - 1 : x - 1);
        const u64 nb = w1[rbase + xn];
        const u64 hz = p ? ((c >> 1) | (nb << 63)) : ((c << 1) | (nb >> 63));
        const u32 yu = (y + 1 == rows) ? 0u : y + 1;
        const u32 yd = (y == 0) ? rows - 1 : y - 1;
        const u64 up = w1[(u64)yu * nw + x], dn = w1[(u64)yd * nw + x];
        ones += __popcll(s) + __popcll(c);
        anti += __popcll(s ^ c) + __popcll(s ^ hz) + __popcll(s ^ up) + __popcll(s ^ dn);
    }
    block_sum2_u64(ones, anti);
    if (threadIdx.x == 0) {
        partials[2u * blockIdx.x] = ones;
        partials[2u * blockIdx.x + 1u] = anti;
    }
}

// ---------------------------------------------------------------------- host launchers ----
// Kernel modes: 0 = queue compaction (k = 10 * cap + R, cap in 3..6 unrolled calls, R in
// {2, 4} x 256 words per workgroup; default cap 4, R 2), 1 = lockstep, 2 = eager lockstep (all
// 16 calls per word, the msc-eager32 ablation; k unused for 1 and 2), 3 = by size: lockstep up
// to MSC_LOCKSTEP_MAX_WORDS words per colour, queue compaction with k above. A half sweep of a
// small lattice is one generation of short-lived workgroups, so its time is the workgroup's
// latency, which the lockstep tile (one word per lane, no queue phases) keeps lower; measured
// on MI355X at T_c (us per half sweep, lockstep / queue): 4.33 / 5.75 at 1024^2, 4.96 / 6.27
// at 4096^2 (2^17 words), 8.36 / 8.24 at 8192^2 (2^19), 22.6 / 19.4 at 16384^2.
constexpr u32 MSC_LOCKSTEP_MAX_WORDS = 1u << 18;
// The process default (mode << 8 | k): one atomic word, so that a launch on one host thread
// never reads a mode without its k while another thread calls ising_msc_config.
static std::atomic<int> g_msc_default{(3 << 8) | 42};

// fused sweeps (msc_fused.hip): 0 not applicable, 1 applicable, 2 applicable and the default's
// choice for this size
int msc_fused_applies(long long rows, int nw);
int msc_fused_sweeps(u64* w0, u64* w1, long long rows, int nw, u32 t0, long long n_sweeps, u32 k0,
                     u32 k1, u32 thr4, u32 thr8, int cap, u32* sync, u32* epoch,
                     unsigned long long* stats, hipStream_t st);

// the concrete mode (0, 1 or 2, or 4 for sweeps with a sync workspace) and k for a lattice of
// n_words words per colour; fused = msc_fused_applies() for the lattice, 0 where the call has
// no sync workspace
static void msc_resolve(int* mode, int* k, u32 n_words, int fused = 0)
{
    if (*mode < 0) {
        const int d = g_msc_default.load(std::memory_order_relaxed);
        *mode = d >> 8;
        *k = d & 0xFF;
    }
    if (*mode == 3) *mode = n_words <= MSC_LOCKSTEP_MAX_WORDS ? 1 : (fused == 2 ? 4 : 0);
    if (*mode == 4 && !fused) *mode = 0;  // a single half sweep: the same tiles, one launch
}

static int msc_check(const void* own, const void* other, long long rows, int nw)
{
    if (!own || !other) return ising_fail(ISING_EINVAL, "msc: null pointer");
    if (rows < 2 || (rows & 1) || nw < 1) return ising_fail(ISING_EINVAL, "msc: bad geometry");
    // the word loops step by at most 2^20 and must not wrap a 32-bit index
    if ((unsigned long long)rows * (unsigned long long)nw + (1ull << 20) > 0xFFFFFFFFull)
        return ising_fail(ISING_ERANGE, "msc: rows * Lx/128 must be < 2^32 words per colour");
    return ISING_OK;
}

static int msc_mode_ok(int mode, int k)
{
    if (mode < 0) return ISING_OK;  // process default
    if (mode > 4)
        return ising_fail(ISING_EINVAL, "msc: mode must be 0 (queue compaction), 1 (lockstep), 2 (eager), 3 (by size) or 4 (fused queue compaction)");
    if (mode == 0 || mode == 3) {
        const int cap = k / 10, rr = k % 10;
        if (cap < 3 || cap > 6 || (rr != 2 && rr != 4))
            return ising_fail(ISING_EINVAL, "msc: queue compaction k = 10 * cap + R with cap in 3..6, R in {2,4}");
    }
    if (mode == 4 && (k % 10 != 2 || k / 10 < 3 || k / 10 > 5))
        return ising_fail(ISING_EINVAL, "msc: fused queue compaction k = 10 * cap + 2 with cap in 3..5");
    return ISING_OK;
}

template <int R, int CAP, int LAY>
static void launch_queue3(u32 n_tiles, hipStream_t st, const MscArgs& a, u64 codes,
                          unsigned long long* stats)
{
    // R = 2: the last min(512, a quarter of the words / 256) tiles are 256-word tiles: about
    // two per CU end the launch (same-box timing at T_c, us per half sweep at 8192^2 / 16384^2:
    // 384 of them 8.83 / 19.63, 512 8.46 / 19.23, 768 8.47 / 19.42, 1024 8.47 / 19.52)
    u32 n_big = n_tiles, grid = n_tiles;
    if (R == 2) {
        u32 small = a.n_words / 1024u;  // 256-word tiles in a quarter of the words
        if (small > 512u) small = 512u;
        small &= ~1u;
        n_big = (a.n_words - small * 256u) / 512u;
        grid = n_big + (a.n_words - n_big * 512u + 255u) / 256u;
    }
    if (stats)
        hipLaunchKernelGGL((msc_half_queue<R, CAP, LAY, true>), dim3(grid), dim3(256), 0, st, a, codes, n_big, stats);
    else
        hipLaunchKernelGGL((msc_half_queue<R, CAP, LAY, false>), dim3(grid), dim3(256), 0, st, a, codes, n_big, stats);
}

template <int R, int CAP>
static void launch_queue2(u32 n_tiles, hipStream_t st, const MscArgs& a, u64 codes,
                          unsigned long long* stats)
{
    const int lay = msc_layout(a.nw, a.n_words);
    if (lay == MSC_ROWW) launch_queue3<R, CAP, MSC_ROWW>(n_tiles, st, a, codes, stats);
    else if (lay == MSC_SEGW) launch_queue3<R, CAP, MSC_SEGW>(n_tiles, st, a, codes, stats);
    else launch_queue3<R, CAP, MSC_GENERAL>(n_tiles, st, a, codes, stats);
}

template <int LAY>
static void launch_lockstep(dim3 grid, hipStream_t st, const MscArgs& a, u64 codes, bool lazy,
                            unsigned long long* stats)
{
    if (lazy) {
        if (stats) hipLaunchKernelGGL((msc_half_lockstep<true, true, LAY>), grid, dim3(256), 0, st, a, codes, stats);
        else hipLaunchKernelGGL((msc_half_lockstep<true, false, LAY>), grid, dim3(256), 0, st, a, codes, stats);
    } else {
        if (stats) hipLaunchKernelGGL((msc_half_lockstep<false, true, LAY>), grid, dim3(256), 0, st, a, codes, stats);
        else hipLaunchKernelGGL((msc_half_lockstep<false, false, LAY>), grid, dim3(256), 0, st, a, codes, stats);
    }
}

template <int R>
static void launch_queue(u32 n_tiles, hipStream_t st, const MscArgs& a, u64 codes, u32 cap,
                         unsigned long long* stats)
{
    if (cap == 3) launch_queue2<R, 3>(n_tiles, st, a, codes, stats);
    else if (cap == 4) launch_queue2<R, 4>(n_tiles, st, a, codes, stats);
    else if (cap == 5) launch_queue2<R, 5>(n_tiles, st, a, codes, stats);
    else launch_queue2<R, 6>(n_tiles, st, a, codes, stats);
}

static int msc_half_impl(u64* own, const u64* other, long long rows, int nw, u32 t, u32 color,
                         u32 k0, u32 k1, u32 thr4, u32 thr8, int mode, int k,
                         unsigned long long* stats, hipStream_t st)
{
    int rc = msc_check(own, other, rows, nw);
    if (rc) return rc;
    if (color > 1) return ising_fail(ISING_EINVAL, "msc_half: colour must be 0 or 1");
    if ((rc = msc_mode_ok(mode, k))) return rc;
    msc_resolve(&mode, &k, (u32)(rows * nw));
    MscArgs a;
    a.own = (gu64*)own;
    a.other = (const gu64*)other;
    a.rows = (u32)rows;
    a.nw = (u32)nw;
    a.n_words = (u32)(rows * nw);
    a.div_nw = make_fastdiv((u32)nw);
    a.t = t;
    a.color = color;
    a.k0 = k0;
    a.k1 = k1;
    const u64 codes = msc_level_codes(thr4, thr8);
    if (mode == 0) {
        const u32 cap = (u32)(k / 10), rr = (u32)(k % 10);
        const u32 n_tiles = (a.n_words + 256u * rr - 1u) / (256u * rr);
        if (rr == 2) launch_queue<2>(n_tiles, st, a, codes, cap, stats);
        else launch_queue<4>(n_tiles, st, a, codes, cap, stats);
    } else {
        const dim3 grid((a.n_words + 255u) / 256u);
        const int lay = msc_layout(a.nw, a.n_words);
        if (lay == MSC_ROWW) launch_lockstep<MSC_ROWW>(grid, st, a, codes, mode == 1, stats);
        else if (lay == MSC_SEGW) launch_lockstep<MSC_SEGW>(grid, st, a, codes, mode == 1, stats);
        else launch_lockstep<MSC_GENERAL>(grid, st, a, codes, mode == 1, stats);
    }
    ISING_LAUNCHED("msc_half");
    return ISING_OK;
}

static int msc_sweeps_impl(u64* w0, u64* w1, long long rows, int nw, u32 t0, long long n_sweeps,
                           u32 k0, u32 k1, u32 thr4, u32 thr8, int mode, int k,
                           unsigned long long* stats, hipStream_t st, u32* sync = nullptr,
                           u32* epoch = nullptr)
{
    if (n_sweeps < 0) return ising_fail(ISING_EINVAL, "msc_sweeps: negative sweep count");
    if (sync) {
        int rc = msc_check(w0, w1, rows, nw);
        if (rc) return rc;
        if ((rc = msc_mode_ok(mode, k))) return rc;
        const int fused = msc_fused_applies(rows, nw);
        if (mode == 4 && !fused)
            return ising_fail(ISING_EINVAL, "msc: the fused form needs Lx/128 in {64, 128, 256, 512}, a power-of-two number of 512-word tiles (>= 16) and round-robin or single-XCD placement");
        int m = mode, kk = k;
        msc_resolve(&m, &kk, (u32)(rows * nw), fused);
        if (m == 4) {
            if (n_sweeps == 0) return ISING_OK;
            const int cap = (mode == 4) ? kk / 10 : 4;
            return msc_fused_sweeps(w0, w1, rows, nw, t0, n_sweeps, k0, k1, thr4, thr8, cap, sync,
                                    epoch, stats, st);
        }
        mode = m;
        k = kk;
    }
    for (long long s = 0; s < n_sweeps; ++s) {
        const u32 t = t0 + (u32)s;
        int rc = msc_half_impl(w0, w1, rows, nw, t, 0u, k0, k1, thr4, thr8, mode, k, stats, st);
        if (rc) return rc;
        rc = msc_half_impl(w1, w0, rows, nw, t, 1u, k0, k1, thr4, thr8, mode, k, stats, st);
        if (rc) return rc;
    }
    return ISING_OK;
}

extern "C" {

int ising_msc_config(int mode, int k)
{
    if (mode < 0) return ising_fail(ISING_EINVAL, "msc_config: mode must be >= 0");
    int rc = msc_mode_ok(mode, k);
    if (rc) return rc;
    int d = g_msc_default.load(std::memory_order_relaxed);
    int next;
    do {  // modes 1 and 2 keep the compaction k set before
        next = (mode << 8) | ((mode == 0 || mode == 3) ? k : (d & 0xFF));
    } while (!g_msc_default.compare_exchange_weak(d, next, std::memory_order_relaxed));
    return ISING_OK;
}

int ising_msc_kernel(long long rows, int nw, int mode, int k, int* out_mode, int* out_k)
{
    if (!out_mode || !out_k) return ising_fail(ISING_EINVAL, "msc_kernel: null output");
    if (rows < 2 || nw < 1 || (unsigned long long)rows * (unsigned long long)nw > 0xFFFFFFFFull)
        return ising_fail(ISING_EINVAL, "msc_kernel: bad geometry");
    int rc = msc_mode_ok(mode, k);
    if (rc) return rc;
    msc_resolve(&mode, &k, (u32)(rows * nw), msc_fused_applies(rows, nw));
    *out_mode = mode;
    *out_k = (mode == 0 || mode == 4) ? (mode == 4 && k % 10 != 2 ? 42 : k) : 0;
    return ISING_OK;
}

int ising_msc_sweeps_sync(uint64_t* w0, uint64_t* w1, long long rows, int nw, uint32_t t0,
                          long long n_sweeps, uint32_t k0, uint32_t k1, uint32_t thr4,
                          uint32_t thr8, int mode, int k, uint32_t* d_sync, uint32_t* epoch,
                          unsigned long long* stats, void* stream)
{
    if (!d_sync || !epoch) return ising_fail(ISING_EINVAL, "msc_sweeps_sync: null workspace or epoch");
    return msc_sweeps_impl(w0, w1, rows, nw, t0, n_sweeps, k0, k1, thr4, thr8, mode, k, stats,
                           as_stream(stream), d_sync, epoch);
}

int ising_msc_half(uint64_t* own, const uint64_t* other, long long rows, int nw, uint32_t t,
                   uint32_t color, uint32_t k0, uint32_t k1, uint32_t thr4, uint32_t thr8,
                   unsigned long long* stats, void* stream)
{
    return msc_half_impl(own, other, rows, nw, t, color, k0, k1, thr4, thr8, -1, 0, stats,
                         as_stream(stream));
}

int ising_msc_half_ex(uint64_t* own, const uint64_t* other, long long rows, int nw, uint32_t t,
                      uint32_t color, uint32_t k0, uint32_t k1, uint32_t thr4, uint32_t thr8,
                      int mode, int k, unsigned long long* stats, void* stream)
{
    return msc_half_impl(own, other, rows, nw, t, color, k0, k1, thr4, thr8, mode, k, stats,
                         as_stream(stream));
}

int ising_msc_sweeps(uint64_t* w0, uint64_t* w1, long long rows, int nw, uint32_t t0,
                     long long n_sweeps, uint32_t k0, uint32_t k1, uint32_t thr4, uint32_t thr8,
                     unsigned long long* stats, void* stream)
{
    return msc_sweeps_impl(w0, w1, rows, nw, t0, n_sweeps, k0, k1, thr4, thr8, -1, 0, stats,
                           as_stream(stream));
}

int ising_msc_sweeps_ex(uint64_t* w0, uint64_t* w1, long long rows, int nw, uint32_t t0,
                        long long n_sweeps, uint32_t k0, uint32_t k1, uint32_t thr4,
                        uint32_t thr8, int mode, int k, unsigned long long* stats, void* stream)
{
    return msc_sweeps_impl(w0, w1, rows, nw, t0, n_sweeps, k0, k1, thr4, thr8, mode, k, stats,
                           as_stream(stream));
}

int ising_msc_obs(const uint64_t* w0, const uint64_t* w1, long long rows, int nw,
                  unsigned long long* scratch2, long long* out2, void* stream)
{
    int rc = msc_check(w0, w1, rows, nw);
    if (rc) return rc;
    if (!scratch2 || !out2) return ising_fail(ISING_EINVAL, "msc_obs: null pointer");
    hipStream_t st = as_stream(stream);
    const u64 n_words = (u64)rows * (u64)nw;
    const unsigned grid = obs_grid(n_words);
    hipLaunchKernelGGL(msc_obs_kernel, dim3(grid), dim3(256), 0, st, w0, w1, (u32)rows, (u32)nw,
                       (u32)n_words, make_fastdiv((u32)nw), scratch2);
    // M = 2 ones - N, E = 2 anti - 2 N
    const long long n_spins = (long long)n_words * 128;
    hipLaunchKernelGGL(obs_finalize, dim3(1), dim3(256), 0, st, scratch2, grid, out2, 2ll,
                       -n_spins, 2ll, -2 * n_spins);
    ISING_LAUNCHED("msc_obs_kernel");
    return ISING_OK;
}

}  // extern "C"

// ------------------------------------------------------ observables without allocation ----
// The partial sums of ising_msc_obs_static: one (ones, anti) pair per reduction workgroup, the
// layout of the caller scratch of ising_msc_obs (ISING_OBS_SCRATCH_BYTES, at most 1024
// workgroups; any lattice size fits, the reduction loops over grid strides). Each GPU holds its
// own copy of this array until the process ends. In host code `scratch` names a host-side
// stand-in, not device memory: host code uses the per-device address from
// hipGetSymbolAddress (static_scratch) instead, never the name.
static __device__ __align__(256) unsigned long long scratch[ISING_OBS_SCRATCH_BYTES / 8];

static unsigned long long* g_scratch_addr[64];  // per device: the address of `scratch`
static std::mutex g_scratch_mu;

// The current device's copy of `scratch`, looked up once per device (with it current).
static int static_scratch(unsigned long long** out)
{
    int dev = 0;
    ISING_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return ising_fail(ISING_ERANGE, "msc_obs_static: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* ptr = nullptr;
        ISING_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return ising_fail(ISING_EHIP, "msc_obs_static: no device address for the scratch array");
        g_scratch_addr[dev] = static_cast<unsigned long long*>(ptr);
    }
    *out = g_scratch_addr[dev];
    return ISING_OK;
}

extern "C" int ising_msc_obs_static(const uint64_t* w0, const uint64_t* w1, long long rows, int nw,
                                    long long* out2, void* stream)
{
    int rc = msc_check(w0, w1, rows, nw);
    if (rc) return rc;
    if (!out2) return ising_fail(ISING_EINVAL, "msc_obs_static: null output");
    unsigned long long* ptr = nullptr;
    if ((rc = static_scratch(&ptr))) return rc;
    return ising_msc_obs(w0, w1, rows, nw, ptr, out2, stream);
}
