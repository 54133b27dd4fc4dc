// Fused multi-spin-coded sweeps: many half sweeps in one launch, ordered by per-tile
// completion counters instead of kernel boundaries.
//
// A half sweep of colour c updates the words of c from the other colour's words in the same
// and the adjacent rows (src/pyising/backends/numpy_backend.py:103-124; ising.cu:102-150). With
// 512-word tiles that cover whole rows (Lx/128 = nw in {64, 128, 256, 512}), tile i of half
// sweep h reads exactly what tiles i-1, i, i+1 (mod the tile count) wrote in half sweep h-1,
// and it overwrites what those tiles read there. So tile i of half h may start as soon as
// those three tiles of half h-1 are done, which is what a kernel boundary waits for
// globally. The launch holds the tiles of H half sweeps in order (workgroup g = h * T + j), each
// workgroup waits on the three counters, runs the unchanged queue-compaction tile
// (msc_queue_tile: same words, same Philox counters, so the result is bit-identical to H
// separate launches) and publishes its counter. Waited-on tiles always have lower workgroup
// ids, and each XCD dispatches its workgroups in id order, so every wait is for a workgroup
// that is already resident or finished: the oldest waiting workgroup can always proceed.
// Every wait is bounded anyway: past FUSED_SPIN_LIMIT polls the workgroup sets a failure bit
// and exits without touching the lattice, later waiters see the bit and exit at once, the
// kernel ends normally and the host reports the failure.
//
// What the launch saves: the gap between back-to-back kernels, the ramp of a fresh grid and
// the drain of its last generation of workgroups, once per launch instead of once per half
// sweep.
//
// Memory visibility. MI355X dispatches the workgroups of a launch round-robin over its 8 XCDs,
// each with its own L2; the starting XCD carries over from earlier dispatches (measured: the
// rotation differs from launch to launch, and is exact within one). A workgroup therefore
// reads its XCD (HW_REG_XCC_ID) and takes the tile of that XCD's band: the tiles of a half
// sweep form eight bands of consecutive rows, band x on XCD x, so a tile and its neighbours
// share one L2 except at band edges, whatever the rotation. Inside an XCD the L2 is the
// coherence point: the producer waits for its stores to be acknowledged (s_waitcnt) before
// publishing, and the consumer reads the tile's words with device-scope loads, which always
// miss in the CU's L1 (a CU may still hold a line it read in an earlier half sweep of the
// launch; measured: ordinary loads then return stale words, the workgroup-scope invalidate
// buffer_inv sc0 does not clear them, and buffer_inv sc1 on every tile also drops the L2 and
// costs 1.8x). A band's first and last tiles talk to the neighbouring XCDs and use the
// device-scope protocol (release: write back the L2, buffer_wbl2 sc1, costly but only 16 tiles
// per half sweep; acquire: buffer_inv sc1). Counters are device-scope atomics. Under any other
// placement an XCD would receive more or fewer than its band's tiles: a tile would go missing,
// which the waits (bounded) or the host's check of the counters after the launch reports.
#include "msc_core.hpp"
#include "placement.hpp"

#include <mutex>
#include <type_traits>
#include <vector>

// Sync workspace (u32 words): [0] failure bits (1: a wait gave up); [FUSED_HDR + i] half
// sweeps completed by tile i, counted from the workspace's zero state (mod 2^32).
constexpr u32 FUSED_HDR = 64;
constexpr u32 FUSED_TILE_WORDS = 512;
constexpr u32 FUSED_SPIN_LIMIT = 1u << 20;
constexpr u32 FUSED_FAIL_WAIT = 1u, FUSED_FAIL_COUNTERS = 2u;
// Half sweeps per launch: longer launches amortize the launch cost further but gain little
// beyond this; a launch must also hold fewer than 2^32 work-items.
constexpr u32 FUSED_MAX_HALVES = 64;

struct FusedArgs {
    u64* w0;
    u64* w1;
    u32 rows, nw, n_words;
    u32 t0, k0, k1;
    FastDiv div_nw;
    u32* sync;
    u32 epoch;      // value every tile counter holds when the launch starts
    u32 tiles_lg;   // tiles per half sweep = 2^tiles_lg (a multiple of 8)
    u32 per_band;   // tiles per XCD band
    u32 one_l2;     // 1: every workgroup runs on one XCD (one L2); 0: round-robin over 8
};

// Wave 0 waits until the tiles this one depends on (lanes 0-2: `dep`; `seen` = the value each
// polling lane loaded first) have completed `target` half sweeps. The tile's own loads are
// device-scope loads that miss in the CU's L1, so no cache invalidation is needed where the
// producers share this XCD's L2; device_scope (band edges) adds the acquire fence that makes
// the other XCDs' released writes visible. Returns false if the wait gave up (bounded) or
// another workgroup's wait did.
__device__ __forceinline__ bool fused_wait(u32* sync, u32 dep, bool poll, u32 seen, u32 target,
                                           bool device_scope)
{
    const u32* p = sync + FUSED_HDR + dep;
    bool ok = !poll || (int)(seen - target) >= 0;
    for (u32 spin = 0; __ballot(!ok) != 0; ++spin) {
        if ((spin & 255u) == 255u) {
            const u32 bad = __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((bad & FUSED_FAIL_WAIT) || spin >= FUSED_SPIN_LIMIT) {
                if ((threadIdx.x & 63u) == 0u)
                    __hip_atomic_fetch_or(sync, FUSED_FAIL_WAIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(2);
        if (!ok) ok = (int)(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) >= 0;
    }
    if (device_scope) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

template <int CAP, bool STATS>
__global__ __launch_bounds__(256) void msc_fused_queue(FusedArgs f, u64 codes,
                                                      unsigned long long* stats)
{
    __shared__ MscQueueLds<2> lds;
    __shared__ u32 go;
    const u32 g = blockIdx.x;
    const u32 h = g >> f.tiles_lg;
    const u32 j = g & ((1u << f.tiles_lg) - 1u);
    const u32 x = f.one_l2 ? (j & 7u) : xcc_id() & 7u;  // band = the XCD this workgroup is on
    const u32 k = j >> 3;
    const u32 tile = x * f.per_band + k;
    const u32 n_tiles = 1u << f.tiles_lg;
    const u32 base = tile * FUSED_TILE_WORDS;
    const bool device_scope = !f.one_l2 && (k == 0u || k + 1u == f.per_band);
    MscArgs a;
    const u32 color = h & 1u;
    a.own = (gu64*)(color ? f.w1 : f.w0);
    a.other = (const gu64*)(color ? f.w0 : f.w1);
    a.rows = f.rows;
    a.nw = f.nw;
    a.n_words = f.n_words;
    a.div_nw = f.div_nw;
    a.t = f.t0 + (h >> 1);
    a.color = color;
    a.k0 = f.k0;
    a.k1 = f.k1;
    // lanes 0-2 of wave 0 load the counters of tiles tile-1, tile, tile+1 (mod the count) first;
    // phase A's first Philox calls need no memory and run while those loads are in flight
    const u32 lane = threadIdx.x & 63u;
    const bool poll = h > 0u && threadIdx.x < 3u;
    const u32 dep = lane == 0u ? (tile + n_tiles - 1u) & (n_tiles - 1u)
                               : (lane == 1u ? tile : (tile + 1u) & (n_tiles - 1u));
    u32 seen = 0;
    if (poll) seen = __hip_atomic_load(f.sync + FUSED_HDR + dep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    MscPhaseACalls<2, CAP> pre;
    msc_phase_a_calls<2, CAP>(a, base, 0, pre);
    if (h > 0u) {
        if (threadIdx.x < 64u) {
            const bool ok = fused_wait(f.sync, dep, poll, seen, f.epoch + h, device_scope);
            if (threadIdx.x == 0) go = ok ? 1u : 0u;
        }
        __syncthreads();
        if (!go) return;  // the whole workgroup: nothing written, counter not advanced
    }
    MscTileLoads<2, MSC_ROWW> T;
    msc_tile_load<2, MSC_ROWW, true, true>(a, base, T);
    msc_phase_a_calls<2, CAP>(a, base, 1, pre);  // covers the latency of the tile's loads
    u32 calls = 0, slots = 0, words = 0;
    msc_queue_tile_loaded<2, CAP, MSC_ROWW, STATS, false, true, MscNoPrefetch, true, true,
                          std::conditional_t<STATS, MscStay, MscLeave>>(
        a, codes, base, lds, T, MscNoPrefetch{}, calls, slots, words, &pre);
    if (STATS) msc_stats(stats, calls, slots, words);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are acknowledged
    __syncthreads();
    if (threadIdx.x == 0) {
        if (device_scope) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // L2 write-back
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // ... completed before publishing
        }
        __hip_atomic_store(f.sync + FUSED_HDR + tile, f.epoch + h + 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Placement probe: the XCD each workgroup of a grid runs on.
__global__ void xcc_probe_kernel(u32* out)
{
    if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

// ------------------------------------------------------------------------------ host ----
// Placement of a dispatch's workgroups, per device: -2 not probed, -1 neither pattern (the
// fused form is not used), 0 every workgroup on one XCD, 8 round-robin over 8 XCDs (any
// starting XCD).
static int g_placement[64];
static std::once_flag g_placement_init;
static std::mutex g_placement_mu;

int fused_placement(int* out)
{
    std::call_once(g_placement_init, [] {
        for (int& v : g_placement) v = -2;
    });
    int dev = 0;
    ISING_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return ising_fail(ISING_ERANGE, "fused msc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_placement_mu);
    if (g_placement[dev] == -2) {
        constexpr int n = 8 * 256;
        u32* d = nullptr;
        ISING_HIP(hipMalloc(&d, n * sizeof(u32)));
        hipLaunchKernelGGL(xcc_probe_kernel, dim3(n), dim3(64), 0, 0, d);
        u32 h[n];
        const hipError_t e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
        (void)hipFree(d);
        ISING_HIP(e);
        bool rr = h[0] < 8u, one = true;
        for (int i = 0; i < n; ++i) {
            rr = rr && h[i] == ((h[0] + (u32)i) & 7u);
            one = one && h[i] == h[0];
        }
        g_placement[dev] = one ? 0 : (rr ? 8 : -1);
    }
    *out = g_placement[dev];
    return ISING_OK;
}

// Tiles per half sweep when the fused form applies to this geometry (whole-row 512-word
// tiles, a power-of-two count of at least 16 so that every XCD band has two or more), else 0.
static u32 fused_tiles(long long rows, int nw)
{
    if (nw != 64 && nw != 128 && nw != 256 && nw != 512) return 0;
    const unsigned long long words = (unsigned long long)rows * (unsigned long long)nw;
    if (words % FUSED_TILE_WORDS) return 0;
    const unsigned long long t = words / FUSED_TILE_WORDS;
    if (t < 16 || (t & (t - 1)) || t > (1ull << 22)) return 0;  // two halves fit a launch
    return (u32)t;
}

static size_t fused_sync_bytes(long long rows, int nw)
{
    const unsigned long long words = (unsigned long long)rows * (unsigned long long)nw;
    const unsigned long long tiles = (words + FUSED_TILE_WORDS - 1) / FUSED_TILE_WORDS;
    return (size_t)(((FUSED_HDR + tiles) * 4ull + 255ull) & ~255ull);
}

// 0: the fused form does not apply to this lattice on the current device (geometry, workgroup
// placement); 1: it applies; 2: it applies and is the by-size default. A half sweep must span
// at least about one generation of workgroups (2048 tiles: 8 per CU) for the fused form to
// win: with fewer, every tile of half h+1 waits out the lifetime of its neighbours in half h,
// and the launch serializes like separate launches, plus the waiting. Measured on MI355X at
// T_c, us per half sweep, one launch per half sweep / fused: 8192^2 (1024 tiles) 8.48 / 15.65,
// 8192 x 16384 (2048) 12.28 / 9.56, 16384^2 (4096) 19.41 / 15.34, 16384 x 32768 (8192)
// 33.47 / 31.12, 32768^2 (16384) 62.69 / 61.51.
constexpr u32 FUSED_DEFAULT_MIN_TILES = 2048;

int msc_fused_applies(long long rows, int nw)
{
    const u32 tiles = fused_tiles(rows, nw);
    if (!tiles) return 0;
    int place = -1;
    if (fused_placement(&place) != ISING_OK) return 0;
    if (place < 0) return 0;
    return tiles >= FUSED_DEFAULT_MIN_TILES ? 2 : 1;
}

int msc_fused_sweeps(u64* w0, u64* w1, long long rows, int nw, u32 t0, long long n_sweeps, u32 k0,
                     u32 k1, u32 thr4, u32 thr8, int cap, u32* sync, u32* epoch,
                     unsigned long long* stats, hipStream_t st)
{
    const u32 n_tiles = fused_tiles(rows, nw);
    int place = -1;
    int rc = fused_placement(&place);
    if (rc) return rc;
    if (!n_tiles || place < 0) return ising_fail(ISING_EINVAL, "fused msc: not applicable here");
    if (cap != 4 && cap != 3 && cap != 5)
        return ising_fail(ISING_EINVAL, "fused msc: cap must be 3, 4 or 5");
    FusedArgs f;
    f.w0 = w0;
    f.w1 = w1;
    f.rows = (u32)rows;
    f.nw = (u32)nw;
    f.n_words = (u32)(rows * nw);
    f.div_nw = make_fastdiv((u32)nw);
    f.k0 = k0;
    f.k1 = k1;
    f.sync = sync;
    f.tiles_lg = (u32)__builtin_ctz(n_tiles);
    f.per_band = n_tiles / 8u;
    f.one_l2 = place == 0 ? 1u : 0u;
    const u64 codes = msc_level_codes(thr4, thr8);
    u32 max_halves = FUSED_MAX_HALVES;
    while ((u64)max_halves * n_tiles * 256u >= (1ull << 32)) max_halves >>= 1;
    const u64 total = 2ull * (u64)n_sweeps;
    for (u64 done = 0; done < total;) {
        const u32 halves = (u32)((total - done) < max_halves ? (total - done) : max_halves);
        f.t0 = t0 + (u32)(done / 2);
        f.epoch = *epoch;
        const dim3 grid(halves * n_tiles);
#define FUSED_LAUNCH(C)                                                                         \
        if (stats) hipLaunchKernelGGL((msc_fused_queue<C, true>), grid, dim3(256), 0, st, f, codes, stats); \
        else hipLaunchKernelGGL((msc_fused_queue<C, false>), grid, dim3(256), 0, st, f, codes, stats)
        if (cap == 3) { FUSED_LAUNCH(3); }
        else if (cap == 5) { FUSED_LAUNCH(5); }
        else { FUSED_LAUNCH(4); }
#undef FUSED_LAUNCH
        ISING_LAUNCHED("msc_fused_queue");
        *epoch += halves;
        done += halves;
    }
    return ISING_OK;
}

extern "C" {

int ising_msc_sync_bytes(long long rows, int nw, size_t* bytes)
{
    if (!bytes || rows < 2 || nw < 1) return ising_fail(ISING_EINVAL, "msc_sync_bytes: bad argument");
    *bytes = fused_sync_bytes(rows, nw);
    return ISING_OK;
}

int ising_msc_sync_status(const uint32_t* sync, long long rows, int nw, uint32_t epoch,
                          int* failed, void* stream)
{
    if (!sync || !failed) return ising_fail(ISING_EINVAL, "msc_sync_status: null pointer");
    if (rows < 2 || nw < 1) return ising_fail(ISING_EINVAL, "msc_sync_status: bad geometry");
    const u32 n_tiles = fused_tiles(rows, nw);
    std::vector<u32> v(FUSED_HDR + n_tiles);
    ISING_HIP(hipMemcpyAsync(v.data(), sync, v.size() * sizeof(u32), hipMemcpyDeviceToHost,
                             as_stream(stream)));
    ISING_HIP(hipStreamSynchronize(as_stream(stream)));
    u32 bits = v[0];
    for (u32 i = 0; i < n_tiles; ++i)
        if (v[FUSED_HDR + i] != epoch) bits |= FUSED_FAIL_COUNTERS;
    *failed = (int)bits;
    return ISING_OK;
}

}  // extern "C"
