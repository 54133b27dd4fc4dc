// Batched temperature scans: one launch per half sweep (or one dependency-ordered launch per
// call) sweeps many lattices of uneven size, dimension, variant and temperature (no reference
// counterpart; the reference runs one lattice per launch sequence,
// src/pyising/backends/base.py:85-90).
//
// Work decomposition. Every lattice is cut into tiles of one workgroup's work for one half
// sweep: int8 = 256 threads x V sites, msc = 512 words (queue compaction) or 256 words
// (lockstep). Colour c + 1 of a lattice may only start when every tile of colour c of that
// lattice is done. Each tile is bit-identical to the single-lattice kernels (same device
// functions, same counters: t = t0 + sweep), so a batched lattice equals the same lattice swept
// alone; a lattice with a sweep limit sits the batch's later sweeps out.
//
// Two forms give identical results. Launch form: one launch per half sweep over all lattices
// (the kernel boundary orders the colours), workgroups taking 1, 2, 4 or 8 tiles of a launch
// order that balances every lattice over the eight XCDs, long tiles first. Dependency-ordered
// form: one launch holds the tiles of many half sweeps, and a tile of half sweep h waits only
// for the tiles of its own lattice that hold its neighbour rows in half sweep h - 1 (per-tile
// completion counters) instead of for the whole grid: the half sweeps of different lattices,
// and of distant parts of one lattice, overlap, and no launch boundary or grid barrier drains
// the GPU between half sweeps.
//
// Binder accumulation. Launch form: while writing a colour, each tile stores its number of +1
// spins in its own slot (slots double-buffered by sweep parity; no atomics, so no contention on
// a per-lattice counter). After a sweep, one workgroup per lattice (in the next sweep's colour-0
// launch or, after a call's last sweep, its own launch) sums its slots, forms
// M = 2 * plus - N and accumulates sum M, sum |M| (64-bit: |M| <= N) and sum M^2, sum M^4 as
// 128-bit (lo, hi) pairs. Dependency-ordered form: each tile stores its count in its slot, and
// a fold entry per lattice at the end of every colour-1 half sweep sums the slots once the
// lattice's tiles are done and folds M into the sums with atomics. The number of sweeps done
// lives in device memory, so a captured graph can be replayed and every replay continues the
// counter stream.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <new>
#include <vector>

#include "int8_core.hpp"
#include "msc_core.hpp"
#include "placement.hpp"

// The C ABI struct's layout, which the Python mirror (pyising_amd/batch.py) repeats.
static_assert(sizeof(IsingBatchLattice) == 96 && offsetof(IsingBatchLattice, t0) == 80 &&
                  offsetof(IsingBatchLattice, vec) == 84 &&
                  offsetof(IsingBatchLattice, max_sweeps) == 88,
              "IsingBatchLattice layout");

// Lattice kinds of the C ABI: int8, msc with the batch's default tile, msc with lockstep
// 256-word tiles, msc with queue-compaction 512-word tiles (the last two let a tuner choose the
// tile per lattice). In a descriptor, kind is int8 or msc and msc_queue says which tile.
enum { BATCH_INT8 = 0, BATCH_MSC = 1, BATCH_MSC_LOCKSTEP = 2, BATCH_MSC_QUEUE = 3 };

struct BatchDesc {
    void* c0;
    void* c1;
    u32 kind, ndim, hb, vec;  // vec: int8 sites per thread (16, 8 or 4)
    u32 msc_queue;            // msc: 1 = queue-compaction 512-word tiles, 0 = lockstep 256
    u32 layout;               // msc: MSC_GENERAL, MSC_ROWW or MSC_SEGW classification
    Int8Geom g;
    Int8Table tb;
    u32 rows, nw, n_words;
    FastDiv div_nw;
    u64 codes;
    u32 k0, k1, t0;
    u32 limit;  // batch sweeps (since create / assign) this lattice takes part in
    u32 tile_begin, n_tiles;
    // dependency-ordered form: tile i of a half sweep needs tiles i - dep_span .. i + dep_span
    // (cyclic) of the previous one, or every tile when dep_all
    u32 dep_span, dep_all;
    long long n_spins;
};

struct BatchAcc {
    long long sum_m;
    // sum M^2 and sum M^4 as 128-bit (lo, hi) pairs: M^2 <= N^2 reaches 2^56 at 16384^2, so a
    // 64-bit sum could wrap after a few hundred ordered sweeps
    unsigned long long sum_abs_m, sum_m2_lo, sum_m2_hi, sum_m4_lo, sum_m4_hi, samples;
};

constexpr u64 BATCH_NO_TILE = ~0ull;

struct BatchKernelArgs {
    const BatchDesc* descs;
    const u32* tile_lat;
    // launch form: workgroup b runs tile (u32)tile_seq[b] of lattice tile_seq[b] >> 32 (one load
    // names both, so the descriptor load is the only dependent one); BATCH_NO_TILE: none
    const u64* tile_seq;
    u32* plus;  // [2 sweep parity][2 colour][total_tiles]: +1 spins written by each tile
    BatchAcc* acc;
    // [0]: sweeps done, read by colour-0 work; [1]: the sweep in progress, written by colour-0
    // launches and read by colour-1 launches (each launch only writes the counter the
    // concurrent workgroups of the same launch do not read)
    unsigned long long* sweeps_done;
    u32 n_lat, total_tiles, n_sweeps, msc_mode, msc_cap;
    u32 moments;  // launch form: 0 = nobody reads the per-tile counts, skip storing them
    u32 n_seq;    // launch form: tile workgroups per half sweep
    u32 per_wg;   // launch form: launch-order entries (tiles) each of them runs in sequence
};

// Per-wave sum of v stored to the wave's own slot, dst[wave] (four slots per tile): no atomics
// and no workgroup barrier, so a wave that is done retires without waiting for the others.
constexpr u32 BATCH_WAVES = 4;  // 256-thread workgroups

__device__ __forceinline__ void wave_store_sum(u32* dst, u32 v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) dst[threadIdx.x >> 6] = v;
}

// One wave: M of lattice `lat` from its tiles' counts of parity `par`, folded into the
// accumulators (sum M, sum |M|, 128-bit sums M^2 and M^4, samples).
// The +1 spins of lattice `lat` in sweep parity `par`: this thread's share of its count
// slots, threads `first`, `first + step`, ...
__device__ __forceinline__ u64 batch_lattice_plus(const BatchKernelArgs& ka, const BatchDesc& d,
                                                  u32 par, u32 first, u32 step)
{
    const size_t slots = (size_t)BATCH_WAVES * ka.total_tiles;  // per (parity, colour)
    const u32* p0 = ka.plus + ((size_t)par * 2 + 0) * slots + (size_t)BATCH_WAVES * d.tile_begin;
    const u32* p1 = ka.plus + ((size_t)par * 2 + 1) * slots + (size_t)BATCH_WAVES * d.tile_begin;
    u64 plus = 0;
    for (u32 i = first; i < BATCH_WAVES * d.n_tiles; i += step) plus += (u64)p0[i] + p1[i];
    return plus;
}

__device__ __forceinline__ void batch_fold_moments(BatchAcc& A, const BatchDesc& d, u64 plus)
{
    const long long m = 2 * (long long)plus - d.n_spins;
    const u64 m2 = (u64)(m * m);
    const u64 lo = m2 * m2, hi = __umul64hi(m2, m2);
    A.sum_m += m;
    A.sum_abs_m += (u64)(m < 0 ? -m : m);
    const u64 n2 = A.sum_m2_lo + m2;
    A.sum_m2_hi += n2 < m2 ? 1u : 0u;
    A.sum_m2_lo = n2;
    const u64 nlo = A.sum_m4_lo + lo;
    A.sum_m4_hi += hi + (nlo < lo ? 1u : 0u);
    A.sum_m4_lo = nlo;
    A.samples += 1;
}

// One 512-word msc tile by queue compaction (4 unrolled calls, then packed rounds), with the
// lattice's wave classification (block-uniform branch). COH: loads that miss in the L1. EX: how
// waves without survivor work leave the tile (MscStay: they stay to its end).
template <bool COH, class EX = MscStay>
__device__ __forceinline__ u32 batch_msc_tile(const MscArgs& a, u64 codes, u32 base, u32 layout,
                                              MscQueueLds<2>& lds, EX ex = EX{})
{
    u32 calls = 0, slots = 0, words = 0;
    if (layout == MSC_ROWW)
        return msc_queue_tile<2, 4, MSC_ROWW, false, true, true, COH, EX>(a, codes, base, lds, calls, slots, words, ex);
    if (layout == MSC_SEGW)
        return msc_queue_tile<2, 4, MSC_SEGW, false, true, true, COH, EX>(a, codes, base, lds, calls, slots, words, ex);
    return msc_queue_tile<2, 4, MSC_GENERAL, false, true, true, COH, EX>(a, codes, base, lds, calls, slots, words, ex);
}

// Early leave of a one-tile workgroup's waves (MscLeave with a hand-over of the wave's +1
// spins): the launch form stores the wave's count in its slot (nothing when moments are off),
// the dependency-ordered form adds it to the tile's LDS total. The waves that stay hand over
// their counts after the tile in the same way, so every wave's count is taken exactly once.
struct BatchLeaveSlot {
    static constexpr bool leave = true;
    u32* dst;  // the tile's per-wave count slots, or null
    __device__ void operator()(u32 ones) const
    {
        if (dst) wave_store_sum(dst, ones);
    }
};
struct BatchLeavePart {
    static constexpr bool leave = true;
    u32* part;  // the tile's count in LDS, or null
    __device__ void operator()(u32 ones) const
    {
        if (part && ones) atomicAdd(part, ones);
    }
};

// One 256-word lockstep msc tile; returns the +1 spins written. With a whole-wave layout the
// word count is a multiple of 64, so the validity test is wave-uniform.
template <bool COH>
__device__ __forceinline__ u32 batch_msc_lockstep(const MscArgs& a, u64 codes, u32 local,
                                                  u32 layout)
{
    const u32 w = local * 256u + threadIdx.x;
    MscState st = {};
    if (w < a.n_words) {
        if (layout == MSC_ROWW) st = msc_classify_wave<MSC_ROWW, COH>(a, w);
        else if (layout == MSC_SEGW) st = msc_classify_wave<MSC_SEGW, COH>(a, w);
        else st = msc_classify_wave<MSC_GENERAL, COH>(a, w);
    }
    u32 calls = 0, slots = 0;
    msc_run<true, false>(a, codes, w, 0u, 16u, st, calls, slots);
    if (w < a.n_words) {
        a.own[w] = (u64)st.pwl | ((u64)st.pwh << 32);
        return __popc(st.pwl) + __popc(st.pwh);
    }
    return 0u;
}

// One int8 tile; thr_lds is this workgroup's threshold table for the lattice (filled by
// batch_int8_v below, which all threads call).
template <int NDIM, int HB, int V, bool COH>
__device__ __forceinline__ u32 batch_int8_tile(const BatchDesc& d, const u32* thr_lds, u32 local,
                                               u32 t, u32 color)
{
    const u32 ci = local * 256u + threadIdx.x;
    if (ci >= d.g.n_chunks) return 0u;
    gi8* own = (gi8*)(color ? d.c1 : d.c0);
    const gi8* other = (const gi8*)(color ? d.c0 : d.c1);
    const u32 minus = int8_update_chunk<NDIM, HB, V, true, COH>(own, other, d.g, thr_lds, t, color,
                                                                 d.k0, d.k1, ci);
    return (u32)V - minus;
}

template <int NDIM, int HB, bool COH>
__device__ __forceinline__ u32 batch_int8_v(const BatchDesc& d, u32* thr_lds, u32 local, u32 t,
                                            u32 color)
{
    int8_fill_table<NDIM, HB>(thr_lds, d.tb);
    __syncthreads();
    if (d.vec == 16) return batch_int8_tile<NDIM, HB, 16, COH>(d, thr_lds, local, t, color);
    if (d.vec == 8) return batch_int8_tile<NDIM, HB, 8, COH>(d, thr_lds, local, t, color);
    return batch_int8_tile<NDIM, HB, 4, COH>(d, thr_lds, local, t, color);
}

// Tile kinds a kernel instantiates: every kind, every kind but the msc queue-compaction tiles
// (whose phase A holds 32 more registers), or only those.
enum : int { TILES_ALL = 0, TILES_LIGHT = 1, TILES_QUEUE = 2 };

// Tile `local` of lattice d, colour `color`, Philox sweep counter t: this thread's count of +1
// spins written. Every thread of the workgroup calls it; callers that ran an earlier tile put a
// barrier first (its threads may still read thr_lds). COH: lattice loads miss in the L1. SET:
// the kinds compiled in (the caller only passes lattices of those kinds).
template <bool COH, int SET = TILES_ALL, class EX = MscStay>
__device__ __forceinline__ u32 batch_tile(const BatchDesc& d, MscQueueLds<2>& lds, u32* thr_lds,
                                          u32 local, u32 t, u32 color, EX ex = EX{})
{
    if (SET != TILES_QUEUE && d.kind == BATCH_INT8) {
        if (d.ndim == 2) return d.hb ? batch_int8_v<2, 1, COH>(d, thr_lds, local, t, color)
                                     : batch_int8_v<2, 0, COH>(d, thr_lds, local, t, color);
        return d.hb ? batch_int8_v<3, 1, COH>(d, thr_lds, local, t, color)
                    : batch_int8_v<3, 0, COH>(d, thr_lds, local, t, color);
    }
    MscArgs a;
    a.own = (gu64*)(color ? d.c1 : d.c0);
    a.other = (const gu64*)(color ? d.c0 : d.c1);
    a.rows = d.rows;
    a.nw = d.nw;
    a.n_words = d.n_words;
    a.div_nw = d.div_nw;
    a.t = t;
    a.color = color;
    a.k0 = d.k0;
    a.k1 = d.k1;
    if constexpr (SET == TILES_QUEUE) return batch_msc_tile<COH>(a, d.codes, local * 512u, d.layout, lds, ex);
    else if constexpr (SET == TILES_LIGHT) return batch_msc_lockstep<COH>(a, d.codes, local, d.layout);
    else return d.msc_queue ? batch_msc_tile<COH>(a, d.codes, local * 512u, d.layout, lds, ex)
                            : batch_msc_lockstep<COH>(a, d.codes, local, d.layout);
}

// ------------------------------------------------------------- launch-per-half-sweep form ----
// The same tiles, one workgroup per tile and one launch per half sweep: the kernel boundary is
// the barrier between colours. The sweep index lives in device memory, so captured graphs
// replay with it advancing: a colour-0 launch reads sweeps_done[0] and copies it to [1], the
// colour-1 launch reads [1] and sets [0] to the next sweep.
// acc_prev (colour 0 only): the n_lat workgroups after the n_seq tile workgroups accumulate the
// moments of the previous sweep, whose count slots (the other parity) this sweep does not write;
// this saves the accumulation launch of every sweep but a call's last.
// K = ka.per_wg tiles per workgroup, a template parameter: the one-tile form keeps the
// register count of a single tile body (a runtime loop costs 17 VGPRs and with them occupancy).
template <u32 K>
__global__ __launch_bounds__(256) void batch_half(BatchKernelArgs ka, u32 color, u32 acc_prev)
{
    __shared__ MscQueueLds<2> lds;
    __shared__ u32 thr_lds[8];
    const u32 sw = (u32)ka.sweeps_done[color];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (color == 0u) ka.sweeps_done[1] = sw;
        else ka.sweeps_done[0] = sw + 1u;
    }
    if (acc_prev && blockIdx.x >= ka.n_seq) {
        const u32 lat = blockIdx.x - ka.n_seq;
        const BatchDesc& dl = ka.descs[lat];
        if (sw - 1u >= dl.limit) return;  // the lattice sat the previous sweep out
        u64 plus = batch_lattice_plus(ka, dl, (sw - 1u) & 1u, threadIdx.x, 256u), unused = 0;
        block_sum2_u64(plus, unused);
        if (threadIdx.x == 0) batch_fold_moments(ka.acc[lat], dl, plus);
        return;
    }
    const u32 par = sw & 1u;
    for (u32 j = 0; j < K; ++j) {  // block-uniform
        const u64 entry = ka.tile_seq[(size_t)blockIdx.x * K + j];
        if (entry == BATCH_NO_TILE) continue;  // padding of a shorter XCD list (whole workgroup)
        const u32 tile = (u32)entry, lat = (u32)(entry >> 32);
        const BatchDesc& d = ka.descs[lat];
        if (sw >= d.limit) continue;  // past the lattice's sweep limit: the whole workgroup idles
        if (j) __syncthreads();  // an earlier tile of this workgroup may still read thr_lds
        const u32 local = tile - d.tile_begin;
        u32* dst = ka.moments ? ka.plus + ((size_t)par * 2 + color) * BATCH_WAVES * ka.total_tiles +
                                    (size_t)BATCH_WAVES * tile
                              : nullptr;
        u32 plus;
        if constexpr (K == 1)  // nothing follows the tile: waves without work may leave
            plus = batch_tile<false>(d, lds, thr_lds, local, d.t0 + sw, color, BatchLeaveSlot{dst});
        else
            plus = batch_tile<false>(d, lds, thr_lds, local, d.t0 + sw, color);
        if (dst) wave_store_sum(dst, plus);
    }
}

// After both colours of a sweep (when moments are on): one workgroup per lattice, spread over
// the CUs, accumulates the sweep that the colour-1 launch just finished.
__global__ __launch_bounds__(256) void batch_accumulate(BatchKernelArgs ka)
{
    const u32 sw = (u32)ka.sweeps_done[1];
    const BatchDesc& d = ka.descs[blockIdx.x];
    if (sw >= d.limit) return;
    u64 plus = batch_lattice_plus(ka, d, sw & 1u, threadIdx.x, 256u), unused = 0;
    block_sum2_u64(plus, unused);
    if (threadIdx.x == 0) batch_fold_moments(ka.acc[blockIdx.x], d, plus);
}

__global__ void batch_zero_u64(unsigned long long* p, u32 n)
{
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ------------------------------------------------------------- dependency-ordered form ----
// One launch holds H half sweeps (H even: whole sweeps) of every lattice. Each half sweep is
// eight lists of tiles, list x for XCD x: every lattice lives on one XCD (all its tiles, every
// half sweep), so a tile and the tiles it waits for share one L2; a lattice too large for one
// XCD's share is cut into eight bands of consecutive tiles, band b on XCD b, and its tiles at a
// band edge use the device-scope protocol below. Workgroup g runs half sweep h = g / 8P and
// entry (g % 8P) / 8 of the list of the XCD it runs on (HW_REG_XCC_ID): the dispatcher places
// workgroup g on XCD (g + r) % 8, so every entry of every list runs exactly once.
//
// Ordering. A tile of lattice half sweep k (k = 2 sweep + colour, counted since the batch was
// made) waits until the tiles of its lattice that hold its neighbour rows have completed k
// half sweeps; a tile publishes k + 1 when it is done. The same tiles read what this one
// overwrites, so the wait covers both directions. Waits are for workgroups with lower ids,
// which each XCD dispatches first: the oldest unfinished workgroup can always proceed. Waits
// are bounded (BF_SPIN_LIMIT polls): a workgroup that gives up sets a failure bit and exits
// without touching the lattice, later waiters see the bit and exit at once, the kernel ends
// normally, and batch_fused_check checks every counter.
//
// Memory. Inside one XCD the L2 is the coherence point: stores are acknowledged (s_waitcnt)
// before the counter is published, every lattice load misses in the CU's L1 (cache policy sc1;
// a CU may hold a line it read in an earlier half sweep of the launch), and a tile that only
// tiles of its own XCD wait for publishes with a plain store, which stays in that L2 (a
// device-scope store costs the poller a trip past the L2: measured 1.4-1.8 us per wait). A
// band-edge tile of a split lattice adds the device-scope fences: release (L2 write-back)
// and a device-scope counter store before publishing, acquire (L2 invalidate) after its wait.
//
// Moments (MOM). Each tile stores its +1 spins in its slot (parity, colour, tile) before
// publishing. The last entries of every list are fold entries, one per lattice (one per band
// and lattice when split): in a colour-1 half sweep it waits for its tiles, sums their slots of
// both colours and folds M into the sums with atomics (folds of consecutive sweeps may
// overlap); a split lattice's eight bands add (sum << 8) + 1 to a per-lattice word first and
// the eighth arrival folds. Then it publishes sweep + 1 in the fold counter of the parity. The
// colour-0 tiles of sweep s + 2, the next to write those slots, wait for the fold of sweep s.
constexpr u32 BF_HDR = 64;  // u32 words before the per-tile counters
constexpr u32 BF_SPIN_LIMIT = 1u << 20;
constexpr u32 BF_FAIL_WAIT = 1u, BF_FAIL_COUNTERS = 2u;
constexpr u64 BF_EDGE = 1ull << 63;   // list-entry flag: band-edge tile of a split lattice
constexpr u64 BF_FOLD = 1ull << 62;   // list-entry flag: fold entry (low word: band)
constexpr u64 BF_SPLIT = 1ull << 61;  // fold entry of a split lattice
constexpr u32 BF_LAT_MASK = (1u << 29) - 1u;
constexpr u32 BF_MAX_SWEEPS = 512;    // sweeps per launch (longer calls take several launches)

struct BatchFusedArgs {
    const BatchDesc* descs;
    const u64* seq;   // [8][P] entries (lattice << 32 | tile, flags), BATCH_NO_TILE padding
    u32* sync;        // [0] failure bits; [BF_HDR + tile] half sweeps completed by the tile
    u32* fold;        // [2 lattice + parity]: 1 + the last sweep of that parity folded
    u32* slots;       // [2 parity][2 colour][tiles]: +1 spins written by each tile
    u64* pacc;        // [2 lattice + parity]: split lattices, (band sums << 8) + arrivals
    BatchAcc* acc;
    const unsigned long long* sweeps_done;
    u32 P, per_half;  // per_half = 8 P workgroups per half sweep
    u32 total_tiles;
    FastDiv div_half;
    u32 one_l2;       // 1: the device runs every workgroup on one XCD
};

// Wave 0: waits until the `count` tiles first, first + 1, ... (cyclic over the lattice's n
// tiles, counters at cnt) have completed `target` half sweeps and, when fold_p is set, until
// *fold_p reaches fold_target. Lanes 0-62 poll one tile each (63 at a time); lane 63 polls the
// fold counter. False if the wait gave up or another workgroup's did.
__device__ bool bf_wait(u32* sync, const u32* cnt, u32 n, u32 first, u32 count, u32 target,
                        const u32* fold_p, u32 fold_target)
{
    const u32 lane = threadIdx.x & 63u;
    u32 spin = 0;
    for (u32 c0 = 0; c0 < count || (c0 == 0u && fold_p); c0 += 63u) {  // wave-uniform
        const u32 i = c0 + lane;
        const bool dep = lane < 63u && i < count;
        const bool fl = lane == 63u && c0 == 0u && fold_p != nullptr;
        u32 j = first + (dep ? i : 0u);
        if (j >= n) j -= n;
        const u32* p = dep ? cnt + j : fold_p;
        const u32 tgt = dep ? target : fold_target;
        bool ok = !(dep || fl);
        while (true) {
            if (!ok) ok = (int)(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - tgt) >= 0;
            if (__ballot(!ok) == 0) break;
            if ((++spin & 255u) == 0u) {
                const u32 bad = __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((bad & BF_FAIL_WAIT) || spin >= BF_SPIN_LIMIT) {
                    if (lane == 0u)
                        __hip_atomic_fetch_or(sync, BF_FAIL_WAIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    return false;
                }
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    return true;
}

// Moments of one sweep folded into A with device-scope atomics (128-bit sums by carry).
__device__ void bf_fold(BatchAcc* A, const BatchDesc& d, u64 plus)
{
    const long long m = 2 * (long long)plus - d.n_spins;
    const u64 m2 = (u64)(m * m);
    const u64 lo = m2 * m2, hi = __umul64hi(m2, m2);
    auto add = [](unsigned long long* p, u64 v) {
        return (u64)__hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    add((unsigned long long*)&A->sum_m, (u64)m);
    add(&A->sum_abs_m, (u64)(m < 0 ? -m : m));
    const u64 o2 = add(&A->sum_m2_lo, m2);
    if (o2 + m2 < o2) add(&A->sum_m2_hi, 1u);
    const u64 o4 = add(&A->sum_m4_lo, lo);
    const u64 c4 = hi + (o4 + lo < o4 ? 1u : 0u);
    if (c4) add(&A->sum_m4_hi, c4);
    add(&A->samples, 1u);
}

// Fold entry of lattice `lat`, band `band` (colour-1 half sweeps; wave 0 only).
__device__ void bf_fold_entry(const BatchFusedArgs& f, const BatchDesc& d, u32 lat, u32 band,
                              bool split, u32 sw, u32 k)
{
    const u32 n = d.n_tiles, par = sw & 1u;
    const u32 t0 = split ? (u32)((u64)n * band / 8u) : 0u;
    const u32 t1 = split ? (u32)((u64)n * (band + 1u) / 8u) : n;
    if (!bf_wait(f.sync, f.sync + BF_HDR + d.tile_begin, n, t0, t1 - t0, k + 1u, nullptr, 0u)) return;
    const u32* s0 = f.slots + (size_t)(2u * par) * f.total_tiles + d.tile_begin;
    const u32* s1 = s0 + f.total_tiles;
    u64 plus = 0;
    for (u32 t = t0 + (threadIdx.x & 63u); t < t1; t += 64u)
        plus += (u64)__hip_atomic_load(s0 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) +
                __hip_atomic_load(s1 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) plus += __shfl_xor(plus, off, 64);
    if (threadIdx.x != 0) return;
    u32* fold = f.fold + 2u * lat + par;
    if (split) {
        unsigned long long* pa = (unsigned long long*)f.pacc + 2u * lat + par;
        const u64 old = (u64)__hip_atomic_fetch_add(pa, (unsigned long long)((plus << 8) + 1u),
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((old & 0xFFu) != 7u) return;  // not the eighth band
        plus += old >> 8;
        __hip_atomic_store(pa, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    bf_fold(f.acc + lat, d, plus);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // folded and cleared before publishing
    // a whole lattice's colour-0 tiles all poll from this XCD: a plain store stays in its L2
    if (split) __hip_atomic_store(fold, sw + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else __hip_atomic_store(fold, sw + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// K: entries per workgroup. The lists hold groups of K entries: up to K consecutive tiles of
// one lattice (padded with BATCH_NO_TILE), or one fold entry and padding. A workgroup waits
// once for the union of its tiles' windows, then runs them in order, publishing each.
// SET: the tile kinds of this kernel. A batch runs two kernels side by side on two streams
// (lattices are independent, and each kernel waits only for its own workgroups): the msc
// lattices with queue-compaction tiles (TILES_QUEUE, 74 VGPRs: 6 workgroups per CU) and the
// rest (TILES_LIGHT, at most 64 VGPRs: 8 per CU, which the latency-bound int8 tiles need).
template <bool MOM, u32 K, int SET>
__global__ __launch_bounds__(256) void batch_fused(BatchFusedArgs f)
{
    __shared__ MscQueueLds<2> lds;
    __shared__ u32 thr_lds[8];
    __shared__ u32 part[1];  // +1 spins of the current tile
    __shared__ u32 go;
    const u32 g = blockIdx.x;
    const u32 h = fast_div(g, f.div_half);
    const u32 q = g - h * f.per_half;
    const u32 x = f.one_l2 ? (q & 7u) : (xcc_id() & 7u);
    const u64* ents = f.seq + (size_t)x * f.P + (size_t)(q >> 3) * K;
    if (MOM && threadIdx.x == 0) part[0] = 0;  // ordered before the first add by a barrier below
    const u64 e0 = ents[0];
    if (e0 == BATCH_NO_TILE) return;  // padding of a shorter list (whole workgroup)
    const u32 lat = (u32)(e0 >> 32) & BF_LAT_MASK;
    const BatchDesc& d = f.descs[lat];
    const u32 sw = (u32)*f.sweeps_done + (h >> 1), color = h & 1u, par = sw & 1u;
    if (sw >= d.limit) return;  // past the lattice's sweep limit (nobody waits for it)
    const u32 k = 2u * sw + color;  // this lattice's half sweeps before this one
    if (e0 & BF_FOLD) {
        if (MOM && color == 1u && threadIdx.x < 64u)
            bf_fold_entry(f, d, lat, (u32)e0, (e0 & BF_SPLIT) != 0, sw, k);
        return;
    }
    u32 kv = 1;  // tiles in this group
    bool edge_any = (e0 & BF_EDGE) != 0;
#pragma unroll
    for (u32 i = 1; i < K; ++i) {
        const u64 e = ents[i];
        if (e != BATCH_NO_TILE) {
            kv = i + 1u;
            edge_any = edge_any || (e & BF_EDGE) != 0;
        }
    }
    const u32 local0 = (u32)e0 - d.tile_begin;
    u32* cnt = f.sync + BF_HDR + d.tile_begin;
    // a launch's first half sweep follows the previous launch (stream order); the fold of
    // sweep sw - 2 is waited for when that sweep ran in this launch
    const bool fold_wait = MOM && color == 0u && h >= 4u;
    if (h > 0u) {
        if (threadIdx.x < 64u) {
            const u32 n = d.n_tiles;
            u32 count = n, first = 0;
            if (!d.dep_all && kv + 2u * d.dep_span < n) {
                count = kv + 2u * d.dep_span;
                first = local0 + n - d.dep_span;
                if (first >= n) first -= n;
            }
            const bool ok = bf_wait(f.sync, cnt, n, first, count, k,
                                    fold_wait ? f.fold + 2u * lat + par : nullptr, sw - 1u);
            if (edge_any) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (threadIdx.x == 0) go = ok ? 1u : 0u;
        }
        __syncthreads();
        if (!go) return;  // the whole workgroup: nothing written, counters not advanced
    }
    if (MOM && h == 0u) __syncthreads();  // part[0] = 0 before any add (h > 0: the wait's barrier)
    for (u32 i = 0; i < (K == 1 ? 1u : kv); ++i) {  // block-uniform
        const u64 e = K == 1 ? e0 : ents[i];
        const u32 tile = (u32)e, local = tile - d.tile_begin;
        if (i) __syncthreads();  // the previous tile's threads may still read thr_lds
        u32 plus;
        if constexpr (K == 1)  // only wave 0 works after the tile: waves without work may leave
            plus = batch_tile<true, SET>(d, lds, thr_lds, local, d.t0 + sw, color,
                                         BatchLeavePart{MOM ? &part[0] : nullptr});
        else
            plus = batch_tile<true, SET>(d, lds, thr_lds, local, d.t0 + sw, color);
        __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are acknowledged
        if (MOM && plus) atomicAdd(&part[0], plus);  // LDS; part[0] is zero here
        __syncthreads();
        if (threadIdx.x == 0) {
            if (MOM) {
                f.slots[(size_t)(2u * par + color) * f.total_tiles + tile] = part[0];
                part[0] = 0;  // for the next tile (its adds follow the loop's barrier)
            }
            if (e & BF_EDGE) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // L2 write-back
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // ... completed before publishing
                __hip_atomic_store(cnt + local, k + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the slot is written before publishing
                __hip_atomic_store(cnt + local, k + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
}

// After a dependency-ordered launch of n sweeps: every tile of a lattice that took part must
// have published its last half sweep, else a tile went missing (failure bit 2). One thread per
// tile; batch_fused_advance then moves the batch's sweep counter on by n.
__global__ __launch_bounds__(256) void batch_fused_check(const BatchDesc* descs, const u32* tile_lat,
                                                         u32 total_tiles, u32* sync,
                                                         const unsigned long long* sweeps_done, u32 n)
{
    const u32 t = blockIdx.x * 256u + threadIdx.x;
    if (t >= total_tiles) return;
    const u32 sw0 = (u32)sweeps_done[0];
    const BatchDesc& d = descs[tile_lat[t]];
    if (sw0 >= d.limit) return;
    const u32 end = (d.limit - sw0 < n) ? d.limit : sw0 + n;
    if (sync[BF_HDR + t] != 2u * end)
        __hip_atomic_fetch_or(sync, BF_FAIL_COUNTERS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void batch_fused_advance(unsigned long long* sweeps_done, u32 n)
{
    if (threadIdx.x == 0) {
        const unsigned long long sw = sweeps_done[0] + n;
        sweeps_done[0] = sw;
        sweeps_done[1] = sw - 1u;
    }
}

// ---------------------------------------------------------------------- host side ----
// IsingBatchLattice, the caller-side description of one lattice, is the C ABI struct of
// ising_hip.h.

struct BatchHandle {
    u32 n_lat = 0, total_tiles = 0, n_seq = 0;  // n_seq: tile workgroups of the launch form
    u32 per_wg = 1;  // launch form: tiles per workgroup (the launch order has n_seq * per_wg)
    u32 cap_lat = 0, cap_tiles = 0, cap_seq = 0;  // device table capacity (reserved: > n_lat)
    std::vector<BatchDesc> host_descs;  // the current lattices (to rebuild the launch order)
    BatchDesc* d_descs = nullptr;
    u32* d_tile_lat = nullptr;
    u64* d_tile_seq = nullptr;
    u32* d_plus = nullptr;
    BatchAcc* d_acc = nullptr;
    unsigned long long* d_done = nullptr;
    int grid = 0, max_blocks = 0, n_cu = 0, per_cu = 0;
    u32 msc_mode = 0, msc_cap = 4;
    int persistent = 0;  // 1: dependency-ordered launches; 0: two launches per sweep
    // dependency-ordered form (batches made by ising_batch_create): the eight lists, P entries
    // each, and the sync words (header, per-tile counters, fold counters, moment words)
    // lists of the two kernels in one buffer: the light kernel's 8 x f_P entries, then the
    // queue kernel's 8 x f_Pq (f_P, f_Pq = 0: no lattices of that kind)
    u64* d_fseq = nullptr;
    u32 f_P = 0, f_Pq = 0;
    hipStream_t f_side = nullptr;  // the queue kernel's stream (default priority)
    hipEvent_t f_fork = nullptr, f_join = nullptr;
    u32* d_fsync = nullptr;
    u32* d_fold = nullptr;
    u32* d_slots = nullptr;
    u64* d_pacc = nullptr;
    size_t fsync_words = 0;  // u64 words of the sync allocation
    int moments = 1;     // launch form: 1 adds the accumulation launch after every sweep
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    long long graph_sweeps = 0;
    // reserved handles: pinned staging of (descs, tile map), double-buffered; the event of a
    // slot marks the end of its last upload, so the host never rewrites a slot in flight
    void* pinned[2] = {nullptr, nullptr};
    hipEvent_t pinned_done[2] = {nullptr, nullptr};
    int next_slot = 0;
};

static void batch_drop_graph(BatchHandle* h)
{
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_sweeps = 0;
}

static void batch_free(BatchHandle* h)
{
    if (!h) return;
    batch_drop_graph(h);
    for (int s = 0; s < 2; ++s) {
        if (h->pinned_done[s]) {
            (void)hipEventSynchronize(h->pinned_done[s]);
            (void)hipEventDestroy(h->pinned_done[s]);
        }
        if (h->pinned[s]) (void)hipHostFree(h->pinned[s]);
    }
    (void)hipFree(h->d_descs);
    (void)hipFree(h->d_tile_lat);
    (void)hipFree(h->d_tile_seq);
    (void)hipFree(h->d_plus);
    (void)hipFree(h->d_acc);
    (void)hipFree(h->d_done);
    (void)hipFree(h->d_fseq);
    (void)hipFree(h->d_fsync);
    if (h->f_fork) (void)hipEventDestroy(h->f_fork);
    if (h->f_join) (void)hipEventDestroy(h->f_join);
    if (h->f_side) (void)hipStreamDestroy(h->f_side);
    delete h;
}

// Dependency window of a lattice's tiles (host): tile i holds units [i per_tile, (i + 1)
// per_tile) of `units` (int8 chunks or msc words), per_row units per row, `rows` rows; a unit
// reads the other colour at rows up to s away (1 in 2D, one plane = ly rows in 3D, cyclic).
// Tile i depends on, and is read by, the tiles holding rows r_first - s .. r_last + s of its
// rows r_first .. r_last: at most `span` tiles away on either side (cyclic). all = 1 when some
// tile's window covers every tile.
static void batch_dep_window(u64 units, u64 per_row, u64 per_tile, u64 rows, u64 s, u32 n,
                             u32& span, u32& all)
{
    span = 0;
    all = 0;
    for (u32 i = 0; i < n; ++i) {
        const u64 c0 = (u64)i * per_tile;
        const u64 c1 = (c0 + per_tile < units ? c0 + per_tile : units) - 1u;
        const u64 ra = c0 / per_row, rb = c1 / per_row;
        if (rb - ra + 2u * s + 1u >= rows) {
            all = 1;
            return;
        }
        const u64 lo = (ra + rows - s) % rows, hi = (rb + s) % rows;
        const u32 t_lo = (u32)(lo * per_row / per_tile);
        const u32 t_hi = (u32)(((hi + 1u) * per_row - 1u) / per_tile);
        const u32 back = (i + n - t_lo) % n, fwd = (t_hi + n - i) % n;
        span = back > span ? back : span;
        span = fwd > span ? fwd : span;
    }
    if (2ull * span + 1u >= n) all = 1;
}

// Host-side descriptors and tile map of n lattices (no device calls). On success `descs` holds
// n entries and `tile_lat` one lattice index per tile.
static int batch_build(const IsingBatchLattice* lats, int n, u32 msc_mode,
                       std::vector<BatchDesc>& descs, std::vector<u32>& tile_lat)
{
    descs.assign((size_t)n, BatchDesc{});
    tile_lat.clear();
    u32 tiles = 0;
    for (int i = 0; i < n; ++i) {
        const IsingBatchLattice& L = lats[i];
        BatchDesc& d = descs[(size_t)i];
        if (!L.c0 || !L.c1) return ising_fail(ISING_EINVAL, "batch: null colour array");
        if (L.lz < 1 || L.ly < 2 || L.lx < 8 || (L.ly & 1) || (L.lx & 1) || (L.ndim == 3 && (L.lz < 2 || (L.lz & 1))))
            return ising_fail(ISING_EINVAL, "batch: bad shape");
        d.c0 = L.c0;
        d.c1 = L.c1;
        d.kind = (u32)(L.kind >= BATCH_MSC ? BATCH_MSC : L.kind);
        d.ndim = (u32)L.ndim;
        d.hb = (u32)(L.heatbath != 0);
        d.k0 = (u32)(L.seed & 0xFFFFFFFFull);
        d.k1 = (u32)(L.seed >> 32);
        d.t0 = L.t0;
        d.limit = L.max_sweeps ? L.max_sweeps : 0xFFFFFFFFu;
        d.n_spins = (long long)L.lz * L.ly * L.lx;
        // exact moments: M^4 <= N^4 = 2^112 at N = 2^28, so the 128-bit sum holds 65536 sweeps
        if (d.n_spins > (1ll << 28))
            return ising_fail(ISING_ERANGE, "batch: at most 2^28 spins per lattice (16384^2), for exact 128-bit moment sums");
        const long long rows = (long long)L.lz * L.ly;
        if (L.kind == BATCH_INT8) {
            if (L.ndim != 2 && L.ndim != 3) return ising_fail(ISING_EINVAL, "batch: int8 ndim 2 or 3");
            if (L.ndim == 2 && L.lz != 1) return ising_fail(ISING_EINVAL, "batch: 2D needs lz = 1");
            const u32 lh = (u32)(L.lx / 2);
            u32 v = 0;
            if (L.vec) {
                if ((L.vec != 16 && L.vec != 8 && L.vec != 4) || L.lx % 8 || lh % (u32)L.vec)
                    return ising_fail(ISING_EINVAL, "batch: int8 vec must be 16, 8 or 4 and divide Lx/2");
                v = (u32)L.vec;
            } else {
                // the widest width that still fills a 256-lane tile, else the narrowest
                for (u32 c = 16; c >= 4; c /= 2) {
                    if (lh % c) continue;
                    v = c;
                    if (rows * (long long)(lh / c) >= 256) break;
                }
            }
            if (!v || L.lx % 8) return ising_fail(ISING_EINVAL, "batch: int8 needs Lx % 8 == 0");
            if (rows * (lh / v) + (1ll << 20) > 0xFFFFFFFFll) return ising_fail(ISING_ERANGE, "batch: int8 lattice too large");
            d.vec = v;
            d.g.lh = lh;
            d.g.ly = (u32)L.ly;
            d.g.lz = (u32)L.lz;
            d.g.chunks_row = lh / v;
            d.g.n_chunks = (u32)(rows * (lh / v));
            d.g.div_chunks = make_fastdiv(d.g.chunks_row);
            d.g.div_ly = make_fastdiv(d.g.ly);
            for (int k = 0; k <= 2 * L.ndim; ++k) d.tb.v[k] = L.table[k];
            d.n_tiles = (d.g.n_chunks + 255u) / 256u;
            batch_dep_window(d.g.n_chunks, d.g.chunks_row, 256u, (u64)rows,
                             L.ndim == 2 ? 1u : (u64)L.ly, d.n_tiles, d.dep_span, d.dep_all);
        } else if (L.kind >= BATCH_MSC && L.kind <= BATCH_MSC_QUEUE) {
            if (L.ndim != 2 || L.lz != 1 || L.lx % 128) return ising_fail(ISING_EINVAL, "batch: msc needs 2D, Lx % 128 == 0");
            if ((long long)L.ly * (L.lx / 128) + (1ll << 20) > 0xFFFFFFFFll)
                return ising_fail(ISING_ERANGE, "batch: msc lattice too large");
            d.rows = (u32)L.ly;
            d.nw = (u32)(L.lx / 128);
            d.n_words = d.rows * d.nw;
            d.div_nw = make_fastdiv(d.nw);
            d.codes = msc_level_codes(L.table[3], L.table[4]);
            d.msc_queue = L.kind == BATCH_MSC_QUEUE || (L.kind == BATCH_MSC && msc_mode == 0);
            d.layout = (u32)msc_layout(d.nw, d.n_words);
            const u32 per = d.msc_queue ? 512u : 256u;
            d.n_tiles = (d.n_words + per - 1u) / per;
            batch_dep_window(d.n_words, d.nw, per, d.rows, 1u, d.n_tiles, d.dep_span, d.dep_all);
        } else {
            return ising_fail(ISING_EINVAL, "batch: kind must be 0 (int8), 1 (msc), 2 (msc lockstep) or 3 (msc queue)");
        }
        if ((unsigned long long)tiles + d.n_tiles > 0x7FFFFFFFull) return ising_fail(ISING_ERANGE, "batch: too many tiles");
        d.tile_begin = tiles;
        tiles += d.n_tiles;
        for (u32 k = 0; k < d.n_tiles; ++k) tile_lat.push_back((u32)i);
    }
    return ISING_OK;
}

// Launch order of the tiles of one half sweep. Workgroup b runs on XCD b % 8, and the XCDs
// have their own L2s, so the schedule is eight lists interleaved: every lattice is cut into
// eight contiguous chunks of tiles, one per list (a chunk's halo rows stay in its XCD's L2 and
// every XCD gets the same share of every lattice, so their work is equal whatever the mix);
// within a list, lattices with longer tiles come first (msc compaction, then msc lockstep,
// then int8), so the last workgroups to start are short ones. Shorter lists are padded with
// BATCH_NO_TILE. The order changes no result: a tile computes the same words wherever it runs.
static u32 batch_tile_cost(const BatchDesc& d)
{
    if (d.kind == BATCH_INT8) return 1u;
    return d.msc_queue ? 3u : 2u;
}

// Launch order of the launch form with k tiles per workgroup: workgroup w (on XCD w % 8) runs
// entries [w k, w k + k) of the order, which are consecutive tiles of XCD (w % 8)'s list.
static void batch_schedule(const std::vector<BatchDesc>& descs, u32 k, std::vector<u64>& seq)
{
    constexpr u32 X = 8;
    std::vector<u32> order(descs.size());
    for (u32 i = 0; i < (u32)descs.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) {
        return batch_tile_cost(descs[a]) > batch_tile_cost(descs[b]);
    });
    std::vector<u64> lists[X];
    for (u32 li : order) {
        const BatchDesc& d = descs[li];
        const u32 q = d.n_tiles / X, r = d.n_tiles % X;
        // lists by length, shortest first: they take the r chunks of q + 1 tiles
        u32 by_len[X];
        for (u32 x = 0; x < X; ++x) by_len[x] = x;
        std::stable_sort(by_len, by_len + X, [&](u32 a, u32 b) { return lists[a].size() < lists[b].size(); });
        u32 next = d.tile_begin;
        for (u32 k = 0; k < X; ++k) {
            const u32 len = q + (k < r ? 1u : 0u);
            for (u32 t = 0; t < len; ++t) lists[by_len[k]].push_back(((u64)li << 32) | (next + t));
            next += len;
        }
    }
    size_t longest = 0;
    for (u32 x = 0; x < X; ++x) longest = lists[x].size() > longest ? lists[x].size() : longest;
    const size_t rows = (longest + k - 1) / k;  // workgroups per XCD
    seq.assign(rows * X * k, BATCH_NO_TILE);
    for (u32 x = 0; x < X; ++x)
        for (size_t i = 0; i < lists[x].size(); ++i)
            seq[((i / k) * X + x) * k + i % k] = lists[x][i];
}

// Lists of the dependency-ordered form: seq[x * P + j] is entry j of XCD x's list for every
// half sweep. Lattices are placed whole on one XCD, largest first onto the least loaded one
// (cost = tiles x an estimate of a tile's time); a lattice whose cost exceeds half an XCD's
// share, whose window is not the whole lattice and whose bands keep interior tiles is cut into
// eight bands of consecutive tiles, band b on XCD b, and its tiles whose window leaves their
// band carry BF_EDGE. Within a list, lattices with longer tiles come first. P = the longest
// list; shorter ones are padded with BATCH_NO_TILE.
static u32 batch_fused_cost(const BatchDesc& d)
{
    if (d.kind == BATCH_INT8) return d.vec / 4u;  // 1024 .. 4096 sites
    return d.msc_queue ? 12u : 8u;                // 512 / 256 words, about 3x / 2x an int8 tile
}

static bool batch_queue_lattice(const BatchDesc& d)
{
    return d.kind != BATCH_INT8 && d.msc_queue;
}

// The lists of one of the two kernels: queue = the msc lattices with queue-compaction tiles,
// else every other lattice. P = 0 when the kernel has no lattices.
static void batch_fused_schedule(const std::vector<BatchDesc>& descs, bool queue, u32 K,
                                 std::vector<u64>& seq, u32& P)
{
    constexpr u32 X = 8;
    u64 total = 0;
    std::vector<u32> order;
    for (u32 i = 0; i < (u32)descs.size(); ++i) {
        if (batch_queue_lattice(descs[i]) != queue) continue;
        order.push_back(i);
        total += (u64)descs[i].n_tiles * batch_fused_cost(descs[i]);
    }
    seq.clear();
    P = 0;
    if (order.empty()) return;
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) {
        return (u64)descs[a].n_tiles * batch_fused_cost(descs[a]) >
               (u64)descs[b].n_tiles * batch_fused_cost(descs[b]);
    });
    u64 load[X] = {};
    std::vector<int> home(descs.size(), -1);  // XCD of a whole lattice, -1 = split into bands
    for (u32 li : order) {
        const BatchDesc& d = descs[li];
        const u64 cost = (u64)d.n_tiles * batch_fused_cost(d);
        const bool split = cost * 2u * X > total && !d.dep_all && d.n_tiles / X >= 2u * d.dep_span + 2u;
        if (split) {
            for (u32 x = 0; x < X; ++x) load[x] += cost / X;
            continue;
        }
        u32 best = 0;
        for (u32 x = 1; x < X; ++x) best = load[x] < load[best] ? x : best;
        home[li] = (int)best;
        load[best] += cost;
    }
    std::vector<u32> by_tile(order);
    std::stable_sort(by_tile.begin(), by_tile.end(), [&](u32 a, u32 b) {
        return batch_fused_cost(descs[a]) > batch_fused_cost(descs[b]);
    });
    std::vector<u64> lists[X];
    auto pad = [&](std::vector<u64>& l) {
        while (l.size() % K) l.push_back(BATCH_NO_TILE);
    };
    for (u32 li : by_tile) {
        const BatchDesc& d = descs[li];
        if (home[li] >= 0) {
            for (u32 t = 0; t < d.n_tiles; ++t) lists[home[li]].push_back(((u64)li << 32) | (d.tile_begin + t));
            pad(lists[home[li]]);
            continue;
        }
        for (u32 b = 0; b < X; ++b) {
            const u32 t0 = (u32)((u64)d.n_tiles * b / X), t1 = (u32)((u64)d.n_tiles * (b + 1) / X);
            for (u32 t = t0; t < t1; ++t) {
                const bool edge = t < t0 + d.dep_span || t + d.dep_span >= t1;
                lists[b].push_back(((u64)li << 32) | (d.tile_begin + t) | (edge ? BF_EDGE : 0ull));
            }
            pad(lists[b]);
        }
    }
    // fold entries last: one per whole lattice on its XCD, one per band of a split lattice
    for (u32 li : order) {
        if (home[li] >= 0) {
            lists[home[li]].push_back(((u64)li << 32) | BF_FOLD);
            pad(lists[home[li]]);
        } else {
            for (u32 b = 0; b < X; ++b) {
                lists[b].push_back(((u64)li << 32) | BF_FOLD | BF_SPLIT | b);
                pad(lists[b]);
            }
        }
    }
    size_t longest = 0;
    for (u32 x = 0; x < X; ++x) longest = lists[x].size() > longest ? lists[x].size() : longest;
    P = (u32)longest;
    seq.assign((size_t)X * P, BATCH_NO_TILE);
    for (u32 x = 0; x < X; ++x)
        for (size_t i = 0; i < lists[x].size(); ++i) seq[(size_t)x * P + i] = lists[x][i];
}

// Capacity of the launch order for cap_lat lattices and cap_tiles tiles: each lattice can
// lengthen the longest list by at most one tile beyond its even share, and k tiles per
// workgroup pad each list by at most 7 more.
static u32 batch_seq_capacity(u32 cap_lat, u32 cap_tiles)
{
    return cap_tiles + 8u * (cap_lat + 8u);
}

// Allocates the device tables for cap_lat lattices and cap_tiles tiles and sets the grid
// parameters (synchronous).
static int batch_alloc(BatchHandle* h, u32 cap_lat, u32 cap_tiles, int blocks_per_cu)
{
    int dev = 0;
    ISING_HIP(hipGetDevice(&dev));
    ISING_HIP(hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    ISING_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&h->max_blocks, (const void*)batch_half<1>, 256, 0));
    if (h->max_blocks < 1) return ising_fail(ISING_EHIP, "batch: kernel does not fit on a CU");
    h->per_cu = (blocks_per_cu <= 0 || blocks_per_cu > h->max_blocks) ? h->max_blocks : blocks_per_cu;
    h->cap_lat = cap_lat;
    h->cap_tiles = cap_tiles;
    h->cap_seq = batch_seq_capacity(cap_lat, cap_tiles);
    ISING_HIP(hipMalloc((void**)&h->d_descs, sizeof(BatchDesc) * (size_t)cap_lat));
    ISING_HIP(hipMalloc((void**)&h->d_tile_lat, sizeof(u32) * (size_t)cap_tiles));
    ISING_HIP(hipMalloc((void**)&h->d_tile_seq, sizeof(u64) * (size_t)h->cap_seq));
    ISING_HIP(hipMalloc((void**)&h->d_plus, sizeof(u32) * 4 * BATCH_WAVES * (size_t)cap_tiles));
    ISING_HIP(hipMalloc((void**)&h->d_acc, sizeof(BatchAcc) * (size_t)cap_lat));
    ISING_HIP(hipMalloc((void**)&h->d_done, 2 * sizeof(unsigned long long)));
    return ISING_OK;
}

static void batch_set_sizes(BatchHandle* h, u32 n_lat, u32 tiles, u32 n_seq)
{
    h->n_lat = n_lat;
    h->total_tiles = tiles;
    h->n_seq = n_seq;
    h->grid = h->per_cu * h->n_cu;
    if ((u32)h->grid > tiles) h->grid = (int)tiles;
}

// Zero the accumulators of the first n_lat lattices and the sweep counter (kernels, so the
// work is stream-ordered and capturable).
static int batch_zero_stats(BatchHandle* h, hipStream_t st)
{
    if (h->n_lat)
        hipLaunchKernelGGL(batch_zero_u64, dim3(1), dim3(256), 0, st, (unsigned long long*)h->d_acc,
                           (u32)(sizeof(BatchAcc) / 8) * h->n_lat);
    hipLaunchKernelGGL(batch_zero_u64, dim3(1), dim3(64), 0, st, h->d_done, 2u);
    ISING_LAUNCHED("batch_zero_u64");
    return ISING_OK;
}

// Lists and sync words of the dependency-ordered form for the handle's lattices (synchronous;
// batches made by ising_batch_create).
// Builds and uploads the two kernels' lists for k tiles per workgroup (synchronous).
static int batch_fused_lists(BatchHandle* h, u32 k)
{
    std::vector<u64> light, queue;
    batch_fused_schedule(h->host_descs, false, k, light, h->f_P);
    batch_fused_schedule(h->host_descs, true, 1u, queue, h->f_Pq);
    light.insert(light.end(), queue.begin(), queue.end());
    if (h->d_fseq) ISING_HIP(hipFree(h->d_fseq));
    h->d_fseq = nullptr;
    ISING_HIP(hipMalloc((void**)&h->d_fseq, sizeof(u64) * light.size()));
    ISING_HIP(hipMemcpy(h->d_fseq, light.data(), sizeof(u64) * light.size(), hipMemcpyHostToDevice));
    return ISING_OK;
}

static int batch_fused_alloc(BatchHandle* h, const std::vector<BatchDesc>& descs, u32 tiles)
{
    const size_t n = descs.size();
    // header + counters, fold counters, slots (u32), split-lattice words (u64), in u64 units
    const size_t cnt_words = (BF_HDR + 5 * (size_t)tiles + 2 * n + 1) / 2;
    h->fsync_words = cnt_words + 2 * n;
    ISING_HIP(hipMalloc((void**)&h->d_fsync, 8 * h->fsync_words));
    h->d_fold = h->d_fsync + BF_HDR + tiles;
    h->d_slots = h->d_fold + 2 * n;
    h->d_pacc = (u64*)h->d_fsync + cnt_words;
    ISING_HIP(hipStreamCreateWithFlags(&h->f_side, hipStreamNonBlocking));
    ISING_HIP(hipEventCreateWithFlags(&h->f_fork, hipEventDisableTiming));
    ISING_HIP(hipEventCreateWithFlags(&h->f_join, hipEventDisableTiming));
    return batch_fused_lists(h, h->per_wg);
}

static void batch_zero_fused(BatchHandle* h, hipStream_t st)
{
    if (!h->d_fsync) return;
    const u32 words = (u32)h->fsync_words;
    const u32 blocks = words / 256u + 1u < 1024u ? words / 256u + 1u : 1024u;
    hipLaunchKernelGGL(batch_zero_u64, dim3(blocks), dim3(256), 0, st, (unsigned long long*)h->d_fsync, words);
}

static void batch_launch_fused_kernel(u32 K, bool mom, bool queue, dim3 grid, hipStream_t st,
                                      const BatchFusedArgs& f)
{
#define BF_LAUNCH(KK, SET)                                                                       \
    if (mom) hipLaunchKernelGGL((batch_fused<true, KK, SET>), grid, dim3(256), 0, st, f);       \
    else hipLaunchKernelGGL((batch_fused<false, KK, SET>), grid, dim3(256), 0, st, f)
    if (queue) {
        BF_LAUNCH(1, TILES_QUEUE);
        return;
    }
    switch (K) {
    case 2: BF_LAUNCH(2, TILES_LIGHT); break;
    case 4: BF_LAUNCH(4, TILES_LIGHT); break;
    case 8: BF_LAUNCH(8, TILES_LIGHT); break;
    default: BF_LAUNCH(1, TILES_LIGHT); break;
    }
#undef BF_LAUNCH
}

// A launch's half sweeps hold fewer than 2^32 work-items.
static u32 batch_fused_per_launch(u32 per_half)
{
    u32 per_launch = BF_MAX_SWEEPS;
    while (per_launch > 1u && 2ull * per_launch * per_half * 256u >= (1ull << 32)) per_launch >>= 1;
    return per_launch;
}

// Each chunk of sweeps: the light kernel on `st` and the queue kernel on the side stream, forked
// and joined with events (captured as two parallel graph nodes), then the counter check.
static int batch_launch_fused(BatchHandle* h, u32 n_sweeps, hipStream_t st)
{
    int place = -1;
    if (int rc = fused_placement(&place)) return rc;
    if (place < 0) return ising_fail(ISING_EINVAL, "batch: the dependency-ordered form needs round-robin or single-XCD placement");
    BatchFusedArgs f;
    f.descs = h->d_descs;
    f.sync = h->d_fsync;
    f.pacc = h->d_pacc;
    f.fold = h->d_fold;
    f.slots = h->d_slots;
    f.total_tiles = h->total_tiles;
    f.acc = h->d_acc;
    f.sweeps_done = h->d_done;
    f.one_l2 = place == 0 ? 1u : 0u;
    BatchFusedArgs fl = f, fq = f;
    fl.seq = h->d_fseq;
    fl.P = h->f_P;
    fl.per_half = 8u * (h->f_P / h->per_wg);
    fq.seq = h->d_fseq + 8ull * h->f_P;
    fq.P = h->f_Pq;
    fq.per_half = 8u * h->f_Pq;
    fl.div_half = make_fastdiv(fl.per_half ? fl.per_half : 1u);
    fq.div_half = make_fastdiv(fq.per_half ? fq.per_half : 1u);
    const u32 per_launch = batch_fused_per_launch(fl.per_half > fq.per_half ? fl.per_half : fq.per_half);
    if (2ull * per_launch * (fl.per_half > fq.per_half ? fl.per_half : fq.per_half) * 256u >= (1ull << 32))
        return ising_fail(ISING_ERANGE, "batch: too many tiles per half sweep for one launch");
    const bool moms = h->moments != 0;
    for (u32 done = 0; done < n_sweeps;) {
        const u32 k = n_sweeps - done < per_launch ? n_sweeps - done : per_launch;
        const bool both = fl.P && fq.P;
        if (both) {
            ISING_HIP(hipEventRecord(h->f_fork, st));
            ISING_HIP(hipStreamWaitEvent(h->f_side, h->f_fork, 0));
        }
        if (fq.P)
            batch_launch_fused_kernel(1u, moms, true, dim3(2u * k * fq.per_half), both ? h->f_side : st, fq);
        if (fl.P) batch_launch_fused_kernel(h->per_wg, moms, false, dim3(2u * k * fl.per_half), st, fl);
        if (both) {
            ISING_HIP(hipEventRecord(h->f_join, h->f_side));
            ISING_HIP(hipStreamWaitEvent(st, h->f_join, 0));
        }
        hipLaunchKernelGGL(batch_fused_check, dim3((h->total_tiles + 255u) / 256u), dim3(256), 0, st,
                           (const BatchDesc*)h->d_descs, (const u32*)h->d_tile_lat, h->total_tiles,
                           h->d_fsync, (const unsigned long long*)h->d_done, k);
        hipLaunchKernelGGL(batch_fused_advance, dim3(1), dim3(64), 0, st, h->d_done, k);
        done += k;
    }
    ISING_LAUNCHED("batch_fused");
    return ISING_OK;
}

static void batch_launch_half(u32 per_wg, dim3 grid, hipStream_t st, const BatchKernelArgs& ka,
                              u32 color, u32 acc_prev)
{
    switch (per_wg) {
    case 2: hipLaunchKernelGGL(batch_half<2>, grid, dim3(256), 0, st, ka, color, acc_prev); break;
    case 4: hipLaunchKernelGGL(batch_half<4>, grid, dim3(256), 0, st, ka, color, acc_prev); break;
    case 8: hipLaunchKernelGGL(batch_half<8>, grid, dim3(256), 0, st, ka, color, acc_prev); break;
    default: hipLaunchKernelGGL(batch_half<1>, grid, dim3(256), 0, st, ka, color, acc_prev); break;
    }
}

static int batch_launch(BatchHandle* h, u32 n_sweeps, hipStream_t st)
{
    if (h->total_tiles == 0) return ISING_OK;  // an empty reserved batch
    BatchKernelArgs ka;
    ka.descs = h->d_descs;
    ka.tile_lat = h->d_tile_lat;
    ka.tile_seq = h->d_tile_seq;
    ka.plus = h->d_plus;
    ka.moments = (u32)h->moments;
    ka.n_seq = h->n_seq;
    ka.per_wg = h->per_wg;
    ka.acc = h->d_acc;
    ka.sweeps_done = h->d_done;
    ka.n_lat = h->n_lat;
    ka.total_tiles = h->total_tiles;
    ka.n_sweeps = n_sweeps;
    ka.msc_mode = h->msc_mode;
    ka.msc_cap = h->msc_cap;
    if (h->persistent) return batch_launch_fused(h, n_sweeps, st);
    for (u32 s = 0; s < n_sweeps; ++s) {
        const u32 acc_prev = (h->moments && s > 0) ? 1u : 0u;
        batch_launch_half(h->per_wg, dim3(h->n_seq + (acc_prev ? h->n_lat : 0u)), st, ka, 0u,
                          acc_prev);
        batch_launch_half(h->per_wg, dim3(h->n_seq), st, ka, 1u, 0u);
    }
    // the last sweep's moments (earlier ones were folded into the next colour-0 launch)
    if (h->moments) hipLaunchKernelGGL(batch_accumulate, dim3(h->n_lat), dim3(256), 0, st, ka);
    ISING_LAUNCHED("batch_half");
    return ISING_OK;
}

static int batch_check_modes(int msc_mode, int msc_cap)
{
    if (msc_mode != 0 && msc_mode != 1) return ising_fail(ISING_EINVAL, "batch: msc_mode must be 0 or 1");
    if (msc_cap != 4) return ising_fail(ISING_EINVAL, "batch: msc_cap must be 4 (unrolled calls of the queue tiles)");
    return ISING_OK;
}

extern "C" {

int ising_batch_create(const IsingBatchLattice* lats, int n, int blocks_per_cu, int msc_mode,
                       int msc_cap, void** out)
{
    if (!lats || n <= 0 || n > (int)BF_LAT_MASK || !out) return ising_fail(ISING_EINVAL, "batch_create: bad argument");
    if (int rc = batch_check_modes(msc_mode, msc_cap)) return rc;
    *out = nullptr;
    std::vector<BatchDesc> descs;
    std::vector<u32> tile_lat;
    if (int rc = batch_build(lats, n, (u32)msc_mode, descs, tile_lat)) return rc;
    const u32 tiles = (u32)tile_lat.size();
    std::vector<u64> seq;
    batch_schedule(descs, 1u, seq);

    BatchHandle* h = new (std::nothrow) BatchHandle();  // no exception crosses the C ABI
    if (!h) return ising_fail(ISING_EINVAL, "batch: out of host memory");
    h->msc_mode = (u32)msc_mode;
    h->msc_cap = (u32)msc_cap;
    int rc = batch_alloc(h, (u32)n, tiles, blocks_per_cu);
    if (rc) {
        batch_free(h);
        return rc;
    }
    batch_set_sizes(h, (u32)n, tiles, (u32)seq.size());
    h->host_descs = descs;
    rc = batch_fused_alloc(h, descs, tiles);
    if (rc) {
        batch_free(h);
        return rc;
    }
    batch_zero_fused(h, 0);
    hipError_t e = hipMemcpy(h->d_descs, descs.data(), sizeof(BatchDesc) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_tile_lat, tile_lat.data(), sizeof(u32) * tile_lat.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_tile_seq, seq.data(), sizeof(u64) * seq.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) rc = batch_zero_stats(h, 0);
    if (e == hipSuccess && !rc) e = hipDeviceSynchronize();
    if (e != hipSuccess || rc) {
        batch_free(h);
        return rc ? rc : ising_check_hip(e, "batch_create");
    }
    *out = h;
    return ISING_OK;
}

// Number of tiles one half sweep of these lattices has (host only, no device call): the
// capacity a reserved batch needs for them.
int ising_batch_tiles(const IsingBatchLattice* lats, int n, int msc_mode, long long* tiles)
{
    if ((!lats && n) || n < 0 || !tiles) return ising_fail(ISING_EINVAL, "batch_tiles: bad argument");
    if (int rc = batch_check_modes(msc_mode, 4)) return rc;
    std::vector<BatchDesc> descs;
    std::vector<u32> tile_lat;
    if (int rc = batch_build(lats, n, (u32)msc_mode, descs, tile_lat)) return rc;
    *tiles = (long long)tile_lat.size();
    return ISING_OK;
}

// A batch with room for max_lattices lattices and max_tiles tiles and no lattices yet; fill it
// with ising_batch_assign. Allocates (synchronously) once; later assigns allocate nothing.
int ising_batch_reserve(int max_lattices, long long max_tiles, int blocks_per_cu, int msc_mode,
                        int msc_cap, void** out)
{
    if (!out || max_lattices <= 0 || max_tiles <= 0 || max_tiles > 0x7FFFFFFFll)
        return ising_fail(ISING_EINVAL, "batch_reserve: bad argument");
    if (int rc = batch_check_modes(msc_mode, msc_cap)) return rc;
    *out = nullptr;
    BatchHandle* h = new (std::nothrow) BatchHandle();  // no exception crosses the C ABI
    if (!h) return ising_fail(ISING_EINVAL, "batch: out of host memory");
    h->msc_mode = (u32)msc_mode;
    h->msc_cap = (u32)msc_cap;
    int rc = batch_alloc(h, (u32)max_lattices, (u32)max_tiles, blocks_per_cu);
    const size_t bytes = sizeof(BatchDesc) * (size_t)max_lattices + sizeof(u32) * (size_t)max_tiles +
                         8 + sizeof(u64) * (size_t)h->cap_seq;
    for (int s = 0; s < 2 && !rc; ++s) {
        hipError_t e = hipHostMalloc(&h->pinned[s], bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->pinned_done[s], hipEventDisableTiming);
        if (e != hipSuccess) rc = ising_check_hip(e, "batch_reserve");
    }
    batch_set_sizes(h, 0, 0, 0);
    if (!rc) rc = batch_zero_stats(h, 0);
    if (!rc) rc = ising_check_hip(hipDeviceSynchronize(), "batch_reserve");
    if (rc) {
        batch_free(h);
        return rc;
    }
    *out = h;
    return ISING_OK;
}

// Replace the lattices of a reserved batch (n may be 0). Stream-ordered on `stream`: launches
// queued there before this call still see the old list, launches after it the new one; the
// per-lattice accumulators and the batch sweep counter restart at zero (lattice i continues
// from its own t0). Allocates nothing; destroys a captured graph. ISING_ERANGE when the
// lattices exceed the reserved capacity.
int ising_batch_assign(void* handle, const IsingBatchLattice* lats, int n, void* stream)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || n < 0 || (n && !lats)) return ising_fail(ISING_EINVAL, "batch_assign: bad argument");
    if (!h->pinned[0]) return ising_fail(ISING_EINVAL, "batch_assign: handle was not reserved");
    if ((u32)n > h->cap_lat) return ising_fail(ISING_ERANGE, "batch_assign: more lattices than reserved");
    std::vector<BatchDesc> descs;
    std::vector<u32> tile_lat;
    if (int rc = batch_build(lats, n, h->msc_mode, descs, tile_lat)) return rc;
    if (tile_lat.size() > h->cap_tiles) return ising_fail(ISING_ERANGE, "batch_assign: more tiles than reserved");
    std::vector<u64> seq;
    batch_schedule(descs, h->per_wg, seq);
    if (seq.size() > h->cap_seq) return ising_fail(ISING_ERANGE, "batch_assign: launch order exceeds its capacity");
    hipStream_t st = as_stream(stream);
    const int s = h->next_slot;
    h->next_slot ^= 1;
    ISING_HIP(hipEventSynchronize(h->pinned_done[s]));  // this slot's previous upload is done
    char* stage = (char*)h->pinned[s];
    const size_t desc_bytes = sizeof(BatchDesc) * (size_t)n;
    const size_t map_bytes = sizeof(u32) * tile_lat.size();
    if (n) {
        memcpy(stage, descs.data(), desc_bytes);
        ISING_HIP(hipMemcpyAsync(h->d_descs, stage, desc_bytes, hipMemcpyHostToDevice, st));
    }
    if (map_bytes) {
        memcpy(stage + desc_bytes, tile_lat.data(), map_bytes);
        ISING_HIP(hipMemcpyAsync(h->d_tile_lat, stage + desc_bytes, map_bytes, hipMemcpyHostToDevice, st));
        const size_t seq_bytes = sizeof(u64) * seq.size();
        const size_t seq_at = (desc_bytes + map_bytes + 7u) & ~(size_t)7u;
        memcpy(stage + seq_at, seq.data(), seq_bytes);
        ISING_HIP(hipMemcpyAsync(h->d_tile_seq, stage + seq_at, seq_bytes, hipMemcpyHostToDevice, st));
    }
    ISING_HIP(hipEventRecord(h->pinned_done[s], st));
    batch_drop_graph(h);
    batch_set_sizes(h, (u32)n, (u32)tile_lat.size(), (u32)(seq.size() / h->per_wg));
    h->host_descs = std::move(descs);
    return batch_zero_stats(h, st);
}

// n_sweeps sweeps of every lattice in the batch's form (stream-ordered, asynchronous).
int ising_batch_sweeps(void* handle, long long n_sweeps, void* stream)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || n_sweeps < 0 || n_sweeps > 0x7FFFFFFFll) return ising_fail(ISING_EINVAL, "batch_sweeps: bad argument");
    if (n_sweeps == 0) return ISING_OK;
    return batch_launch(h, (u32)n_sweeps, as_stream(stream));
}

// Capture the launches of `sweeps_per_replay` sweeps into a hipGraph (stream capture); replays
// continue the device-side sweep counter.
int ising_batch_graph(void* handle, long long sweeps_per_replay, void* stream)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || sweeps_per_replay <= 0 || sweeps_per_replay > 0x7FFFFFFFll)
        return ising_fail(ISING_EINVAL, "batch_graph: bad argument");
    hipStream_t st = as_stream(stream);
    if (!st) return ising_fail(ISING_EINVAL, "batch_graph: capture needs a created stream");
    if (!h->total_tiles) return ising_fail(ISING_EINVAL, "batch_graph: the batch has no lattices");
    batch_drop_graph(h);
    ISING_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = batch_launch(h, (u32)sweeps_per_replay, st);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return ising_check_hip(e, "hipStreamEndCapture");
    h->graph = g;
    ISING_HIP(hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
    h->graph_sweeps = sweeps_per_replay;
    return ISING_OK;
}

int ising_batch_replay(void* handle, int replays, void* stream)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || !h->graph_exec || replays < 0) return ising_fail(ISING_EINVAL, "batch_replay: no graph");
    for (int r = 0; r < replays; ++r) ISING_HIP(hipGraphLaunch(h->graph_exec, as_stream(stream)));
    return ISING_OK;
}

// Accumulators: n x 7 int64 (sum M, sum |M|, sum M^2 lo, sum M^2 hi, sum M^4 lo, sum M^4 hi,
// samples) and the number of sweeps done. Synchronous.
int ising_batch_stats(void* handle, long long* out, long long* sweeps_done, void* stream)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || !out) return ising_fail(ISING_EINVAL, "batch_stats: bad argument");
    ISING_HIP(hipStreamSynchronize(as_stream(stream)));
    ISING_HIP(hipMemcpy(out, h->d_acc, sizeof(BatchAcc) * h->n_lat, hipMemcpyDeviceToHost));
    if (sweeps_done) {
        unsigned long long d = 0;
        ISING_HIP(hipMemcpy(&d, h->d_done, sizeof(d), hipMemcpyDeviceToHost));
        *sweeps_done = (long long)d;
    }
    return ISING_OK;
}

int ising_batch_reset_stats(void* handle, void* stream)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h) return ising_fail(ISING_EINVAL, "batch_reset_stats: bad handle");
    if (!h->n_lat) return ISING_OK;
    hipLaunchKernelGGL(batch_zero_u64, dim3(1), dim3(256), 0, as_stream(stream),
                       (unsigned long long*)h->d_acc, (u32)(sizeof(BatchAcc) / 8) * h->n_lat);
    ISING_LAUNCHED("batch_zero_u64");
    return ISING_OK;
}

// 1 = dependency-ordered launches (up to 512 sweeps per launch, tiles ordered by per-tile
// completion counters; batches made by ising_batch_create, on a device whose dispatches place
// workgroups round-robin over its XCDs or all on one), 0 = one launch per half sweep plus one
// accumulation launch per sweep (the default). Both give identical lattices and sums. Destroys
// a captured graph.
int ising_batch_set_persistent(void* handle, int persistent)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || (persistent != 0 && persistent != 1))
        return ising_fail(ISING_EINVAL, "batch_set_persistent: bad argument");
    if (persistent) {
        if (!h->d_fsync)
            return ising_fail(ISING_EINVAL, "batch_set_persistent: the dependency-ordered form needs a batch made by ising_batch_create");
        int place = -1;  // probed here, outside any stream capture
        if (int rc = fused_placement(&place)) return rc;
        if (place < 0)
            return ising_fail(ISING_EINVAL, "batch_set_persistent: this device places workgroups neither round-robin nor on one XCD");
    }
    batch_drop_graph(h);
    h->persistent = persistent;
    return ISING_OK;
}

// Failure bits of the dependency-ordered launches since the batch was made (synchronizes the
// stream): 1 = a tile's wait gave up, 2 = a tile did not publish its last half sweep. The
// lattices and sums are undefined after a failure.
int ising_batch_status(void* handle, int* failed, void* stream)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || !failed) return ising_fail(ISING_EINVAL, "batch_status: bad argument");
    *failed = 0;
    if (!h->d_fsync) return ISING_OK;
    u32 bits = 0;
    ISING_HIP(hipMemcpyAsync(&bits, h->d_fsync, sizeof(bits), hipMemcpyDeviceToHost, as_stream(stream)));
    ISING_HIP(hipStreamSynchronize(as_stream(stream)));
    *failed = (int)bits;
    return ISING_OK;
}

// 1 (default) = accumulate the moments of M after every sweep; 0 = skip that work (one launch
// per sweep fewer in the launch form, no per-tile atomics in the dependency-ordered form).
// Lattices are the same either way. Destroys a captured graph.
int ising_batch_set_moments(void* handle, int on)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || (on != 0 && on != 1)) return ising_fail(ISING_EINVAL, "batch_set_moments: bad argument");
    batch_drop_graph(h);
    h->moments = on;
    return ISING_OK;
}

// Each tile workgroup runs k (1, 2, 4 or 8) consecutive tiles of its XCD's list in sequence
// (fewer, longer-lived workgroups; one tile each by default); in the dependency-ordered form
// the k tiles are consecutive tiles of one lattice and share one wait. Rebuilds and uploads the
// launch orders after a device synchronisation (no launch may be in flight on any stream);
// results do not depend on k. Destroys a captured graph.
int ising_batch_set_tiles_per_workgroup(void* handle, int k)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h || (k != 1 && k != 2 && k != 4 && k != 8))
        return ising_fail(ISING_EINVAL, "batch_set_tiles_per_workgroup: k must be 1, 2, 4 or 8");
    if ((u32)k == h->per_wg) return ISING_OK;
    batch_drop_graph(h);
    std::vector<u64> seq;
    if (h->n_lat) {
        batch_schedule(h->host_descs, (u32)k, seq);
        if (seq.size() > h->cap_seq)
            return ising_fail(ISING_ERANGE, "batch_set_tiles_per_workgroup: launch order exceeds its capacity");
        ISING_HIP(hipDeviceSynchronize());  // launches queued earlier read the old order
        ISING_HIP(hipMemcpy(h->d_tile_seq, seq.data(), sizeof(u64) * seq.size(), hipMemcpyHostToDevice));
    }
    h->per_wg = (u32)k;
    if (h->n_lat) batch_set_sizes(h, h->n_lat, h->total_tiles, (u32)(seq.size() / (u32)k));
    if (h->d_fseq) return batch_fused_lists(h, (u32)k);  // the dependency-ordered form's lists
    return ISING_OK;
}

int ising_batch_info(void* handle, int* grid, int* total_tiles, int* max_blocks_per_cu)
{
    BatchHandle* h = (BatchHandle*)handle;
    if (!h) return ising_fail(ISING_EINVAL, "batch_info: bad handle");
    if (grid) *grid = h->grid;
    if (total_tiles) *total_tiles = (int)h->total_tiles;
    if (max_blocks_per_cu) *max_blocks_per_cu = h->max_blocks;
    return ISING_OK;
}

int ising_batch_destroy(void* handle)
{
    batch_free((BatchHandle*)handle);
    return ISING_OK;
}

}  // extern "C"
