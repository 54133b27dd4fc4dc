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
// (src/pyising/philox.py:67-78). The comparison is decided MSB-first and stops as soon as no
// lane of the word is undecided, so the number of Philox calls per word is data dependent
// (about 3.8 on average at T_c, up to 16).
//
// Two kernels produce identical words:
//  * msc_half_lockstep: one word per lane; a wave pays the maximum call count of its 64 words.
//  * msc_half_queue: a wave owns a tile of 64*K words. Phase 1 computes (a0, a1, own ^ a2) of
//    every word with coalesced loads into LDS. Phase 2 is a wave-local work queue: a lane whose
//    word is decided takes the next undecided word (ballot + mbcnt rank, wave-uniform queue
//    head in a scalar register), so every lane issues a useful Philox call on every iteration
//    until the tile runs dry; the cost is the tile's mean, not its per-wave maximum. Phase 3
//    writes the tile back with coalesced stores. Any lane may evaluate any (word, level) call
//    because the stream is counter-based.
#include "common.hpp"

struct MscArgs {
    u64* own;
    const u64* other;
    u32 rows, nw;
    u32 n_words;  // rows * nw (< 2^32, checked on the host)
    u32 t, color, k0, k1, thr4, thr8;
    FastDiv div_nw;
};

struct MscWord {
    u64 o, a0, a1, a2;
};

// Neighbour classification of word w (global index), loads through the cache hierarchy.
__device__ __forceinline__ MscWord msc_classify(const MscArgs& a, u32 w)
{
    const u32 y = fast_div(w, a.div_nw);
    const u32 x = w - y * a.nw;
    const u64 rbase = (u64)y * a.nw;
    const u32 p = (y + a.color) & 1u;
    const u64 o = a.own[w];
    const u64 c = a.other[w];
    const u32 xn = p ? ((x + 1 == a.nw) ? 0u : x + 1) : ((x == 0) ? a.nw - 1 : x - 1);
    const u64 nb = a.other[rbase + xn];
    const u64 hz = p ? ((c >> 1) | (nb << 63)) : ((c << 1) | (nb >> 63));
    const u32 yu = (y + 1 == a.rows) ? 0u : y + 1;
    const u32 yd = (y == 0) ? a.rows - 1 : y - 1;
    const u64 up = a.other[(u64)yu * a.nw + x];
    const u64 dn = a.other[(u64)yd * a.nw + x];
    const u64 x1 = o ^ c, x2 = o ^ hz, x3 = o ^ up, x4 = o ^ dn;
    const u64 anyx = x1 | x2 | x3 | x4;
    MscWord r;
    r.o = o;
    r.a2 = (x1 & x2) | (x3 & x4) | ((x1 ^ x2) & (x3 ^ x4));
    r.a0 = ~anyx;
    r.a1 = anyx & ~r.a2;
    return r;
}

// Two lazy comparison levels from one Philox call (levels 2 lp and 2 lp + 1).
__device__ __forceinline__ void msc_lazy_pair(const PhiloxOut& r, u32 lp, u32 thr4, u32 thr8,
                                              u64 a0, u64 a1, u64& acc, u64& und)
{
    const u64 lv0 = (u64)r.x | ((u64)r.y << 32);
    const u64 lv1 = (u64)r.z | ((u64)r.w << 32);
    const u32 b0 = 31u - 2u * lp;
    const u64 t4a = ((thr4 >> b0) & 1u) ? ~0ull : 0ull;
    const u64 t8a = ((thr8 >> b0) & 1u) ? ~0ull : 0ull;
    const u64 tma = (a1 & t4a) | (a0 & t8a);
    acc |= und & ~lv0 & tma;
    und &= ~(lv0 ^ tma);
    const u32 b1 = b0 - 1u;
    const u64 t4b = ((thr4 >> b1) & 1u) ? ~0ull : 0ull;
    const u64 t8b = ((thr8 >> b1) & 1u) ? ~0ull : 0ull;
    const u64 tmb = (a1 & t4b) | (a0 & t8b);
    acc |= und & ~lv1 & tmb;
    und &= ~(lv1 ^ tmb);
}

// stats (optional, diagnostics only): [0] += Philox calls issued for real words,
// [1] += 64 * per-wave iterations (the lane-slots the SIMD executes), [2] += words.
template <bool LAZY>
__global__ __launch_bounds__(256) void msc_half_lockstep(MscArgs a, unsigned long long* stats)
{
    const u32 blk = xcd_band_block(blockIdx.x, gridDim.x);
    const u32 w = blk * blockDim.x + threadIdx.x;
    const bool valid = w < a.n_words;
    u32 calls = 0;
    if (valid) {
        const MscWord m = msc_classify(a, w);
        u64 acc = 0, und = m.a0 | m.a1;
        for (u32 lp = 0; lp < 16u && (!LAZY || und); ++lp) {
            const PhiloxOut r = philox4x32_10(w, a.t, a.color, 1u + lp, a.k0, a.k1);
            ++calls;
            msc_lazy_pair(r, lp, a.thr4, a.thr8, m.a0, m.a1, acc, und);
        }
        a.own[w] = m.o ^ (m.a2 | acc);
    }
    if (stats) {
        u32 wmax = calls;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) wmax = max(wmax, (u32)__shfl_xor((int)wmax, off, 64));
        u64 tot = calls;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) tot += __shfl_xor(tot, off, 64);
        const u64 nval = __popcll(__ballot(valid));
        if ((threadIdx.x & 63) == 0 && nval) {
            atomicAdd(stats + 0, (unsigned long long)tot);
            atomicAdd(stats + 1, (unsigned long long)wmax * 64ull);
            atomicAdd(stats + 2, (unsigned long long)nval);
        }
    }
}

__device__ __forceinline__ u32 lane_rank(u64 mask)  // set bits of mask below this lane
{
    return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

template <int K>
__global__ __launch_bounds__(256) void msc_half_queue(MscArgs a, u32 n_tiles,
                                                     unsigned long long* stats)
{
    constexpr u32 T = 64u * K;  // words per wave tile
    __shared__ u64 lds[4][3][T];
    const u32 wave = threadIdx.x >> 6;
    const u32 lane = threadIdx.x & 63u;
    u64* L0 = lds[wave][0];
    u64* L1 = lds[wave][1];
    u64* LP = lds[wave][2];
    const u32 blk = xcd_band_block(blockIdx.x, gridDim.x);
    const u32 tile = blk * 4u + wave;
    const bool tile_ok = tile < n_tiles;
    const u32 base = tile * T;

    // phase 1: classify every word of the tile (coalesced: word base + i * 64 + lane)
    if (tile_ok) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const u32 s = (u32)i * 64u + lane;
            const u32 w = base + s;
            if (w < a.n_words) {
                const MscWord m = msc_classify(a, w);
                L0[s] = m.a0;
                L1[s] = m.a1;
                LP[s] = m.o ^ m.a2;
            } else {
                L0[s] = 0;
                L1[s] = 0;
            }
        }
    }
    __syncthreads();

    // phase 2: wave-local queue over the tile's undecided words
    u64 iters = 0, calls = 0;
    if (tile_ok) {
        u32 head = 0;  // wave-uniform queue head
        u32 cur = 0, lp = 0;
        u64 a0 = 0, a1 = 0, acc = 0, und = 0;
        bool live = true;
        const u32 wbase = base;  // Philox word counter
        while (true) {
            bool need = live && (und == 0);
            u64 m = __ballot(need);
            while (m) {
                const u32 pos = head + lane_rank(m);
                head += (u32)__popcll(m);
                if (need) {
                    if (pos < T) {
                        cur = pos;
                        a0 = L0[pos];
                        a1 = L1[pos];
                        und = a0 | a1;
                        acc = 0;
                        lp = 0;
                        need = (und == 0);  // word decided without random numbers
                    } else {
                        live = false;
                        need = false;
                    }
                }
                m = __ballot(need);
            }
            if (__ballot(und != 0) == 0) break;
            ++iters;
            if (und) {
                const PhiloxOut r = philox4x32_10(wbase + cur, a.t, a.color, 1u + lp, a.k0, a.k1);
                ++calls;
                msc_lazy_pair(r, lp, a.thr4, a.thr8, a0, a1, acc, und);
                ++lp;
                if (lp == 16u) und = 0;  // all 32 bits compared: u == thr, not accepted
                if (und == 0) LP[cur] ^= acc;
            }
        }
    }
    __syncthreads();

    // phase 3: coalesced write-back
    if (tile_ok) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const u32 s = (u32)i * 64u + lane;
            const u32 w = base + s;
            if (w < a.n_words) a.own[w] = LP[s];
        }
    }
    if (stats && tile_ok) {
        u64 tot = calls;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) tot += __shfl_xor(tot, off, 64);
        if (lane == 0) {
            const u32 nw = (a.n_words - base < T) ? (a.n_words - base) : T;
            atomicAdd(stats + 0, (unsigned long long)tot);
            atomicAdd(stats + 1, (unsigned long long)iters * 64ull);
            atomicAdd(stats + 2, (unsigned long long)nw);
        }
    }
}

// Two lazy levels with per-lane level index lp (the work-queue kernels): the threshold bit of
// level b is a sign-extended one-bit field, 0 or ~0, of the scalar threshold.
__device__ __forceinline__ void msc_lazy_pair_lane(const PhiloxOut& r, u32 lp, u32 thr4, u32 thr8,
                                                   u64 a0, u64 a1, u64& acc, u64& und)
{
    const u32 b0 = 31u - 2u * lp;
    const u32 lvs[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int lev = 0; lev < 2; ++lev) {
        const u32 b = b0 - (u32)lev;
        const u32 t4 = (u32)__builtin_amdgcn_sbfe((int)thr4, b, 1);
        const u32 t8 = (u32)__builtin_amdgcn_sbfe((int)thr8, b, 1);
        u32 ac[2] = {(u32)acc, (u32)(acc >> 32)};
        u32 un[2] = {(u32)und, (u32)(und >> 32)};
        const u32 x0[2] = {(u32)a0, (u32)(a0 >> 32)};
        const u32 x1[2] = {(u32)a1, (u32)(a1 >> 32)};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const u32 lv = lvs[2 * lev + h];
            const u32 tm = (x1[h] & t4) | (x0[h] & t8);
            ac[h] |= un[h] & ~lv & tm;
            un[h] &= ~(lv ^ tm);
        }
        acc = (u64)ac[0] | ((u64)ac[1] << 32);
        und = (u64)un[0] | ((u64)un[1] << 32);
    }
}

// Streaming work queue. Each wave owns `groups` consecutive groups of 64 words. A 128-slot
// ring in LDS holds classified, unclaimed words (a0, a1, own ^ a2); slot counter c maps to
// word wbase + c, so no index is stored. A lane claims the next slot when its word is
// decided (ballot + mbcnt rank), copies the slot into registers (the slot is free at once),
// and stores the finished word straight to `own`. The ring is topped up with the next group
// (coalesced loads by all 64 lanes) whenever it holds <= 64 words, so lanes only idle in the
// final drain of the wave's range. 3 KB of LDS per wave keeps 8 waves per SIMD resident.
__global__ __launch_bounds__(256) void msc_half_stream(MscArgs a, u32 n_groups, u32 groups,
                                                      unsigned long long* stats)
{
    __shared__ u64 ring[4][3][128];
    const u32 wave = threadIdx.x >> 6;
    const u32 lane = threadIdx.x & 63u;
    u64* R0 = ring[wave][0];
    u64* R1 = ring[wave][1];
    u64* RP = ring[wave][2];
    const u32 blk = xcd_band_block(blockIdx.x, gridDim.x);
    const u32 wid = blk * 4u + wave;
    const u32 g_begin = wid * groups;
    if (g_begin >= n_groups) return;  // whole wave exits; no block barriers below
    const u32 g_end = min(g_begin + groups, n_groups);
    const u32 wbase = g_begin * 64u;

    u32 gnext = g_begin;
    u32 head = 0, tail = 0;  // wave-uniform ring counters
    bool has = false;
    u32 word = 0, lp = 0;
    u64 a0 = 0, a1 = 0, pw = 0, acc = 0, und = 0;
    u64 iters = 0, calls = 0;

    while (true) {
        // claim words for idle lanes, topping the ring up when it runs dry
        while (true) {
            const bool need = !has;
            const u64 m = __ballot(need);
            if (m == 0) break;
            if (tail == head) {
                if (gnext >= g_end) break;
                const u32 w = gnext * 64u + lane;
                const u32 s = tail & 127u;
                if (w < a.n_words) {
                    const MscWord cw = msc_classify(a, w);
                    R0[s] = cw.a0;
                    R1[s] = cw.a1;
                    RP[s] = cw.o ^ cw.a2;
                } else {
                    R0[s] = 0;
                    R1[s] = 0;
                }
                __builtin_amdgcn_wave_barrier();
                tail += 64u;
                ++gnext;
            }
            const u32 avail = tail - head;
            const u32 rank = lane_rank(m);
            if (need && rank < avail) {
                const u32 c = head + rank;
                const u32 s = c & 127u;
                word = wbase + c;
                a0 = R0[s];
                a1 = R1[s];
                pw = RP[s];
                und = a0 | a1;
                acc = 0;
                lp = 0;
                has = true;
                if (und == 0) {  // decided without random numbers
                    if (word < a.n_words) a.own[word] = pw;
                    has = false;
                }
            }
            const u32 taken = (u32)__popcll(m);
            head += taken < avail ? taken : avail;
        }
        // keep one group queued ahead
        if (tail - head <= 64u && gnext < g_end) {
            const u32 w = gnext * 64u + lane;
            const u32 s = tail & 127u;
            if (w < a.n_words) {
                const MscWord cw = msc_classify(a, w);
                R0[s] = cw.a0;
                R1[s] = cw.a1;
                RP[s] = cw.o ^ cw.a2;
            } else {
                R0[s] = 0;
                R1[s] = 0;
            }
            __builtin_amdgcn_wave_barrier();
            tail += 64u;
            ++gnext;
        }
        if (__ballot(has) == 0) break;
        ++iters;
        if (has) {
            const PhiloxOut r = philox4x32_10(word, a.t, a.color, 1u + lp, a.k0, a.k1);
            ++calls;
            msc_lazy_pair_lane(r, lp, a.thr4, a.thr8, a0, a1, acc, und);
            ++lp;
            if (lp == 16u) und = 0;  // all 32 bits compared: u == thr, not accepted
            if (und == 0) {
                a.own[word] = pw ^ acc;
                has = false;
            }
        }
    }
    if (stats) {
        u64 tot = calls;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) tot += __shfl_xor(tot, off, 64);
        if (lane == 0) {
            const u32 wend = min(g_end * 64u, a.n_words);
            atomicAdd(stats + 0, (unsigned long long)tot);
            atomicAdd(stats + 1, (unsigned long long)iters * 64ull);
            atomicAdd(stats + 2, (unsigned long long)(wend - wbase));
        }
    }
}

// ------------------------------------------------------------------------ observables ----
// Colour-0 word s with other-colour neighbours (c, hz, up, dn): ones += popc(s) + popc(c),
// anti += popc(s ^ nb) over the four; M = 2 ones - N, E = 2 anti - 2 N
// (src/pyising/backends/observables_torch.py:85-111).
__global__ __launch_bounds__(256) void msc_obs_kernel(const u64* __restrict__ w0,
                                                      const u64* __restrict__ w1, u32 rows, u32 nw,
                                                      u32 n_words, FastDiv div_nw,
                                                      unsigned long long* acc)
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
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        ones += __shfl_xor(ones, off, 64);
        anti += __shfl_xor(anti, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(acc + 0, (unsigned long long)ones);
        atomicAdd(acc + 1, (unsigned long long)anti);
    }
}

__global__ void msc_obs_finalize(unsigned long long* acc, long long* out, long long n_spins)
{
    const long long ones = (long long)acc[0], anti = (long long)acc[1];
    out[0] = 2 * ones - n_spins;
    out[1] = 2 * anti - 2 * n_spins;
    acc[0] = 0;
    acc[1] = 0;
}

__global__ void msc_zero2(unsigned long long* p)
{
    if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------- host launchers ----
// Kernel modes: 0 = tile work queue (tile of 64*K words per wave, K in {2,4,8,16}),
// 1 = lockstep (one word per lane), 2 = eager lockstep (all 16 calls, ablation),
// 3 = streaming work queue (K = groups of 64 words per wave, 1..64).
static int g_msc_mode = 0;
static int g_msc_k = 8;

static int msc_check(const void* own, const void* other, long long rows, int nw)
{
    if (!own || !other) return ising_fail(ISING_EINVAL, "msc: null pointer");
    if (rows < 2 || (rows & 1) || nw < 1) return ising_fail(ISING_EINVAL, "msc: bad geometry");
    if ((unsigned long long)rows * (unsigned long long)nw + 65536ull > 0xFFFFFFFFull)
        return ising_fail(ISING_ERANGE, "msc: rows * Lx/128 must be < 2^32 words per colour");
    return ISING_OK;
}

static int msc_mode_ok(int mode, int k)
{
    if (mode < 0) return ISING_OK;  // process default
    if (mode > 3)
        return ising_fail(ISING_EINVAL, "msc: mode must be 0 (tile queue), 1 (lockstep), 2 (eager) or 3 (stream)");
    if (mode == 0 && k != 2 && k != 4 && k != 8 && k != 16)
        return ising_fail(ISING_EINVAL, "msc: tile K must be 2, 4, 8 or 16");
    if (mode == 3 && (k < 1 || k > 64))
        return ising_fail(ISING_EINVAL, "msc: stream groups per wave must be in 1..64");
    return ISING_OK;
}

static int msc_half_impl(u64* own, const u64* other, long long rows, int nw, u32 t, u32 color,
                         u32 k0, u32 k1, u32 thr4, u32 thr8, int mode, int k,
                         unsigned long long* stats, hipStream_t st)
{
    int rc = msc_check(own, other, rows, nw);
    if (rc) return rc;
    if (color > 1) return ising_fail(ISING_EINVAL, "msc_half: colour must be 0 or 1");
    if ((rc = msc_mode_ok(mode, k))) return rc;
    if (mode < 0) {
        mode = g_msc_mode;
        k = g_msc_k;
    }
    MscArgs a;
    a.own = own;
    a.other = other;
    a.rows = (u32)rows;
    a.nw = (u32)nw;
    a.n_words = (u32)(rows * nw);
    a.div_nw = make_fastdiv((u32)nw);
    a.t = t;
    a.color = color;
    a.k0 = k0;
    a.k1 = k1;
    a.thr4 = thr4;
    a.thr8 = thr8;
    if (mode == 3) {
        const u32 n_groups = (a.n_words + 63u) / 64u;
        const u32 waves = (n_groups + (u32)k - 1u) / (u32)k;
        const u32 blocks = (waves + 3u) / 4u;
        hipLaunchKernelGGL(msc_half_stream, dim3(blocks), dim3(256), 0, st, a, n_groups, (u32)k, stats);
    } else if (mode != 0) {
        const u64 blocks = (a.n_words + 255) / 256;
        if (blocks > 0x7FFFFFFFull) return ising_fail(ISING_ERANGE, "msc: lattice too large");
        if (mode == 1)
            hipLaunchKernelGGL(msc_half_lockstep<true>, dim3((unsigned)blocks), dim3(256), 0, st, a, stats);
        else
            hipLaunchKernelGGL(msc_half_lockstep<false>, dim3((unsigned)blocks), dim3(256), 0, st, a, stats);
    } else {
        const u64 T = 64ull * (u64)k;
        const u64 n_tiles = (a.n_words + T - 1) / T;
        const u64 blocks = (n_tiles + 3) / 4;
        if (n_tiles > 0xFFFFFFFFull) return ising_fail(ISING_ERANGE, "msc: lattice too large");
        const dim3 grid((unsigned)blocks);
        switch (k) {
        case 2: hipLaunchKernelGGL(msc_half_queue<2>, grid, dim3(256), 0, st, a, (u32)n_tiles, stats); break;
        case 4: hipLaunchKernelGGL(msc_half_queue<4>, grid, dim3(256), 0, st, a, (u32)n_tiles, stats); break;
        case 16: hipLaunchKernelGGL(msc_half_queue<16>, grid, dim3(256), 0, st, a, (u32)n_tiles, stats); break;
        default: hipLaunchKernelGGL(msc_half_queue<8>, grid, dim3(256), 0, st, a, (u32)n_tiles, stats); break;
        }
    }
    ISING_LAUNCHED("msc_half");
    return ISING_OK;
}

static int msc_sweeps_impl(u64* w0, u64* w1, long long rows, int nw, u32 t0, long long n_sweeps,
                           u32 k0, u32 k1, u32 thr4, u32 thr8, int mode, int k,
                           unsigned long long* stats, hipStream_t st)
{
    if (n_sweeps < 0) return ising_fail(ISING_EINVAL, "msc_sweeps: negative sweep count");
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
    g_msc_mode = mode;
    g_msc_k = (mode == 0 || mode == 3) ? k : g_msc_k;
    return ISING_OK;
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
    hipLaunchKernelGGL(msc_zero2, dim3(1), dim3(64), 0, st, scratch2);
    const u64 blocks = (n_words + 255) / 256;
    hipLaunchKernelGGL(msc_obs_kernel, dim3((unsigned)(blocks < 2048 ? blocks : 2048)), dim3(256), 0,
                       st, w0, w1, (u32)rows, (u32)nw, (u32)n_words, make_fastdiv((u32)nw), scratch2);
    hipLaunchKernelGGL(msc_obs_finalize, dim3(1), dim3(1), 0, st, scratch2, out2,
                       (long long)n_words * 128);
    ISING_LAUNCHED("msc_obs_kernel");
    return ISING_OK;
}

}  // extern "C"
