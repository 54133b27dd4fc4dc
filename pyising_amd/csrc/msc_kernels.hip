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
// Cost model (measured with SQ counters on MI355X): Philox4x32-10 is 36 VALU ops, every op
// including v_mad_u64_u32 at the full wave64 rate, so the kernel is VALU-issue bound and its
// time is (lane-slots per word) x (ops per slot). Three things keep both factors low:
//  * the level index lp is wave-uniform everywhere, so the threshold bits of a level are
//    known to the whole wave: the host packs the 4-bit case code of every level pair into one
//    64-bit argument and each call costs one scalar switch into straight-line v_bitop3 code
//    (0-3 ops per 32-bit half and level) instead of a masked general update;
//  * the loop is wave-uniform (no per-lane exit), so there is no exec-mask bookkeeping; a
//    decided lane's updates are no-ops;
//  * survivor compaction (mode 0): lockstep pays the slowest of 64 words; after `cap` levels
//    only ~18% of words are undecided at T_c, so those are regrouped through LDS into full
//    waves that start at the same level (still uniform) and run to completion.
#include "common.hpp"

struct MscArgs {
    u64* own;
    const u64* other;
    u32 rows, nw;
    u32 n_words;  // rows * nw (< 2^32, checked on the host)
    u32 t, color, k0, k1;
    FastDiv div_nw;
};

// Lazy state of one word as 32-bit halves: a1 (lanes with exactly one anti-aligned
// neighbour), und (undecided lanes, a subset of a0 | a1, so a0 = ~a1 on und) and pw (the word
// being built: own ^ a2 ^ lanes accepted so far; accepted sets of successive levels are
// disjoint, so they are XORed in).
struct MscState {
    u32 a1l, a1h, pwl, pwh, ul, uh;
};

// Neighbour classification of word w: five loads through the cache hierarchy and bit-sliced
// counting of anti-aligned neighbours.
__device__ __forceinline__ MscState msc_classify(const MscArgs& a, u32 w)
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
    const u64 a2 = (x1 & x2) | (x3 & x4) | ((x1 ^ x2) & (x3 ^ x4));
    const u64 a1 = anyx & ~a2;
    const u64 und = ~a2;  // a0 | a1
    const u64 pw = o ^ a2;
    MscState s;
    s.a1l = (u32)a1;
    s.a1h = (u32)(a1 >> 32);
    s.pwl = (u32)pw;
    s.pwh = (u32)(pw >> 32);
    s.ul = (u32)und;
    s.uh = (u32)(und >> 32);
    return s;
}

// ------------------------------------------------------------------ lazy comparison ----
// Three-input boolean functions as single v_bitop3_b32 ops; the immediate is the function
// evaluated on the operand patterns (0xF0, 0xCC, 0xAA).
template <typename F>
constexpr u32 b3imm(F f)
{
    return f(0xF0u, 0xCCu, 0xAAu) & 0xFFu;
}
constexpr u32 B3_A_ANDN_B = b3imm([](u32 x, u32 y, u32) { return x & ~y; });
constexpr u32 B3_A_XOR_B_ANDN_C = b3imm([](u32 x, u32 y, u32 z) { return x ^ (y & ~z); });
constexpr u32 B3_A_B_ANDN_C = b3imm([](u32 x, u32 y, u32 z) { return x & y & ~z; });
constexpr u32 B3_A_XNOR_BC = b3imm([](u32 x, u32 y, u32 z) { return x & ~(y ^ z); });
constexpr u32 B3_A_ANDN_B_ANDN_C = b3imm([](u32 x, u32 y, u32 z) { return x & ~y & ~z; });
constexpr u32 B3_A_AND_B_XOR_C = b3imm([](u32 x, u32 y, u32 z) { return x & (y ^ z); });
static_assert(B3_A_ANDN_B == 0x30u && B3_A_XOR_B_ANDN_C == 0xB4u && B3_A_B_ANDN_C == 0x40u &&
                  B3_A_XNOR_BC == 0x90u && B3_A_ANDN_B_ANDN_C == 0x10u && B3_A_AND_B_XOR_C == 0x60u,
              "bitop3 truth tables");

#define bop3(a, b, c, imm) ((u32)__builtin_amdgcn_bitop3_b32((a), (b), (c), (imm)))

// One level on a 32-bit half for threshold bits (B4 of thr4, B8 of thr8): a1 lanes compare
// their uniform's bit with B4, a0 lanes with B8; a lane whose bit differs is decided (u bit 0
// under a threshold bit 1 accepts, the other way rejects).
template <u32 B4, u32 B8>
__device__ __forceinline__ void msc_half_level(u32 lv, u32 a1, u32& pw, u32& und)
{
    if constexpr (B4 && B8) {  // every undecided lane sees a 1 bit
        pw = bop3(pw, und, lv, B3_A_XOR_B_ANDN_C);
        und &= lv;
    } else if constexpr (B4) {  // a1 lanes see 1, a0 lanes 0
        pw ^= bop3(und, a1, lv, B3_A_B_ANDN_C);
        und = bop3(und, lv, a1, B3_A_XNOR_BC);
    } else if constexpr (B8) {  // a0 lanes (und & ~a1) see 1, a1 lanes 0
        pw ^= bop3(und, a1, lv, B3_A_ANDN_B_ANDN_C);
        und = bop3(und, lv, a1, B3_A_AND_B_XOR_C);
    } else {  // every undecided lane sees a 0 bit: a 1 in u rejects
        und = bop3(und, lv, 0u, B3_A_ANDN_B);
    }
}

// Both levels of one Philox call. CODE: bit 0 / 1 = thr4 / thr8 bit of level 2 lp, bits 2 / 3
// the same for level 2 lp + 1.
template <u32 CODE>
__device__ __forceinline__ void msc_pair(const PhiloxOut& r, MscState& s)
{
    msc_half_level<CODE & 1u, (CODE >> 1) & 1u>(r.x, s.a1l, s.pwl, s.ul);
    msc_half_level<CODE & 1u, (CODE >> 1) & 1u>(r.y, s.a1h, s.pwh, s.uh);
    msc_half_level<(CODE >> 2) & 1u, (CODE >> 3) & 1u>(r.z, s.a1l, s.pwl, s.ul);
    msc_half_level<(CODE >> 2) & 1u, (CODE >> 3) & 1u>(r.w, s.a1h, s.pwh, s.uh);
}

__device__ __forceinline__ void msc_pair_switch(u32 code, const PhiloxOut& r, MscState& s)
{
    switch (code) {  // code is wave-uniform: a scalar branch tree
    case 0: msc_pair<0>(r, s); break;
    case 1: msc_pair<1>(r, s); break;
    case 2: msc_pair<2>(r, s); break;
    case 3: msc_pair<3>(r, s); break;
    case 4: msc_pair<4>(r, s); break;
    case 5: msc_pair<5>(r, s); break;
    case 6: msc_pair<6>(r, s); break;
    case 7: msc_pair<7>(r, s); break;
    case 8: msc_pair<8>(r, s); break;
    case 9: msc_pair<9>(r, s); break;
    case 10: msc_pair<10>(r, s); break;
    case 11: msc_pair<11>(r, s); break;
    case 12: msc_pair<12>(r, s); break;
    case 13: msc_pair<13>(r, s); break;
    case 14: msc_pair<14>(r, s); break;
    default: msc_pair<15>(r, s); break;
    }
}

// Nibble lp of the result = case code of levels 2 lp and 2 lp + 1 (host side, once per call).
static inline u64 msc_level_codes(u32 thr4, u32 thr8)
{
    u64 codes = 0;
    for (u32 lp = 0; lp < 16; ++lp) {
        const u32 b0 = 31u - 2u * lp, b1 = b0 - 1u;
        const u64 c = ((thr4 >> b0) & 1u) | (((thr8 >> b0) & 1u) << 1) |
                      (((thr4 >> b1) & 1u) << 2) | (((thr8 >> b1) & 1u) << 3);
        codes |= c << (4u * lp);
    }
    return codes;
}

// Wave-uniform lockstep loop over levels lp .. lp_end-1. LAZY stops once every lane of the
// wave is decided; the eager form always runs to lp_end. Returns the level reached.
template <bool LAZY, bool STATS>
__device__ __forceinline__ u32 msc_run(const MscArgs& a, u64 codes, u32 w, u32 lp, u32 lp_end,
                                       MscState& s, u32& calls, u32& slots)
{
    while (lp < lp_end && (!LAZY || __ballot((s.ul | s.uh) != 0u) != 0)) {
        if (STATS) {
            calls += (!LAZY || (s.ul | s.uh) != 0u) ? 1u : 0u;
            ++slots;
        }
        const PhiloxOut r = philox4x32_10(w, a.t, a.color, 1u + lp, a.k0, a.k1);
        msc_pair_switch((u32)(codes >> (4u * lp)) & 15u, r, s);
        ++lp;
    }
    return lp;
}

__device__ __forceinline__ void msc_stats(unsigned long long* stats, u32 calls, u32 slots,
                                          u32 words)
{
    u64 tot = calls, nwd = words;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        tot += __shfl_xor(tot, off, 64);
        nwd += __shfl_xor(nwd, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && nwd) {
        atomicAdd(stats + 0, (unsigned long long)tot);
        atomicAdd(stats + 1, (unsigned long long)slots * 64ull);
        atomicAdd(stats + 2, (unsigned long long)nwd);
    }
}

// ---------------------------------------------------------------------------- kernels ----
// stats (optional, diagnostics only): [0] += Philox calls the words needed, [1] += lane-slots
// the SIMDs executed (64 x wave iterations), [2] += words.

// Lockstep: one word per lane, the wave pays the maximum call count of its 64 words.
// LAZY = false is the eager ablation (all 16 calls, src/pyising/backends/cuda/ising.cu:135).
template <bool LAZY, bool STATS>
__global__ __launch_bounds__(256) void msc_half_lockstep(MscArgs a, u64 codes,
                                                        unsigned long long* stats)
{
    const u32 w = xcd_band_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool valid = w < a.n_words;
    MscState s = {};
    if (valid) s = msc_classify(a, w);
    u32 calls = 0, slots = 0;
    msc_run<LAZY, STATS>(a, codes, w, 0u, 16u, s, calls, slots);
    if (valid) a.own[w] = (u64)s.pwl | ((u64)s.pwh << 32);
    if (STATS) msc_stats(stats, calls, slots, valid ? 1u : 0u);
}

// Lockstep + survivor compaction (the default). A 256-thread workgroup owns W = 256 * R
// consecutive words.
// Phase A: R rounds of one word per lane, lockstep levels 0 .. cap-1 (nearly every word needs
// them at T_c); decided words are stored, the others' state (a1, und, pw) parked in LDS and
// their slot appended to a queue (one LDS atomic per wave). One barrier.
// Phase B: waves take 64 queued survivors at a time (LDS batch counter) and run them lockstep
// from level cap to completion; all start at the same level, so the level index stays
// uniform, and the wave pays only the maximum of the survivors' remaining calls.
// R = 2 uses 13 KB of LDS: 8 workgroups (32 waves) per CU.
template <int R, bool STATS>
__global__ __launch_bounds__(256) void msc_half_compact(MscArgs a, u64 codes, u32 cap,
                                                       unsigned long long* stats)
{
    constexpr u32 W = 256u * R;
    __shared__ u64 s_a1[W], s_und[W], s_pw[W];
    __shared__ uint16_t s_q[W];
    __shared__ u32 s_qn, s_next;
    const u32 tid = threadIdx.x;
    const u32 lane = tid & 63u;
    const u32 base = xcd_band_block(blockIdx.x, gridDim.x) * W;
    if (tid == 0) {
        s_qn = 0;
        s_next = 0;
    }
    __syncthreads();
    u32 calls = 0, slots = 0, words = 0;
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
        const u32 slot = (u32)rr * 256u + tid;
        const u32 w = base + slot;
        const bool valid = w < a.n_words;
        MscState s = {};
        if (valid) {
            s = msc_classify(a, w);
            ++words;
        }
        msc_run<true, STATS>(a, codes, w, 0u, cap, s, calls, slots);
        const bool park = valid && (s.ul | s.uh) != 0u;
        if (valid && !park) a.own[w] = (u64)s.pwl | ((u64)s.pwh << 32);
        const u64 pm = __ballot(park);
        u32 qb = 0;
        if (lane == 0 && pm) qb = atomicAdd(&s_qn, (u32)__popcll(pm));
        qb = __shfl(qb, 0, 64);
        if (park) {
            s_a1[slot] = (u64)s.a1l | ((u64)s.a1h << 32);
            s_und[slot] = (u64)s.ul | ((u64)s.uh << 32);
            s_pw[slot] = (u64)s.pwl | ((u64)s.pwh << 32);
            s_q[qb + lane_rank(pm)] = (uint16_t)slot;
        }
    }
    __syncthreads();
    const u32 n = s_qn;
    while (true) {
        u32 b = 0;
        if (lane == 0) b = atomicAdd(&s_next, 64u);
        b = __shfl(b, 0, 64);
        if (b >= n) break;  // wave-uniform
        const u32 i = b + lane;
        MscState s = {};
        u32 slot = 0;
        if (i < n) {
            slot = s_q[i];
            const u64 a1 = s_a1[slot], und = s_und[slot], pw = s_pw[slot];
            s.a1l = (u32)a1;
            s.a1h = (u32)(a1 >> 32);
            s.ul = (u32)und;
            s.uh = (u32)(und >> 32);
            s.pwl = (u32)pw;
            s.pwh = (u32)(pw >> 32);
        }
        msc_run<true, STATS>(a, codes, base + slot, cap, 16u, s, calls, slots);
        if (i < n) a.own[base + slot] = (u64)s.pwl | ((u64)s.pwh << 32);
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

// out = (M, E); also clears the accumulator for the next call.
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
// Kernel modes: 0 = lockstep + survivor compaction (k = 10 * cap + R, cap in 1..15,
// R in {1, 2, 4}; default cap 4, R 2), 1 = lockstep, 2 = eager lockstep (all 16 calls per word,
// the msc-eager32 ablation; k unused for 1 and 2).
static int g_msc_mode = 0;
static int g_msc_k = 42;

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
    if (mode > 2)
        return ising_fail(ISING_EINVAL, "msc: mode must be 0 (compaction), 1 (lockstep) or 2 (eager)");
    if (mode == 0) {
        const int cap = k / 10, rr = k % 10;
        if (cap < 1 || cap > 15 || (rr != 1 && rr != 2 && rr != 4))
            return ising_fail(ISING_EINVAL,
                              "msc: compaction k = 10 * cap + R with cap in 1..15, R in {1,2,4}");
    }
    return ISING_OK;
}

template <int R>
static void launch_compact(dim3 grid, hipStream_t st, const MscArgs& a, u64 codes, u32 cap,
                           unsigned long long* stats)
{
    if (stats)
        hipLaunchKernelGGL((msc_half_compact<R, true>), grid, dim3(256), 0, st, a, codes, cap, stats);
    else
        hipLaunchKernelGGL((msc_half_compact<R, false>), grid, dim3(256), 0, st, a, codes, cap, stats);
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
    const u64 codes = msc_level_codes(thr4, thr8);
    if (mode == 0) {
        const u32 cap = (u32)(k / 10), rr = (u32)(k % 10);
        const dim3 grid((a.n_words + 256u * rr - 1u) / (256u * rr));
        if (rr == 1) launch_compact<1>(grid, st, a, codes, cap, stats);
        else if (rr == 2) launch_compact<2>(grid, st, a, codes, cap, stats);
        else launch_compact<4>(grid, st, a, codes, cap, stats);
    } else {
        const dim3 grid((a.n_words + 255u) / 256u);
        if (mode == 1) {
            if (stats) hipLaunchKernelGGL((msc_half_lockstep<true, true>), grid, dim3(256), 0, st, a, codes, stats);
            else hipLaunchKernelGGL((msc_half_lockstep<true, false>), grid, dim3(256), 0, st, a, codes, stats);
        } else {
            if (stats) hipLaunchKernelGGL((msc_half_lockstep<false, true>), grid, dim3(256), 0, st, a, codes, stats);
            else hipLaunchKernelGGL((msc_half_lockstep<false, false>), grid, dim3(256), 0, st, a, codes, stats);
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
    if (mode == 0) g_msc_k = k;
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
