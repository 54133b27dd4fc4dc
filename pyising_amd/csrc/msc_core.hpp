// Device-side building blocks of the multi-spin-coded half sweep, shared by the single-lattice
// kernels (msc_kernels.hip) and the batched persistent kernel (batch_kernels.hip). See
// msc_kernels.hip for the layout, the decision rule and the cost model.
#pragma once

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


// LDS of one compaction tile: survivor state and the survivor queue of W = 256 * R words.
template <int R>
struct MscCompactLds {
    u64 a1[256 * R], und[256 * R], pw[256 * R];
    uint16_t q[256 * R];
    u32 qn, next;
};

// Lockstep + survivor compaction over the W = 256 * R words starting at `base`, executed by
// one 256-thread workgroup (all threads must call it; it contains block barriers).
// Phase A: R rounds of one word per lane, lockstep levels 0 .. cap-1 (nearly every word needs
// them at T_c); decided words are stored, the others' state (a1, und, pw) parked in LDS and
// their slot appended to a queue (one LDS atomic per wave). One barrier.
// Phase B: waves take 64 queued survivors at a time (LDS batch counter) and run them lockstep
// from level cap to completion; all start at the same level, so the level index stays
// uniform, and the wave pays only the maximum of the survivors' remaining calls.
// Returns, when COUNT, this thread's share of the number of +1 spins written.
template <int R, bool STATS, bool COUNT>
__device__ __forceinline__ u32 msc_compact_tile(const MscArgs& a, u64 codes, u32 cap, u32 base,
                                                MscCompactLds<R>& L, u32& calls, u32& slots,
                                                u32& words)
{
    const u32 tid = threadIdx.x;
    const u32 lane = tid & 63u;
    u32 ones = 0;
    if (tid == 0) {
        L.qn = 0;
        L.next = 0;
    }
    __syncthreads();
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
        if (valid && !park) {
            a.own[w] = (u64)s.pwl | ((u64)s.pwh << 32);
            if (COUNT) ones += __popc(s.pwl) + __popc(s.pwh);
        }
        const u64 pm = __ballot(park);
        u32 qb = 0;
        if (lane == 0 && pm) qb = atomicAdd(&L.qn, (u32)__popcll(pm));
        qb = __shfl(qb, 0, 64);
        if (park) {
            L.a1[slot] = (u64)s.a1l | ((u64)s.a1h << 32);
            L.und[slot] = (u64)s.ul | ((u64)s.uh << 32);
            L.pw[slot] = (u64)s.pwl | ((u64)s.pwh << 32);
            L.q[qb + lane_rank(pm)] = (uint16_t)slot;
        }
    }
    __syncthreads();
    const u32 n = L.qn;
    while (true) {
        u32 b = 0;
        if (lane == 0) b = atomicAdd(&L.next, 64u);
        b = __shfl(b, 0, 64);
        if (b >= n) break;  // wave-uniform
        const u32 i = b + lane;
        MscState s = {};
        u32 slot = 0;
        if (i < n) {
            slot = L.q[i];
            const u64 a1 = L.a1[slot], und = L.und[slot], pw = L.pw[slot];
            s.a1l = (u32)a1;
            s.a1h = (u32)(a1 >> 32);
            s.ul = (u32)und;
            s.uh = (u32)(und >> 32);
            s.pwl = (u32)pw;
            s.pwh = (u32)(pw >> 32);
        }
        msc_run<true, STATS>(a, codes, base + slot, cap, 16u, s, calls, slots);
        if (i < n) {
            a.own[base + slot] = (u64)s.pwl | ((u64)s.pwh << 32);
            if (COUNT) ones += __popc(s.pwl) + __popc(s.pwh);
        }
    }
    __syncthreads();  // the LDS of this tile may be reused by the caller's next tile
    return ones;
}
