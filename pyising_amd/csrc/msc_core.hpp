// Device-side building blocks of the multi-spin-coded half sweep, shared by the single-lattice
// kernels (msc_kernels.hip) and the batched persistent kernel (batch_kernels.hip). See
// msc_kernels.hip for the layout, the decision rule and the cost model.
#pragma once

#include "common.hpp"

struct MscArgs {
    gu64* own;  // global memory (kernel arguments and descriptor pointers are cast to it)
    const gu64* other;
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

// ------------------------------------------------------------------ classification ----
// Three-input boolean functions as single v_bitop3_b32 ops; the immediate is the function
// evaluated on the operand patterns (0xF0, 0xCC, 0xAA).
template <typename F>
constexpr u32 b3imm(F f)
{
    return f(0xF0u, 0xCCu, 0xAAu) & 0xFFu;
}
constexpr u32 B3_MAJ = b3imm([](u32 x, u32 y, u32 z) { return (x & y) | (x & z) | (y & z); });
static_assert(B3_MAJ == 0xE8u, "bitop3 truth table");

#define bop3(a, b, c, imm) ((u32)__builtin_amdgcn_bitop3_b32((a), (b), (c), (imm)))

// One 32-bit half: own spins o, other-colour neighbours n1..n4 = (c, hz, up, dn). With
// k = the number of neighbours anti-aligned with o: und = k <= 1 (lanes that need a
// comparison), a1 = k == 1, pw = o ^ (k >= 2) (those lanes flip unconditionally). Counting
// n1..n3 with a full adder, cy = maj(n1, n2, n3) and sm = n1 ^ n2 ^ n3, k3 <= 1 exactly when
// cy == o, and k3 == 0 exactly when also sm == o; adding n4: und = (cy == o) & ~((sm ^ o) &
// (n4 ^ o)) and a1 = und & (sm ^ n4). Six VALU ops (checked against the count over all 32
// inputs on the host).
constexpr u32 B3_XOR3 = b3imm([](u32 x, u32 y, u32 z) { return x ^ y ^ z; });
constexpr u32 B3_BOTH_DIFFER = b3imm([](u32 x, u32 y, u32 z) { return (x ^ z) & (y ^ z); });
constexpr u32 B3_EQ_AND_NOT = b3imm([](u32 x, u32 y, u32 z) { return ~(x ^ y) & ~z; });
constexpr u32 B3_AND_XOR = b3imm([](u32 x, u32 y, u32 z) { return x & (y ^ z); });
static_assert(B3_XOR3 == 0x96u && B3_BOTH_DIFFER == 0x42u && B3_EQ_AND_NOT == 0x41u && B3_AND_XOR == 0x60u,
              "bitop3 truth tables");

__device__ __forceinline__ void msc_bits(u32 o, u32 c, u32 hz, u32 up, u32 dn, u32& a1, u32& pw,
                                         u32& und)
{
    const u32 cy = bop3(c, hz, up, B3_MAJ);
    const u32 sm = bop3(c, hz, up, B3_XOR3);
    const u32 g = bop3(sm, dn, o, B3_BOTH_DIFFER);
    und = bop3(cy, o, g, B3_EQ_AND_NOT);
    a1 = bop3(und, sm, dn, B3_AND_XOR);
    pw = ~(o ^ und);
}

__device__ __forceinline__ MscState msc_state(u64 o, u64 c, u64 hz, u64 up, u64 dn)
{
    MscState s;
    msc_bits((u32)o, (u32)c, (u32)hz, (u32)up, (u32)dn, s.a1l, s.pwl, s.ul);
    msc_bits((u32)(o >> 32), (u32)(c >> 32), (u32)(hz >> 32), (u32)(up >> 32), (u32)(dn >> 32),
             s.a1h, s.pwh, s.uh);
    return s;
}

// Classification of the 64 consecutive words w0 .. w0 + 63 of one row by the calling wave
// (requires nw % 64 == 0 and w0 % 64 == 0; w0 wave-uniform). Row, parity and the row pointers
// are scalar; the four same-column loads use one scalar base and the lane offset. The
// horizontal neighbour's boundary bit comes from the adjacent lane's word through a DPP
// wavefront shift; the wave's edge lane takes it from one uniform load of the word beyond.
// Split into the loads and the arithmetic so that callers can issue loads early.
struct MscRowLoad {
    u64 o, c, up, dn;
    u32 e;      // boundary half-word of the word beyond the wave's edge (uniform)
    u32 right;  // 1: colour sites at odd columns, horizontal neighbour j + 1 (uniform)
};

// Load of one packed word. COH: a device-scope (agent-coherent) load, which always misses in
// the CU's L1: for kernels in which another workgroup of the same launch may have rewritten
// the word after this CU cached it (the fused sweeps); otherwise an ordinary cached load.
template <bool COH>
__device__ __forceinline__ u64 msc_ld(const gu64* p)
{
    if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}

template <bool COH = false>
__device__ __forceinline__ MscRowLoad msc_load_row(const MscArgs& a, u32 w0, u32 lane)
{
    w0 = uni(w0);
    const u32 y = fast_div(w0, a.div_nw);
    const u32 x0 = w0 - y * a.nw;
    const u32 yu = (y + 1u == a.rows) ? 0u : y + 1u;
    const u32 yd = (y == 0u) ? a.rows - 1u : y - 1u;
    const gu64* crow = a.other + (u64)y * a.nw;
    MscRowLoad L;
    L.o = msc_ld<COH>(a.own + w0 + lane);
    L.c = msc_ld<COH>(crow + x0 + lane);
    L.up = msc_ld<COH>(a.other + (u64)yu * a.nw + x0 + lane);
    L.dn = msc_ld<COH>(a.other + (u64)yd * a.nw + x0 + lane);
    L.right = (y + a.color) & 1u;
    // one load of the word beyond the edge (x0 + 64 or x0 - 1, wrapped), its half selected
    // after the load, so that no wait for it sits between this and the next round's loads
    const u32 xr = (x0 + 64u == a.nw) ? 0u : x0 + 64u;
    const u32 xl = (x0 == 0u) ? a.nw - 1u : x0 - 1u;
    const u64 ev = msc_ld<COH>(crow + (L.right ? xr : xl));
    L.e = L.right ? (u32)ev : (u32)(ev >> 32);
    return L;
}

__device__ __forceinline__ MscState msc_classify_row(const MscRowLoad& L)
{
    const u32 cl = (u32)L.c, ch = (u32)(L.c >> 32);
    u32 hzl, hzh;
    if (L.right) {
        // hz = (c >> 1) | (next word << 63); wave_shl:1 -- lane i reads lane i + 1, lane 63
        // has no source and keeps e
        const u32 nb = (u32)__builtin_amdgcn_update_dpp((int)L.e, (int)cl, 0x130, 0xF, 0xF, false);
        hzl = __builtin_amdgcn_alignbit(ch, cl, 1u);
        hzh = __builtin_amdgcn_alignbit(nb, ch, 1u);
    } else {
        // hz = (c << 1) | (previous word >> 63); wave_shr:1 -- lane i reads lane i - 1, lane 0
        // has no source and keeps e
        const u32 nb = (u32)__builtin_amdgcn_update_dpp((int)L.e, (int)ch, 0x138, 0xF, 0xF, false);
        hzh = __builtin_amdgcn_alignbit(ch, cl, 31u);
        hzl = __builtin_amdgcn_alignbit(cl, nb, 31u);
    }
    return msc_state(L.o, L.c, (u64)hzl | ((u64)hzh << 32), L.up, L.dn);
}

// Classification when nw divides 64 and the lattice has a multiple of 64 words: the calling
// wave's 64 words w0 .. w0 + 63 (w0 % 64 == 0, wave-uniform) are 64 / nw whole rows, lane l
// holding word x = l % nw of row y0 + l / nw. Rows alternate parity inside the wave, so the
// side of the horizontal neighbour is per lane; the boundary half-word comes from the lane
// holding the word beside this one in the same row (wrapping within the row: the torus), by
// one ds_bpermute. Every lane of the wave must execute it.
struct MscSegLoad {
    u64 o, c, up, dn;
    u32 right;  // per lane: 1 = colour sites at odd columns, neighbour j + 1
};

template <bool COH = false>
__device__ __forceinline__ MscSegLoad msc_load_seg(const MscArgs& a, u32 w0, u32 lane)
{
    w0 = uni(w0);
    const u32 lg = (u32)__builtin_ctz(a.nw);  // nw = 2^lg <= 32
    const u32 y = (w0 >> lg) + (lane >> lg);
    const u32 x = lane & (a.nw - 1u);
    const u32 yu = (y + 1u == a.rows) ? 0u : y + 1u;
    const u32 yd = (y == 0u) ? a.rows - 1u : y - 1u;
    MscSegLoad L;
    L.o = msc_ld<COH>(a.own + w0 + lane);
    L.c = msc_ld<COH>(a.other + w0 + lane);
    L.up = msc_ld<COH>(a.other + (yu << lg) + x);
    L.dn = msc_ld<COH>(a.other + (yd << lg) + x);
    L.right = (y + a.color) & 1u;
    return L;
}

__device__ __forceinline__ MscState msc_classify_seg(const MscSegLoad& L, u32 nw, u32 lane)
{
    const u32 cl = (u32)L.c, ch = (u32)(L.c >> 32);
    const u32 x = lane & (nw - 1u);
    const u32 src = (lane - x) + ((x + (L.right ? 1u : nw - 1u)) & (nw - 1u));
    // right: bit 0 of the next word (its low half); left: bit 63 of the previous word (its
    // high half). The source lane is in the same row, so it has the same side.
    const u32 nb = (u32)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)(L.right ? cl : ch));
    const u32 hzl = L.right ? __builtin_amdgcn_alignbit(ch, cl, 1u) : __builtin_amdgcn_alignbit(cl, nb, 31u);
    const u32 hzh = L.right ? __builtin_amdgcn_alignbit(nb, ch, 1u) : __builtin_amdgcn_alignbit(ch, cl, 31u);
    return msc_state(L.o, L.c, (u64)hzl | ((u64)hzh << 32), L.up, L.dn);
}

// Which classification a lattice's waves use: whole-row waves (nw % 64 == 0), row-segment
// waves (nw divides 64 and the word count is a multiple of 64), or per-lane general loads.
enum : int { MSC_GENERAL = 0, MSC_ROWW = 1, MSC_SEGW = 2 };

static inline int msc_layout(u32 nw, u32 n_words)
{
    if (nw % 64u == 0u) return MSC_ROWW;
    if (nw && 64u % nw == 0u && n_words % 64u == 0u) return MSC_SEGW;
    return MSC_GENERAL;
}

// Neighbour loads of word w (any geometry): five loads through the cache hierarchy, the
// horizontal neighbour already shifted into place.
struct MscLoad {
    u64 o, c, hz, up, dn;
};

template <bool COH = false>
__device__ __forceinline__ MscLoad msc_load(const MscArgs& a, u32 w)
{
    const u32 y = fast_div(w, a.div_nw);
    const u32 x = w - y * a.nw;
    const u64 rbase = (u64)y * a.nw;
    const u32 p = (y + a.color) & 1u;
    MscLoad L;
    L.o = msc_ld<COH>(a.own + w);
    L.c = msc_ld<COH>(a.other + w);
    const u32 xn = p ? ((x + 1 == a.nw) ? 0u : x + 1) : ((x == 0) ? a.nw - 1 : x - 1);
    const u64 nb = msc_ld<COH>(a.other + rbase + xn);
    L.hz = p ? ((L.c >> 1) | (nb << 63)) : ((L.c << 1) | (nb >> 63));
    const u32 yu = (y + 1 == a.rows) ? 0u : y + 1;
    const u32 yd = (y == 0) ? a.rows - 1 : y - 1;
    L.up = msc_ld<COH>(a.other + (u64)yu * a.nw + x);
    L.dn = msc_ld<COH>(a.other + (u64)yd * a.nw + x);
    return L;
}

__device__ __forceinline__ MscState msc_classify(const MscLoad& L)
{
    return msc_state(L.o, L.c, L.hz, L.up, L.dn);
}

__device__ __forceinline__ MscState msc_classify(const MscArgs& a, u32 w)
{
    return msc_classify(msc_load(a, w));
}

// Classification of word w in a kernel whose waves hold 64 consecutive words starting at a
// multiple of 64 (all lanes of the wave must call it when LAY != MSC_GENERAL; with n_words a
// multiple of 64 a wave is wholly valid or wholly past the end, so callers branch per wave).
template <int LAY, bool COH = false>
__device__ __forceinline__ MscState msc_classify_wave(const MscArgs& a, u32 w)
{
    const u32 lane = threadIdx.x & 63u;
    if constexpr (LAY == MSC_ROWW) return msc_classify_row(msc_load_row<COH>(a, w - lane, lane));
    else if constexpr (LAY == MSC_SEGW) return msc_classify_seg(msc_load_seg<COH>(a, w - lane, lane), a.nw, lane);
    else return msc_classify(msc_load<COH>(a, w));
}

// Philox call of word w at level pair lp (lp wave-uniform): counter (w, t, colour, 1 + lp).
struct MscCounter {
    u32 hi0, lo0;  // w * M0, the per-lane half of Philox round 0
};

__device__ __forceinline__ MscCounter msc_counter(u32 w)
{
    const u64 p = (u64)w * PHILOX_M0;
    return MscCounter{(u32)(p >> 32), (u32)p};
}

__device__ __forceinline__ PhiloxOut msc_philox(const MscArgs& a, const MscCounter& c, u32 lp)
{
    return philox_lane(c.hi0, c.lo0, a.t, a.color, 1u + lp, a.k0, a.k1);
}

// The same call with a per-lane level pair.
__device__ __forceinline__ PhiloxOut msc_philox_lanes(const MscArgs& a, const MscCounter& c,
                                                      u32 lp)
{
    return philox_lane<true>(c.hi0, c.lo0, a.t, a.color, 1u + lp, a.k0, a.k1);
}

// ------------------------------------------------------------------ lazy comparison ----
// Both levels of one Philox call in closed form. The threshold bit a lane compares against at
// one level is 0, 1, a1 (thr4 bit 1, thr8 bit 0) or ~a1 (the reverse): kinds 0-3. With x, y
// the uniform bits of the two levels: m = und & (x == t_A) stays undecided after level A,
// und' = m & (y == t_B); the lanes accepted are those decided at a level whose threshold bit
// is 1 (their uniform bit is then 0).
template <u32 B4, u32 B8>
constexpr u32 msc_tkind()
{
    return B4 ? (B8 ? 1u : 2u) : (B8 ? 3u : 0u);
}

// The definition, one level after the other (also the reference the forms below are proved
// against at compile time).
constexpr u32 msc_teq(u32 k, u32 x, u32 a)  // lanes whose bit equals the threshold bit
{
    return k == 0 ? ~x : k == 1 ? x : k == 2 ? ~(x ^ a) : (x ^ a);
}
constexpr u32 msc_tbit(u32 k, u32 a)  // the threshold bit
{
    return k == 0 ? 0u : k == 1 ? ~0u : k == 2 ? a : ~a;
}
constexpr void msc_two_levels_def(u32 ka, u32 kb, u32 x, u32 y, u32 a, u32& pw, u32& und)
{
    const u32 m = und & msc_teq(ka, x, a);
    const u32 u2 = m & msc_teq(kb, y, a);
    pw ^= ((und ^ m) & msc_tbit(ka, a)) ^ ((m ^ u2) & msc_tbit(kb, a));
    und = u2;
}

// The same pair as explicit v_bitop3 ops. Per 32-bit half, with u = und: und' = u & G(x, y, a)
// and pw' = pw ^ (u & T(x, y, a)), where G = (x == t_A) & (y == t_B) keeps the lanes both
// levels leave undecided and T = t_A & ~x | (x == t_A) & t_B & ~y accepts. G and T are
// three-input functions, so each is one op, and so are the products with u and the xor into
// pw; ops fold where G or T ignore an input: und' in one op when G ignores a, pw' in one op
// when T depends on a single input, none when T = 0 (no level accepts), xor3(pw, u, und')
// when T = ~G (every decided lane is accepted). Two to eight ops per level pair for both
// halves, against two to sixteen in the compiler's form of the definition.
struct MscB3Emu {  // v_bitop3_b32 on the host: bit i of the result is bit (a_i b_i c_i) of IMM
    template <u32 IMM>
    static constexpr u32 op(u32 a, u32 b, u32 c)
    {
        u32 r = 0;
        for (u32 i = 0; i < 32; ++i) {
            const u32 sel = (((a >> i) & 1u) << 2) | (((b >> i) & 1u) << 1) | ((c >> i) & 1u);
            r |= ((IMM >> sel) & 1u) << i;
        }
        return r;
    }
};
struct MscB3Dev {
    template <u32 IMM>
    static __device__ __forceinline__ u32 op(u32 a, u32 b, u32 c) { return bop3(a, b, c, IMM); }
};

template <u32 KA, u32 KB>
struct MscPairForm {
    // truth tables over (x, y, a) = (0xF0, 0xCC, 0xAA)
    static constexpr u32 G = (msc_teq(KA, 0xF0u, 0xAAu) & msc_teq(KB, 0xCCu, 0xAAu)) & 0xFFu;
    static constexpr u32 T = ((msc_tbit(KA, 0xAAu) & ~0xF0u) |
                              (msc_teq(KA, 0xF0u, 0xAAu) & msc_tbit(KB, 0xAAu) & ~0xCCu)) & 0xFFu;
    static constexpr bool dep_x(u32 f) { return (((f >> 4) ^ f) & 0x0Fu) != 0u; }
    static constexpr bool dep_y(u32 f) { return (((f >> 2) ^ f) & 0x33u) != 0u; }
    static constexpr bool dep_a(u32 f) { return (((f >> 1) ^ f) & 0x55u) != 0u; }
    // und' = bop3(u, x, y) when G ignores a: G restricted to (x, y) at a = 0, as (0xCC, 0xAA)
    static constexpr u32 G_UXY = (0xF0u & msc_teq(KA, 0xCCu, 0u) & msc_teq(KB, 0xAAu, 0u)) & 0xFFu;
    static constexpr int T_VARS = (dep_x(T) ? 1 : 0) + (dep_y(T) ? 1 : 0) + (dep_a(T) ? 1 : 0);
    // pw' = bop3(pw, u, v) when T depends on v alone: T(v) as the pattern 0xAA
    static constexpr u32 T_V = dep_x(T) ? ((msc_tbit(KA, 0u) & ~0xAAu) | (msc_teq(KA, 0xAAu, 0u) & msc_tbit(KB, 0u) & ~0u))
                             : dep_y(T) ? ((msc_tbit(KA, 0u) & ~0u) | (msc_teq(KA, 0u, 0u) & msc_tbit(KB, 0u) & ~0xAAu))
                                        : ((msc_tbit(KA, 0xAAu) & ~0u) | (msc_teq(KA, 0u, 0xAAu) & msc_tbit(KB, 0xAAu) & ~0u));
    static constexpr u32 PW_UV = (0xF0u ^ (0xCCu & T_V)) & 0xFFu;      // pw ^ (u & T(v))
    static constexpr u32 PW_UT = (0xF0u ^ (0xCCu & 0xAAu)) & 0xFFu;    // pw ^ (u & t)
    static constexpr u32 XOR3 = (0xF0u ^ 0xCCu ^ 0xAAu) & 0xFFu;
    static constexpr u32 AND = (0xF0u & 0xCCu) & 0xFFu;                // first two operands

    template <class B>
    static constexpr void apply(u32 x, u32 y, u32 a, u32& pw, u32& und)
    {
        const u32 u = und;
        u32 u2 = 0;
        if constexpr (!dep_a(G)) u2 = B::template op<G_UXY>(u, x, y);
        else u2 = B::template op<AND>(u, B::template op<G>(x, y, a), 0u);
        if constexpr (T == 0u) {
        } else if constexpr (T == (~G & 0xFFu)) {
            pw = B::template op<XOR3>(pw, u, u2);
        } else if constexpr (T_VARS == 1) {
            const u32 v = dep_x(T) ? x : dep_y(T) ? y : a;
            pw = B::template op<PW_UV>(pw, u, v);
        } else {
            pw = B::template op<PW_UT>(pw, u, B::template op<T>(x, y, a));
        }
        und = u2;
    }

    // every combination of (u, x, y, a, pw) bits, one per bit position
    static constexpr bool proved()
    {
        const u32 u = 0xFFFF0000u, x = 0xFF00FF00u, y = 0xF0F0F0F0u, a = 0xCCCCCCCCu, p = 0xAAAAAAAAu;
        u32 pw0 = p, u0 = u, pw1 = p, u1 = u;
        msc_two_levels_def(KA, KB, x, y, a, pw0, u0);
        apply<MscB3Emu>(x, y, a, pw1, u1);
        return pw0 == pw1 && u0 == u1;
    }
};

template <u32 CODE>
constexpr bool msc_pair_proved()
{
    return MscPairForm<msc_tkind<CODE & 1u, (CODE >> 1) & 1u>(),
                       msc_tkind<(CODE >> 2) & 1u, (CODE >> 3) & 1u>()>::proved();
}
static_assert(msc_pair_proved<0>() && msc_pair_proved<1>() && msc_pair_proved<2>() && msc_pair_proved<3>() &&
                  msc_pair_proved<4>() && msc_pair_proved<5>() && msc_pair_proved<6>() && msc_pair_proved<7>() &&
                  msc_pair_proved<8>() && msc_pair_proved<9>() && msc_pair_proved<10>() && msc_pair_proved<11>() &&
                  msc_pair_proved<12>() && msc_pair_proved<13>() && msc_pair_proved<14>() && msc_pair_proved<15>(),
              "level-pair forms equal the level-by-level definition on every input");

// CODE: bit 0 / 1 = thr4 / thr8 bit of level 2 lp, bits 2 / 3 the same for level 2 lp + 1.
// Level 2 lp uses r.x (low half) and r.y (high half), level 2 lp + 1 uses r.z and r.w.
template <u32 CODE>
__device__ __forceinline__ void msc_pair(const PhiloxOut& r, MscState& s)
{
    using F = MscPairForm<msc_tkind<CODE & 1u, (CODE >> 1) & 1u>(), msc_tkind<(CODE >> 2) & 1u, (CODE >> 3) & 1u>()>;
    F::template apply<MscB3Dev>(r.x, r.z, s.a1l, s.pwl, s.ul);
    F::template apply<MscB3Dev>(r.y, r.w, s.a1h, s.pwh, s.uh);
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

// The same level pair applied to N words at once (one per round of a tile, all at the same
// level): the case switch, a chain of scalar compares and branches, runs once for the N.
template <u32 CODE, int N>
__device__ __forceinline__ void msc_pair_n(const PhiloxOut* const (&r)[N], MscState (&s)[N])
{
#pragma unroll
    for (int i = 0; i < N; ++i) msc_pair<CODE>(*r[i], s[i]);
}

template <int N>
__device__ __forceinline__ void msc_pair_switch_n(u32 code, const PhiloxOut* const (&r)[N],
                                                  MscState (&s)[N])
{
    switch (code) {
    case 0: msc_pair_n<0>(r, s); break;
    case 1: msc_pair_n<1>(r, s); break;
    case 2: msc_pair_n<2>(r, s); break;
    case 3: msc_pair_n<3>(r, s); break;
    case 4: msc_pair_n<4>(r, s); break;
    case 5: msc_pair_n<5>(r, s); break;
    case 6: msc_pair_n<6>(r, s); break;
    case 7: msc_pair_n<7>(r, s); break;
    case 8: msc_pair_n<8>(r, s); break;
    case 9: msc_pair_n<9>(r, s); break;
    case 10: msc_pair_n<10>(r, s); break;
    case 11: msc_pair_n<11>(r, s); break;
    case 12: msc_pair_n<12>(r, s); break;
    case 13: msc_pair_n<13>(r, s); break;
    case 14: msc_pair_n<14>(r, s); break;
    default: msc_pair_n<15>(r, s); break;
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
    const MscCounter ctr = msc_counter(w);
    while (lp < lp_end && (!LAZY || __ballot((s.ul | s.uh) != 0u) != 0)) {
        if (STATS) {
            calls += (!LAZY || (s.ul | s.uh) != 0u) ? 1u : 0u;
            ++slots;
        }
        const PhiloxOut r = msc_philox(a, ctr, lp);
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


// Store of a finished word. NT: non-temporal (streaming) store, so the new spins leave the
// L2 during the kernel instead of in the write-back at the kernel boundary.
template <bool NT>
__device__ __forceinline__ void msc_store(gu64* p, const MscState& s)
{
    const u64 v = (u64)s.pwl | ((u64)s.pwh << 32);
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// LDS of one queue-compaction tile (W = 256 * R words): survivor state by slot, two survivor
// queues used alternately, and the queue length after phase A and after every phase-B round.
template <int R>
struct MscQueueLds {
    u64 a1[256 * R], und[256 * R], pw[256 * R];
    uint16_t q[2][256 * R];
    u32 qn[18];
};

template <int R>
__device__ __forceinline__ void msc_park(MscQueueLds<R>& L, u32 slot, const MscState& s)
{
    L.a1[slot] = (u64)s.a1l | ((u64)s.a1h << 32);
    L.und[slot] = (u64)s.ul | ((u64)s.uh << 32);
    L.pw[slot] = (u64)s.pwl | ((u64)s.pwh << 32);
}

template <int R>
__device__ __forceinline__ MscState msc_unpark(const MscQueueLds<R>& L, u32 slot)
{
    const u64 a1 = L.a1[slot], und = L.und[slot], pw = L.pw[slot];
    MscState s;
    s.a1l = (u32)a1;
    s.a1h = (u32)(a1 >> 32);
    s.ul = (u32)und;
    s.uh = (u32)(und >> 32);
    s.pwl = (u32)pw;
    s.pwh = (u32)(pw >> 32);
    return s;
}

// Append the lanes of `keep` (km = its ballot, taken by the caller before any branch so that
// it stays a scalar mask) to queue `q` (one LDS atomic per wave) and park their state.
template <int R>
__device__ __forceinline__ void msc_enqueue(MscQueueLds<R>& L, uint16_t* q, u32* qn, u64 km,
                                            bool keep, u32 slot, const MscState& s,
                                            bool state_a1)
{
    u32 qb = 0;
    if ((threadIdx.x & 63u) == 0u && km) qb = atomicAdd(qn, (u32)__popcll(km));
    qb = __builtin_amdgcn_readlane(qb, 0);
    if (keep) {
        if (state_a1) L.a1[slot] = (u64)s.a1l | ((u64)s.a1h << 32);
        L.und[slot] = (u64)s.ul | ((u64)s.uh << 32);
        L.pw[slot] = (u64)s.pwl | ((u64)s.pwh << 32);
        q[qb + lane_rank(km)] = (uint16_t)slot;
    }
}

// Early end of the waves of a queue tile with no survivor work left: MscStay keeps every wave
// to the end of the tile, MscLeave ends them (with nothing to hand over).
struct MscStay {
    static constexpr bool leave = false;
    __device__ void operator()(u32) const {}
};
struct MscLeave {
    static constexpr bool leave = true;
    __device__ void operator()(u32) const {}
};

// End of the calling wave, once its memory operations have completed (wave-uniform call
// sites only). S_BARRIER waits only for the waves of the workgroup that have not ended (CDNA
// ISA), so the remaining waves go on using __syncthreads; a barrier they pass afterwards also
// orders the ended wave's stores (complete before it ended) before what follows it.
__device__ __forceinline__ void msc_wave_end()
{
    __builtin_amdgcn_s_waitcnt(0);  // vmcnt, expcnt and lgkmcnt all zero
    __builtin_amdgcn_endpgm();
}

// Queue compaction over the W = 256 * R words starting at `base`, one 256-thread workgroup
// (all threads must call it; it contains block barriers).
// Phase A: R rounds of one word per lane, CAP lockstep Philox calls unrolled without exit
//   tests (at T_c nearly every word needs them); decided words are stored, survivors parked.
//   JOINT (implied by PRE): the calls of all R rounds are computed first, then every round is
//   classified and each level pair's case switch (a chain of scalar compares and branches)
//   runs once for the R words of a lane instead of once per round; the calls of every round
//   are then live together (more VGPRs).
// Phase B: rounds of exactly one call per survivor. Survivors are packed into full waves, all
//   at the same level (so the level's case code stays wave-uniform); those still undecided
//   are queued for the next round. A wave slot is wasted only in the last partial wave of a
//   round, instead of on every lane that finished before the slowest of its 64 words.
// Tail: once at most 64 survivors remain, wave 0 finishes them, lockstep, or two level pairs
// per call when at most 32 remain.
// EX::leave: a wave ends (msc_wave_end) as soon as no survivor work is left for it -- waves
//   1..3 once the queue is shorter than their first slot, and before the tail -- instead of
//   idling at the round barriers while other waves finish; its slot then serves another
//   workgroup. The hook ex(ones) runs first (all lanes of the wave), e.g. to hand over the
//   wave's count of +1 spins. Only for kernels whose code after the tile needs no wave but
//   wave 0 (it runs in the waves that are left; wave 0 always is).
// ROWW: nw % 64 == 0 (every wave's 64 words lie in one row, see msc_classify_row).
// Returns, when COUNT, this thread's share of the number of +1 spins written.
// Phase A's loads of one tile, issued before any of its arithmetic (the words of every round
// are independent), so that memory latency hides behind earlier work.
template <int R, int LAY>
struct MscTileLoads {
    MscRowLoad lr[LAY == MSC_ROWW ? R : 1];
    MscSegLoad ls[LAY == MSC_SEGW ? R : 1];
    MscLoad lg[LAY == MSC_GENERAL ? R : 1];
};

// Branch-free: words past the end load a clamped (valid) address instead, and their results
// are discarded by the caller. FULL: every word of the tile exists (no clamp).
template <int R, int LAY, bool COH = false, bool FULL = false>
__device__ __forceinline__ void msc_tile_load(const MscArgs& a, u32 base, MscTileLoads<R, LAY>& T)
{
    const u32 tid = threadIdx.x;
    const u32 wave_w = uni(tid & ~63u);  // the wave's first word in a round, scalar
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
        if (LAY != MSC_GENERAL) {
            // n_words is a multiple of 64: clamp the wave's first word
            const u32 w0 = base + (u32)rr * 256u + wave_w;
            const u32 wc = (FULL || w0 < a.n_words) ? w0 : a.n_words - 64u;
            if (LAY == MSC_ROWW) T.lr[rr] = msc_load_row<COH>(a, wc, tid & 63u);
            else T.ls[rr] = msc_load_seg<COH>(a, wc, tid & 63u);
        } else {
            const u32 w = base + (u32)rr * 256u + tid;
            T.lg[rr] = msc_load<COH>(a, w < a.n_words ? w : a.n_words - 1u);
        }
    }
}

struct MscNoPrefetch {
    __device__ void operator()() const {}
};

// Phase A's Philox calls of a tile, computed by the caller ahead of the tile (PRE): level
// pairs 0 .. CAP-1 of the word slot r * 256 + threadIdx.x for every round r.
template <int R, int CAP>
struct MscPhaseACalls {
    PhiloxOut r[R][CAP];
};

template <int R, int CAP>
__device__ __forceinline__ void msc_phase_a_calls(const MscArgs& a, u32 base, int rr,
                                                  MscPhaseACalls<R, CAP>& P)
{
    const MscCounter ctr = msc_counter(base + (u32)rr * 256u + threadIdx.x);
#pragma unroll
    for (int lp = 0; lp < CAP; ++lp) {
        P.r[rr][lp] = msc_philox(a, ctr, (u32)lp);
        // pinned where it is written, so that it covers the memory latency that follows
        asm volatile("" : "+v"(P.r[rr][lp].x), "+v"(P.r[rr][lp].y), "+v"(P.r[rr][lp].z), "+v"(P.r[rr][lp].w));
    }
}

// FULL: every word of the tile exists (base + 256 R <= n_words), so no word needs the
// past-the-end guards.
template <int R, int CAP, int LAY, bool STATS, bool COUNT, bool NT, class PF, bool PRE = false,
          bool FULL = false, class EX = MscStay, bool JOINT = false>
__device__ __forceinline__ u32 msc_queue_tile_loaded(const MscArgs& a, u64 codes, u32 base,
                                                     MscQueueLds<R>& L,
                                                     const MscTileLoads<R, LAY>& T, PF prefetch,
                                                     u32& calls, u32& slots, u32& words,
                                                     const MscPhaseACalls<R, CAP>* pre = nullptr,
                                                     EX ex = EX{})
{
    static_assert(CAP >= 1 && CAP <= 15, "phase A levels");
    const u32 tid = threadIdx.x;
    const u32 lane = tid & 63u;
    // wave index and queue lengths are wave-uniform: held in scalar registers, the phase-B
    // level lp stays scalar too and each call's case switch is a scalar branch, not a
    // per-lane branch tree under exec masks
    const u32 wave = uni(tid >> 6);
    u32 ones = 0;
    if (tid < 18u) L.qn[tid] = 0u;
    __syncthreads();
    // end of a phase-A word: stored if decided, else parked and queued
    auto finish = [&](int rr, const MscState& s, bool valid) {
        const u32 slot = (u32)rr * 256u + tid;
        const bool park = (s.ul | s.uh) != 0u;  // invalid lanes have und = 0
        const u64 km = __ballot(park);
        if (valid && !park) {
            msc_store<NT>(a.own + base + slot, s);
            if (COUNT) ones += __popc(s.pwl) + __popc(s.pwh);
        }
        msc_enqueue(L, L.q[0], &L.qn[0], km, park, slot, s, true);
    };
    auto classify = [&](int rr, MscState& s, bool& valid) {
        valid = FULL || base + (u32)rr * 256u + tid < a.n_words;
        if constexpr (LAY == MSC_ROWW) s = msc_classify_row(T.lr[rr]);
        else if constexpr (LAY == MSC_SEGW) s = msc_classify_seg(T.ls[rr], a.nw, lane);
        else s = msc_classify(T.lg[rr]);
        if (!valid) {
            s.ul = 0u;  // nothing to decide, store or park
            s.uh = 0u;
        }
        if (valid) ++words;
    };
    if constexpr (PRE || JOINT) {
        // Every round is classified first, then each level pair's case switch (a chain of
        // scalar compares and branches) runs once for the words of all R rounds instead of once
        // per round. Without PRE the calls of all rounds are computed here first (pinned where
        // written, so that they cover the latency of the tile's loads).
        MscPhaseACalls<R, CAP> own_calls;
        if constexpr (!PRE) {
#pragma unroll
            for (int rr = 0; rr < R; ++rr) msc_phase_a_calls<R, CAP>(a, base, rr, own_calls);
            __builtin_amdgcn_sched_barrier(0);
        }
        const MscPhaseACalls<R, CAP>* calls_in = PRE ? pre : &own_calls;
        MscState s[R];
        bool valid[R];
#pragma unroll
        for (int rr = 0; rr < R; ++rr) classify(rr, s[rr], valid[rr]);
#pragma unroll
        for (int lp = 0; lp < CAP; ++lp) {
            const PhiloxOut* rp[R];
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                rp[rr] = &calls_in->r[rr][lp];
                if (STATS) {
                    calls += (s[rr].ul | s[rr].uh) != 0u ? 1u : 0u;
                    ++slots;
                }
            }
            msc_pair_switch_n<R>((u32)(codes >> (4u * (u32)lp)) & 15u, rp, s);
        }
#pragma unroll
        for (int rr = 0; rr < R; ++rr) finish(rr, s[rr], valid[rr]);
    } else {
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            // the CAP Philox calls depend on the counter only: computed before the spin loads
            // are consumed, they cover the memory latency (at the start of every workgroup
            // above all)
            const MscCounter ctr = msc_counter(base + (u32)rr * 256u + tid);
            PhiloxOut r[CAP];
#pragma unroll
            for (int lp = 0; lp < CAP; ++lp) {
                r[lp] = msc_philox(a, ctr, (u32)lp);
                // an empty asm that "modifies" the outputs pins the call here: the optimiser
                // would otherwise sink it below the classification and its wait for the loads
                asm volatile("" : "+v"(r[lp].x), "+v"(r[lp].y), "+v"(r[lp].z), "+v"(r[lp].w));
            }
            __builtin_amdgcn_sched_barrier(0);
            MscState s;
            bool valid;
            classify(rr, s, valid);
#pragma unroll
            for (int lp = 0; lp < CAP; ++lp) {
                if (STATS) {
                    calls += (s.ul | s.uh) != 0u ? 1u : 0u;
                    ++slots;
                }
                msc_pair_switch((u32)(codes >> (4u * (u32)lp)) & 15u, r[lp], s);
            }
            finish(rr, s, valid);
        }
    }
    __syncthreads();
    prefetch();  // e.g. the next tile's loads, which then complete while phase B runs
    u32 n = uni(L.qn[0]);
    u32 cur = 0;
    u32 lp = CAP;
    while (n > 64u && lp < 16u) {  // block-uniform
        if (EX::leave && wave != 0u && wave * 64u >= n) {  // the queue only shrinks
            ex(ones);
            msc_wave_end();
        }
        const u32 code = (u32)(codes >> (4u * lp)) & 15u;
        for (u32 i0 = wave * 64u; i0 < n; i0 += 256u) {  // wave-uniform
            // lanes past the queue's end (i >= n >= 1) repeat its last entry with nothing
            // undecided, so they neither store nor re-queue
            const u32 i = i0 + lane;
            const u32 slot = L.q[cur][i < n ? i : n - 1u];
            MscState s = msc_unpark(L, slot);
            if (i >= n) {
                s.ul = 0u;
                s.uh = 0u;
            }
            if (STATS) {
                calls += (s.ul | s.uh) != 0u ? 1u : 0u;
                ++slots;
            }
            const PhiloxOut r = msc_philox(a, msc_counter(base + slot), lp);
            msc_pair_switch(code, r, s);
            const bool still = (s.ul | s.uh) != 0u;
            const u64 km = __ballot(still);
            if (i < n && !still) {
                msc_store<NT>(a.own + base + slot, s);
                if (COUNT) ones += __popc(s.pwl) + __popc(s.pwh);
            }
            msc_enqueue(L, L.q[cur ^ 1u], &L.qn[lp + 1u - CAP], km, still, slot, s, false);
        }
        __syncthreads();
        n = uni(L.qn[lp + 1u - CAP]);
        cur ^= 1u;
        ++lp;
    }
    if (EX::leave && wave != 0u && (n <= 32u || wave * 64u >= n)) {
        ex(ones);
        msc_wave_end();
    }
    if (n <= 32u) {
        // Tail of at most 32 survivors, run by wave 0. Lane j < 32 owns survivor j and computes
        // its level pair lp; lane j + 32 computes level pair lp + 1 of the same word (any lane
        // may evaluate any (word, level) call: the stream is counter-based) and hands its four
        // words to the owner through ds_bpermute; the owner applies both pairs in order. One
        // wave-call thus covers two level pairs where lockstep would spend two.
        if (wave == 0u && n > 0u) {
            const u32 j = lane & 31u;
            const u32 slot = L.q[cur][j < n ? j : n - 1u];
            MscState s = msc_unpark(L, slot);
            if (lane >= n) {  // helpers (lane >= 32) and owners past the queue's end
                s.ul = 0u;
                s.uh = 0u;
            }
            const MscCounter ctr = msc_counter(base + slot);
            const int from = (int)(((lane + 32u) & 63u) << 2);
            while (lp < 16u && __ballot((s.ul | s.uh) != 0u) != 0) {  // wave-uniform
                if (STATS) {
                    calls += (s.ul | s.uh) != 0u ? 1u : 0u;
                    ++slots;
                }
                if (lp + 1u < 16u) {
                    const PhiloxOut r = msc_philox_lanes(a, ctr, lp + (lane >> 5));
                    PhiloxOut h;
                    h.x = (u32)__builtin_amdgcn_ds_bpermute(from, (int)r.x);
                    h.y = (u32)__builtin_amdgcn_ds_bpermute(from, (int)r.y);
                    h.z = (u32)__builtin_amdgcn_ds_bpermute(from, (int)r.z);
                    h.w = (u32)__builtin_amdgcn_ds_bpermute(from, (int)r.w);
                    msc_pair_switch((u32)(codes >> (4u * lp)) & 15u, r, s);
                    if (STATS) calls += (s.ul | s.uh) != 0u ? 1u : 0u;
                    msc_pair_switch((u32)(codes >> (4u * (lp + 1u))) & 15u, h, s);
                    lp += 2u;
                } else {
                    msc_pair_switch((u32)(codes >> (4u * lp)) & 15u, msc_philox(a, ctr, lp), s);
                    ++lp;
                }
            }
            if (lane < n) {
                msc_store<NT>(a.own + base + slot, s);
                if (COUNT) ones += __popc(s.pwl) + __popc(s.pwh);
            }
        }
    } else {
        // 33..64 survivors (or more once every level is spent): wave 0 runs them lockstep
        for (u32 i0 = wave * 64u; i0 < n; i0 += 256u) {
            const u32 i = i0 + lane;
            const u32 slot = L.q[cur][i < n ? i : n - 1u];
            MscState s = msc_unpark(L, slot);
            if (i >= n) {
                s.ul = 0u;
                s.uh = 0u;
            }
            msc_run<true, STATS>(a, codes, base + slot, lp, 16u, s, calls, slots);
            if (i < n) {
                msc_store<NT>(a.own + base + slot, s);
                if (COUNT) ones += __popc(s.pwl) + __popc(s.pwh);
            }
        }
    }
    __syncthreads();  // the LDS of this tile may be reused by the caller's next tile
    return ones;
}

template <int R, int CAP, int LAY, bool STATS, bool COUNT, bool NT, bool COH = false,
          class EX = MscStay, bool JOINT = false>
__device__ __forceinline__ u32 msc_queue_tile(const MscArgs& a, u64 codes, u32 base,
                                              MscQueueLds<R>& L, u32& calls, u32& slots,
                                              u32& words, EX ex = EX{})
{
    MscTileLoads<R, LAY> T;
    msc_tile_load<R, LAY, COH>(a, base, T);
    return msc_queue_tile_loaded<R, CAP, LAY, STATS, COUNT, NT, MscNoPrefetch, false, false, EX, JOINT>(
        a, codes, base, L, T, MscNoPrefetch{}, calls, slots, words, nullptr, ex);
}
