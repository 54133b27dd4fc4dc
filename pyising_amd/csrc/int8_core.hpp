// Device-side building blocks of the int8 compressed-checkerboard half sweep, shared by the
// single-lattice kernels (int8_kernels.hip) and the batched persistent kernel
// (batch_kernels.hip). See int8_kernels.hip for the layout and the decision rules.
#pragma once

#include "common.hpp"

struct Int8Table {
    u32 v[7];  // table[0 .. 2d]; unused entries zero
};

struct Int8Geom {
    u32 lh;          // colour sites per row
    u32 ly, lz;      // rows = lz * ly (lz = 1 in 2D)
    u32 chunks_row;  // lh / V
    u32 n_chunks;    // rows * lh / V (< 2^32, checked on the host)
    FastDiv div_chunks, div_ly;
};

template <int V>
struct Vec;
template <>
struct Vec<4> {
    u32 w[1];
};
template <>
struct Vec<8> {
    u32 w[2];
};
template <>
struct Vec<16> {
    u32 w[4];
};

// B: int8_t (generic pointer: LDS, or a kernel argument, which the compiler knows is global) or
// gi8 (a global pointer read from memory). The accesses use clang vector types, which copy by
// value in any address space (HIP's uint4 has a copy constructor, which binds generic
// references only).
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
typedef u32 u32x2 __attribute__((ext_vector_type(2)));

template <int V, class B = int8_t>
__device__ __forceinline__ Vec<V> load_vec(const B* p)
{
    Vec<V> r;
    if constexpr (V == 16) {
        const u32x4 q = *reinterpret_cast<const typename in_as<B, u32x4>::type*>(p);
        r.w[0] = q.x; r.w[1] = q.y; r.w[2] = q.z; r.w[3] = q.w;
    } else if constexpr (V == 8) {
        const u32x2 q = *reinterpret_cast<const typename in_as<B, u32x2>::type*>(p);
        r.w[0] = q.x; r.w[1] = q.y;
    } else {
        r.w[0] = *reinterpret_cast<const typename in_as<B, u32>::type*>(p);
    }
    return r;
}

// Loads that always miss in the CU's L1 (device-scope cache policy sc1, served by the L2): for
// kernels in which another workgroup of the same launch may have rewritten the bytes after this
// CU cached them (the dependency-ordered batch). base is wave-uniform; off < 2^32.
constexpr int COH_SC1 = 16;  // buffer-instruction cache policy bit of gfx950: sc1

__device__ __forceinline__ __amdgpu_buffer_rsrc_t coh_rsrc(const void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, -1, 0x00020000);
}

template <int V, bool COH, class B>
__device__ __forceinline__ Vec<V> load_vec_at(const B* base, u64 off)
{
    if constexpr (!COH) {
        return load_vec<V>(base + off);
    } else {
        const __amdgpu_buffer_rsrc_t r = coh_rsrc((const void*)base);
        Vec<V> v;
        if constexpr (V == 16) {
            const auto q = __builtin_amdgcn_raw_buffer_load_b128(r, (u32)off, 0, COH_SC1);
            v.w[0] = q[0]; v.w[1] = q[1]; v.w[2] = q[2]; v.w[3] = q[3];
        } else if constexpr (V == 8) {
            const auto q = __builtin_amdgcn_raw_buffer_load_b64(r, (u32)off, 0, COH_SC1);
            v.w[0] = q[0]; v.w[1] = q[1];
        } else {
            v.w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, (u32)off, 0, COH_SC1);
        }
        return v;
    }
}

template <bool COH, class B>
__device__ __forceinline__ u32 load_byte_at(const B* base, u64 off)
{
    if constexpr (!COH) return (u32)(uint8_t)base[off];
    else return (u32)__builtin_amdgcn_raw_buffer_load_b8(coh_rsrc((const void*)base), (u32)off, 0, COH_SC1);
}

template <int V, class B = int8_t>
__device__ __forceinline__ void store_vec(B* p, const Vec<V>& r)
{
    if constexpr (V == 16) {
        *reinterpret_cast<typename in_as<B, u32x4>::type*>(p) = u32x4{r.w[0], r.w[1], r.w[2], r.w[3]};
    } else if constexpr (V == 8) {
        *reinterpret_cast<typename in_as<B, u32x2>::type*>(p) = u32x2{r.w[0], r.w[1]};
    } else {
        *reinterpret_cast<typename in_as<B, u32>::type*>(p) = r.w[0];
    }
}

#define LOW_BITS 0x01010101u

// 2 * bits + (u < thr ? 1 : 0) as one compare into VCC and one add with carry-in (the compiler
// otherwise selects 0/1 into a register and shifts and ORs the bits together).
__device__ __forceinline__ u32 shift_in_lt(u32 bits, u32 u, u32 thr)
{
    u32 r;
    asm("v_cmp_lt_u32 vcc, %1, %2\n\tv_addc_co_u32 %0, vcc, %3, %3, vcc"
        : "=v"(r)
        : "v"(u), "v"(thr), "v"(bits)
        : "vcc");
    return r;
}

__device__ __forceinline__ u32 byte_sum(u32 v)  // sum of the 4 bytes (each <= 63)
{
    return (v * LOW_BITS) >> 24;
}

// Per-workgroup threshold table in LDS, indexed by the count byte 4k (byte offset = 4k):
// Metropolis, k = anti-aligned neighbours: table[2d - k] for k < d (k >= d always flips);
// heat bath, k = -1 neighbours: table[2d - k]. Threads 0..7 write it; callers put a barrier
// between this and the first int8_update_chunk.
template <int NDIM, int HB>
__device__ __forceinline__ void int8_fill_table(u32* thr_lds, const Int8Table& tb)
{
    constexpr u32 Z = 2 * NDIM;
    const u32 k = threadIdx.x;
    if (k < 8u) {
        u32 v = 0;
#pragma unroll
        for (u32 i = 0; i <= Z; ++i)
            if (Z - i == k && (HB || k < (u32)NDIM)) v = tb.v[i];
        thr_lds[k] = v;
    }
}

// Horizontal neighbours of V consecutive colour sites j0 .. j0 + V - 1, from the other colour's
// vector C at the same sites and the edge byte e beyond them: p = 1 takes site j + 1 (e = site
// j0 + V), p = 0 takes site j - 1 (e = site j0 - 1).
template <int V>
__device__ __forceinline__ Vec<V> int8_horizontal(const Vec<V>& C, u32 e, u32 p)
{
    constexpr int NWd = V / 4;
    Vec<V> H;
    if (p) {
#pragma unroll
        for (int i = 0; i < NWd; ++i) {
            const u32 nxt = (i + 1 < NWd) ? C.w[i + 1 < NWd ? i + 1 : 0] : e;
            H.w[i] = __builtin_amdgcn_alignbit(nxt, C.w[i], 8u);
        }
    } else {
#pragma unroll
        for (int i = 0; i < NWd; ++i) {
            const u32 prv = (i > 0) ? C.w[i > 0 ? i - 1 : 0] : (e << 24);
            H.w[i] = __builtin_amdgcn_alignbit(C.w[i], prv, 24u);
        }
    }
    return H;
}

// The Philox outputs of the V / 4 quads of V consecutive colour sites whose first site has
// colour index n0: quad q = n0 / 4 + i takes counter (q, t, colour, 0) (truncated to 32 bits).
// They depend on no spin, so callers compute them while the spins are still being loaded.
template <int V>
struct Int8Draws {
    PhiloxOut r[V / 4];
};

// q0: the first quad's index, n0 / 4 truncated to 32 bits (the counter wraps with it).
template <int V>
__device__ __forceinline__ Int8Draws<V> int8_draws_q(u32 q0, u32 t, u32 color, u32 k0, u32 k1)
{
    Int8Draws<V> D;
#pragma unroll
    for (int i = 0; i < V / 4; ++i) {
        const u32 q = q0 + (u32)i;
        const u64 qm = (u64)q * PHILOX_M0;
        D.r[i] = philox_lane((u32)(qm >> 32), (u32)qm, t, color, 0u, k0, k1);
        // pinned where it is written, so that it covers the latency of the loads before it
        asm volatile("" : "+v"(D.r[i].x), "+v"(D.r[i].y), "+v"(D.r[i].z), "+v"(D.r[i].w));
    }
    return D;
}

template <int V>
__device__ __forceinline__ Int8Draws<V> int8_draws(u64 n0, u32 t, u32 color, u32 k0, u32 k1)
{
    return int8_draws_q<V>((u32)(n0 >> 2), t, color, k0, k1);
}

// The new spins of V consecutive colour sites of colour `color` whose first site has colour
// index n0 (own spins S; the other colour at the same sites C, beside them H, in the rows
// either side U and D and, in 3D, in the planes either side ZU and ZD).
// Per 4-site word: one Philox call (scalar-uniform rounds: counter words 1-3 are sweep, colour
// and 0); neighbour counts as bytes 4k from (s ^ n) & 0x04040404 (Metropolis: anti-aligned;
// spins are 0x01 / 0xFF, so bit 2 differs exactly for opposite spins) or n & 0x04040404 (heat
// bath: -1 neighbours); each byte addresses its site's threshold in the LDS table; the four
// borrows of u - threshold accumulate into a 4-bit vector, spread to bytes by one multiply
// ((b * 0x204081) & 0x01010101 moves bit b to byte b). Metropolis adds the forced flips
// (count >= d) by a bytewise compare and flips with s ^ 0xFE; heat bath writes 0x01 or 0xFF.
// minus (when COUNT): the number of -1 spins written.
// One 4-site word: own spins s, the other colour's neighbours (c, h, u, d and in 3D zu, zd),
// the quad's Philox output r; returns the new word.
template <int NDIM, int HB>
__device__ __forceinline__ u32 int8_decide_quad(u32 s, u32 c, u32 h, u32 u, u32 d, u32 zu, u32 zd,
                                                const PhiloxOut& r, const uint8_t* thr_bytes)
{
    constexpr u32 K4 = 0x04040404u;
    constexpr u32 FORCE = (128u - 4u * NDIM) * LOW_BITS;  // byte >= 4d sets its bit 7
    const u32 uu[4] = {r.x, r.y, r.z, r.w};
    u32 cnt;
    if constexpr (HB) {
        cnt = (c & K4) + (h & K4) + (u & K4) + (d & K4);
        if constexpr (NDIM == 3) cnt += (zu & K4) + (zd & K4);
    } else {
        cnt = ((s ^ c) & K4) + ((s ^ h) & K4) + ((s ^ u) & K4) + ((s ^ d) & K4);
        if constexpr (NDIM == 3) cnt += ((s ^ zu) & K4) + ((s ^ zd) & K4);
    }
    u32 bits = 0;
#pragma unroll
    for (int b = 3; b >= 0; --b) {
        const u32 thr = *reinterpret_cast<const u32*>(thr_bytes + __builtin_amdgcn_ubfe(cnt, 8u * b, 8u));
        bits = shift_in_lt(bits, uu[b], thr);
    }
    u32 f01 = __umul24(bits, 0x204081u) & LOW_BITS;  // bits < 16
    if constexpr (!HB) f01 |= ((cnt + FORCE) >> 7) & LOW_BITS;
    const u32 m = (f01 << 8) - (f01 << 1);  // 0xFE in every flagged byte
    return HB ? ~m : (s ^ m);
}

// R: the quads' Philox outputs (int8_draws).
template <int NDIM, int HB, int V, bool COUNT>
__device__ __forceinline__ Vec<V> int8_decide_drawn(const Vec<V>& S, const Vec<V>& C,
                                                    const Vec<V>& H, const Vec<V>& U,
                                                    const Vec<V>& D, const Vec<V>& ZU,
                                                    const Vec<V>& ZD, const Int8Draws<V>& R,
                                                    const u32* thr_lds, u32& minus)
{
    const uint8_t* thr_bytes = reinterpret_cast<const uint8_t*>(thr_lds);
    Vec<V> O;
    minus = 0;
#pragma unroll
    for (int i = 0; i < V / 4; ++i) {
        O.w[i] = int8_decide_quad<NDIM, HB>(S.w[i], C.w[i], H.w[i], U.w[i], D.w[i],
                                            NDIM == 3 ? ZU.w[i] : 0u, NDIM == 3 ? ZD.w[i] : 0u,
                                            R.r[i], thr_bytes);
        if constexpr (COUNT) minus += __popc(O.w[i] & 0x80808080u);
    }
    return O;
}

// The same with each quad's Philox call made beside its decision.
template <int NDIM, int HB, int V, bool COUNT>
__device__ __forceinline__ Vec<V> int8_decide(const Vec<V>& S, const Vec<V>& C, const Vec<V>& H,
                                              const Vec<V>& U, const Vec<V>& D, const Vec<V>& ZU,
                                              const Vec<V>& ZD, u64 n0, const u32* thr_lds,
                                              u32 t, u32 color, u32 k0, u32 k1, u32& minus)
{
    const uint8_t* thr_bytes = reinterpret_cast<const uint8_t*>(thr_lds);
    Vec<V> O;
    minus = 0;
#pragma unroll
    for (int i = 0; i < V / 4; ++i) {
        const u32 q = (u32)((n0 >> 2) + (u64)i);  // quad index, truncated to 32 bits
        const u64 qm = (u64)q * PHILOX_M0;
        const PhiloxOut r = philox_lane((u32)(qm >> 32), (u32)qm, t, color, 0u, k0, k1);
        O.w[i] = int8_decide_quad<NDIM, HB>(S.w[i], C.w[i], H.w[i], U.w[i], D.w[i],
                                            NDIM == 3 ? ZU.w[i] : 0u, NDIM == 3 ? ZD.w[i] : 0u,
                                            r, thr_bytes);
        if constexpr (COUNT) minus += __popc(O.w[i] & 0x80808080u);
    }
    return O;
}

// Update the V consecutive colour sites of chunk ci (one row segment) of colour `color` in
// global memory (int8_decide on the loaded neighbours). Returns the number of -1 spins written
// (for magnetisation sums when COUNT).
// COH: the other colour and the own sites are read with loads that miss in the L1.
// EARLY: the quads' Philox outputs are computed while the loads are in flight (their 4 V
// registers then live across the loads: for kernels with registers to spare).
// B: the byte type of the lattice pointers (int8_t or gi8, see load_vec).
template <int NDIM, int HB, int V, bool COUNT, bool COH = false, bool EARLY = false, class B = int8_t>
__device__ __forceinline__ u32 int8_update_chunk(B* __restrict__ own,
                                                 const B* __restrict__ other, const Int8Geom& g,
                                                 const u32* thr_lds, u32 t, u32 color, u32 k0,
                                                 u32 k1, u32 ci)
{
    const u32 row = fast_div(ci, g.div_chunks);
    const u32 j0 = (ci - row * g.chunks_row) * V;
    const u32 z = fast_div(row, g.div_ly);
    const u32 y = row - z * g.ly;
    const u32 p = (y + z + color) & 1u;
    const u64 rbase = (u64)row * g.lh;
    const u64 n0 = rbase + j0;

    const Vec<V> S = load_vec_at<V, COH>(own, n0);
    const Vec<V> C = load_vec_at<V, COH>(other, n0);
    const u32 yu = (y + 1 == g.ly) ? 0u : y + 1;
    const u32 yd = (y == 0) ? g.ly - 1 : y - 1;
    const Vec<V> U = load_vec_at<V, COH>(other, ((u64)z * g.ly + yu) * g.lh + j0);
    const Vec<V> D = load_vec_at<V, COH>(other, ((u64)z * g.ly + yd) * g.lh + j0);
    Vec<V> ZU, ZD;  // 3D only
    if constexpr (NDIM == 3) {
        const u32 zu = (z + 1 == g.lz) ? 0u : z + 1;
        const u32 zd = (z == 0) ? g.lz - 1 : z - 1;
        ZU = load_vec_at<V, COH>(other, ((u64)zu * g.ly + y) * g.lh + j0);
        ZD = load_vec_at<V, COH>(other, ((u64)zd * g.ly + y) * g.lh + j0);
    }
    // horizontal neighbour: j + 1 when p == 1, j - 1 when p == 0 (one edge byte from memory)
    const u32 je = p ? ((j0 + V == g.lh) ? 0u : j0 + V) : ((j0 == 0) ? g.lh - 1 : j0 - 1);
    const u32 e = load_byte_at<COH>(other, rbase + je);
    u32 minus = 0;
    Vec<V> O;
    if constexpr (EARLY) {
        // the draws need no spin: computed while the loads above are in flight
        const Int8Draws<V> R = int8_draws<V>(n0, t, color, k0, k1);
        __builtin_amdgcn_sched_barrier(0);
        const Vec<V> H = int8_horizontal<V>(C, e, p);
        O = int8_decide_drawn<NDIM, HB, V, COUNT>(S, C, H, U, D, ZU, ZD, R, thr_lds, minus);
    } else {
        const Vec<V> H = int8_horizontal<V>(C, e, p);
        O = int8_decide<NDIM, HB, V, COUNT>(S, C, H, U, D, ZU, ZD, n0, thr_lds, t, color, k0, k1,
                                            minus);
    }
    store_vec<V>(own + n0, O);
    return minus;
}

// ------------------------------------------------------------- whole-row waves ----
// Layouts in which a wave's 64 chunks (w0 .. w0 + 63, w0 a multiple of 64) share their row
// arithmetic: I8_ROWW, chunks_row % 64 == 0 (the wave lies in one row: row, plane, parity
// and the neighbour rows are scalar); I8_SEGW, chunks_row divides 64 and 64 / chunks_row rows
// divide ly (the wave is whole rows of one plane: the plane and its neighbours are scalar,
// lane l holds chunk l % chunks_row of row y0 + l / chunks_row). Loads take a scalar base
// plus a 32-bit lane offset (a plane is below 4 GiB), and the horizontal edge byte comes from
// the lane holding the neighbouring chunk (DPP shift in a row wave, ds_bpermute in a segment
// wave) instead of a load. I8_GENERAL: one chunk's arithmetic per lane (int8_update_chunk).
enum : int { I8_GENERAL = 0, I8_ROWW = 1, I8_SEGW = 2 };

static inline int int8_layout(const Int8Geom& g)
{
    const u32 cr = g.chunks_row;
    if (g.n_chunks % 64u || (unsigned long long)g.ly * g.lh >= (1ull << 32)) return I8_GENERAL;
    if (cr % 64u == 0) return I8_ROWW;
    if (64u % cr == 0 && g.ly % (64u / cr) == 0) return I8_SEGW;
    return I8_GENERAL;
}

template <int NDIM, int HB, int V, bool COUNT, bool COH, bool EARLY, int LAY>
__device__ __forceinline__ u32 int8_update_wave(int8_t* __restrict__ own,
                                                const int8_t* __restrict__ other,
                                                const Int8Geom& g, const u32* thr_lds, u32 t,
                                                u32 color, u32 k0, u32 k1, u32 w0)
{
    static_assert(LAY == I8_ROWW || LAY == I8_SEGW, "whole-row wave layouts only");
    constexpr int NWd = V / 4;
    const u32 lane = threadIdx.x & 63u;
    w0 = uni(w0);
    u64 pbase;       // first byte of the wave's plane (SEGW) or row (ROWW): scalar
    u32 off;         // this lane's chunk, bytes from pbase
    u32 uoff, doff;  // the rows above and below, bytes from pbase (SEGW: within the plane)
    u32 p;           // 1: colour sites at odd columns, horizontal neighbour j + 1
    u32 e;           // the edge byte beyond this lane's vector
    u32 z;
    u32 c0 = 0;           // ROWW: the wave's first chunk in its row
    const int8_t* ubase;  // ROWW: base pointers of the rows above and below
    const int8_t* dbase;
    if constexpr (LAY == I8_ROWW) {
        const u32 row = fast_div(w0, g.div_chunks);
        c0 = w0 - row * g.chunks_row;
        z = fast_div(row, g.div_ly);
        const u32 y = row - z * g.ly;
        p = (y + z + color) & 1u;
        pbase = (u64)row * g.lh;
        const u32 yu = (y + 1u == g.ly) ? 0u : y + 1u;
        const u32 yd = (y == 0u) ? g.ly - 1u : y - 1u;
        ubase = other + ((u64)z * g.ly + yu) * g.lh;
        dbase = other + ((u64)z * g.ly + yd) * g.lh;
        off = (c0 + lane) * (u32)V;
        uoff = off;
        doff = off;
    } else {
        const u32 lg = (u32)__builtin_ctz(g.chunks_row);
        const u32 row0 = w0 >> lg;
        z = fast_div(row0, g.div_ly);
        const u32 y0 = row0 - z * g.ly;
        pbase = (u64)z * g.ly * g.lh;
        const u32 c = lane & (g.chunks_row - 1u);
        const u32 y = y0 + (lane >> lg);
        p = (y + z + color) & 1u;
        const u32 yoff = y * g.lh;
        off = yoff + c * (u32)V;
        uoff = (y + 1u == g.ly) ? c * (u32)V : off + g.lh;
        doff = (y == 0u) ? (g.ly - 1u) * g.lh + c * (u32)V : off - g.lh;
        ubase = other + pbase;
        dbase = other + pbase;
    }
    int8_t* own_p = own + pbase;
    const int8_t* oth_p = other + pbase;
    const Vec<V> S = load_vec_at<V, COH>(own_p, off);
    const Vec<V> C = load_vec_at<V, COH>(oth_p, off);
    const Vec<V> U = load_vec_at<V, COH>(ubase, uoff);
    const Vec<V> D = load_vec_at<V, COH>(dbase, doff);
    Vec<V> ZU, ZD;  // 3D only
    if constexpr (NDIM == 3) {
        // the same row of the planes either side: pbase is a plane (SEGW) or row (ROWW) start,
        // and the planes' offsets are whole multiples of a plane
        const u64 plane = (u64)g.ly * g.lh;
        const u64 zu = (z + 1u == g.lz) ? 0u : (u64)(z + 1u);
        const u64 zd = (z == 0u) ? (u64)(g.lz - 1u) : (u64)(z - 1u);
        const u64 in_plane = pbase - (u64)z * plane;  // 0 for SEGW
        ZU = load_vec_at<V, COH>(other + zu * plane + in_plane, off);
        ZD = load_vec_at<V, COH>(other + zd * plane + in_plane, off);
    }
    u32 ew = 0;  // ROWW: the word holding the edge byte of the wave's outer lane
    if constexpr (LAY == I8_ROWW) {
        const u32 eo = p ? (((c0 + 64u) * (u32)V == g.lh) ? 0u : (c0 + 64u) * (u32)V)
                         : (c0 == 0u ? g.lh - 4u : c0 * (u32)V - 4u);
        ew = load_vec_at<4, COH>(oth_p, eo).w[0];
    }
    const Int8Draws<V> R = EARLY ? int8_draws_q<V>((u32)(pbase >> 2) + (off >> 2), t, color, k0, k1)
                                 : Int8Draws<V>{};
    if constexpr (EARLY) __builtin_amdgcn_sched_barrier(0);
    if constexpr (LAY == I8_ROWW) {
        // lane i takes lane i + 1's first word (p = 1) or lane i - 1's last word (p = 0); the
        // wave's outer lane keeps the loaded edge word
        const u32 nb = p ? (u32)__builtin_amdgcn_update_dpp((int)ew, (int)C.w[0], 0x130, 0xF, 0xF, false)
                         : (u32)__builtin_amdgcn_update_dpp((int)ew, (int)C.w[NWd - 1], 0x138, 0xF, 0xF, false);
        e = p ? (nb & 0xFFu) : (nb >> 24);
    } else {
        // the lane holding the next (p = 1) or previous (p = 0) chunk of the same row, wrapping
        // within the row (the torus); that lane has the same parity
        const u32 cr = g.chunks_row;
        const u32 c = lane & (cr - 1u);
        const u32 src = (lane - c) + ((c + (p ? 1u : cr - 1u)) & (cr - 1u));
        const u32 nb = (u32)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)(p ? C.w[0] : C.w[NWd - 1]));
        e = p ? (nb & 0xFFu) : (nb >> 24);
    }
    const Vec<V> H = int8_horizontal<V>(C, e, p);
    u32 minus = 0;
    Vec<V> O;
    if constexpr (EARLY)
        O = int8_decide_drawn<NDIM, HB, V, COUNT>(S, C, H, U, D, ZU, ZD, R, thr_lds, minus);
    else
        O = int8_decide<NDIM, HB, V, COUNT>(S, C, H, U, D, ZU, ZD, pbase + off, thr_lds, t, color,
                                            k0, k1, minus);
    store_vec<V>(own_p + off, O);
    return minus;
}
