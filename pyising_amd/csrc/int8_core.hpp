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

template <int V>
__device__ __forceinline__ Vec<V> load_vec(const int8_t* p)
{
    Vec<V> r;
    if constexpr (V == 16) {
        const uint4 q = *reinterpret_cast<const uint4*>(p);
        r.w[0] = q.x; r.w[1] = q.y; r.w[2] = q.z; r.w[3] = q.w;
    } else if constexpr (V == 8) {
        const uint2 q = *reinterpret_cast<const uint2*>(p);
        r.w[0] = q.x; r.w[1] = q.y;
    } else {
        r.w[0] = *reinterpret_cast<const u32*>(p);
    }
    return r;
}

template <int V>
__device__ __forceinline__ void store_vec(int8_t* p, const Vec<V>& r)
{
    if constexpr (V == 16) {
        *reinterpret_cast<uint4*>(p) = make_uint4(r.w[0], r.w[1], r.w[2], r.w[3]);
    } else if constexpr (V == 8) {
        *reinterpret_cast<uint2*>(p) = make_uint2(r.w[0], r.w[1]);
    } else {
        *reinterpret_cast<u32*>(p) = r.w[0];
    }
}

#define LOW_BITS 0x01010101u

__device__ __forceinline__ u32 byte_sum(u32 v)  // sum of the 4 bytes (each <= 63)
{
    return (v * LOW_BITS) >> 24;
}

template <int NDIM>
__device__ __forceinline__ u32 metro_thr(u32 a, const Int8Table& tb)
{
    // only a < d can be accepted by the comparison; a >= d always flips
    if constexpr (NDIM == 2) {
        return a == 0 ? tb.v[4] : tb.v[3];
    } else {
        return a == 0 ? tb.v[6] : (a == 1 ? tb.v[5] : tb.v[4]);
    }
}

template <int NDIM>
__device__ __forceinline__ u32 heat_thr(u32 m, const Int8Table& tb)
{
    constexpr int Z = 2 * NDIM;
    u32 thr = tb.v[Z];
#pragma unroll
    for (int k = 1; k <= Z; ++k) thr = (m == (u32)k) ? tb.v[Z - k] : thr;
    return thr;
}


// Update the V consecutive colour sites of chunk ci (one row segment) of colour `color`.
// Returns the number of -1 spins written (for magnetisation sums when COUNT).
template <int NDIM, int HB, int V, bool COUNT>
__device__ __forceinline__ u32 int8_update_chunk(int8_t* __restrict__ own,
                                                 const int8_t* __restrict__ other, const Int8Geom& g,
                                                 const Int8Table& tb, u32 t, u32 color, u32 k0,
                                                 u32 k1, u32 ci)
{
    constexpr int NWd = V / 4;
    const u32 row = fast_div(ci, g.div_chunks);
    const u32 j0 = (ci - row * g.chunks_row) * V;
    const u32 z = fast_div(row, g.div_ly);
    const u32 y = row - z * g.ly;
    const u32 p = (y + z + color) & 1u;
    const u64 rbase = (u64)row * g.lh;
    const u64 n0 = rbase + j0;

    const Vec<V> S = load_vec<V>(own + n0);
    const Vec<V> C = load_vec<V>(other + n0);
    const u32 yu = (y + 1 == g.ly) ? 0u : y + 1;
    const u32 yd = (y == 0) ? g.ly - 1 : y - 1;
    const Vec<V> U = load_vec<V>(other + ((u64)z * g.ly + yu) * g.lh + j0);
    const Vec<V> D = load_vec<V>(other + ((u64)z * g.ly + yd) * g.lh + j0);
    Vec<V> ZU, ZD;
    if constexpr (NDIM == 3) {
        const u32 zu = (z + 1 == g.lz) ? 0u : z + 1;
        const u32 zd = (z == 0) ? g.lz - 1 : z - 1;
        ZU = load_vec<V>(other + ((u64)zu * g.ly + y) * g.lh + j0);
        ZD = load_vec<V>(other + ((u64)zd * g.ly + y) * g.lh + j0);
    }
    // horizontal neighbour: j + 1 when p == 1, j - 1 when p == 0 (one edge byte from memory)
    const u32 je = p ? ((j0 + V == g.lh) ? 0u : j0 + V) : ((j0 == 0) ? g.lh - 1 : j0 - 1);
    const u32 e = (u32)(uint8_t)other[rbase + je];
    Vec<V> H;
    if (p) {
#pragma unroll
        for (int i = 0; i < NWd; ++i) {
            const u32 nxt = (i + 1 < NWd) ? C.w[i + 1 < NWd ? i + 1 : 0] : e;
            H.w[i] = (C.w[i] >> 8) | (nxt << 24);
        }
    } else {
#pragma unroll
        for (int i = 0; i < NWd; ++i) {
            const u32 prv = (i > 0) ? (C.w[i > 0 ? i - 1 : 0] >> 24) : e;
            H.w[i] = (C.w[i] << 8) | prv;
        }
    }

    Vec<V> O;
    u32 minus = 0;
#pragma unroll
    for (int i = 0; i < NWd; ++i) {
        const u32 q = (u32)((n0 >> 2) + (u64)i);  // quad index, truncated to 32 bits
        const PhiloxOut r = philox4x32_10(q, t, color, 0u, k0, k1);
        const u32 uu[4] = {r.x, r.y, r.z, r.w};
        u32 cnt;
        if constexpr (HB) {
            cnt = ((C.w[i] >> 1) & LOW_BITS) + ((H.w[i] >> 1) & LOW_BITS) +
                  ((U.w[i] >> 1) & LOW_BITS) + ((D.w[i] >> 1) & LOW_BITS);
            if constexpr (NDIM == 3)
                cnt += ((ZU.w[i] >> 1) & LOW_BITS) + ((ZD.w[i] >> 1) & LOW_BITS);
        } else {
            const u32 s = S.w[i];
            cnt = (((s ^ C.w[i]) >> 1) & LOW_BITS) + (((s ^ H.w[i]) >> 1) & LOW_BITS) +
                  (((s ^ U.w[i]) >> 1) & LOW_BITS) + (((s ^ D.w[i]) >> 1) & LOW_BITS);
            if constexpr (NDIM == 3)
                cnt += (((s ^ ZU.w[i]) >> 1) & LOW_BITS) + (((s ^ ZD.w[i]) >> 1) & LOW_BITS);
        }
        u32 out = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const u32 c = (cnt >> (8 * b)) & 0xFFu;
            if constexpr (HB) {
                const bool up = uu[b] < heat_thr<NDIM>(c, tb);
                out |= (up ? 0x01u : 0xFFu) << (8 * b);
            } else {
                const bool flip = (c >= (u32)NDIM) || (uu[b] < metro_thr<NDIM>(c, tb));
                out |= (flip ? 0xFEu : 0x00u) << (8 * b);
            }
        }
        O.w[i] = HB ? out : (S.w[i] ^ out);
        if constexpr (COUNT) minus += byte_sum((O.w[i] >> 1) & LOW_BITS);
    }
    store_vec<V>(own + n0, O);
    return minus;
}
