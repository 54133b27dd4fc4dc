// Device-side conversions between a full lattice (rows, Lx) of int8 spins and the two colour
// arrays (src/pyising/lattice.py:49-90 of the reference): int8 compressed colours, or packed
// uint64 words for multi-spin coding. A full lattice is uploaded once and split on the GPU,
// so a multi-GB lattice never needs host-side gather temporaries.
//
// Input spins: any value > 0 is +1, anything else is -1. This accepts both +-1 lattices and
// raw {0, 1} draws (numpy default_rng(seed).integers(0, 2) before the "* 2 - 1" of
// src/pyising/lattice.py:95-96 maps 0 -> -1, 1 -> +1: the same spins).
#include "common.hpp"

// row r = z * ly + y has parity (y + z) & 1; colour-c site j at x = 2 j + ((parity + c) & 1)
__device__ __forceinline__ u32 row_par(u32 r, u32 ly)
{
    const u32 z = r / ly;
    return ((r - z * ly) + z) & 1u;
}

__device__ __forceinline__ u32 spin_byte(u32 v)  // v: input byte -> 0x01 / 0xFF
{
    return ((int8_t)(uint8_t)v > 0) ? 0x01u : 0xFFu;
}

// Grid-stride loops throughout: a dispatch may not hold 2^32 work-items, and a large lattice
// (msc 131072^2: 2^27 words x 64 lanes) needs more than that.

// one thread per 8 full-lattice bytes (4 sites of each colour); lx % 8 == 0
__global__ __launch_bounds__(256) void split_int8_kernel(const uint8_t* __restrict__ full,
                                                         int8_t* __restrict__ c0,
                                                         int8_t* __restrict__ c1, u32 rows, u32 ly,
                                                         u32 lx)
{
    const u32 per_row = lx / 8;
    const u64 n = (u64)rows * per_row;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32 r = (u32)(i / per_row);
        const u32 k = (u32)(i - (u64)r * per_row);
        const uint2 v = *reinterpret_cast<const uint2*>(full + (u64)r * lx + 8ull * k);
        u32 ev = 0, od = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const u32 src = b < 2 ? v.x : v.y;
            const u32 sh = 16 * (b & 1);
            ev |= spin_byte((src >> sh) & 0xFFu) << (8 * b);
            od |= spin_byte((src >> (sh + 8)) & 0xFFu) << (8 * b);
        }
        const u32 par = row_par(r, ly);
        const u64 o = (u64)r * (lx / 2) + 4ull * k;
        *reinterpret_cast<u32*>(c0 + o) = par ? od : ev;
        *reinterpret_cast<u32*>(c1 + o) = par ? ev : od;
    }
}

__global__ __launch_bounds__(256) void merge_int8_kernel(const int8_t* __restrict__ c0,
                                                         const int8_t* __restrict__ c1,
                                                         uint8_t* __restrict__ full, u32 rows,
                                                         u32 ly, u32 lx)
{
    const u32 per_row = lx / 8;
    const u64 n = (u64)rows * per_row;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32 r = (u32)(i / per_row);
        const u32 k = (u32)(i - (u64)r * per_row);
        const u64 o = (u64)r * (lx / 2) + 4ull * k;
        const u32 a = *reinterpret_cast<const u32*>(c0 + o);
        const u32 b = *reinterpret_cast<const u32*>(c1 + o);
        const u32 par = row_par(r, ly);
        const u32 ev = par ? b : a, od = par ? a : b;
        u32 lo = 0, hi = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const u32 pair = ((ev >> (8 * q)) & 0xFFu) | (((od >> (8 * q)) & 0xFFu) << 8);
            if (q < 2) lo |= pair << (16 * q);
            else hi |= pair << (16 * (q - 2));
        }
        *reinterpret_cast<uint2*>(full + (u64)r * lx + 8ull * k) = make_uint2(lo, hi);
    }
}

// one wave per packed word pair (row r, word x): lane b handles colour site j = 64 x + b,
// reads the two bytes at x_full = 2 j, 2 j + 1 and the wave ballot is the word. 2D: ly = rows.
// The word index is uniform across the wave, so every wave runs the loop in lockstep.
__global__ __launch_bounds__(256) void split_pack_kernel(const uint8_t* __restrict__ full,
                                                         u64* __restrict__ w0, u64* __restrict__ w1,
                                                         u32 rows, u32 nw)
{
    const u32 lane = threadIdx.x & 63u;
    const u64 n = (u64)rows * nw;
    const u64 stride = ((u64)gridDim.x * blockDim.x) >> 6;
    const u32 lx = nw * 128u;
    for (u64 wi = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6; wi < n; wi += stride) {
        const u32 r = (u32)(wi / nw);
        const u32 x = (u32)(wi - (u64)r * nw);
        const uint16_t pair =
            *reinterpret_cast<const uint16_t*>(full + (u64)r * lx + 2ull * (64ull * x + lane));
        const bool ev = (int8_t)(pair & 0xFF) > 0;
        const bool od = (int8_t)(pair >> 8) > 0;
        const u64 bev = __ballot(ev), bod = __ballot(od);
        if (lane == 0) {
            const bool par = r & 1u;
            w0[wi] = par ? bod : bev;
            w1[wi] = par ? bev : bod;
        }
    }
}

__global__ __launch_bounds__(256) void unpack_merge_kernel(const u64* __restrict__ w0,
                                                           const u64* __restrict__ w1,
                                                           uint8_t* __restrict__ full, u32 rows,
                                                           u32 nw)
{
    const u32 lane = threadIdx.x & 63u;
    const u64 n = (u64)rows * nw;
    const u64 stride = ((u64)gridDim.x * blockDim.x) >> 6;
    for (u64 wi = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6; wi < n; wi += stride) {
        const u32 r = (u32)(wi / nw);
        const u32 x = (u32)(wi - (u64)r * nw);
        const bool par = r & 1u;
        const u64 a = w0[wi], b = w1[wi];
        const u64 ev = par ? b : a, od = par ? a : b;
        const u32 se = ((ev >> lane) & 1ull) ? 0x01u : 0xFFu;
        const u32 so = ((od >> lane) & 1ull) ? 0x01u : 0xFFu;
        *reinterpret_cast<uint16_t*>(full + (u64)r * (nw * 128ull) + 2ull * (64ull * x + lane)) =
            (uint16_t)(se | (so << 8));
    }
}

__global__ __launch_bounds__(256) void fill_u64_kernel(u64* __restrict__ p, u64 n, u64 v)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// ----------------------------------------------------------------------------- hot start ----
// numpy's default_rng(seed).integers(0, 2, n, int8) on the device (the reference's hot start,
// src/pyising/lattice.py:93-96). The bit generator is PCG64, PCG XSL-RR 128/64 (O'Neill 2014):
// state <- state * M + inc mod 2^128, output rotr64(hi ^ lo, state >> 122) of the new state.
// For range 2, numpy's int8 path takes bytes of the 64-bit outputs in little-endian order and
// keeps byte >> 7, so output byte i is bit 8 (i % 8) + 7 of draw i / 8. The host passes the
// seeded (state, inc) numpy computes; each thread jumps ahead (LCG advance by squaring, Brown
// 1994) to its first draw and writes HOT_DRAWS draws as 8 bytes each.
#define HOT_DRAWS 16u

struct U128 {
    u64 hi, lo;
};

__device__ __forceinline__ U128 mul128(U128 a, U128 b)  // mod 2^128
{
    return U128{__umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo, a.lo * b.lo};
}

__device__ __forceinline__ U128 add128(U128 a, U128 b)
{
    const u64 lo = a.lo + b.lo;
    return U128{a.hi + b.hi + (lo < a.lo ? 1ull : 0ull), lo};
}

__device__ __forceinline__ u64 pcg_output(U128 s)
{
    const u64 x = s.hi ^ s.lo;
    const u32 r = (u32)(s.hi >> 58);
    return (x >> r) | (x << ((64u - r) & 63u));
}

__global__ __launch_bounds__(256) void hot_start_kernel(u64* __restrict__ out, u64 n_draws,
                                                        U128 state, U128 inc)
{
    const U128 mult = U128{0x2360ED051FC65DA4ull, 0x4385DF649FCCF645ull};
    const u64 first = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * HOT_DRAWS;
    if (first >= n_draws) return;
    // advance `state` by `first` steps
    U128 cur_mult = mult, cur_plus = inc, acc_mult = U128{0, 1}, acc_plus = U128{0, 0};
    for (u64 d = first; d; d >>= 1) {
        if (d & 1) {
            acc_mult = mul128(acc_mult, cur_mult);
            acc_plus = add128(mul128(acc_plus, cur_mult), cur_plus);
        }
        cur_plus = mul128(add128(cur_mult, U128{0, 1}), cur_plus);
        cur_mult = mul128(cur_mult, cur_mult);
    }
    U128 s = add128(mul128(acc_mult, state), acc_plus);
    const u64 last = first + HOT_DRAWS < n_draws ? first + HOT_DRAWS : n_draws;
    for (u64 k = first; k < last; ++k) {
        s = add128(mul128(s, mult), inc);
        out[k] = (pcg_output(s) >> 7) & 0x0101010101010101ull;
    }
}

// Blocks of 256 for `threads` work-items of a grid-stride layout kernel, at most 8192 blocks
// (32 per CU: several generations of workgroups, enough to stream at full bandwidth).
static unsigned grid_for(u64 threads)
{
    const u64 blocks = (threads + 255) / 256;
    return (unsigned)(blocks < 8192 ? blocks : 8192);
}

extern "C" {

int ising_split_int8(const void* full, int8_t* c0, int8_t* c1, long long rows, int ly, int lx,
                     void* stream)
{
    if (!full || !c0 || !c1) return ising_fail(ISING_EINVAL, "split_int8: null pointer");
    if (rows <= 0 || ly <= 0 || rows % ly || lx <= 0 || lx % 8)
        return ising_fail(ISING_EINVAL, "split_int8: bad geometry (Lx % 8 == 0, rows % ly == 0)");
    const u64 n = (u64)rows * (u64)(lx / 8);
    hipLaunchKernelGGL(split_int8_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream),
                       (const uint8_t*)full, c0, c1, (u32)rows, (u32)ly, (u32)lx);
    ISING_LAUNCHED("split_int8_kernel");
    return ISING_OK;
}

int ising_merge_int8(const int8_t* c0, const int8_t* c1, void* full, long long rows, int ly, int lx,
                     void* stream)
{
    if (!full || !c0 || !c1) return ising_fail(ISING_EINVAL, "merge_int8: null pointer");
    if (rows <= 0 || ly <= 0 || rows % ly || lx <= 0 || lx % 8)
        return ising_fail(ISING_EINVAL, "merge_int8: bad geometry (Lx % 8 == 0, rows % ly == 0)");
    const u64 n = (u64)rows * (u64)(lx / 8);
    hipLaunchKernelGGL(merge_int8_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), c0,
                       c1, (uint8_t*)full, (u32)rows, (u32)ly, (u32)lx);
    ISING_LAUNCHED("merge_int8_kernel");
    return ISING_OK;
}

int ising_split_pack(const void* full, uint64_t* w0, uint64_t* w1, long long rows, int lx,
                     void* stream)
{
    if (!full || !w0 || !w1) return ising_fail(ISING_EINVAL, "split_pack: null pointer");
    if (rows <= 0 || lx <= 0 || lx % 128)
        return ising_fail(ISING_EINVAL, "split_pack: Lx must be a multiple of 128");
    const u64 nwords = (u64)rows * (u64)(lx / 128);
    hipLaunchKernelGGL(split_pack_kernel, dim3(grid_for(nwords * 64)), dim3(256), 0,
                       as_stream(stream), (const uint8_t*)full, w0, w1, (u32)rows, (u32)(lx / 128));
    ISING_LAUNCHED("split_pack_kernel");
    return ISING_OK;
}

int ising_unpack_merge(const uint64_t* w0, const uint64_t* w1, void* full, long long rows, int lx,
                       void* stream)
{
    if (!full || !w0 || !w1) return ising_fail(ISING_EINVAL, "unpack_merge: null pointer");
    if (rows <= 0 || lx <= 0 || lx % 128)
        return ising_fail(ISING_EINVAL, "unpack_merge: Lx must be a multiple of 128");
    const u64 nwords = (u64)rows * (u64)(lx / 128);
    hipLaunchKernelGGL(unpack_merge_kernel, dim3(grid_for(nwords * 64)), dim3(256), 0,
                       as_stream(stream), w0, w1, (uint8_t*)full, (u32)rows, (u32)(lx / 128));
    ISING_LAUNCHED("unpack_merge_kernel");
    return ISING_OK;
}

// n bytes (n % 8 == 0) of numpy default_rng(seed).integers(0, 2, n, int8), from the seeded
// PCG64 (state, inc) that numpy reports (PCG64(seed).state). Stream-ordered.
int ising_hot_start(void* d_out, long long n, uint64_t state_hi, uint64_t state_lo,
                    uint64_t inc_hi, uint64_t inc_lo, void* stream)
{
    if (!d_out || n < 0 || (n & 7)) return ising_fail(ISING_EINVAL, "hot_start: null buffer or n % 8 != 0");
    if (((uintptr_t)d_out & 7u) != 0) return ising_fail(ISING_EINVAL, "hot_start: buffer not 8-byte aligned");
    if (n == 0) return ISING_OK;
    const u64 draws = (u64)n / 8u;
    const u64 threads = (draws + HOT_DRAWS - 1u) / HOT_DRAWS;
    // one thread per HOT_DRAWS draws, each jumping ahead once: a single generation of threads
    if (threads > 0xFFFFFF00ull) return ising_fail(ISING_ERANGE, "hot_start: n too large");
    hipLaunchKernelGGL(hot_start_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       as_stream(stream),
                       (u64*)d_out, draws, U128{state_hi, state_lo}, U128{inc_hi, inc_lo});
    ISING_LAUNCHED("hot_start_kernel");
    return ISING_OK;
}

int ising_fill_u64(uint64_t* dst, long long n, uint64_t value, void* stream)
{
    if (!dst || n < 0) return ising_fail(ISING_EINVAL, "fill_u64: bad argument");
    if (n == 0) return ISING_OK;
    const u64 blocks = ((u64)n + 255) / 256;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0,
                       as_stream(stream), dst, (u64)n, (u64)value);
    ISING_LAUNCHED("fill_u64_kernel");
    return ISING_OK;
}

}  // extern "C"
