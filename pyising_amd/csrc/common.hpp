// Shared definitions of the gfx950 Ising library: status handling for the C ABI, the
// in-register Philox4x32-10, and launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ising_hip.h"  // the C ABI: every entry point here is checked against its declaration

#include <string>

typedef uint32_t u32;
typedef uint64_t u64;

// ---------------------------------------------------------------- C-ABI status handling ----
// Every entry point returns 0 on success or a negative code; the message of the last failure
// on the calling thread is available from ising_last_error().
enum {
    ISING_OK = 0,
    ISING_EINVAL = -1,   // bad argument (shape, size, pointer)
    ISING_EHIP = -2,     // a HIP runtime call failed
    ISING_ERANGE = -3,   // a size beyond what the entry point supports
};

void ising_set_error(const std::string& msg);
int ising_fail(int code, const std::string& msg);
int ising_check_hip(hipError_t e, const char* what);

#define ISING_HIP(call)                                              \
    do {                                                             \
        hipError_t _e = (call);                                      \
        if (_e != hipSuccess) return ising_check_hip(_e, #call);     \
    } while (0)

#define ISING_LAUNCHED(what)                                         \
    do {                                                             \
        hipError_t _e = hipGetLastError();                           \
        if (_e != hipSuccess) return ising_check_hip(_e, what);      \
    } while (0)

// ------------------------------------------------------------------------------ Philox ----
// Philox4x32-10 (Salmon et al., SC'11). Round: (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2,
// c <- (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0), then the key is bumped by (W0, W1). The
// 32x32->64 products compile to one v_mad_u64_u32 each; where a counter word is uniform across
// the wave (sweep, colour) the compiler keeps the early rounds' products on the scalar unit.
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

struct PhiloxOut {
    u32 x, y, z, w;
};

// a ^ b ^ c in one VALU op: gfx950's v_bitop3_b32 with truth table 0x96.
__device__ __forceinline__ u32 xor3(u32 a, u32 b, u32 c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// On CDNA the 32x32->64 multiply (v_mad_u64_u32) issues at the full 64-bit rate, so the
// round's XORs dominate; rounds 1-9 use one v_bitop3 per output word instead of two XORs.
// Round 0 keeps plain XORs so that its wave-uniform half (counter words 1 and 2 are the sweep
// and the colour) stays on the scalar unit.
__device__ __forceinline__ PhiloxOut philox4x32_10(u32 c0, u32 c1, u32 c2, u32 c3, u32 k0, u32 k1)
{
    {
        const u64 p0 = (u64)c0 * PHILOX_M0;
        const u64 p1 = (u64)c2 * PHILOX_M1;
        const u32 n0 = (u32)(p1 >> 32) ^ c1 ^ k0;
        const u32 n2 = (u32)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (u32)p1;
        c2 = n2;
        c3 = (u32)p0;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
#pragma unroll
    for (int r = 1; r < 10; ++r) {
        const u64 p0 = (u64)c0 * PHILOX_M0;
        const u64 p1 = (u64)c2 * PHILOX_M1;
        const u32 n0 = xor3((u32)(p1 >> 32), c1, k0);
        const u32 n2 = xor3((u32)(p0 >> 32), c3, k1);
        c0 = n0;
        c1 = (u32)p1;
        c2 = n2;
        c3 = (u32)p0;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    return PhiloxOut{c0, c1, c2, c3};
}

// Wave-uniform value forced into a scalar register.
__device__ __forceinline__ u32 uni(u32 x) { return __builtin_amdgcn_readfirstlane(x); }

// Device (global) memory as an explicit address space. A pointer that a kernel reads from
// memory (a batch descriptor) is otherwise generic, and every access through it becomes a flat
// instruction: it counts against both the vector-memory and the LDS wait counters, so a wait
// for LDS also waits for it, and it takes no scalar base address. Kernel arguments are known to
// be global, so only descriptor pointers need the cast; accesses must go through the typed
// pointer (a cast back to a generic pointer loses the address space again).
typedef __attribute__((address_space(1))) u64 gu64;
typedef __attribute__((address_space(1))) int8_t gi8;

// T in the address space of the byte type B (int8_t: generic, gi8: global), keeping const.
template <class B, class T>
struct in_as {
    using type = T;
};
template <class T>
struct in_as<gi8, T> {
    using type = __attribute__((address_space(1))) T;
};
template <class B, class T>
struct in_as<const B, T> {
    using type = const typename in_as<B, T>::type;
};

// Philox4x32-10 of the counter (c0, c1, c2, c3) where only c0 varies across the wave and c1,
// c2, c3, k0, k1 are wave-uniform (REQUIRED: callers pass sweep, colour, level and key). The
// caller passes (hi0, lo0) = c0 * M0, which is loop-invariant over levels. The uniform halves
// of rounds 0-2 stay on the scalar unit and every per-lane XOR with two uniform inputs becomes
// one VALU XOR: 36 VALU ops per call. Same outputs as philox4x32_10.
// C3_LANE: c3 may vary across the wave too (it only enters round 0's XOR; same op count).
template <bool C3_LANE = false>
__device__ __forceinline__ PhiloxOut philox_lane(u32 hi0, u32 lo0, u32 c1, u32 c2, u32 c3, u32 k0,
                                                 u32 k1)
{
    // round 0: the c2 product is uniform
    const u64 q1 = (u64)c2 * PHILOX_M1;
    const u32 u0 = uni((u32)(q1 >> 32) ^ c1 ^ k0);
    const u32 u1 = (u32)q1;
    const u32 x2 = C3_LANE ? xor3(hi0, c3, uni(k1)) : hi0 ^ uni(c3 ^ k1);
    const u32 x3 = lo0;
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
    // round 1: the first product is uniform
    const u64 q0 = (u64)u0 * PHILOX_M0;
    const u64 p1 = (u64)x2 * PHILOX_M1;
    const u32 y0 = (u32)(p1 >> 32) ^ uni(u1 ^ k0);
    const u32 y1 = (u32)p1;
    const u32 y2 = x3 ^ uni((u32)(q0 >> 32) ^ k1);
    const u32 y3 = uni((u32)q0);
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
    // round 2: counter word 3 is still uniform
    u32 c0;
    {
        const u64 p0 = (u64)y0 * PHILOX_M0;
        const u64 pp = (u64)y2 * PHILOX_M1;
        c0 = xor3((u32)(pp >> 32), y1, k0);
        c1 = (u32)pp;
        c2 = (u32)(p0 >> 32) ^ uni(y3 ^ k1);
        c3 = (u32)p0;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
#pragma unroll
    for (int r = 3; r < 10; ++r) {
        const u64 p0 = (u64)c0 * PHILOX_M0;
        const u64 p1r = (u64)c2 * PHILOX_M1;
        const u32 n0 = xor3((u32)(p1r >> 32), c1, k0);
        const u32 n2 = xor3((u32)(p0 >> 32), c3, k1);
        c0 = n0;
        c1 = (u32)p1r;
        c2 = n2;
        c3 = (u32)p0;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    return PhiloxOut{c0, c1, c2, c3};
}

// ---------------------------------------------------------- division by a runtime constant ----
// q = n / d for 32-bit n with a host-precomputed magic number (Granlund & Montgomery 1994,
// round-up variant): t = mulhi(n, m); q = (t + ((n - t) >> s1)) >> s2. d == 1 uses m = 0,
// s1 = s2 = 0. Replaces the ~40-instruction integer division in per-site index math.
struct FastDiv {
    u32 d, m, s1, s2;
};

static inline FastDiv make_fastdiv(u32 d)
{
    FastDiv f;
    f.d = d;
    if (d <= 1) {
        f.m = 0;
        f.s1 = 0;
        f.s2 = 0;
        return f;
    }
    u32 l = 0;
    while ((1ull << l) < d) ++l;  // ceil(log2 d)
    f.m = (u32)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    f.s1 = 1;
    f.s2 = l - 1;
    return f;
}

__device__ __forceinline__ u32 fast_div(u32 n, const FastDiv& f)
{
    const u32 t = __umulhi(n, f.m);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Number of set bits of a wave-wide mask below the calling lane (wave64).
__device__ __forceinline__ u32 lane_rank(u64 mask)
{
    return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

// --------------------------------------------------------------------- launch helpers ----
// MI355X dispatches workgroup i to XCD i % 8. For stencil kernels whose neighbouring blocks
// share halo rows, remap so that each XCD walks one contiguous band of blocks: rows then
// stay in that XCD's L2. Valid when nblocks % 8 == 0 (otherwise identity).
__device__ __forceinline__ u32 xcd_band_block(u32 b, u32 nblocks)
{
    if (nblocks & 7u) return b;
    const u32 per = nblocks >> 3;
    return (b & 7u) * per + (b >> 3);
}

// ------------------------------------------------------------- exact reductions ----
// Observables reduce without atomics: each workgroup (256 threads) of a grid of at most
// OBS_MAX_BLOCKS writes its two integer partial sums to its own slots of the caller's scratch
// (ISING_OBS_SCRATCH_BYTES), and one finalize workgroup adds them up. Device-scope atomics on
// one address serialize at the memory side (one per wave cost ~0.2 ms on a 256^3 lattice).
constexpr u32 OBS_MAX_BLOCKS = ISING_OBS_SCRATCH_BYTES / 16u;

// Sum of a and b over the 256 threads of the workgroup (wave shuffles, then LDS); every thread
// must call it; thread 0 receives the totals.
__device__ __forceinline__ void block_sum2_u64(u64& a, u64& b)
{
    __shared__ u64 part[2][4];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if ((threadIdx.x & 63u) == 0u) {
        part[0][threadIdx.x >> 6] = a;
        part[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        b = part[1][0] + part[1][1] + part[1][2] + part[1][3];
    }
}

// out[0] = ka * A + ca, out[1] = kb * B + cb with (A, B) the sums of the nblocks partial pairs
// (one 256-thread workgroup). Nothing needs clearing between calls: every partial is rewritten.
// Static (one copy per source file that launches it); unused in the others.
[[maybe_unused]] static __global__ void __launch_bounds__(256) obs_finalize(const unsigned long long* partials,
                                                          u32 nblocks, long long* out,
                                                          long long ka, long long ca, long long kb,
                                                          long long cb)
{
    u64 a = 0, b = 0;
    for (u32 i = threadIdx.x; i < nblocks; i += 256u) {
        a += partials[2u * i];
        b += partials[2u * i + 1u];
    }
    block_sum2_u64(a, b);
    if (threadIdx.x == 0) {
        out[0] = ka * (long long)a + ca;
        out[1] = kb * (long long)b + cb;
    }
}

static inline unsigned obs_grid(u64 work_items)
{
    const u64 blocks = (work_items + 255u) / 256u;
    return (unsigned)(blocks < OBS_MAX_BLOCKS ? (blocks ? blocks : 1u) : OBS_MAX_BLOCKS);
}

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
