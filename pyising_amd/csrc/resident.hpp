// Resident sweeps of small 2D lattices, shared by the int8 (int8_resident.hip) and packed msc
// (msc_resident.hip) kernels: one cooperative launch holds every half sweep of a call, each
// workgroup keeping a band of rows of both colours in LDS for the whole launch.
//
// Decomposition. B workgroups (B | rows, B a multiple of 8, at most one per CU, launched
// cooperatively so that all are resident) own bands of R = rows / B consecutive rows. A
// workgroup holds its band plus T halo rows on each side (T <= R), both colours, in LDS, and
// runs T half sweeps per round without talking to anybody: half sweep i of a round updates the
// rows within T - 1 - i of its band (redundantly with its neighbours: the same sites, the same
// Philox counters, so the same bits), which needs the other colour within T - i rows, which the
// previous half sweep (or the round's start) left valid. After a round the band's T edge rows
// of both colours are current; they go to the neighbour bands through a double-buffered
// workspace, and the next round starts once the neighbours' rows have arrived.
//
// Exchange, two forms (the plan picks one per lattice shape):
// * tagged words (small exchanges, at most one word per thread and part): every 4 bytes of a
//   published row travel in one 8-byte word together with the round's tag, (tag << 32) | bytes,
//   written and read as single-copy-atomic device-scope (agent) accesses: a reader that sees
//   the tag it waits for has the bytes of that round, so the receiver polls the rows
//   themselves. No acknowledgement wait, counter or cache maintenance sits between the
//   writer's stores and the reader's loads, and device-scope atomics of one location are
//   coherent across the XCDs' L2s, so bands on different XCDs exchange the same way. Tags are
//   epoch + round + 1, unique over the workspace's life (the epoch advances by the rounds of
//   every launch), so a word left from an earlier round or launch never matches.
// * rows and a counter (larger exchanges, where the tagged form's doubled bytes cost more than
//   its shorter handshake saves): the rows are copied out with plain stores, the writer waits
//   for their acknowledgement and then publishes its round count; the reader waits for the
//   neighbours' counts and copies the rows in with loads that miss in the L1. Within an XCD
//   the L2 is the coherence point; the bands at an XCD boundary add the device-scope release
//   (L2 write-back) before publishing and the acquire (L2 invalidate) after waiting.
// Either way a writer reuses a buffer (parity) two rounds later, after it has received the
// reader's rows of the round in between, which the reader published after consuming the
// buffer. Waits are bounded. Before writing its rows back to the lattice a band waits until
// both neighbours have finished the launch (their counters): their initial halo loads, which
// read this band's rows, are done.
//
// Placement. Bands are numbered so that the B / 8 bands of one XCD are consecutive (band =
// XCD * B / 8 + workgroup / 8, from HW_REG_XCC_ID): neighbours share an L2 except at the eight
// XCD boundaries.
//
// A row is row_bytes bytes (int8: lh; msc: 8 nw), a multiple of 16.
#pragma once

#include "int8_core.hpp"
#include "placement.hpp"

// Workspace (u32 words): [0] failure bits (1: a wait gave up), [RS_HDR + b] rounds completed by
// band b (from the workspace's zero state, mod 2^32: after every round in the row form, after a
// launch's last round in the tagged form), then the exchange area, [parity][band][side: 0 =
// first T rows, 1 = last T rows][colour][T rows]: row_bytes bytes per row (row form) or
// row_bytes / 4 words of 8 bytes, (tag << 32) | 4 row bytes (tagged form).
constexpr u32 RS_HDR = 64;
constexpr u32 RS_SPIN_LIMIT = 1u << 20;
constexpr u32 RS_THREADS = 1024;
constexpr u32 RS_MAX_LDS = 64u * 1024u;  // both colours, band and halos
// tagged exchange words a thread receives per round at most: one per part (the plan uses the
// tagged form only while a part has at most RS_THREADS words)
constexpr u32 RS_RX = 4;

struct ResidentBands {
    u32 rows, row_bytes, R, T, B;
    u32 n_halves;  // half sweeps of this launch, starting with colour 0
    u32 epoch;     // rounds every band counter holds when the launch starts
    u32 one_l2;    // 1: every workgroup on one XCD
    u32 tagged;    // 1: tagged-word exchange, 0: rows and a counter
    u32* ws;
};

__device__ __forceinline__ u32* rs_flag(const ResidentBands& a, u32 band) { return a.ws + RS_HDR + band; }

// Exchange words of (parity, band, side, colour): T rows of row_bytes / 4 words. The header
// (RS_HDR + B words, B a multiple of 8) keeps them 8-byte aligned.
__device__ __forceinline__ u64* rs_xwords(const ResidentBands& a, u32 par, u32 band, u32 side,
                                          u32 color)
{
    u64* base = reinterpret_cast<u64*>(a.ws + RS_HDR + a.B);
    const size_t words = (size_t)a.T * a.row_bytes / 4u;
    return base + ((((size_t)par * a.B + band) * 2u + side) * 2u + color) * words;
}

// Exchange rows of (parity, band, side, colour) in the row form: T rows of row_bytes bytes.
__device__ __forceinline__ int8_t* rs_xrows(const ResidentBands& a, u32 par, u32 band, u32 side,
                                            u32 color)
{
    int8_t* base = reinterpret_cast<int8_t*>(a.ws + RS_HDR + a.B);
    const size_t rows_t = (size_t)a.T * a.row_bytes;
    return base + ((((size_t)par * a.B + band) * 2u + side) * 2u + color) * rows_t;
}

// Copy n_rows rows, 16 bytes per thread and step. COH_SRC: the source may have been written by
// another workgroup of this launch (L1-missing loads).
template <bool COH_SRC>
__device__ __forceinline__ void rs_copy_rows(int8_t* dst, const int8_t* src, u32 n_rows,
                                             u32 row_bytes)
{
    const u32 n16 = n_rows * row_bytes / 16u;
    for (u32 i = threadIdx.x; i < n16; i += RS_THREADS) {
        const Vec<16> v = load_vec_at<16, COH_SRC>(src, (u64)i * 16u);
        store_vec<16>(dst + (size_t)i * 16u, v);
    }
}

// LDS rows 0 .. n_rows - 1 from global rows first, first + 1, ... (mod rows; first < rows and
// n_rows <= rows).
__device__ __forceinline__ void rs_load_global(const ResidentBands& a, int8_t* lds_c,
                                               const int8_t* g, u32 first, u32 n_rows)
{
    const u32 per_row = a.row_bytes / 16u;
    for (u32 i = threadIdx.x; i < n_rows * per_row; i += RS_THREADS) {
        const u32 r = i / per_row, c = i - r * per_row;
        u32 y = first + r;
        if (y >= a.rows) y -= a.rows;
        const Vec<16> v = load_vec<16>(g + (size_t)y * a.row_bytes + c * 16u);
        store_vec<16>(lds_c + (size_t)r * a.row_bytes + c * 16u, v);
    }
}

// A round's exchange has four parts q = (colour q & 1, side q >> 1) of nq = T x row_bytes / 4
// words each, the T rows of the part. Side 0 = the band's first T rows (LDS rows T .. 2T - 1),
// side 1 = its last T rows (LDS rows R .. R + T - 1); at the neighbours they fill the halo
// rows: the upper neighbour's side 1 -> LDS rows 0 .. T - 1, the lower neighbour's side 0 ->
// LDS rows T + R .. 2T + R - 1. Thread t moves words t, t + RS_THREADS, ... of every part.
constexpr u32 RS_RX_PART = RS_RX / 4u;  // words per part and thread

// Publish this band's edge rows of the round with tag `tag` into buffer `par`.
__device__ __forceinline__ void rs_publish(const ResidentBands& a, const int8_t* lds, u32 nl,
                                           u32 band, u32 par, u32 tag)
{
    const u32 nq = a.T * a.row_bytes / 4u;
#pragma unroll
    for (u32 q = 0; q < 4u; ++q) {
        const u32 c = q & 1u, side = q >> 1;
        const u32* src = reinterpret_cast<const u32*>(lds + ((size_t)c * nl + (side ? a.R : a.T)) * a.row_bytes);
        u64* dst = rs_xwords(a, par, band, side, c);
        for (u32 i = threadIdx.x; i < nq; i += RS_THREADS)
            __hip_atomic_store(dst + i, ((u64)tag << 32) | src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Receive the neighbours' edge rows of the round with tag `tag` from buffer `par` into this
// band's halo rows. Every thread issues its loads (at most RS_RX) at once, then polls the words
// whose tag has not arrived (bounded: RS_SPIN_LIMIT polls, or another band's failure). Returns
// false if this thread's wait gave up (it also sets the failure bit).
__device__ __forceinline__ bool rs_receive(const ResidentBands& a, int8_t* lds, u32 nl, u32 up,
                                           u32 dn, u32 par, u32 tag)
{
    const u32 nq = a.T * a.row_bytes / 4u;
    const u64* src[4];
    u32 dst[4];  // LDS byte offsets of the parts' first words
#pragma unroll
    for (u32 q = 0; q < 4u; ++q) {
        const u32 c = q & 1u, side = q >> 1;
        src[q] = rs_xwords(a, par, side ? up : dn, side, c);
        dst[q] = (c * nl + (side ? 0u : a.T + a.R)) * a.row_bytes;
    }
    u64 v[4][RS_RX_PART];
    u32 pending = 0;  // bit q * RS_RX_PART + j: word j of part q still to arrive
#pragma unroll
    for (u32 q = 0; q < 4u; ++q) {
#pragma unroll
        for (u32 j = 0; j < RS_RX_PART; ++j) {
            const u32 i = threadIdx.x + j * RS_THREADS;
            v[q][j] = 0u;
            if (i < nq) {
                v[q][j] = __hip_atomic_load(src[q] + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                pending |= 1u << (q * RS_RX_PART + j);
            }
        }
    }
    for (u32 spin = 0;; ++spin) {
#pragma unroll
        for (u32 q = 0; q < 4u; ++q) {
#pragma unroll
            for (u32 j = 0; j < RS_RX_PART; ++j) {
                const u32 bit = 1u << (q * RS_RX_PART + j);
                if ((pending & bit) && (u32)(v[q][j] >> 32) == tag) {
                    const u32 i = threadIdx.x + j * RS_THREADS;
                    *reinterpret_cast<u32*>(lds + dst[q] + 4u * i) = (u32)v[q][j];
                    pending &= ~bit;
                }
            }
        }
        if (!pending) return true;
        if ((spin & 255u) == 255u) {
            const u32 bad = __hip_atomic_load(a.ws, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((bad & 1u) || spin >= RS_SPIN_LIMIT) {
                __hip_atomic_fetch_or(a.ws, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(2);
#pragma unroll
        for (u32 q = 0; q < 4u; ++q) {
#pragma unroll
            for (u32 j = 0; j < RS_RX_PART; ++j) {
                if (pending & (1u << (q * RS_RX_PART + j)))
                    v[q][j] = __hip_atomic_load(src[q] + threadIdx.x + j * RS_THREADS, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Wave 0 waits until both neighbour bands' counters reach `target`.
__device__ bool rs_wait(const ResidentBands& a, u32 up, u32 dn, u32 target)
{
    const u32 lane = threadIdx.x & 63u;
    const u32* p = rs_flag(a, lane == 0u ? up : dn);
    bool ok = lane >= 2u;
    for (u32 spin = 0;; ++spin) {
        if (!ok) ok = (int)(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) >= 0;
        if (__ballot(!ok) == 0) return true;
        if ((spin & 255u) == 255u) {
            const u32 bad = __hip_atomic_load(a.ws, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((bad & 1u) || spin >= RS_SPIN_LIMIT) {
                if (lane == 0u) __hip_atomic_fetch_or(a.ws, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// The band of the calling workgroup and its neighbours, plus the round loop.
// UPDATE(own, other, hh, ext, first) runs half sweep hh of the launch (colour hh & 1, sweep
// hh >> 1 of the call) on LDS rows T - ext .. T + R + ext - 1, i.e. the band and ext rows either
// side (own, other = the LDS arrays of colour hh & 1 and of the other colour; LDS row r = global
// row (first + r) mod rows), and ends with a workgroup barrier. Returns after writing the band
// back, or early if a wait gave up.
// The two pointers are offsets from the LDS base, not entries of an array of pointers: an
// array indexed by the colour hides their address space from the compiler, which then reaches
// the band through flat loads and stores (slower, and waiting on both memory counters).
template <class UPDATE>
__device__ __forceinline__ void rs_run(const ResidentBands& a, int8_t* lds, int8_t* g0,
                                       int8_t* g1, u32* go, UPDATE update)
{
    const u32 per_xcd = a.B / 8u;
    const u32 x = a.one_l2 ? (blockIdx.x & 7u) : (xcc_id() & 7u);
    const u32 slot = blockIdx.x >> 3;
    const u32 band = x * per_xcd + slot;
    const u32 up = band == 0u ? a.B - 1u : band - 1u;
    const u32 dn = band + 1u == a.B ? 0u : band + 1u;
    const u32 R = a.R, T = a.T, rb = a.row_bytes;
    const u32 nl = R + 2u * T;  // LDS rows per colour
    const size_t csz = (size_t)nl * rb;  // LDS bytes per colour
    const u32 first = (band * R + a.rows - T) % a.rows;  // global row of LDS row 0
    rs_load_global(a, lds, g0, first, nl);
    rs_load_global(a, lds + csz, g1, first, nl);
    __syncthreads();
    // a band at an XCD boundary talks to another L2 (row form)
    const bool edge = !a.one_l2 && (slot == 0u || slot + 1u == per_xcd);
    // waits for both neighbours' round count (wave 0, then the whole workgroup)
    auto neighbours_at = [&](u32 target) {
        if (threadIdx.x < 64u) {
            const bool ok = rs_wait(a, up, dn, target);
            if (edge) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (threadIdx.x == 0) *go = ok ? 1u : 0u;
        }
        __syncthreads();
        return *go != 0u;
    };
    const u32 rounds = (a.n_halves + T - 1u) / T;
    for (u32 r = 0; r < rounds; ++r) {
        const u32 h0 = r * T;
        const u32 m = a.n_halves - h0 < T ? a.n_halves - h0 : T;
        if (r > 0u) {
            // halos: the neighbours' edge rows after round r - 1
            const u32 par = (r - 1u) & 1u;
            if (a.tagged) {
                const bool ok = rs_receive(a, lds, nl, up, dn, par, a.epoch + r);
                if (__syncthreads_or(!ok)) return;  // a wait gave up (the status reports it)
            } else {
                if (!neighbours_at(a.epoch + r)) return;
#pragma unroll
                for (u32 c = 0; c < 2; ++c) {
                    rs_copy_rows<true>(lds + c * csz, rs_xrows(a, par, up, 1u, c), T, rb);
                    rs_copy_rows<true>(lds + c * csz + (size_t)(T + R) * rb, rs_xrows(a, par, dn, 0u, c), T, rb);
                }
                __syncthreads();
            }
        }
        for (u32 i = 0; i < m; ++i) {
            const u32 hh = h0 + i;
            const size_t own_off = (size_t)(hh & 1u) * csz;
            update(lds + own_off, lds + (csz - own_off), hh, m - 1u - i, first);
        }
        // publish this band's first and last T rows of both colours for the next round's halos
        // (the update's final barrier ordered them): none after the last round, where the row
        // form still publishes its count
        if (a.tagged) {
            if (r + 1u < rounds) rs_publish(a, lds, nl, band, r & 1u, a.epoch + r + 1u);
        } else {
            const u32 par = r & 1u;
            if (r + 1u < rounds) {
#pragma unroll
                for (u32 c = 0; c < 2; ++c) {
                    rs_copy_rows<false>(rs_xrows(a, par, band, 0u, c), lds + c * csz + (size_t)T * rb, T, rb);
                    rs_copy_rows<false>(rs_xrows(a, par, band, 1u, c), lds + c * csz + (size_t)R * rb, T, rb);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are acknowledged
            __syncthreads();
            if (threadIdx.x == 0) {
                u32* fl = rs_flag(a, band);
                if (edge) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // L2 write-back
                    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                    __hip_atomic_store(fl, a.epoch + r + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    __hip_atomic_store(fl, a.epoch + r + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }
    // the band's own rows back to the lattice, once both neighbours have finished the launch:
    // their initial halo loads, which read this band's rows, are done
    if (a.tagged && threadIdx.x == 0)
        __hip_atomic_store(rs_flag(a, band), a.epoch + rounds, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!neighbours_at(a.epoch + rounds)) return;
    for (u32 c = 0; c < 2; ++c)
        rs_copy_rows<false>((c ? g1 : g0) + (size_t)band * R * rb, lds + c * csz + (size_t)T * rb, R, rb);
}

// ------------------------------------------------------------------------------ host ----
// Plan for a 2D lattice of `rows` rows of row_bytes bytes: B bands (a power of two from 8
// dividing rows, at most the CU count; `bands` > 0 forces one) and T = min(R, 8) half sweeps per
// round; B = 0 when the form does not apply (row_bytes % 16, LDS beyond RS_MAX_LDS, rows < 16,
// placement). Default B: the largest that
// keeps R >= 4 rows per band (else R >= 2): with R = 2 a round holds two half sweeps and the
// exchange dominates (measured with int8 at L = 256 .. 2048 and a counter-and-copy exchange of
// about 2.3 us per round: R = 4 beat R = 2 by 1.3-2x; more bands beat fewer as long as R >= 4).
struct ResidentPlan {
    u32 B = 0, R = 0, T = 0, tagged = 0;
    size_t lds = 0, ws_bytes = 0;
    int place = -1;
};

static inline int resident_plan(long long rows, long long row_bytes, int bands, ResidentPlan& pl)
{
    pl = ResidentPlan{};
    if (rows < 16 || row_bytes < 16 || row_bytes % 16 || rows > (1 << 20) || row_bytes > (1 << 20))
        return ISING_OK;
    int dev = 0, n_cu = 0;
    ISING_HIP(hipGetDevice(&dev));
    ISING_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    if (int rc = fused_placement(&pl.place)) return rc;
    if (pl.place < 0) return ISING_OK;
    const u32 min_rows = bands ? 2u : (rows >= 32 ? 4u : 2u);
    for (u32 B = 256; B >= 8; B /= 2) {
        if (B > (u32)n_cu || rows % B || rows / B < min_rows || (bands && B != (u32)bands)) continue;
        const u32 R = (u32)(rows / B);
        const u32 T = R < 8u ? R : 8u;
        const size_t lds = 2ull * (R + 2u * T) * (size_t)row_bytes;
        if (lds > RS_MAX_LDS) continue;
        pl.B = B;
        pl.R = R;
        pl.T = T;
        pl.lds = lds;
        // tagged words while a part (T rows of one colour and side) is at most one word per
        // thread (same box, us per sweep, row form -> tagged: int8 1024^2 (T x row_bytes =
        // 2048) 3.31 -> 3.08, msc 1024^2 5.36 -> 5.03, msc 4096^2 (2048) 7.51 -> 7.30; int8
        // 2048^2 (8192) 5.94 -> 6.35)
        pl.tagged = (unsigned long long)T * (unsigned long long)row_bytes / 4u <= RS_THREADS ? 1u : 0u;
        // header, then 2 parities x B bands x 2 sides x 2 colours x T rows of row_bytes bytes,
        // twice that as tagged words
        pl.ws_bytes = 4ull * (RS_HDR + B) + (pl.tagged ? 16ull : 8ull) * B * (size_t)T * row_bytes;
        break;
    }
    return ISING_OK;
}

// Cooperative launches of a resident kernel for n_sweeps sweeps (at most 2^19 per launch), the
// workspace's epoch advanced by the rounds each publishes. ARGS: the kernel's argument struct,
// whose first member `bands` is the ResidentBands.
template <class ARGS>
static int resident_launch(const void* fn, ARGS a, const ResidentPlan& pl, u32 t0,
                           long long n_sweeps, u32* epoch, hipStream_t st)
{
    for (long long done = 0; done < n_sweeps;) {
        const long long k = n_sweeps - done < (1ll << 19) ? n_sweeps - done : (1ll << 19);
        a.bands.n_halves = (u32)(2 * k);
        a.t0 = t0 + (u32)done;
        a.bands.epoch = *epoch;
        void* args[] = {&a};
        ISING_HIP(hipLaunchCooperativeKernel(fn, dim3(pl.B), dim3(RS_THREADS), args, (unsigned)pl.lds, st));
        *epoch += (a.bands.n_halves + pl.T - 1u) / pl.T;  // rounds published
        done += k;
    }
    return ISING_OK;
}

// Failure bits of the resident launches since the workspace was zeroed (synchronizes the stream).
static inline int resident_status(const uint32_t* ws, int* failed, hipStream_t st)
{
    u32 v = 0;
    ISING_HIP(hipMemcpyAsync(&v, ws, sizeof(v), hipMemcpyDeviceToHost, st));
    ISING_HIP(hipStreamSynchronize(st));
    *failed = (int)v;
    return ISING_OK;
}
