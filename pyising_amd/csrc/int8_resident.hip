// Resident int8 sweeps of small 2D lattices: one launch holds every half sweep of a call, each
// workgroup keeping a band of rows of both colours in LDS for the whole launch.
//
// Why: a 2D lattice of L <= 2048 is too small to fill the GPU for long; one launch per half
// sweep then costs more in launch gaps and ramp than in work (L = 1024: 3.5 us per half sweep,
// of which the sweep itself is a fraction). Here the lattice never leaves the CUs between half
// sweeps.
//
// Decomposition. B workgroups (B | rows, B a multiple of 8, at most one per CU, launched
// cooperatively so that all are resident) own bands of R = rows / B consecutive rows. A
// workgroup holds its band plus T halo rows on each side (T <= R), both colours, in LDS, and
// runs T half sweeps per round without talking to anybody: half sweep i of a round updates the
// rows within T - 1 - i of its band (redundantly with its neighbours: the same sites, the same
// Philox counters, so the same bits), which needs the other colour within T - i rows, which the
// previous half sweep (or the round's start) left valid. After a round the band's T edge rows
// of both colours are current; they are exchanged with the neighbour bands through a
// double-buffered workspace (flags per band, bounded waits), and the next round starts.
// Updates are int8_decide of the global kernels (int8_core.hpp) with the global row index in
// the Philox counter, so every site draws the numbers it draws in the one-launch-per-half-sweep
// kernels: bit-identical lattices.
//
// Placement. Bands are numbered so that the B / 8 bands of one XCD are consecutive (band =
// XCD * B / 8 + workgroup / 8, from HW_REG_XCC_ID): neighbours share an L2, where plain stores
// acknowledged before the flag and L1-missing loads are coherent; the two bands at each XCD
// boundary add the device-scope release (L2 write-back) before publishing and acquire (L2
// invalidate) after waiting.
#include "int8_core.hpp"
#include "placement.hpp"

// Workspace (u32 words): [0] failure bits (1: a wait gave up), [RS_HDR + b] rounds published
// by band b (from the workspace's zero state, mod 2^32), then the exchange rows:
// [parity][band][side: 0 = first T rows, 1 = last T rows][colour][T rows][lh bytes].
constexpr u32 RS_HDR = 64;
constexpr u32 RS_SPIN_LIMIT = 1u << 20;
constexpr u32 RS_THREADS = 1024;
// Sites per thread and step: 8 (measured at L = 512 .. 2048 against 16 with 512 threads and 4
// with 1024: more waves per CU hide the Philox and LDS latency of the short half sweeps).
constexpr u32 RS_V = 8;
constexpr u32 RS_MAX_LDS = 64u * 1024u;  // both colours, band and halos

struct ResidentArgs {
    int8_t* c0;
    int8_t* c1;
    u32 rows, lh, R, T, B;
    FastDiv div_row;  // by lh / RS_V, the chunks of a row
    u32 n_halves;   // half sweeps of this launch, starting with colour 0
    u32 t0, k0, k1;
    u32 epoch;      // rounds every band flag holds when the launch starts
    u32 one_l2;     // 1: every workgroup on one XCD
    u32* ws;
    Int8Table tb;
};

__device__ __forceinline__ u32* rs_flag(const ResidentArgs& a, u32 band) { return a.ws + RS_HDR + band; }

__device__ __forceinline__ int8_t* rs_xrows(const ResidentArgs& a, u32 par, u32 band, u32 side,
                                            u32 color)
{
    int8_t* base = reinterpret_cast<int8_t*>(a.ws + RS_HDR + a.B);
    const size_t rows_t = (size_t)a.T * a.lh;
    return base + ((((size_t)par * a.B + band) * 2u + side) * 2u + color) * rows_t;
}

// Copy n_rows rows of lh bytes, 16 bytes per thread and step (lh % 16 == 0). COH_SRC: the source
// may have been written by another workgroup of this launch (L1-missing loads).
template <bool COH_SRC>
__device__ __forceinline__ void rs_copy_rows(int8_t* dst, const int8_t* src, u32 n_rows, u32 lh)
{
    const u32 n16 = n_rows * lh / 16u;
    for (u32 i = threadIdx.x; i < n16; i += RS_THREADS) {
        const Vec<16> v = load_vec_at<16, COH_SRC>(src, (u64)i * 16u);
        store_vec<16>(dst + (size_t)i * 16u, v);
    }
}

// Rows of the global colour array: LDS row r of the band holds global row (first + r) mod rows.
__device__ __forceinline__ void rs_load_global(const ResidentArgs& a, int8_t* lds_c,
                                               const int8_t* g, u32 first, u32 n_rows,
                                               u32 lds_row0)
{
    const u32 per_row = a.lh / 16u;
    for (u32 i = threadIdx.x; i < n_rows * per_row; i += RS_THREADS) {
        const u32 r = i / per_row, c = i - r * per_row;
        u32 y = first + r;
        if (y >= a.rows) y -= a.rows;
        const Vec<16> v = load_vec<16>(g + (size_t)y * a.lh + c * 16u);
        store_vec<16>(lds_c + (size_t)(lds_row0 + r) * a.lh + c * 16u, v);
    }
}

// Wave 0 waits until both neighbour bands have published `target` rounds.
__device__ bool rs_wait(const ResidentArgs& a, u32 up, u32 dn, u32 target)
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

template <int HB>
__global__ __launch_bounds__(RS_THREADS) void int8_resident_kernel(ResidentArgs a)
{
    extern __shared__ __attribute__((aligned(16))) int8_t rs_lds[];
    __shared__ u32 thr_lds[8];
    __shared__ u32 go;
    const u32 per_xcd = a.B / 8u;
    const u32 x = a.one_l2 ? (blockIdx.x & 7u) : (xcc_id() & 7u);
    const u32 band = x * per_xcd + (blockIdx.x >> 3);
    const u32 slot = blockIdx.x >> 3;
    const u32 up = band == 0u ? a.B - 1u : band - 1u;
    const u32 dn = band + 1u == a.B ? 0u : band + 1u;
    // a band at an XCD boundary talks to another L2
    const bool edge = !a.one_l2 && (slot == 0u || slot + 1u == per_xcd);
    const u32 R = a.R, T = a.T, lh = a.lh;
    const u32 nl = R + 2u * T;  // LDS rows per colour
    int8_t* lc[2] = {rs_lds, rs_lds + (size_t)nl * lh};
    const u32 first = band * R + a.rows - T;  // global row of LDS row 0, plus rows
    int8_fill_table<2, HB>(thr_lds, a.tb);
    rs_load_global(a, lc[0], a.c0, first % a.rows, nl, 0u);
    rs_load_global(a, lc[1], a.c1, first % a.rows, nl, 0u);
    __syncthreads();

    const u32 per_row = lh / RS_V;
    const u32 rounds = (a.n_halves + T - 1u) / T;
    // waits for both neighbours' round count (wave 0, then the whole workgroup); false if a
    // wait gave up
    auto neighbours_at = [&](u32 target) {
        if (threadIdx.x < 64u) {
            const bool ok = rs_wait(a, up, dn, target);
            if (edge) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (threadIdx.x == 0) go = ok ? 1u : 0u;
        }
        __syncthreads();
        return go != 0u;
    };
    for (u32 r = 0; r < rounds; ++r) {
        const u32 h0 = r * T;
        const u32 m = a.n_halves - h0 < T ? a.n_halves - h0 : T;
        if (r > 0u) {
            // halos: the neighbours' edge rows after round r - 1
            if (!neighbours_at(a.epoch + r)) return;  // a wait gave up (the status reports it)
            const u32 par = (r - 1u) & 1u;
#pragma unroll
            for (u32 c = 0; c < 2; ++c) {
                rs_copy_rows<true>(lc[c], rs_xrows(a, par, up, 1u, c), T, lh);
                rs_copy_rows<true>(lc[c] + (size_t)(T + R) * lh, rs_xrows(a, par, dn, 0u, c), T, lh);
            }
            __syncthreads();
        }
        for (u32 i = 0; i < m; ++i) {
            const u32 hh = h0 + i;
            const u32 color = hh & 1u;
            const u32 t = a.t0 + (hh >> 1);
            const u32 ext = m - 1u - i;  // rows beyond the band updated in this half sweep
            const u32 row_lo = T - ext, n_rows = R + 2u * ext;
            int8_t* own = lc[color];
            const int8_t* other = lc[color ^ 1u];
            for (u32 it = threadIdx.x; it < n_rows * per_row; it += RS_THREADS) {
                const u32 rr = fast_div(it, a.div_row);
                const u32 row = row_lo + rr;  // LDS row, 1 <= row <= nl - 2
                const u32 j0 = (it - rr * per_row) * RS_V;
                u32 y = first + row;
                y = y >= a.rows ? (y >= 2u * a.rows ? y - 2u * a.rows : y - a.rows) : y;
                const u32 p = (y + color) & 1u;
                const size_t at = (size_t)row * lh + j0;
                const Vec<RS_V> S = load_vec<RS_V>(own + at);
                const Vec<RS_V> C = load_vec<RS_V>(other + at);
                const Vec<RS_V> U = load_vec<RS_V>(other + at + lh);  // row y + 1
                const Vec<RS_V> D = load_vec<RS_V>(other + at - lh);  // row y - 1
                const u32 je = p ? ((j0 + RS_V == lh) ? 0u : j0 + RS_V) : ((j0 == 0u) ? lh - 1u : j0 - 1u);
                const u32 e = (u32)(uint8_t)other[(size_t)row * lh + je];
                const Vec<RS_V> H = int8_horizontal<RS_V>(C, e, p);
                u32 minus = 0;
                const Vec<RS_V> O = int8_decide<2, HB, RS_V, false>(S, C, H, U, D, U, D,
                                                                    (u64)y * lh + j0, thr_lds, t,
                                                                    color, a.k0, a.k1, minus);
                store_vec<RS_V>(own + at, O);
            }
            __syncthreads();
        }
        // publish the round: this band's first and last T rows of both colours (for the next
        // round's halos), or, after the last round, only the count (the neighbours then know
        // that this band's initial loads are long done before they write their rows back)
        {
            const u32 par = r & 1u;
            if (r + 1u < rounds) {
#pragma unroll
                for (u32 c = 0; c < 2; ++c) {
                    rs_copy_rows<false>(rs_xrows(a, par, band, 0u, c), lc[c] + (size_t)T * lh, T, lh);
                    rs_copy_rows<false>(rs_xrows(a, par, band, 1u, c), lc[c] + (size_t)R * lh, T, lh);
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
    // the band's own rows back to the lattice, once both neighbours have loaded their halos
    // from them (they published all rounds, the first after their initial loads)
    if (!neighbours_at(a.epoch + rounds)) return;
    const u32 own_first = band * R;
    for (u32 c = 0; c < 2; ++c) {
        int8_t* g = c ? a.c1 : a.c0;
        rs_copy_rows<false>(g + (size_t)own_first * lh, lc[c] + (size_t)T * lh, R, lh);
    }
}

// ------------------------------------------------------------------------------ host ----
// Plan for a 2D lattice (rows x 2 lh): B bands (a power of two from 8 dividing rows, at most the
// CU count) and T = min(R, 8) half sweeps per round; 0 bands when the resident form does not
// apply (lh % 16, LDS beyond RS_MAX_LDS, rows < 16, placement). Default B: the largest that
// keeps R >= 4 rows per band (else R >= 2): with R = 2 a round holds two half sweeps, and the
// exchange, about 2.3 us per round, dominates (measured at L = 256 .. 2048: R = 4 beats R = 2
// by 1.3-2x; more bands beat fewer as long as R >= 4).
struct ResidentPlan {
    u32 B = 0, R = 0, T = 0;
    size_t lds = 0, ws_bytes = 0;
};

static int resident_plan(long long rows, int lh, int bands, ResidentPlan& pl)
{
    pl = ResidentPlan{};
    if (rows < 16 || lh < 16 || lh % 16 || rows > (1 << 20)) return ISING_OK;
    int dev = 0, n_cu = 0;
    ISING_HIP(hipGetDevice(&dev));
    ISING_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    int place = -1;
    if (int rc = fused_placement(&place)) return rc;
    if (place < 0) return ISING_OK;
    const u32 min_rows = bands ? 2u : (rows >= 32 ? 4u : 2u);
    for (u32 B = 256; B >= 8; B /= 2) {
        if (B > (u32)n_cu || rows % B || rows / B < min_rows || (bands && B != (u32)bands)) continue;
        const u32 R = (u32)(rows / B);
        const u32 T = R < 8u ? R : 8u;
        const size_t lds = 2ull * (R + 2u * T) * (size_t)lh;
        if (lds > RS_MAX_LDS) continue;
        pl.B = B;
        pl.R = R;
        pl.T = T;
        pl.lds = lds;
        pl.ws_bytes = 4ull * (RS_HDR + B) + 2ull * B * 2 * 2 * (size_t)T * lh;
        break;
    }
    return ISING_OK;
}

extern "C" {

// Workspace bytes of the resident sweeps of a 2D lattice (rows x 2 lh) with `bands` bands (0:
// the default choice), 0 when they do not apply here.
int ising_int8_resident_bytes(long long rows, int lh, int bands, size_t* bytes)
{
    if (!bytes || rows < 1 || lh < 1 || bands < 0) return ising_fail(ISING_EINVAL, "int8_resident_bytes: bad argument");
    ResidentPlan pl;
    if (int rc = resident_plan(rows, lh, bands, pl)) return rc;
    *bytes = pl.ws_bytes;
    return ISING_OK;
}

// n_sweeps sweeps of a 2D int8 lattice in one cooperative launch (colour 0 then 1 per sweep,
// Philox counters t0, t0 + 1, ...): bit-identical to 2 n_sweeps calls of ising_int8_half.
// ws: a workspace of ising_int8_resident_bytes zeroed before first use and kept between calls;
// *epoch: the rounds its flags hold (0 after zeroing), advanced by this call.
int ising_int8_sweeps_resident(int8_t* d_c0, int8_t* d_c1, long long rows, int lh, int heatbath,
                               const uint32_t* thr, uint32_t t0, long long n_sweeps, uint32_t k0,
                               uint32_t k1, int bands, uint32_t* ws, uint32_t* epoch, void* stream)
{
    if (!d_c0 || !d_c1 || !thr || !ws || !epoch || n_sweeps < 0 || n_sweeps > (1ll << 30) || bands < 0)
        return ising_fail(ISING_EINVAL, "int8_sweeps_resident: bad argument");
    if (n_sweeps == 0) return ISING_OK;
    ResidentPlan pl;
    if (int rc = resident_plan(rows, lh, bands, pl)) return rc;
    if (!pl.B) return ising_fail(ISING_EINVAL, "int8_sweeps_resident: not applicable to this lattice here");
    int place = -1;
    if (int rc = fused_placement(&place)) return rc;
    ResidentArgs a;
    a.c0 = d_c0;
    a.c1 = d_c1;
    a.rows = (u32)rows;
    a.lh = (u32)lh;
    a.R = pl.R;
    a.T = pl.T;
    a.B = pl.B;
    a.div_row = make_fastdiv((u32)lh / RS_V);
    a.t0 = t0;
    a.k0 = k0;
    a.k1 = k1;
    a.one_l2 = place == 0 ? 1u : 0u;
    a.ws = ws;
    for (int k = 0; k < 7; ++k) a.tb.v[k] = k <= 4 ? thr[k] : 0u;
    hipStream_t st = as_stream(stream);
    // launches of at most 2^20 half sweeps (flags advance by one per round)
    for (long long done = 0; done < n_sweeps;) {
        const long long k = n_sweeps - done < (1ll << 19) ? n_sweeps - done : (1ll << 19);
        a.n_halves = (u32)(2 * k);
        a.t0 = t0 + (u32)done;
        a.epoch = *epoch;
        void* args[] = {&a};
        const void* fn = heatbath ? (const void*)int8_resident_kernel<1> : (const void*)int8_resident_kernel<0>;
        ISING_HIP(hipLaunchCooperativeKernel(fn, dim3(pl.B), dim3(RS_THREADS), args, (unsigned)pl.lds, st));
        *epoch += (a.n_halves + pl.T - 1u) / pl.T;  // rounds published
        done += k;
    }
    return ISING_OK;
}

// Failure bits of the resident launches since the workspace was zeroed (synchronizes the
// stream): 1 = a wait gave up (the lattice is then invalid).
int ising_int8_resident_status(const uint32_t* ws, int* failed, void* stream)
{
    if (!ws || !failed) return ising_fail(ISING_EINVAL, "int8_resident_status: null pointer");
    u32 v = 0;
    ISING_HIP(hipMemcpyAsync(&v, ws, sizeof(v), hipMemcpyDeviceToHost, as_stream(stream)));
    ISING_HIP(hipStreamSynchronize(as_stream(stream)));
    *failed = (int)v;
    return ISING_OK;
}

}  // extern "C"
