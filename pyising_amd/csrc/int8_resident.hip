// Resident int8 sweeps of small 2D lattices: one launch holds every half sweep of a call, each
// workgroup keeping a band of rows of both colours in LDS for the whole launch (the band,
// round and exchange machinery is resident.hpp's).
//
// Why: a 2D lattice of L <= 2048 is too small to fill the GPU for long; one launch per half
// sweep then costs more in launch gaps and ramp than in work (L = 1024: 3.5 us per half sweep,
// of which the sweep itself is a fraction). Here the lattice never leaves the CUs between half
// sweeps. Updates are int8_decide of the global kernels (int8_core.hpp) with the global row
// index in the Philox counter, so every site draws the numbers it draws in the
// one-launch-per-half-sweep kernels: bit-identical lattices.
#include "resident.hpp"

// Sites per thread and step: 8 (measured at L = 512 .. 2048 against 16 with 512 threads and 4
// with 1024: more waves per CU hide the Philox and LDS latency of the short half sweeps; 4 was
// measured again with the tagged exchange: 3.27 against 3.18 us per sweep at 1024^2, 7.56
// against 6.17 at 2048^2).
constexpr u32 RS_V = 8;

struct Int8ResidentArgs {
    ResidentBands bands;  // row_bytes = lh
    int8_t* c0;
    int8_t* c1;
    FastDiv div_row;  // by lh / RS_V, the chunks of a row
    u32 t0, k0, k1;
    Int8Table tb;
};

// Whole-row waves (lh = 64 RS_V times a power of two, e.g. L = 1024, 2048): a wave holds 64
// consecutive chunks of one row, so the row, its parity and its LDS offsets are wave-uniform
// and the horizontal edge byte comes from the lane holding the neighbouring chunk (DPP). With
// one wave per row (ONE) the row wraps inside the wave (a rotate: the torus); else the wave's
// outer lane takes the edge word from one uniform LDS load. own, other: the LDS colour arrays;
// rows row_lo .. row_lo + n_rows - 1 of the band.
template <int HB, bool ONE>
__device__ __forceinline__ void int8_rs_rows(const Int8ResidentArgs& a, int8_t* own,
                                             const int8_t* other, const u32* thr_lds, u32 row_lo,
                                             u32 n_rows, u32 first, u32 color, u32 t)
{
    const u32 lh = a.bands.row_bytes, rows = a.bands.rows;
    const u32 per_row = lh / RS_V;
    const u32 lg = ONE ? 0u : (u32)__builtin_ctz(per_row / 64u);
    const u32 wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    for (u32 rw = wave; rw < (n_rows << lg); rw += RS_THREADS / 64u) {
        const u32 rr = ONE ? rw : uni(rw >> lg);
        const u32 c0 = ONE ? 0u : uni((rw - (rr << lg)) * 64u);  // the wave's first chunk
        const u32 row = row_lo + rr;
        u32 y = first + row;
        if (y >= rows) y -= rows;
        const u32 p = uni((y + color) & 1u);
        const u32 j0 = (c0 + lane) * RS_V;
        const size_t at = (size_t)row * lh + j0;
        const Vec<RS_V> S = load_vec<RS_V>(own + at);
        const Vec<RS_V> C = load_vec<RS_V>(other + at);
        const Vec<RS_V> U = load_vec<RS_V>(other + at + lh);  // row y + 1
        const Vec<RS_V> D = load_vec<RS_V>(other + at - lh);  // row y - 1
        // p = 1: first byte of the next chunk (lane i reads lane i + 1); p = 0: last byte of the
        // previous chunk (lane i reads lane i - 1)
        u32 nb;
        if constexpr (ONE) {
            nb = p ? (u32)__builtin_amdgcn_update_dpp(0, (int)C.w[0], 0x134, 0xF, 0xF, false)  // wave_rol:1
                   : (u32)__builtin_amdgcn_update_dpp(0, (int)C.w[RS_V / 4 - 1], 0x13C, 0xF, 0xF, false);  // wave_ror:1
        } else {
            const u32 eo = p ? ((c0 + 64u == per_row) ? 0u : (c0 + 64u) * RS_V)
                             : (c0 == 0u ? lh - 4u : c0 * RS_V - 4u);
            const u32 ew = *reinterpret_cast<const u32*>(other + (size_t)row * lh + eo);
            nb = p ? (u32)__builtin_amdgcn_update_dpp((int)ew, (int)C.w[0], 0x130, 0xF, 0xF, false)  // wave_shl:1
                   : (u32)__builtin_amdgcn_update_dpp((int)ew, (int)C.w[RS_V / 4 - 1], 0x138, 0xF, 0xF, false);  // wave_shr:1
        }
        const u32 e = p ? (nb & 0xFFu) : (nb >> 24);
        const Vec<RS_V> H = int8_horizontal<RS_V>(C, e, p);
        u32 minus = 0;
        const Vec<RS_V> O = int8_decide<2, HB, RS_V, false>(S, C, H, U, D, U, D, (u64)y * lh + j0,
                                                            thr_lds, t, color, a.k0, a.k1, minus);
        store_vec<RS_V>(own + at, O);
    }
}

template <int HB>
__global__ __launch_bounds__(RS_THREADS) void int8_resident_kernel(Int8ResidentArgs a)
{
    extern __shared__ __attribute__((aligned(16))) int8_t rs_lds[];
    __shared__ u32 thr_lds[8];
    __shared__ u32 go;
    int8_fill_table<2, HB>(thr_lds, a.tb);  // ordered before its first use by rs_run's barrier
    const u32 lh = a.bands.row_bytes, R = a.bands.R, T = a.bands.T, rows = a.bands.rows;
    const u32 per_row = lh / RS_V;
    rs_run(a.bands, rs_lds, a.c0, a.c1, &go,
           [&](int8_t* own, const int8_t* other, u32 hh, u32 ext, u32 first) {
        const u32 color = hh & 1u;
        const u32 t = a.t0 + (hh >> 1);
        const u32 row_lo = T - ext, n_rows = R + 2u * ext;
        const u32 wpr = per_row / 64u;  // waves per row, when whole
        if (per_row % 64u == 0u && (wpr & (wpr - 1u)) == 0u) {
            if (wpr == 1u) int8_rs_rows<HB, true>(a, own, other, thr_lds, row_lo, n_rows, first, color, t);
            else int8_rs_rows<HB, false>(a, own, other, thr_lds, row_lo, n_rows, first, color, t);
            __syncthreads();
            return;
        }
        for (u32 it = threadIdx.x; it < n_rows * per_row; it += RS_THREADS) {
            const u32 rr = fast_div(it, a.div_row);
            const u32 row = row_lo + rr;  // LDS row, 1 <= row <= R + 2T - 2
            const u32 j0 = (it - rr * per_row) * RS_V;
            u32 y = first + row;
            if (y >= rows) y -= rows;
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
    });
}

// ------------------------------------------------------------------------------ host ----
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
// *epoch: the rounds its counters hold (0 after zeroing), advanced by this call.
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
    Int8ResidentArgs a;
    a.bands.rows = (u32)rows;
    a.bands.row_bytes = (u32)lh;
    a.bands.R = pl.R;
    a.bands.T = pl.T;
    a.bands.B = pl.B;
    a.bands.one_l2 = pl.place == 0 ? 1u : 0u;
    a.bands.tagged = pl.tagged;
    a.bands.ws = ws;
    a.c0 = d_c0;
    a.c1 = d_c1;
    a.div_row = make_fastdiv((u32)lh / RS_V);
    a.k0 = k0;
    a.k1 = k1;
    for (int k = 0; k < 7; ++k) a.tb.v[k] = k <= 4 ? thr[k] : 0u;
    const void* fn = heatbath ? (const void*)int8_resident_kernel<1> : (const void*)int8_resident_kernel<0>;
    return resident_launch(fn, a, pl, t0, n_sweeps, epoch, as_stream(stream));
}

// Failure bits of the resident launches since the workspace was zeroed (synchronizes the
// stream): 1 = a wait gave up (the lattice is then invalid).
int ising_int8_resident_status(const uint32_t* ws, int* failed, void* stream)
{
    if (!ws || !failed) return ising_fail(ISING_EINVAL, "int8_resident_status: null pointer");
    return resident_status(ws, failed, as_stream(stream));
}

}  // extern "C"
