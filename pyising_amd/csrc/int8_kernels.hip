// int8 compressed-checkerboard half sweeps (2D/3D, Metropolis and heat bath) and the exact
// integer observables of that layout, for gfx950.
//
// Layout (restates src/pyising/lattice.py:1-14 of the reference): colour c of the lattice is
// a row-major (rows, lh) int8 array, lh = Lx / 2, rows = Ly (2D) or Lz * Ly (3D); in row
// r = z * Ly + y the colour-c site j sits at x = 2 j + ((y + z + c) & 1). Spins are the bytes
// 0x01 (+1) and 0xFF (-1).
//
// Work split: one thread owns V consecutive colour sites of one row (V = 16, 8 or 4, the
// largest that divides lh) and loads them as whole 16/8/4-byte vectors: own, the other colour
// at the same j, the other colour at y +- 1 (and z +- 1 in 3D), plus one edge byte for the
// horizontal neighbour. Neighbour counting is SWAR over 4 sites per 32-bit word: bit 1 of a
// spin byte is 1 exactly for -1, so ((a ^ b) >> 1) & 0x01010101 marks anti-aligned pairs and
// (b >> 1) & 0x01010101 marks -1 neighbours, 4 sites per instruction, no carries between
// bytes (counts <= 6).
//
// Random numbers (src/pyising/philox.py:54-64): colour-site n takes word n % 4 of
// Philox4x32-10((u32)(n / 4), t, colour, 0; k0, k1), so each 4-site word of a thread's vector
// costs exactly one Philox call.
//
// Decision rules (src/pyising/backends/numpy_backend.py:35-44, src/pyising/lattice.py:128-142):
//   Metropolis, a = anti-aligned neighbours: flip iff a >= d or u < table[2d - a]
//   heat bath, m = -1 neighbours:            new spin +1 iff u < table[2d - m]
// The table is passed by value and copied into an 8-entry LDS table per workgroup, indexed
// directly by the SWAR count byte (see int8_update_chunk).
#include "int8_core.hpp"

// One thread: sites j0 .. j0 + V - 1 of one row. LAY: I8_GENERAL (each lane's own row
// arithmetic), I8_ROWW or I8_SEGW (a wave's row arithmetic shared, int8_update_wave).
template <int NDIM, int HB, int V, int LAY>
__global__ __launch_bounds__(256) void int8_half_kernel(int8_t* __restrict__ own,
                                                        const int8_t* __restrict__ other,
                                                        Int8Geom g, Int8Table tb, u32 t, u32 color,
                                                        u32 k0, u32 k1)
{
    __shared__ u32 thr_lds[8];
    int8_fill_table<NDIM, HB>(thr_lds, tb);
    __syncthreads();
    const u32 blk = xcd_band_block(blockIdx.x, gridDim.x);
    if constexpr (LAY == I8_GENERAL) {
        const u32 ci = blk * blockDim.x + threadIdx.x;
        if (ci >= g.n_chunks) return;
        int8_update_chunk<NDIM, HB, V, false, false, true>(own, other, g, thr_lds, t, color, k0,
                                                           k1, ci);
    } else {
        const u32 w0 = blk * blockDim.x + uni(threadIdx.x & ~63u);  // n_chunks % 64 == 0
        if (w0 >= g.n_chunks) return;                                // the whole wave
        int8_update_wave<NDIM, HB, V, false, false, true, LAY>(own, other, g, thr_lds, t, color,
                                                               k0, k1, w0);
    }
}


// ------------------------------------------------------------------------ observables ----
// M = sum of all spins = N - 2 * (#-1); E = -sum over colour-0 sites of s * h =
// 2 * (#anti-aligned bonds) - d * N (every bond has exactly one colour-0 end,
// src/pyising/backends/observables_torch.py:1-4). Accumulated as exact integers: per-lane
// counts, wave reduction with cross-lane adds, LDS across the workgroup's waves, one partial
// pair per workgroup, summed by obs_finalize.

template <int NDIM, int V>
__global__ __launch_bounds__(256) void int8_obs_kernel(const int8_t* __restrict__ c0,
                                                       const int8_t* __restrict__ c1, Int8Geom g,
                                                       unsigned long long* __restrict__ partials)
{
    constexpr int NWd = V / 4;
    u64 minus = 0, anti = 0;
    const u32 stride = gridDim.x * blockDim.x;
    for (u32 ci = blockIdx.x * blockDim.x + threadIdx.x; ci < g.n_chunks; ci += stride) {
        const u32 row = fast_div(ci, g.div_chunks);
        const u32 j0 = (ci - row * g.chunks_row) * V;
        const u32 z = fast_div(row, g.div_ly);
        const u32 y = row - z * g.ly;
        const u32 p = (y + z) & 1u;  // colour 0
        const u64 rbase = (u64)row * g.lh;
        const u64 n0 = rbase + j0;
        const Vec<V> S = load_vec<V>(c0 + n0);
        const Vec<V> C = load_vec<V>(c1 + n0);
        const u32 yu = (y + 1 == g.ly) ? 0u : y + 1;
        const u32 yd = (y == 0) ? g.ly - 1 : y - 1;
        const Vec<V> U = load_vec<V>(c1 + ((u64)z * g.ly + yu) * g.lh + j0);
        const Vec<V> D = load_vec<V>(c1 + ((u64)z * g.ly + yd) * g.lh + j0);
        Vec<V> ZU, ZD;
        if constexpr (NDIM == 3) {
            const u32 zu = (z + 1 == g.lz) ? 0u : z + 1;
            const u32 zd = (z == 0) ? g.lz - 1 : z - 1;
            ZU = load_vec<V>(c1 + ((u64)zu * g.ly + y) * g.lh + j0);
            ZD = load_vec<V>(c1 + ((u64)zd * g.ly + y) * g.lh + j0);
        }
        const u32 je = p ? ((j0 + V == g.lh) ? 0u : j0 + V) : ((j0 == 0) ? g.lh - 1 : j0 - 1);
        const u32 e = (u32)(uint8_t)c1[rbase + je];
        u32 cm = 0, ca = 0;
#pragma unroll
        for (int i = 0; i < NWd; ++i) {
            u32 h;
            if (p) {
                const u32 nxt = (i + 1 < NWd) ? C.w[i + 1 < NWd ? i + 1 : 0] : e;
                h = (C.w[i] >> 8) | (nxt << 24);
            } else {
                const u32 prv = (i > 0) ? (C.w[i > 0 ? i - 1 : 0] >> 24) : e;
                h = (C.w[i] << 8) | prv;
            }
            const u32 s = S.w[i];
            cm += byte_sum(((s >> 1) & LOW_BITS) + ((C.w[i] >> 1) & LOW_BITS));
            u32 a = (((s ^ C.w[i]) >> 1) & LOW_BITS) + (((s ^ h) >> 1) & LOW_BITS) +
                    (((s ^ U.w[i]) >> 1) & LOW_BITS) + (((s ^ D.w[i]) >> 1) & LOW_BITS);
            if constexpr (NDIM == 3)
                a += (((s ^ ZU.w[i]) >> 1) & LOW_BITS) + (((s ^ ZD.w[i]) >> 1) & LOW_BITS);
            ca += byte_sum(a);
        }
        minus += cm;
        anti += ca;
    }
    block_sum2_u64(minus, anti);
    if (threadIdx.x == 0) {
        partials[2u * blockIdx.x] = minus;
        partials[2u * blockIdx.x + 1u] = anti;
    }
}

// ---------------------------------------------------------------------- host launchers ----
// Sites per lane: the widest of 16, 8, 4 dividing lh that still gives >= 2^17 lanes (512
// workgroups, two per CU), else the narrowest that divides lh. A half sweep of a small lattice
// is one generation of workgroups and takes a lane's latency, which grows with the sites it
// updates; large lattices stream at full bandwidth with the widest vectors. V only tiles the
// sites: the Philox counter of a site is its quad index, so every V gives the same spins.
static int pick_vec(u32 lh, long long rows)
{
    int narrow = 0;
    for (int v = 16; v >= 4; v /= 2) {
        if (lh % (u32)v) continue;
        if ((unsigned long long)rows * (lh / (u32)v) >= (1ull << 17)) return v;
        narrow = v;
    }
    return narrow;
}

static int make_geom(long long rows, int lh, int ly, int lz, int ndim, Int8Geom* g, int* V)
{
    if (rows <= 0 || lh <= 0 || ly <= 0 || lz <= 0 || (ndim != 2 && ndim != 3))
        return ising_fail(ISING_EINVAL, "int8: bad geometry");
    if ((long long)ly * lz != rows) return ising_fail(ISING_EINVAL, "int8: rows != ly * lz");
    if (ndim == 2 && lz != 1) return ising_fail(ISING_EINVAL, "int8: 2D requires lz == 1");
    const int v = pick_vec((u32)lh, rows);
    if (!v) return ising_fail(ISING_EINVAL, "int8: lh must be a multiple of 4 (Lx % 8 == 0)");
    if (ly < 2 || lh < 4 || (ndim == 3 && lz < 2)) return ising_fail(ISING_EINVAL, "int8: extents too small");
    g->lh = (u32)lh;
    g->ly = (u32)ly;
    g->lz = (u32)lz;
    const u64 n_chunks = (u64)rows * (u64)(lh / v);
    if (n_chunks + (1ull << 20) > 0xFFFFFFFFull)  // chunk loops step by at most 2^20
        return ising_fail(ISING_ERANGE, "int8: lattice too large (rows * Lx/2 / V must be < 2^32)");
    g->chunks_row = (u32)(lh / v);
    g->n_chunks = (u32)n_chunks;
    g->div_chunks = make_fastdiv(g->chunks_row);
    g->div_ly = make_fastdiv(g->ly);
    *V = v;
    return ISING_OK;
}

template <int NDIM, int HB, int LAY>
static void launch_half_lay(int V, dim3 grid, hipStream_t st, int8_t* own, const int8_t* other,
                            const Int8Geom& g, const Int8Table& tb, u32 t, u32 color, u32 k0,
                            u32 k1)
{
    if (V == 16)
        hipLaunchKernelGGL((int8_half_kernel<NDIM, HB, 16, LAY>), grid, dim3(256), 0, st, own,
                           other, g, tb, t, color, k0, k1);
    else if (V == 8)
        hipLaunchKernelGGL((int8_half_kernel<NDIM, HB, 8, LAY>), grid, dim3(256), 0, st, own,
                           other, g, tb, t, color, k0, k1);
    else
        hipLaunchKernelGGL((int8_half_kernel<NDIM, HB, 4, LAY>), grid, dim3(256), 0, st, own,
                           other, g, tb, t, color, k0, k1);
}

template <int NDIM, int HB>
static void launch_half(int V, dim3 grid, hipStream_t st, int8_t* own, const int8_t* other,
                        const Int8Geom& g, const Int8Table& tb, u32 t, u32 color, u32 k0, u32 k1)
{
    const int lay = int8_layout(g);
    if (lay == I8_ROWW)
        launch_half_lay<NDIM, HB, I8_ROWW>(V, grid, st, own, other, g, tb, t, color, k0, k1);
    else if (lay == I8_SEGW)
        launch_half_lay<NDIM, HB, I8_SEGW>(V, grid, st, own, other, g, tb, t, color, k0, k1);
    else
        launch_half_lay<NDIM, HB, I8_GENERAL>(V, grid, st, own, other, g, tb, t, color, k0, k1);
}

static int int8_half_impl(int8_t* own, const int8_t* other, long long rows, int lh, int ly, int lz,
                          int ndim, int heatbath, const u32* table, u32 t, u32 color, u32 k0,
                          u32 k1, hipStream_t st)
{
    if (!own || !other || !table) return ising_fail(ISING_EINVAL, "int8_half: null pointer");
    if (color > 1) return ising_fail(ISING_EINVAL, "int8_half: colour must be 0 or 1");
    Int8Geom g;
    int V;
    int rc = make_geom(rows, lh, ly, lz, ndim, &g, &V);
    if (rc) return rc;
    Int8Table tb = {};
    for (int i = 0; i <= 2 * ndim; ++i) tb.v[i] = table[i];
    const dim3 grid((unsigned)((g.n_chunks + 255) / 256));
    if (ndim == 2) {
        if (heatbath) launch_half<2, 1>(V, grid, st, own, other, g, tb, t, color, k0, k1);
        else launch_half<2, 0>(V, grid, st, own, other, g, tb, t, color, k0, k1);
    } else {
        if (heatbath) launch_half<3, 1>(V, grid, st, own, other, g, tb, t, color, k0, k1);
        else launch_half<3, 0>(V, grid, st, own, other, g, tb, t, color, k0, k1);
    }
    ISING_LAUNCHED("int8_half_kernel");
    return ISING_OK;
}

extern "C" {

int ising_int8_half(int8_t* own, const int8_t* other, long long rows, int lh, int ly, int lz,
                    int ndim, int heatbath, const uint32_t* table, uint32_t t, uint32_t color,
                    uint32_t k0, uint32_t k1, void* stream)
{
    return int8_half_impl(own, other, rows, lh, ly, lz, ndim, heatbath, table, t, color, k0, k1,
                          as_stream(stream));
}

int ising_int8_kernel(long long rows, int lh, int ly, int lz, int ndim, int* vec, int* layout)
{
    if (!vec || !layout) return ising_fail(ISING_EINVAL, "int8_kernel: null pointer");
    Int8Geom g;
    int V;
    int rc = make_geom(rows, lh, ly, lz, ndim, &g, &V);
    if (rc) return rc;
    *vec = V;
    *layout = int8_layout(g);
    return ISING_OK;
}

int ising_int8_sweeps(int8_t* c0, int8_t* c1, long long rows, int lh, int ly, int lz, int ndim,
                      int heatbath, const uint32_t* table, uint32_t t0, long long n_sweeps,
                      uint32_t k0, uint32_t k1, void* stream)
{
    if (n_sweeps < 0) return ising_fail(ISING_EINVAL, "int8_sweeps: negative sweep count");
    for (long long s = 0; s < n_sweeps; ++s) {
        const u32 t = t0 + (u32)s;
        int rc = int8_half_impl(c0, c1, rows, lh, ly, lz, ndim, heatbath, table, t, 0u, k0, k1,
                                as_stream(stream));
        if (rc) return rc;
        rc = int8_half_impl(c1, c0, rows, lh, ly, lz, ndim, heatbath, table, t, 1u, k0, k1,
                            as_stream(stream));
        if (rc) return rc;
    }
    return ISING_OK;
}

int ising_int8_obs(const int8_t* c0, const int8_t* c1, long long rows, int lh, int ly, int lz,
                   int ndim, unsigned long long* scratch2, long long* out2, void* stream)
{
    if (!c0 || !c1 || !scratch2 || !out2) return ising_fail(ISING_EINVAL, "int8_obs: null pointer");
    Int8Geom g;
    int V;
    int rc = make_geom(rows, lh, ly, lz, ndim, &g, &V);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    const dim3 grid(obs_grid(g.n_chunks));
#define OBS_LAUNCH(ND, VV) \
    hipLaunchKernelGGL((int8_obs_kernel<ND, VV>), grid, dim3(256), 0, st, c0, c1, g, scratch2)
    if (ndim == 2) {
        if (V == 16) OBS_LAUNCH(2, 16);
        else if (V == 8) OBS_LAUNCH(2, 8);
        else OBS_LAUNCH(2, 4);
    } else {
        if (V == 16) OBS_LAUNCH(3, 16);
        else if (V == 8) OBS_LAUNCH(3, 8);
        else OBS_LAUNCH(3, 4);
    }
#undef OBS_LAUNCH
    // M = N - 2 minus, E = 2 anti - d N
    const long long n_spins = 2 * rows * (long long)lh;
    hipLaunchKernelGGL(obs_finalize, dim3(1), dim3(256), 0, st, scratch2, grid.x, out2, -2ll,
                       n_spins, 2ll, -(long long)ndim * n_spins);
    ISING_LAUNCHED("int8_obs_kernel");
    return ISING_OK;
}

}  // extern "C"
