// Resident multi-spin-coded sweeps of small 2D lattices: one launch holds every half sweep of a
// call, each workgroup keeping a band of rows of both packed colours in LDS for the whole
// launch (the band, round and exchange machinery is resident.hpp's).
//
// Why: at L <= 4096 a half sweep is one generation of short workgroups or less, and one launch
// per half sweep spends most of its time in launch gaps, ramp and drain. Here the words never
// leave the CUs between half sweeps. Measured (flips/ns, best per-half-sweep kernel against
// resident, same box): 1024^2 113 / 176, 2048^2 429 / 672, 4096^2 1514 / 2036; at 8192^2 the
// per-half-sweep compaction kernel is as fast (3722 / 3683). A band's half sweep holds few
// words (L = 1024: at most 80 per workgroup), so with one lane per word each lockstep wave's
// chain of Philox calls, not the VALU throughput, sets its time; bands of at most
// RS_LANES8_WORDS words per half sweep therefore give each word eight lanes (below).
//
// One lane per word: own, the other colour at the same word, the other colour's words in the
// rows either side and the horizontal neighbour word from LDS, the same classification
// (msc_state) and lockstep lazy comparison (msc_run) as the per-half-sweep lockstep kernel,
// with the global word index in the Philox counter (word, sweep, colour, 1 + level pair):
// bit-identical words (src/pyising/backends/numpy_backend.py:103-124 is the reference's half
// sweep). Eight lanes per word: the same classification on every lane of the word's segment,
// one level pair per lane, combined as described at rs_lanes8_combine.
#include "msc_core.hpp"
#include "resident.hpp"

struct MscResidentArgs {
    ResidentBands bands;  // row_bytes = 8 nw
    u64* w0;
    u64* w1;
    u32 nw;
    FastDiv div_nw;
    u32 t0, k0, k1;
    u64 codes;    // case code of every level pair (msc_level_codes)
    u32 lanes8;   // 1: eight lanes per word (small bands), 0: one lane per word, lockstep
};

// ------------------------------------------------------------ eight lanes per word ----
// When a band's half sweep holds few words (at most RS_LANES8_WORDS), a lockstep wave is a
// chain of up to 16 dependent Philox calls and the VALUs sit idle. Eight lanes then take one
// word: lane j of the word's segment evaluates level pair lp0 + j (any lane may evaluate any
// (word, level) call: the stream is counter-based) and, for a word undecided in every bit, the
// pair's masks G_j (bits its two levels leave undecided) and T_j (bits they accept). A bit
// undecided on entry is decided at the first pair whose G is 0 and accepted iff that pair's T
// is 1, and T_j lies inside ~G_j, so with E_j = G_0 & .. & G_{j-1} (E_0 all ones):
//   A = OR_j (E_j & T_j),  pw ^= und & A,  und &= G_0 & .. & G_7,
// the same bits as the pairs applied one after the other (checked by the oracle tests). The
// scan and the reductions stay inside the segment (two quads of a 16-lane DPP row).
constexpr u32 RS_LANES8_WORDS = 512;

template <u32 CTRL, u32 BANKS = 0xFu>
__device__ __forceinline__ u32 rs_dpp(u32 old, u32 v)
{
    return (u32)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xF, BANKS, false);
}

// Segment combine of one 32-bit half: A (accepted bits) and Qall (AND of every G) on all eight
// lanes of the segment.
__device__ __forceinline__ void rs_lanes8_combine(u32 G, u32 T, u32 lane8, u32& A, u32& Qall)
{
    // quad-local inclusive AND-scan: quad_perm [0,0,1,2] then [0,1,0,1] (AND is idempotent, so
    // a lane that reads itself is unchanged)
    u32 q = G & rs_dpp<0x90>(0u, G);
    q &= rs_dpp<0x44>(0u, q);
    // every lane's quad total (butterfly [1,0,3,2], [2,3,0,1]); the lower quad's total moves to
    // the upper quad (row_shr:4, banks 1 and 3 only; banks 0 and 2 keep all ones)
    u32 qt = G & rs_dpp<0xB1>(0u, G);
    qt &= rs_dpp<0x4E>(0u, qt);
    const u32 Q = q & rs_dpp<0x114, 0xAu>(~0u, qt);
    // exclusive scan: row_shr:1 (lane 0 of a row keeps all ones), segment starts all ones
    u32 E = rs_dpp<0x111>(~0u, Q);
    if (lane8 == 0u) E = ~0u;
    const u32 acc = E & T;
    u32 a = acc | rs_dpp<0xB1>(0u, acc);
    a |= rs_dpp<0x4E>(0u, a);
    A = a | rs_dpp<0x141>(0u, a);  // row_half_mirror: the other quad of the segment
    Qall = qt & rs_dpp<0x141>(0u, qt);
}

// G and T of the level pair whose case code is `code`, for one half: x, y the two levels' bits,
// a1 the bits with one anti-aligned neighbour (threshold bit: thr4's on a1, thr8's elsewhere).
__device__ __forceinline__ void rs_pair_masks(u32 code, u32 x, u32 y, u32 a1, u32& G, u32& T)
{
    const u32 m4a = 0u - (code & 1u), m8a = 0u - ((code >> 1) & 1u);
    const u32 m4b = 0u - ((code >> 2) & 1u), m8b = 0u - ((code >> 3) & 1u);
    constexpr u32 SEL = b3imm([](u32 s, u32 p, u32 q) { return (s & p) | (~s & q); });
    constexpr u32 BOTH_EQ = b3imm([](u32 e, u32 yy, u32 t) { return ~e & ~(yy ^ t); });
    constexpr u32 EQ_ACC = b3imm([](u32 e, u32 t, u32 yy) { return ~e & t & ~yy; });
    constexpr u32 ACC_OR = b3imm([](u32 t, u32 xx, u32 r) { return (t & ~xx) | r; });
    const u32 ta = bop3(a1, m4a, m8a, SEL), tb = bop3(a1, m4b, m8b, SEL);
    const u32 e = x ^ ta;
    G = bop3(e, y, tb, BOTH_EQ);
    T = bop3(ta, x, bop3(e, tb, y, EQ_ACC), ACC_OR);
}

__global__ __launch_bounds__(RS_THREADS) void msc_resident_kernel(MscResidentArgs a)
{
    extern __shared__ __attribute__((aligned(16))) int8_t rs_lds[];
    __shared__ u32 go;
    const u32 nw = a.nw, R = a.bands.R, T = a.bands.T, rows = a.bands.rows;
    const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    rs_run(a.bands, rs_lds, (int8_t*)a.w0, (int8_t*)a.w1, &go,
           [&](int8_t* own_lds, const int8_t* other_lds, u32 hh, u32 ext, u32 first) {
        MscArgs m;
        m.color = hh & 1u;
        m.t = a.t0 + (hh >> 1);
        m.k0 = a.k0;
        m.k1 = a.k1;
        u64* own = reinterpret_cast<u64*>(own_lds);
        const u64* other = reinterpret_cast<const u64*>(other_lds);
        const u32 row_lo = T - ext, n = (R + 2u * ext) * nw;
        if (a.lanes8) {
            const u32 seg = threadIdx.x >> 3, j = threadIdx.x & 7u;
            for (u32 base = 0; base < n; base += RS_THREADS / 8u) {
                const u32 it = base + seg;
                const bool valid = it < n;
                const u32 itc = valid ? it : n - 1u;
                const u32 rr = fast_div(itc, a.div_nw);
                const u32 x = itc - rr * nw;
                const u32 row = row_lo + rr;
                u32 y = first + row;
                if (y >= rows) y -= rows;
                const u32 p = (y + m.color) & 1u;
                const u64* crow = other + (size_t)row * nw;
                const u64 c = crow[x];
                const u32 xn = p ? ((x + 1u == nw) ? 0u : x + 1u) : ((x == 0u) ? nw - 1u : x - 1u);
                const u64 nb = crow[xn];
                const u64 hz = p ? ((c >> 1) | (nb << 63)) : ((c << 1) | (nb >> 63));
                MscState st = msc_state(own[(size_t)row * nw + x], c, hz,
                                        other[(size_t)(row + 1u) * nw + x],
                                        other[(size_t)(row - 1u) * nw + x]);
                if (!valid) {
                    st.ul = 0u;
                    st.uh = 0u;
                }
                const MscCounter ctr = msc_counter(y * nw + x);
                for (u32 lp0 = 0; lp0 < 16u && __ballot((st.ul | st.uh) != 0u) != 0; lp0 += 8u) {
                    const u32 lp = lp0 + j;  // this lane's level pair
                    const PhiloxOut r = msc_philox_lanes(m, ctr, lp);
                    const u32 code = (u32)(a.codes >> (4u * lp)) & 15u;
                    u32 gl, tl, gh, th, al, ql, ah, qh;
                    rs_pair_masks(code, r.x, r.z, st.a1l, gl, tl);  // levels 2 lp, 2 lp + 1
                    rs_pair_masks(code, r.y, r.w, st.a1h, gh, th);
                    rs_lanes8_combine(gl, tl, j, al, ql);
                    rs_lanes8_combine(gh, th, j, ah, qh);
                    st.pwl ^= st.ul & al;
                    st.pwh ^= st.uh & ah;
                    st.ul &= ql;
                    st.uh &= qh;
                }
                if (valid && j == 0u) own[(size_t)row * nw + x] = (u64)st.pwl | ((u64)st.pwh << 32);
            }
            __syncthreads();
            return;
        }
        // whole waves (the lazy loop ballots across the wave); lanes past the end idle
        for (u32 base = wave * 64u; base < n; base += RS_THREADS) {
            const u32 it = base + lane;
            const bool valid = it < n;
            const u32 itc = valid ? it : n - 1u;
            const u32 rr = fast_div(itc, a.div_nw);
            const u32 x = itc - rr * nw;
            const u32 row = row_lo + rr;  // LDS row, 1 <= row <= R + 2T - 2
            u32 y = first + row;
            if (y >= rows) y -= rows;
            const u32 p = (y + m.color) & 1u;
            const u64* crow = other + (size_t)row * nw;
            const u64 o = own[(size_t)row * nw + x];
            const u64 c = crow[x];
            const u32 xn = p ? ((x + 1u == nw) ? 0u : x + 1u) : ((x == 0u) ? nw - 1u : x - 1u);
            const u64 nb = crow[xn];
            const u64 hz = p ? ((c >> 1) | (nb << 63)) : ((c << 1) | (nb >> 63));
            const u64 up = other[(size_t)(row + 1u) * nw + x];  // row y + 1
            const u64 dn = other[(size_t)(row - 1u) * nw + x];  // row y - 1
            MscState st = msc_state(o, c, hz, up, dn);
            if (!valid) {
                st.ul = 0u;
                st.uh = 0u;
            }
            u32 calls = 0, slots = 0;
            msc_run<true, false>(m, a.codes, y * nw + x, 0u, 16u, st, calls, slots);
            if (valid) own[(size_t)row * nw + x] = (u64)st.pwl | ((u64)st.pwh << 32);
        }
        __syncthreads();
    });
}

extern "C" {

// Workspace bytes of the resident sweeps of a 2D packed lattice (rows x 128 nw) with `bands`
// bands (0: the default choice), 0 when they do not apply here (nw odd, LDS, placement).
int ising_msc_resident_bytes(long long rows, int nw, int bands, size_t* bytes)
{
    if (!bytes || rows < 1 || nw < 1 || bands < 0) return ising_fail(ISING_EINVAL, "msc_resident_bytes: bad argument");
    ResidentPlan pl;
    if (int rc = resident_plan(rows, 8ll * nw, bands, pl)) return rc;
    *bytes = pl.ws_bytes;
    return ISING_OK;
}

// n_sweeps sweeps of a 2D packed lattice in one cooperative launch (colour 0 then 1 per sweep,
// Philox counters t0, t0 + 1, ...): bit-identical to the per-half-sweep msc kernels.
// ws: a workspace of ising_msc_resident_bytes zeroed before first use and kept between calls;
// *epoch: the rounds its counters hold (0 after zeroing), advanced by this call.
int ising_msc_sweeps_resident(uint64_t* d_w0, uint64_t* d_w1, long long rows, int nw, uint32_t t0,
                              long long n_sweeps, uint32_t k0, uint32_t k1, uint32_t thr4,
                              uint32_t thr8, int bands, uint32_t* ws, uint32_t* epoch, void* stream)
{
    if (!d_w0 || !d_w1 || !ws || !epoch || n_sweeps < 0 || n_sweeps > (1ll << 30) || bands < 0 || nw < 1)
        return ising_fail(ISING_EINVAL, "msc_sweeps_resident: bad argument");
    if (n_sweeps == 0) return ISING_OK;
    ResidentPlan pl;
    if (int rc = resident_plan(rows, 8ll * nw, bands, pl)) return rc;
    if (!pl.B) return ising_fail(ISING_EINVAL, "msc_sweeps_resident: not applicable to this lattice here");
    MscResidentArgs a;
    a.bands.rows = (u32)rows;
    a.bands.row_bytes = 8u * (u32)nw;
    a.bands.R = pl.R;
    a.bands.T = pl.T;
    a.bands.B = pl.B;
    a.bands.one_l2 = pl.place == 0 ? 1u : 0u;
    a.bands.tagged = pl.tagged;
    a.bands.ws = ws;
    a.w0 = d_w0;
    a.w1 = d_w1;
    a.nw = (u32)nw;
    a.div_nw = make_fastdiv((u32)nw);
    a.k0 = k0;
    a.k1 = k1;
    a.codes = msc_level_codes(thr4, thr8);
    // eight lanes per word while a half sweep's words per band (band and halo rows) fit four
    // passes of the workgroup; larger bands keep one lane per word (their waves are many
    // already, and eight calls per word make them VALU-bound). Measured, us per sweep, one lane
    // / eight lanes per word: 1024^2 (80 words) 5.0-5.3 / 2.92, 2048^2 (352) 5.57 / 5.33,
    // 4096^2 (960) 7.3 / 13.4.
    a.lanes8 = (unsigned long long)(pl.R + 2u * pl.T - 2u) * (unsigned long long)nw <= RS_LANES8_WORDS ? 1u : 0u;
    return resident_launch((const void*)msc_resident_kernel, a, pl, t0, n_sweeps, epoch,
                           as_stream(stream));
}

// Failure bits of the resident launches since the workspace was zeroed (synchronizes the
// stream): 1 = a wait gave up (the lattice is then invalid).
int ising_msc_resident_status(const uint32_t* ws, int* failed, void* stream)
{
    if (!ws || !failed) return ising_fail(ISING_EINVAL, "msc_resident_status: null pointer");
    return resident_status(ws, failed, as_stream(stream));
}

}  // extern "C"
