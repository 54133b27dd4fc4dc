/*
 * Plain C restatement of pyising's checkerboard half sweeps -- the oracle for large and long
 * runs. TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the benchmark's CPU
 * baseline as the checker, never by the product library.
 *
 * Pinned to the NumPy restatement (oracle/ising_oracle.py) bit for bit on small cases by
 * tests/test_oracle_c.py; the NumPy restatement is in turn pinned by known-answer vectors and
 * full-lattice definitions.
 *
 * Reference algorithm (paths relative to the reference repository):
 *   philox4x32_10       src/pyising/philox.py:26-51, src/pyising/backends/cuda/ising.cu:10-21
 *   int8 half sweep     src/pyising/backends/numpy_backend.py:26-44, ising.cu:25-87
 *   msc half sweep      src/pyising/backends/numpy_backend.py:54-124, ising.cu:92-150
 *   observables         src/pyising/backends/observables_torch.py:46-111
 * Every site is computed independently (reads only the other colour), so the OpenMP row
 * loop gives the same result for any thread count.
 */
#include <stdint.h>
#include <stddef.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)c[0] * 0xD2511F53u;
        uint64_t p1 = (uint64_t)c[2] * 0xCD9E8D57u;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c[4] = {ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]};
        philox10(c, key[2 * i], key[2 * i + 1]);
        for (int j = 0; j < 4; ++j) out[4 * i + j] = c[j];
    }
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* row r = z * ly + y; row parity (y + z) % 2 (src/pyising/lattice.py:43-46). */
static inline void row_coords(int64_t r, int64_t ly, int64_t* y, int64_t* z)
{
    *z = r / ly;
    *y = r - *z * ly;
}

/* Neighbour sum h of colour site (r, j) of colour `color`; other = (rows, lh). */
static inline int neighbour_sum(const int8_t* other, int64_t r, int64_t j, int64_t lh, int64_t ly,
                                int64_t lz, int ndim, uint32_t color)
{
    int64_t y, z;
    row_coords(r, ly, &y, &z);
    const int p = (int)((y + z + color) & 1);
    const int8_t* row = other + r * lh;
    int h = row[j];
    h += p ? row[(j + 1) % lh] : row[(j + lh - 1) % lh];
    const int64_t yu = (y + 1) % ly, yd = (y + ly - 1) % ly;
    h += other[(z * ly + yu) * lh + j];
    h += other[(z * ly + yd) * lh + j];
    if (ndim == 3) {
        const int64_t zu = (z + 1) % lz, zd = (z + lz - 1) % lz;
        h += other[(zu * ly + y) * lh + j];
        h += other[(zd * ly + y) * lh + j];
    }
    return h;
}

/* One row r = z * ly + y of colour `color` (row parity (y + z) % 2, src/pyising/lattice.py:
 * 43-46): h = other[j] + other[j +- 1] (+1 when (y + z + color) is odd, else -1, wrapping in
 * the row) + the same column in rows y +- 1 (and planes z +- 1 in 3D), periodic. Colour site
 * n = r * lh + j takes word n % 4 of Philox((n / 4, t, colour, 0), key); lh % 4 == 0, so a
 * quad never straddles two rows. */
static void int8_row(int8_t* own, const int8_t* other, int64_t r, int64_t lh, int64_t ly,
                     int64_t lz, int ndim, int heatbath, const uint32_t* table, uint32_t t,
                     uint32_t color, uint32_t k0, uint32_t k1)
{
    const int64_t z = r / ly, y = r - z * ly;
    const int p = (int)((y + z + color) & 1);
    const int8_t* row = other + r * lh;
    const int8_t* up = other + (z * ly + (y + 1 == ly ? 0 : y + 1)) * lh;
    const int8_t* dn = other + (z * ly + (y == 0 ? ly - 1 : y - 1)) * lh;
    const int8_t* fw = row;
    const int8_t* bk = row;
    if (ndim == 3) {
        fw = other + ((z + 1 == lz ? 0 : z + 1) * ly + y) * lh;
        bk = other + ((z == 0 ? lz - 1 : z - 1) * ly + y) * lh;
    }
    int8_t* o = own + r * lh;
    for (int64_t j0 = 0; j0 < lh; j0 += 4) {
        uint32_t c[4] = {(uint32_t)((r * lh + j0) / 4), t, color, 0u};
        philox10(c, k0, k1);
        for (int i = 0; i < 4; ++i) {
            const int64_t j = j0 + i;
            const int64_t jn = p ? (j + 1 == lh ? 0 : j + 1) : (j == 0 ? lh - 1 : j - 1);
            int h = row[j] + row[jn] + up[j] + dn[j];
            if (ndim == 3) h += fw[j] + bk[j];
            const int s = o[j];
            if (heatbath) {
                o[j] = (c[i] < table[(h + 2 * ndim) / 2]) ? 1 : -1;
            } else {
                const int x = s * h;
                o[j] = (x <= 0 || c[i] < table[(x + 2 * ndim) / 2]) ? (int8_t)-s : (int8_t)s;
            }
        }
    }
}

void oracle_int8_half(int8_t* own, const int8_t* other, int64_t rows, int64_t lh, int64_t ly,
                      int64_t lz, int ndim, int heatbath, const uint32_t* table, uint32_t t,
                      uint32_t color, uint32_t k0, uint32_t k1)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; ++r)
        int8_row(own, other, r, lh, ly, lz, ndim, heatbath, table, t, color, k0, k1);
}

void oracle_int8_sweeps(int8_t* c0, int8_t* c1, int64_t rows, int64_t lh, int64_t ly, int64_t lz,
                        int ndim, int heatbath, const uint32_t* table, uint32_t t0, int64_t n,
                        uint32_t k0, uint32_t k1)
{
    for (int64_t s = 0; s < n; ++s) {
        oracle_int8_half(c0, c1, rows, lh, ly, lz, ndim, heatbath, table, t0 + (uint32_t)s, 0u,
                         k0, k1);
        oracle_int8_half(c1, c0, rows, lh, ly, lz, ndim, heatbath, table, t0 + (uint32_t)s, 1u,
                         k0, k1);
    }
}

/* msc: own/other (rows, nw) uint64, 2D only; ly = rows. */
void oracle_msc_half(uint64_t* own, const uint64_t* other, int64_t rows, int64_t nw, uint32_t t,
                     uint32_t color, uint32_t k0, uint32_t k1, uint32_t thr4, uint32_t thr8,
                     uint64_t* calls_out)
{
    uint64_t calls_total = 0;
#pragma omp parallel for schedule(static) reduction(+ : calls_total)
    for (int64_t y = 0; y < rows; ++y) {
        const int p = (int)((y + color) & 1);
        const uint64_t* orow = other + y * nw;
        const uint64_t* up = other + ((y + 1) % rows) * nw;
        const uint64_t* dn = other + ((y + rows - 1) % rows) * nw;
        for (int64_t x = 0; x < nw; ++x) {
            const int64_t w = y * nw + x;
            const uint64_t o = own[w], c = orow[x];
            uint64_t hz;
            if (p) hz = (c >> 1) | (orow[(x + 1) % nw] << 63);
            else hz = (c << 1) | (orow[(x + nw - 1) % nw] >> 63);
            const uint64_t x1 = o ^ c, x2 = o ^ hz, x3 = o ^ up[x], x4 = o ^ dn[x];
            const uint64_t anyx = x1 | x2 | x3 | x4;
            const uint64_t a0 = ~anyx;
            const uint64_t a2 = (x1 & x2) | (x3 & x4) | ((x1 ^ x2) & (x3 ^ x4));
            const uint64_t a1 = anyx & ~a2;
            uint64_t acc = 0, und = a0 | a1;
            for (int lp = 0; lp < 16 && und; ++lp) {
                uint32_t cc[4] = {(uint32_t)w, t, color, 1u + (uint32_t)lp};
                philox10(cc, k0, k1);
                ++calls_total;
                const uint64_t lv[2] = {(uint64_t)cc[0] | ((uint64_t)cc[1] << 32),
                                        (uint64_t)cc[2] | ((uint64_t)cc[3] << 32)};
                for (int i = 0; i < 2; ++i) {
                    const int bit = 31 - 2 * lp - i;
                    const uint64_t t4 = ((thr4 >> bit) & 1u) ? ~0ull : 0ull;
                    const uint64_t t8 = ((thr8 >> bit) & 1u) ? ~0ull : 0ull;
                    const uint64_t tm = (a1 & t4) | (a0 & t8);
                    acc |= und & ~lv[i] & tm;
                    und &= ~(lv[i] ^ tm);
                }
            }
            own[w] = o ^ (a2 | acc);
        }
    }
    if (calls_out) *calls_out += calls_total;
}

void oracle_msc_sweeps(uint64_t* w0, uint64_t* w1, int64_t rows, int64_t nw, uint32_t t0,
                       int64_t n, uint32_t k0, uint32_t k1, uint32_t thr4, uint32_t thr8,
                       uint64_t* calls_out)
{
    for (int64_t s = 0; s < n; ++s) {
        oracle_msc_half(w0, w1, rows, nw, t0 + (uint32_t)s, 0u, k0, k1, thr4, thr8, calls_out);
        oracle_msc_half(w1, w0, rows, nw, t0 + (uint32_t)s, 1u, k0, k1, thr4, thr8, calls_out);
    }
}

/* (M, E): M = sum of spins; E = -sum over colour-0 sites of s * h. */
void oracle_int8_obs(const int8_t* c0, const int8_t* c1, int64_t rows, int64_t lh, int64_t ly,
                     int64_t lz, int ndim, int64_t* out)
{
    int64_t m = 0, e = 0;
#pragma omp parallel for schedule(static) reduction(+ : m, e)
    for (int64_t r = 0; r < rows; ++r) {
        for (int64_t j = 0; j < lh; ++j) {
            const int64_t n = r * lh + j;
            m += c0[n] + c1[n];
            e += (int64_t)c0[n] * neighbour_sum(c1, r, j, lh, ly, lz, ndim, 0u);
        }
    }
    out[0] = m;
    out[1] = -e;
}

static inline int popc64(uint64_t v) { return __builtin_popcountll(v); }

void oracle_msc_obs(const uint64_t* w0, const uint64_t* w1, int64_t rows, int64_t nw, int64_t* out)
{
    int64_t ones = 0, anti = 0;
#pragma omp parallel for schedule(static) reduction(+ : ones, anti)
    for (int64_t y = 0; y < rows; ++y) {
        const int p = (int)(y & 1);
        const uint64_t* orow = w1 + y * nw;
        const uint64_t* up = w1 + ((y + 1) % rows) * nw;
        const uint64_t* dn = w1 + ((y + rows - 1) % rows) * nw;
        for (int64_t x = 0; x < nw; ++x) {
            const uint64_t s = w0[y * nw + x], c = orow[x];
            uint64_t hz;
            if (p) hz = (c >> 1) | (orow[(x + 1) % nw] << 63);
            else hz = (c << 1) | (orow[(x + nw - 1) % nw] >> 63);
            ones += popc64(s) + popc64(c);
            anti += popc64(s ^ c) + popc64(s ^ hz) + popc64(s ^ up[x]) + popc64(s ^ dn[x]);
        }
    }
    const int64_t n = rows * nw * 128;
    out[0] = 2 * ones - n;
    out[1] = -(2 * n - 2 * anti);
}
