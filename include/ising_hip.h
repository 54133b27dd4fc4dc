/*
 * ising_hip.h -- C ABI of libising_hip.so, the gfx950 (MI355X) kernels of the pyising `hip`
 * backend.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative status (ISING_EINVAL = -1 bad
 *     argument, ISING_EHIP = -2 HIP runtime failure, ISING_ERANGE = -3 size beyond what the
 *     entry point supports); ising_last_error() gives the message of the calling thread's last
 *     failure. No exceptions cross the ABI.
 *   - Pointers named d_* / spin arrays / scratch / out are device pointers owned by the
 *     caller; `table` and the Philox known-answer arrays are host pointers.
 *   - `stream` is a hipStream_t (NULL = the null stream). Launches are stream-ordered and
 *     asynchronous unless stated otherwise; one host thread per simulation.
 *   - Layouts follow the reference (src/pyising/lattice.py:1-14): colour c is a row-major
 *     (rows, Lx/2) int8 array (+1 = 0x01, -1 = 0xFF), rows = Ly (2D) or Lz*Ly (3D); the packed
 *     msc layout is (rows, Lx/128) uint64 with bit b of word x = colour site 64x + b, 1 = +1.
 *   - Philox key (k0, k1) = (low, high) 32 bits of the 64-bit seed (src/pyising/philox.py:20-23);
 *     `t` is the sweep counter (counter word 1), colour is counter word 2.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   IsingSimulation._half_sweep          src/pyising/backends/base.py:93-103
 *   CupyInt8._half_sweep / int8_half     src/pyising/backends/cupy_backend.py:68-72,
 *                                        src/pyising/backends/cuda/ising.cu:35-87
 *   CupyMultiSpin._half_sweep / msc_half src/pyising/backends/cupy_backend.py:129-135,
 *                                        src/pyising/backends/cuda/ising.cu:102-150
 *   IsingSimulation.sweep                src/pyising/backends/base.py:85-90
 *   observables_int8 / observables_packed src/pyising/backends/observables_torch.py:46-111
 *   to/from_compressed, pack/unpack_bits src/pyising/lattice.py:49-90
 *   philox_calls_per_word (stats buffer) src/pyising/backends/cupy_backend.py:137-144
 *   philox4x32_10                        src/pyising/backends/cuda/ising.cu:10-21
 *   copy_bandwidth, philox_rate probes   bench/roofline.py:24-78
 */
#ifndef ISING_HIP_H
#define ISING_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scratch a caller passes to the observable reductions (ising_int8_obs, ising_msc_obs): one
 * pair of integer partial sums per workgroup, any contents on entry. */
#define ISING_OBS_SCRATCH_BYTES 16384

/* ---------------------------------------------------------------- status, runtime ---- */
/* Message of the calling thread's last failure (empty if none). No reference counterpart. */
const char* ising_last_error(void);
/* Number of visible GPUs; the reference gates GPU entries on torch.cuda.is_available()
 * (src/pyising/backends/_gpu.py:39-46). */
int ising_device_count(int* n);
/* Architecture name, compute-unit count and HBM bytes of the current device
 * (bench/run.py:82-106 machine_info). */
int ising_device_info(char* name, int name_len, int* n_cu, long long* total_bytes);
/* Device memory and copies: the reference allocates through CuPy (cupy_backend.py:62,106,150). */
int ising_malloc(void** ptr, size_t bytes);
int ising_free(void* ptr);
int ising_mem_info(size_t* free_bytes, size_t* total_bytes);
int ising_memcpy_h2d(void* d_dst, const void* src, size_t bytes, void* stream); /* synchronous */
int ising_memcpy_d2h(void* dst, const void* d_src, size_t bytes, void* stream); /* synchronous */
int ising_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, void* stream);
/* Streams (default priority, non-blocking) and events; synchronize() of base.py:130-131. */
int ising_stream_create(void** stream);
int ising_stream_destroy(void* stream);
int ising_stream_sync(void* stream);
int ising_device_sync(void);
int ising_event_create(void** ev);
int ising_event_destroy(void* ev);
int ising_event_record(void* ev, void* stream);
int ising_event_query(void* ev); /* 1 = complete, 0 = pending, < 0 error */
int ising_event_elapsed_ms(float* ms, void* start, void* stop);
int ising_event_sync(void* ev);
/* Work queued on `stream` after this call waits for the last record of `ev`. */
int ising_stream_wait_event(void* stream, void* ev);

/* ----------------------------------------------------------------- layouts ---- */
/* Full lattice (rows, lx) bytes -> two int8 colour arrays (rows, lx/2); any byte > 0 is +1,
 * anything else -1 (so raw {0,1} hot-start draws are accepted). ly = rows per z-plane
 * (ly = rows in 2D). Replaces to_compressed, src/pyising/lattice.py:49-60. lx % 8 == 0. */
int ising_split_int8(const void* d_full, int8_t* d_c0, int8_t* d_c1, long long rows, int ly, int lx,
                     void* stream);
/* Inverse: from_compressed, src/pyising/lattice.py:63-72. */
int ising_merge_int8(const int8_t* d_c0, const int8_t* d_c1, void* d_full, long long rows, int ly,
                     int lx, void* stream);
/* Full 2D lattice -> two packed colour arrays (rows, lx/128): to_compressed + pack_bits,
 * src/pyising/lattice.py:49-82. lx % 128 == 0. */
int ising_split_pack(const void* d_full, uint64_t* d_w0, uint64_t* d_w1, long long rows, int lx,
                     void* stream);
/* Inverse: unpack_bits + from_compressed, src/pyising/lattice.py:63-72,85-90. */
int ising_unpack_merge(const uint64_t* d_w0, const uint64_t* d_w1, void* d_full, long long rows,
                       int lx, void* stream);
/* Hot start on the device: n bytes (n % 8 == 0, 8-byte aligned) equal to numpy
 * default_rng(seed).integers(0, 2, n, int8), given the seeded PCG64 (state, inc) that numpy
 * reports for default_rng(seed) (split into 64-bit halves). Replaces the host draws of
 * random_lattice, src/pyising/lattice.py:93-96; the split kernels then map 0/1 to -1/+1. */
int ising_hot_start(void* d_out, long long n, uint64_t state_hi, uint64_t state_lo,
                    uint64_t inc_hi, uint64_t inc_lo, void* stream);
/* Fill n uint64 words with `value` (cold starts, counters): _fill_cold of
 * src/pyising/backends/base.py:119-120, cupy_backend.py:154-156. */
int ising_fill_u64(uint64_t* d_dst, long long n, uint64_t value, void* stream);

/* ----------------------------------------------------------- int8 checkerboard ---- */
/* One half sweep of colour `color`: own (rows, lh) updated in place from other.
 * rows = ly * lz (lz = 1 in 2D); ndim 2 or 3; heatbath 0/1; table = host array of 2*ndim+1
 * uint32 thresholds (src/pyising/lattice.py:128-142). lh % 4 == 0.
 * Replaces CupyInt8._half_sweep (cupy_backend.py:68-72) / int8_half (ising.cu:35-87). */
int ising_int8_half(int8_t* d_own, const int8_t* d_other, long long rows, int lh, int ly, int lz,
                    int ndim, int heatbath, const uint32_t* table, uint32_t t, uint32_t color,
                    uint32_t k0, uint32_t k1, void* stream);
/* The per-half-sweep kernel's choice for a lattice: *vec = sites per lane (16, 8 or 4) and
 * *layout = 0 (each lane's own row arithmetic), 1 (row waves: a wave's 64 lanes in one row) or
 * 2 (segment waves: a wave is whole rows of one plane). Host only. No reference counterpart
 * (kernel selection). */
int ising_int8_kernel(long long rows, int lh, int ly, int lz, int ndim, int* vec, int* layout);
/* n_sweeps full sweeps (colour 0 then 1, t = t0, t0+1, ...): IsingSimulation.sweep
 * (base.py:85-90) with all launches issued from C. */
int ising_int8_sweeps(int8_t* d_c0, int8_t* d_c1, long long rows, int lh, int ly, int lz, int ndim,
                      int heatbath, const uint32_t* table, uint32_t t0, long long n_sweeps,
                      uint32_t k0, uint32_t k1, void* stream);
/* Resident sweeps of a 2D lattice (rows x 2 lh): the same n_sweeps sweeps as ising_int8_sweeps
 * (bit-identical lattices) in one cooperative launch, each workgroup keeping a band of rows of
 * both colours plus halo rows in LDS and exchanging edge rows with its neighbours every few
 * half sweeps; for lattices too small to fill the GPU between per-half-sweep launches. Same
 * entry point the reference's sweep (base.py:85-90) would bind.
 * ising_int8_resident_bytes: the workspace bytes for `bands` bands (0: the library's choice),
 * 0 when the form does not apply (lh % 16, rows, LDS, workgroup placement). The workspace is
 * zeroed before its first use and kept between calls with the same `bands`; *epoch (0 after
 * zeroing) is advanced by each call. Calls on one workspace must not overlap. */
int ising_int8_resident_bytes(long long rows, int lh, int bands, size_t* bytes);
int ising_int8_sweeps_resident(int8_t* d_c0, int8_t* d_c1, long long rows, int lh, int heatbath,
                               const uint32_t* table, uint32_t t0, long long n_sweeps, uint32_t k0,
                               uint32_t k1, int bands, uint32_t* d_ws, uint32_t* epoch,
                               void* stream);
/* Failure bits of the resident launches since the workspace was zeroed (synchronizes the
 * stream): 1 = a bounded wait for a neighbour band gave up; the lattice is then invalid. The
 * bits are the workspace's first 32-bit word, so a caller may instead copy that word on the
 * stream together with other results and read both with one host wait. */
int ising_int8_resident_status(const uint32_t* d_ws, int* failed, void* stream);
/* Exact (M, E) into out2[0..1] (device int64); d_scratch2 = ISING_OBS_SCRATCH_BYTES of device
 * memory (any contents).
 * observables_int8, src/pyising/backends/observables_torch.py:46-69. */
int ising_int8_obs(const int8_t* d_c0, const int8_t* d_c1, long long rows, int lh, int ly, int lz,
                   int ndim, unsigned long long* d_scratch2, long long* d_out2, void* stream);

/* ------------------------------------------------------ multi-spin coding (2D) ---- */
/* Process default msc kernel: mode 0 queue compaction (k = 10 * cap + R: cap in 3..6 unrolled
 * Philox calls per word, then packed one-call rounds for the undecided words; R in {2,4} x 256
 * words per workgroup), 1 lockstep, 2 eager (all 16 calls; the msc-eager32 ablation of
 * cupy_backend.py:176-180), 3 by size (default, k = 42: lockstep up to 2^18 words per colour,
 * e.g. 4096 x 8192, queue compaction with k above, fused as in mode 4 in ising_msc_sweeps_sync
 * where it applies and a half sweep has at least 2048 tiles, e.g. 8192 x 16384), 4 fused queue
 * compaction (k = 10 * cap + 2, cap in 3..5: the
 * compaction tiles of up to 64 half sweeps in one launch, ordered by per-tile completion
 * counters; only in ising_msc_sweeps_sync, a single half sweep runs mode 0). All modes give
 * identical words. No reference counterpart (kernel selection). */
int ising_msc_config(int mode, int k);
/* The concrete kernel (mode 0, 1 or 2, and k for mode 0, else 0) that a half sweep of a
 * rows x (128 nw) lattice runs with `mode`, `k` (mode < 0: the process default). Host only.
 * No reference counterpart (kernel selection). */
int ising_msc_kernel(long long rows, int nw, int mode, int k, int* out_mode, int* out_k);
/* One half sweep with the process default kernel. thr4 = table[3], thr8 = table[4].
 * stats (optional device uint64[3], accumulated): Philox calls, lane-slots executed, words.
 * Replaces CupyMultiSpin._half_sweep (cupy_backend.py:129-135) / msc_half (ising.cu:102-150)
 * and the stats buffer of cupy_backend.py:137-144. */
int ising_msc_half(uint64_t* d_own, const uint64_t* d_other, long long rows, int nw, uint32_t t,
                   uint32_t color, uint32_t k0, uint32_t k1, uint32_t thr4, uint32_t thr8,
                   unsigned long long* d_stats, void* stream);
/* Same with an explicit kernel mode and k (mode < 0: process default). */
int ising_msc_half_ex(uint64_t* d_own, const uint64_t* d_other, long long rows, int nw, uint32_t t,
                      uint32_t color, uint32_t k0, uint32_t k1, uint32_t thr4, uint32_t thr8,
                      int mode, int k, unsigned long long* d_stats, void* stream);
/* n_sweeps full sweeps: IsingSimulation.sweep (base.py:85-90). */
int ising_msc_sweeps(uint64_t* d_w0, uint64_t* d_w1, long long rows, int nw, uint32_t t0,
                     long long n_sweeps, uint32_t k0, uint32_t k1, uint32_t thr4, uint32_t thr8,
                     unsigned long long* d_stats, void* stream);
int ising_msc_sweeps_ex(uint64_t* d_w0, uint64_t* d_w1, long long rows, int nw, uint32_t t0,
                        long long n_sweeps, uint32_t k0, uint32_t k1, uint32_t thr4,
                        uint32_t thr8, int mode, int k, unsigned long long* d_stats,
                        void* stream);
/* n_sweeps full sweeps like ising_msc_sweeps_ex, with a sync workspace for the fused form
 * (mode 4, or mode 3 / the default where it applies: Lx/128 in {64, 128, 256, 512}, a
 * power-of-two number >= 16 of 512-word tiles, and the device places workgroups round-robin
 * over its 8 XCDs or all on one). d_sync = ising_msc_sync_bytes of device memory, zero before its
 * first use and kept by the caller between calls together with *epoch (host, in/out: the
 * workspace's counter value, 0 for a zeroed workspace; advanced by the half sweeps run fused).
 * The first call on a device probes its workgroup placement once (synchronous). A failed
 * dependency wait leaves the lattice unswept from that tile on and sets a failure bit that
 * ising_msc_sync_status reports; the workspace must then be zeroed (and *epoch reset to 0).
 * Replaces IsingSimulation.sweep (base.py:85-90) for msc with one launch per up to 32 sweeps
 * instead of one per half sweep. */
int ising_msc_sweeps_sync(uint64_t* d_w0, uint64_t* d_w1, long long rows, int nw, uint32_t t0,
                          long long n_sweeps, uint32_t k0, uint32_t k1, uint32_t thr4,
                          uint32_t thr8, int mode, int k, uint32_t* d_sync, uint32_t* epoch,
                          unsigned long long* d_stats, void* stream);
/* Resident sweeps of a 2D packed lattice (rows x 128 nw, nw even): the same n_sweeps sweeps as
 * the per-half-sweep kernels (bit-identical words) in one cooperative launch, each workgroup
 * keeping a band of rows of both packed colours plus halo rows in LDS and exchanging edge rows
 * with its neighbours every few half sweeps; for lattices up to 4096^2, whose half sweeps are
 * too short to amortize a launch (8192^2 fits but runs no faster). The lazy comparison is the
 * lockstep one. Same
 * entry point the reference's sweep (base.py:85-90) would bind.
 * ising_msc_resident_bytes: the workspace bytes for `bands` bands (0: the library's choice),
 * 0 when the form does not apply (nw odd, rows, LDS, workgroup placement); workspace and *epoch
 * as for ising_int8_sweeps_resident. */
int ising_msc_resident_bytes(long long rows, int nw, int bands, size_t* bytes);
int ising_msc_sweeps_resident(uint64_t* d_w0, uint64_t* d_w1, long long rows, int nw, uint32_t t0,
                              long long n_sweeps, uint32_t k0, uint32_t k1, uint32_t thr4,
                              uint32_t thr8, int bands, uint32_t* d_ws, uint32_t* epoch,
                              void* stream);
/* Failure bits of the resident launches since the workspace was zeroed (synchronizes the
 * stream): 1 = a bounded wait for a neighbour band gave up; the lattice is then invalid. As for
 * int8, the bits are the workspace's first 32-bit word. */
int ising_msc_resident_status(const uint32_t* d_ws, int* failed, void* stream);
/* Bytes of the sync workspace of a rows x (128 nw) lattice (host only). No reference
 * counterpart. */
int ising_msc_sync_bytes(long long rows, int nw, size_t* bytes);
/* Failure bits of the sync workspace of a rows x (128 nw) lattice whose counters should all
 * hold `epoch` (synchronous read on `stream`): 0 = none, 1 = a dependency wait gave up,
 * 2 = a tile counter differs from epoch (a tile was skipped: workgroups were not placed
 * round-robin over the XCDs). No reference counterpart. */
int ising_msc_sync_status(const uint32_t* d_sync, long long rows, int nw, uint32_t epoch,
                          int* failed, void* stream);
/* Exact (M, E) of a packed lattice: observables_packed, observables_torch.py:85-111.
 * d_scratch2 = ISING_OBS_SCRATCH_BYTES of device memory (any contents). */
int ising_msc_obs(const uint64_t* d_w0, const uint64_t* d_w1, long long rows, int nw,
                  unsigned long long* d_scratch2, long long* d_out2, void* stream);
/* Exact (M, E) of a packed lattice like ising_msc_obs, allocating nothing and taking no
 * scratch: the partial sums go to a fixed 16 KiB device array of the library (one copy per GPU,
 * kept until the process ends), so any lattice size works. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Calls that overlap on one GPU race on the array; the array
 * keeps the last call's partial sums, and every call rewrites the ones it reads. */
int ising_msc_obs_static(const uint64_t* d_w0, const uint64_t* d_w1, long long rows, int nw,
                         long long* d_out2, void* stream);

/* ---------------------------------------------------------- batched scans ---- */
/* No reference counterpart: the reference sweeps one lattice per launch sequence
 * (src/pyising/backends/base.py:85-90). A batch is a list of lattices (int8 2D/3D either
 * algorithm, or msc 2D), each with its own shape, thresholds, seed and sweep counter, whose
 * device colour arrays stay owned by the caller. Every half sweep of all of them is one launch
 * (default), or up to 512 sweeps are one dependency-ordered launch in which each tile waits only
 * for the tiles of its own lattice that hold its neighbour rows in the previous half sweep
 * (ising_batch_set_persistent); every lattice's result is bit-identical to sweeping it alone. After every sweep M = sum of spins is accumulated per lattice: sum M, sum |M|,
 * sum M^2 and sum M^4 (128-bit) for Binder cumulants; a batched lattice holds at most 2^28
 * spins (ISING_ERANGE beyond), so the sums stay exact for 65536 sweeps
 * (src/pyising/observables.py:39-42). */
typedef struct {
    int kind;            /* 0 = int8 compressed checkerboard; msc packed: 1 = the batch's
                          * default tile (msc_mode), 2 = lockstep 256-word tiles, 3 = queue
                          * compaction 512-word tiles */
    int ndim;            /* 2 or 3 (msc: 2) */
    int heatbath;        /* int8 only: 0 Metropolis, 1 heat bath */
    int lz, ly, lx;      /* shape; lz = 1 in 2D */
    unsigned long long seed;
    uint32_t table[7];   /* thresholds (2*ndim+1 used; msc uses table[3], table[4]) */
    void* c0;            /* device colour arrays (int8 (rows, lx/2) or uint64 (rows, lx/128)) */
    void* c1;
    uint32_t t0;         /* sweep counter at batch start */
    int vec;             /* int8: sites per lane, 16, 8 or 4 dividing lx/2; 0 = the widest that
                          * still fills one 256-lane tile (else the narrowest). Any width gives
                          * the same spins. */
    uint32_t max_sweeps; /* sweeps this lattice takes part in, counted from the batch's create
                          * or assign (0: no limit); past it the lattice's tiles and moment
                          * sums are skipped, so it holds exactly t0 + max_sweeps sweeps while
                          * the others go on (continuous batching: no step is cut short to
                          * the lattice that needs the fewest sweeps) */
} IsingBatchLattice;
/* blocks_per_cu: accepted and unused (launch sizes follow the tile counts); msc_mode 0 =
 * queue-compaction tiles (512 words), 1 = lockstep tiles (256 words); msc_cap = unrolled calls
 * of the queue tiles (must be 4). */
int ising_batch_create(const IsingBatchLattice* lats, int n, int blocks_per_cu, int msc_mode,
                       int msc_cap, void** handle);
/* n_sweeps sweeps of every lattice (asynchronous; launch form or dependency-ordered form). */
int ising_batch_sweeps(void* handle, long long n_sweeps, void* stream);
/* Capture the launches of sweeps_per_replay sweeps into a hipGraph on `stream` (a created
 * stream); ISING_EHIP if the runtime cannot capture them. Replays continue the sweep counter. */
int ising_batch_graph(void* handle, long long sweeps_per_replay, void* stream);
int ising_batch_replay(void* handle, int replays, void* stream);
/* Synchronous read of n x 7 int64 accumulators (sum M, sum |M|, sum M^2 lo, hi, sum M^4 lo,
 * hi, samples; the 128-bit sums as unsigned (lo, hi) words) and the number of sweeps done. */
int ising_batch_stats(void* handle, long long* out, long long* sweeps_done, void* stream);
int ising_batch_reset_stats(void* handle, void* stream);
int ising_batch_info(void* handle, int* grid, int* total_tiles, int* max_blocks_per_cu);
/* Form of a sweep call: 0 (default) = one launch per half sweep (+ one accumulation launch per
 * sweep while moments are on); 1 = dependency-ordered launches of up to 512 sweeps each (per-tile
 * completion counters; each lattice runs on one XCD, or in eight bands, one per XCD), followed
 * by a check of the counters. 1 needs a batch made by ising_batch_create and a
 * device that places workgroups round-robin over its XCDs or all on one (ISING_EINVAL
 * otherwise). */
int ising_batch_set_persistent(void* handle, int persistent);
/* Failure bits of the dependency-ordered launches since create (synchronizes `stream`): 1 = a
 * tile's bounded wait gave up, 2 = a tile did not publish its last half sweep. After a failure
 * the lattices and sums are undefined. */
int ising_batch_status(void* handle, int* failed, void* stream);
/* 1 (default): accumulate the moments of M after every sweep (one more launch per sweep in the
 * launch form); 0: skip it, e.g. in a serving loop that reads each lattice's own (M, E). */
int ising_batch_set_moments(void* handle, int on);
/* Launch form: tiles each workgroup runs in sequence, k = 1 (default), 2, 4 or 8 (the tiles of
 * one workgroup come from one XCD's list, so locality is kept). Synchronous: waits for the
 * device, then rebuilds the launch order. Results do not depend on k. */
int ising_batch_set_tiles_per_workgroup(void* handle, int k);
int ising_batch_destroy(void* handle);
/* Tiles of one half sweep of these lattices (host only, no device call): the tile capacity a
 * reserved batch needs for them. */
int ising_batch_tiles(const IsingBatchLattice* lats, int n, int msc_mode, long long* tiles);
/* Continuous batching (no reference counterpart): a batch with device tables for up to
 * max_lattices lattices and max_tiles tiles and no lattices yet. Allocates once (synchronous). */
int ising_batch_reserve(int max_lattices, long long max_tiles, int blocks_per_cu, int msc_mode,
                        int msc_cap, void** handle);
/* Replace the lattices of a reserved batch (n >= 0), stream-ordered on `stream` (pinned,
 * double-buffered uploads; allocates nothing). Sweeps queued before see the old list, sweeps
 * after it the new one; accumulators and the batch sweep counter restart at zero, lattice i
 * continues from its own t0. ISING_ERANGE beyond the reserved capacity. */
int ising_batch_assign(void* handle, const IsingBatchLattice* lats, int n, void* stream);

/* ------------------------------------------------------------------ probes ---- */
/* Philox4x32-10 of n (counter[4], key[2]) pairs, host arrays in and out: the known-answer
 * path of philox4x32_10 (ising.cu:10-21; tests/test_philox.py:7-18). Synchronous. */
int ising_philox_kat(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int n);
/* Throughput probe: blocks x 256 threads x `calls` Philox calls (bench/roofline.py:24-36). */
int ising_philox_rate(uint32_t* d_out, int blocks, int calls, uint32_t k0, uint32_t k1,
                      void* stream);
/* Streaming copy kernel, bytes % 16 == 0 (bench/roofline.py:39-56 copy bandwidth). */
int ising_copy(void* d_dst, const void* d_src, size_t bytes, int blocks, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ISING_HIP_H */
