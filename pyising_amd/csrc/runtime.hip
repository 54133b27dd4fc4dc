// C-ABI runtime glue: error state, device memory, copies, streams, events, and the
// measurement probes (Philox known-answer kernel, Philox throughput, copy bandwidth).
#include <string.h>

#include <string>

#include "common.hpp"

static thread_local std::string g_last_error;

void ising_set_error(const std::string& msg) { g_last_error = msg; }

int ising_fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int ising_check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return ISING_OK;
    g_last_error = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    return ISING_EHIP;
}

// ------------------------------------------------------------------------------ probes ----
__global__ void philox_kat_kernel(const u32* __restrict__ ctr, const u32* __restrict__ key,
                                  u32* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PhiloxOut r = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3],
                                      key[2 * i], key[2 * i + 1]);
    out[4 * i] = r.x;
    out[4 * i + 1] = r.y;
    out[4 * i + 2] = r.z;
    out[4 * i + 3] = r.w;
}

// Throughput probe: `calls` Philox calls per thread with counter (tid, i, 0, 0), results
// XOR-folded so none is dead code.
__global__ __launch_bounds__(256) void philox_rate_kernel(u32* __restrict__ out, int calls, u32 k0,
                                                          u32 k1)
{
    const u32 tid = blockIdx.x * blockDim.x + threadIdx.x;
    u32 acc = 0;
    for (int i = 0; i < calls; ++i) {
        const PhiloxOut r = philox4x32_10(tid, (u32)i, 0u, 0u, k0, k1);
        acc ^= r.x ^ r.y ^ r.z ^ r.w;
    }
    out[tid] = acc;
}

// Streaming copy of n16 16-byte vectors, grid-stride.
__global__ __launch_bounds__(256) void copy_kernel(uint4* __restrict__ dst,
                                                   const uint4* __restrict__ src, u64 n16)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

extern "C" {

const char* ising_last_error(void) { return g_last_error.c_str(); }

int ising_device_count(int* n)
{
    if (!n) return ising_fail(ISING_EINVAL, "device_count: null pointer");
    ISING_HIP(hipGetDeviceCount(n));
    return ISING_OK;
}

int ising_device_info(char* name, int name_len, int* n_cu, long long* total_bytes)
{
    int dev = 0;
    ISING_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    ISING_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
        strncpy(name, p.gcnArchName, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (n_cu) *n_cu = p.multiProcessorCount;
    if (total_bytes) *total_bytes = (long long)p.totalGlobalMem;
    return ISING_OK;
}

int ising_malloc(void** ptr, size_t bytes)
{
    if (!ptr) return ising_fail(ISING_EINVAL, "malloc: null pointer");
    *ptr = nullptr;
    ISING_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return ISING_OK;
}

int ising_free(void* ptr)
{
    if (ptr) ISING_HIP(hipFree(ptr));
    return ISING_OK;
}

int ising_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    ISING_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return ISING_OK;
}

int ising_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!bytes) return ISING_OK;
    ISING_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    ISING_HIP(hipStreamSynchronize(as_stream(stream)));
    return ISING_OK;
}

int ising_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!bytes) return ISING_OK;
    ISING_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    ISING_HIP(hipStreamSynchronize(as_stream(stream)));
    return ISING_OK;
}

int ising_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!bytes) return ISING_OK;
    ISING_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return ISING_OK;
}

int ising_stream_create(void** stream)
{
    if (!stream) return ising_fail(ISING_EINVAL, "stream_create: null pointer");
    hipStream_t s;
    ISING_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return ISING_OK;
}

int ising_stream_destroy(void* stream)
{
    if (stream) ISING_HIP(hipStreamDestroy(as_stream(stream)));
    return ISING_OK;
}

int ising_stream_sync(void* stream)
{
    ISING_HIP(hipStreamSynchronize(as_stream(stream)));
    return ISING_OK;
}

int ising_device_sync(void)
{
    ISING_HIP(hipDeviceSynchronize());
    return ISING_OK;
}

int ising_event_create(void** ev)
{
    if (!ev) return ising_fail(ISING_EINVAL, "event_create: null pointer");
    hipEvent_t e;
    ISING_HIP(hipEventCreate(&e));
    *ev = e;
    return ISING_OK;
}

int ising_event_destroy(void* ev)
{
    if (ev) ISING_HIP(hipEventDestroy((hipEvent_t)ev));
    return ISING_OK;
}

int ising_event_record(void* ev, void* stream)
{
    ISING_HIP(hipEventRecord((hipEvent_t)ev, as_stream(stream)));
    return ISING_OK;
}

int ising_event_query(void* ev)  // 1 = complete, 0 = pending
{
    hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return ising_check_hip(e, "hipEventQuery");
}

int ising_event_elapsed_ms(float* ms, void* start, void* stop)
{
    ISING_HIP(hipEventSynchronize((hipEvent_t)stop));
    ISING_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return ISING_OK;
}

int ising_event_sync(void* ev)
{
    ISING_HIP(hipEventSynchronize((hipEvent_t)ev));
    return ISING_OK;
}

// Work queued on `stream` after this call waits for the last record of `ev`.
int ising_stream_wait_event(void* stream, void* ev)
{
    ISING_HIP(hipStreamWaitEvent(as_stream(stream), (hipEvent_t)ev, 0));
    return ISING_OK;
}

// Host arrays in, host array out (n counters of 4 words, n keys of 2 words).
int ising_philox_kat(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int n)
{
    if (n <= 0 || !ctr || !key || !out) return ising_fail(ISING_EINVAL, "philox_kat: bad argument");
    u32 *dc = nullptr, *dk = nullptr, *dout = nullptr;
    int rc = ISING_OK;
    hipError_t e = hipMalloc((void**)&dc, sizeof(u32) * 4 * n);
    if (e == hipSuccess) e = hipMalloc((void**)&dk, sizeof(u32) * 2 * n);
    if (e == hipSuccess) e = hipMalloc((void**)&dout, sizeof(u32) * 4 * n);
    if (e == hipSuccess) e = hipMemcpy(dc, ctr, sizeof(u32) * 4 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dk, key, sizeof(u32) * 2 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(philox_kat_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, dc, dk, dout, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(u32) * 4 * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = ising_check_hip(e, "philox_kat");
    (void)hipFree(dc);
    (void)hipFree(dk);
    (void)hipFree(dout);
    return rc;
}

int ising_philox_rate(uint32_t* out, int blocks, int calls, uint32_t k0, uint32_t k1, void* stream)
{
    if (!out || blocks <= 0 || calls <= 0) return ising_fail(ISING_EINVAL, "philox_rate: bad argument");
    hipLaunchKernelGGL(philox_rate_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), out, calls,
                       k0, k1);
    ISING_LAUNCHED("philox_rate_kernel");
    return ISING_OK;
}

int ising_copy(void* dst, const void* src, size_t bytes, int blocks, void* stream)
{
    if (!dst || !src || bytes % 16 || blocks <= 0)
        return ising_fail(ISING_EINVAL, "copy: pointers, 16-byte multiple and blocks > 0 required");
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), (uint4*)dst,
                       (const uint4*)src, (u64)(bytes / 16));
    ISING_LAUNCHED("copy_kernel");
    return ISING_OK;
}

}  // extern "C"
