// HIP runtime services behind the internal C ABI: one default-priority stream
// per process, device memory, copies, the uploaded constant tables and
// per-kernel timing with HIP events on the launch stream.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include "tcg.h"
#include "tc_dev.h"

namespace {
hipStream_t g_stream = nullptr;
TcgConsts* g_consts = nullptr;
int g_wide = 0;
int g_timing = 0;
hipEvent_t g_ev[8];
int g_ev_ready = 0;
int g_last_kind = 0;  // 1 encode, 2 decode
thread_local char g_err[256];

int32_t fail(const char* what, hipError_t e)
{
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    return -2;
}
}  // namespace

hipStream_t tcg_stream() { return g_stream; }
const TcgConsts* tcg_consts_dev() { return g_consts; }
int tcg_wide_mode() { return g_wide; }

int32_t tcg_check(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { return fail(what, e); }
    return 0;
}

void TcgTimer::mark(hipStream_t s, int i)
{
    if (g_timing && g_ev_ready && i < 8) { (void)hipEventRecord(g_ev[i], s); }
}
void TcgTimer::finish_encode() { g_last_kind = 1; }
void TcgTimer::finish_decode() { g_last_kind = 2; }

// plain 16-byte-vector device copy: the reference point for HBM byte counters
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) { dst[i] = src[i]; }
}

extern "C" {

int32_t tcg_copy(uint64_t dst, uint64_t src, size_t bytes)
{
    const size_t n16 = bytes / 16u;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, g_stream, (const uint4*)(uintptr_t)src,
                       (uint4*)(uintptr_t)dst, n16);
    const hipError_t e = hipStreamSynchronize(g_stream);
    return e == hipSuccess ? 0 : fail("copy", e);
}

int32_t tcg_init(void)
{
    if (g_stream != nullptr) { return 0; }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        snprintf(g_err, sizeof(g_err), "no HIP device available (%s)",
                 e != hipSuccess ? hipGetErrorString(e) : "count 0");
        return -16;
    }
    e = hipSetDevice(0);
    if (e != hipSuccess) { return fail("hipSetDevice", e); }
    e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
    if (e != hipSuccess) { return fail("hipStreamCreate", e); }
    e = hipMalloc((void**)&g_consts, sizeof(TcgConsts));
    if (e != hipSuccess) { return fail("hipMalloc consts", e); }
    for (int i = 0; i < 8; ++i) { (void)hipEventCreate(&g_ev[i]); }
    g_ev_ready = 1;
    return 0;
}

const char* tcg_last_error(void) { return g_err; }

int32_t tcg_malloc(uint64_t* dptr, size_t bytes)
{
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes ? bytes : 1u);
    if (e != hipSuccess) { (void)hipGetLastError(); return fail("hipMalloc", e); }
    *dptr = (uint64_t)(uintptr_t)p;
    return 0;
}

int32_t tcg_free(uint64_t dptr)
{
    if (dptr == 0u) { return 0; }
    const hipError_t e = hipFree((void*)(uintptr_t)dptr);
    return e == hipSuccess ? 0 : fail("hipFree", e);
}

int32_t tcg_host_alloc(void** hptr, size_t bytes)
{
    const hipError_t e = hipHostMalloc(hptr, bytes ? bytes : 1u, hipHostMallocDefault);
    return e == hipSuccess ? 0 : fail("hipHostMalloc", e);
}

int32_t tcg_host_free(void* hptr)
{
    if (hptr == nullptr) { return 0; }
    const hipError_t e = hipHostFree(hptr);
    return e == hipSuccess ? 0 : fail("hipHostFree", e);
}

int32_t tcg_h2d(uint64_t dst, const void* src, size_t bytes)
{
    if (bytes == 0u) { return 0; }
    const hipError_t e = hipMemcpyAsync((void*)(uintptr_t)dst, src, bytes, hipMemcpyHostToDevice, g_stream);
    return e == hipSuccess ? 0 : fail("hipMemcpy H2D", e);
}

int32_t tcg_d2h(void* dst, uint64_t src, size_t bytes)
{
    if (bytes == 0u) { return 0; }
    hipError_t e = hipMemcpyAsync(dst, (const void*)(uintptr_t)src, bytes, hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) { e = hipStreamSynchronize(g_stream); }
    return e == hipSuccess ? 0 : fail("hipMemcpy D2H", e);
}

int32_t tcg_h2d_2d(uint64_t dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows)
{
    if (width == 0u || rows == 0u) { return 0; }
    const hipError_t e = hipMemcpy2DAsync((void*)(uintptr_t)dst, dpitch, src, spitch, width, rows,
                                          hipMemcpyHostToDevice, g_stream);
    return e == hipSuccess ? 0 : fail("hipMemcpy2D H2D", e);
}

int32_t tcg_d2h_2d(void* dst, size_t dpitch, uint64_t src, size_t spitch, size_t width, size_t rows)
{
    if (width == 0u || rows == 0u) { return 0; }
    hipError_t e = hipMemcpy2DAsync(dst, dpitch, (const void*)(uintptr_t)src, spitch, width, rows,
                                    hipMemcpyDeviceToHost, g_stream);
    if (e == hipSuccess) { e = hipStreamSynchronize(g_stream); }
    return e == hipSuccess ? 0 : fail("hipMemcpy2D D2H", e);
}

int32_t tcg_sync(void)
{
    const hipError_t e = hipStreamSynchronize(g_stream);
    return e == hipSuccess ? 0 : fail("hipStreamSynchronize", e);
}

int32_t tcg_set_tables(const double* cost4097, const uint32_t* crc256, const uint32_t* x2n32)
{
    TcgConsts h;
    memcpy(h.cost, cost4097, sizeof(h.cost));
    memcpy(h.crc, crc256, sizeof(h.crc));
    memcpy(h.x2n, x2n32, sizeof(h.x2n));
    hipError_t e = hipMemcpyAsync(g_consts, &h, sizeof(h), hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) { e = hipStreamSynchronize(g_stream); }
    return e == hipSuccess ? 0 : fail("upload tables", e);
}

void tcg_set_wide(int wide) { g_wide = wide; }

void tcg_set_timing(int on) { g_timing = on; }

void tcg_get_timing(tcg_timing* t)
{
    memset(t, 0, sizeof(*t));
    if (!g_timing || !g_ev_ready) { return; }
    (void)hipStreamSynchronize(g_stream);
    float ms[4] = { 0, 0, 0, 0 };
    const int nmarks = g_last_kind == 1 ? 4 : 3;
    for (int i = 0; i < nmarks; ++i) {
        if (hipEventElapsedTime(&ms[i], g_ev[i], g_ev[i + 1]) != hipSuccess) { ms[i] = -1.0f; }
    }
    if (g_last_kind == 1) {
        t->fdct_ms = ms[0]; t->slice_enc_ms = ms[1]; t->scan_ms = ms[2]; t->emit_ms = ms[3];
    } else if (g_last_kind == 2) {
        t->slice_dec_ms = ms[0]; t->idct_ms = ms[1]; t->deblock_ms = ms[2];
    }
}

}  // extern "C"
