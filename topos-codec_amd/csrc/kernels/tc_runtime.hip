// HIP runtime services behind the internal C ABI: device initialisation and
// the uploaded constant tables (process-wide), contexts (one default-priority
// stream each, with timing events), device memory, copies and per-kernel
// timing with HIP events on the launch stream.
#include <hip/hip_runtime.h>
#include <pthread.h>

#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tcg.h"
#include "tc_math.h"
#include "tc_dev.h"

namespace {
int g_ready = 0;
TcgConsts* g_consts = nullptr;
int g_timing = 0;
uint32_t g_cus = 0;
uint32_t g_clock_khz = 0;
int g_graphs = 1;
thread_local char g_err[256];

int32_t fail(const char* what, hipError_t e)
{
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    return -2;
}
}  // namespace

const TcgConsts* tcg_consts_dev() { return g_consts; }

int32_t tcg_check(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { return fail(what, e); }
    return 0;
}

int tcg_use_graphs() { return g_graphs && !g_timing; }

// Device memory of the no-allocation decode entry point: coefficients,
// descriptors and slice status of its launches live here instead of in
// buffers the library allocates. Every GPU the library runs on holds its own
// copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives per device (the array's name is a host-side
// stand-in, not a device address).
static __device__ __align__(256) uint8_t scratch[TCG_NOALLOC_BYTES];
static uint64_t g_scratch_ptr[TCG_MAX_DEVICES];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

static int32_t scratch_setup_current_device(uint64_t* out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { return fail("hipGetDevice", e); }
    if (dev < 0 || dev >= TCG_MAX_DEVICES) {
        snprintf(g_err, sizeof(g_err), "device %d beyond the %d the no-allocation scratch covers", dev,
                 TCG_MAX_DEVICES);
        return -8;
    }
    pthread_mutex_lock(&g_scratch_mu);
    if (g_scratch_ptr[dev] == 0u) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            pthread_mutex_unlock(&g_scratch_mu);
            return fail("hipGetSymbolAddress(scratch)", e);
        }
        g_scratch_ptr[dev] = (uint64_t)(uintptr_t)p;
    }
    *out = g_scratch_ptr[dev];
    pthread_mutex_unlock(&g_scratch_mu);
    return 0;
}

uint32_t tcg_cu_count() { return g_cus; }
uint32_t tcg_clock_khz() { return g_clock_khz; }

hipGraphExec_t tcg_graph_find(tcg_ctx* c, const uint64_t* key)
{
    for (int i = 0; i < c->ngraphs; ++i) {
        if (memcmp(c->graphs[i].key, key, sizeof(c->graphs[i].key)) == 0) {
            c->graphs[i].used = ++c->clock;
            return c->graphs[i].exec;
        }
    }
    return nullptr;
}

int32_t tcg_graph_store(tcg_ctx* c, const uint64_t* key, hipGraph_t g, hipGraphExec_t* out)
{
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { return fail("hipGraphInstantiate", e); }
    int slot = c->ngraphs;
    if (slot == TCG_GRAPHS_PER_CTX) {
        slot = 0;
        for (int i = 1; i < c->ngraphs; ++i) {
            if (c->graphs[i].used < c->graphs[slot].used) { slot = i; }
        }
        (void)hipGraphExecDestroy(c->graphs[slot].exec);
    } else {
        c->ngraphs++;
    }
    memcpy(c->graphs[slot].key, key, sizeof(c->graphs[slot].key));
    c->graphs[slot].exec = exec;
    c->graphs[slot].used = ++c->clock;
    *out = exec;
    return 0;
}

void tcg_mark(tcg_ctx* c, int i)
{
    if (i == 0) { c->nmarks = 0; c->pairs_valid = 0; }
    if (g_timing && i < 8) {
        (void)hipEventRecord(c->ev[i], c->stream);
        c->nmarks = i + 1;
    }
}

unsigned long long* tcg_pairs_counter(tcg_ctx* c)
{
    // only sequences launched eagerly with timing events count (never under
    // graph capture, never on the no-allocation context)
    if (c->nmarks == 0 || c->eager) { return nullptr; }
    if (c->pairs_dev == nullptr) {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(unsigned long long)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        c->pairs_dev = (unsigned long long*)p;
    }
    if (hipMemsetAsync(c->pairs_dev, 0, sizeof(unsigned long long), c->stream) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    c->pairs_valid = 1;
    return c->pairs_dev;
}

// plain 16-byte-vector device copy: the reference point for HBM byte counters
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) { dst[i] = src[i]; }
}

extern "C" {

int32_t tcg_init(void)
{
    if (g_ready) { return 0; }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        snprintf(g_err, sizeof(g_err), "no HIP device available (%s)",
                 e != hipSuccess ? hipGetErrorString(e) : "count 0");
        return -16;
    }
    const char* gv = getenv("TOPOS_HIP_GRAPHS");
    if (gv != nullptr && gv[0] == '0') { g_graphs = 0; }
    e = hipSetDevice(0);
    if (e != hipSuccess) { return fail("hipSetDevice", e); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.multiProcessorCount > 0) {
        g_cus = (uint32_t)prop.multiProcessorCount;
        g_clock_khz = prop.clockRate > 0 ? (uint32_t)prop.clockRate : 0u;
    } else {
        g_cus = 1u;
    }
    e = hipMalloc((void**)&g_consts, sizeof(TcgConsts));
    if (e != hipSuccess) { return fail("hipMalloc consts", e); }
    uint64_t sp = 0u;
    if (scratch_setup_current_device(&sp) != 0) { return -2; }
    g_ready = 1;
    return 0;
}

int32_t tcg_noalloc_scratch(uint64_t* ptr)
{
    return scratch_setup_current_device(ptr);
}

void tcg_ctx_set_eager(tcg_ctx* c, int eager) { c->eager = eager; }

const char* tcg_last_error(void) { return g_err; }

int32_t tcg_ctx_create(tcg_ctx** out)
{
    tcg_ctx* c = (tcg_ctx*)calloc(1, sizeof(tcg_ctx));
    if (c == nullptr) {
        snprintf(g_err, sizeof(g_err), "context allocation failed");
        return -2;
    }
    // default priority, non-blocking with respect to the null stream
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { free(c); return fail("hipStreamCreate", e); }
    for (int i = 0; i < 8; ++i) {
        e = hipEventCreate(&c->ev[i]);
        if (e != hipSuccess) {
            for (int j = 0; j < i; ++j) { (void)hipEventDestroy(c->ev[j]); }
            (void)hipStreamDestroy(c->stream);
            free(c);
            return fail("hipEventCreate", e);
        }
    }
    *out = c;
    return 0;
}

void tcg_ctx_destroy(tcg_ctx* c)
{
    if (c == nullptr) { return; }
    (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < c->ngraphs; ++i) { (void)hipGraphExecDestroy(c->graphs[i].exec); }
    if (c->pin != nullptr) { (void)hipHostFree(c->pin); }
    if (c->pairs_dev != nullptr) { (void)hipFree(c->pairs_dev); }
    for (int i = 0; i < 8; ++i) { (void)hipEventDestroy(c->ev[i]); }
    (void)hipStreamDestroy(c->stream);
    free(c);
}

int32_t tcg_copy(tcg_ctx* c, uint64_t dst, uint64_t src, size_t bytes)
{
    const size_t n16 = bytes / 16u;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, c->stream, (const uint4*)(uintptr_t)src,
                       (uint4*)(uintptr_t)dst, n16);
    const hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? 0 : fail("copy", e);
}

int32_t tcg_d2d(tcg_ctx* c, uint64_t dst, uint64_t src, size_t bytes)
{
    const hipError_t e = hipMemcpyAsync((void*)(uintptr_t)dst, (const void*)(uintptr_t)src, bytes,
                                        hipMemcpyDeviceToDevice, c->stream);
    return e == hipSuccess ? 0 : fail("device copy", e);
}

void* tcg_pinned_alloc(size_t bytes)
{
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

void tcg_pinned_free(void* p)
{
    if (p != nullptr) { (void)hipHostFree(p); }
}

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

int32_t tcg_h2d(tcg_ctx* c, uint64_t dst, const void* src, size_t bytes)
{
    if (bytes == 0u) { return 0; }
    const hipError_t e = hipMemcpyAsync((void*)(uintptr_t)dst, src, bytes, hipMemcpyHostToDevice, c->stream);
    return e == hipSuccess ? 0 : fail("hipMemcpy H2D", e);
}

int32_t tcg_d2h(tcg_ctx* c, void* dst, uint64_t src, size_t bytes)
{
    if (bytes == 0u) { return 0; }
    const hipError_t e = hipMemcpyAsync(dst, (const void*)(uintptr_t)src, bytes, hipMemcpyDeviceToHost, c->stream);
    return e == hipSuccess ? 0 : fail("hipMemcpy D2H", e);
}

int32_t tcg_d2h_sync(tcg_ctx* c, void* dst, uint64_t src, size_t bytes)
{
    if (bytes == 0u) { return 0; }
    hipError_t e = hipMemcpyAsync(dst, (const void*)(uintptr_t)src, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) { e = hipStreamSynchronize(c->stream); }
    return e == hipSuccess ? 0 : fail("hipMemcpy D2H", e);
}

// Host copies: rows that are tight on both sides go as one plain copy (the
// runtime streams pageable memory at full PCIe rate); strided ones go through
// the context's pinned staging buffer: the rows are packed on the CPU and
// cross PCIe as one DMA (a strided copy of pageable memory is row by row).
// The stream is drained before the buffer is refilled, so the previous copy
// out of it has finished.
static int pin_need(tcg_ctx* c, size_t bytes)
{
    if (c->pin_cap >= bytes) { return 1; }
    if (c->pin != nullptr) { (void)hipHostFree(c->pin); }
    c->pin = nullptr;
    c->pin_cap = 0;
    const size_t cap = (bytes + ((size_t)1 << 20) - 1u) & ~(((size_t)1 << 20) - 1u);
    if (hipHostMalloc(&c->pin, cap, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->pin = nullptr;
        return 0;
    }
    c->pin_cap = cap;
    return 1;
}

int32_t tcg_h2d_2d(tcg_ctx* c, uint64_t dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                   size_t rows)
{
    if (width == 0u || rows == 0u) { return 0; }
    if (spitch == width && dpitch == width) {
        const hipError_t e1 = hipMemcpyAsync((void*)(uintptr_t)dst, src, width * rows, hipMemcpyHostToDevice,
                                             c->stream);
        return e1 == hipSuccess ? 0 : fail("hipMemcpy H2D", e1);
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && pin_need(c, width * rows)) {
        uint8_t* p = (uint8_t*)c->pin;
        for (size_t r = 0; r < rows; ++r) { memcpy(p + r * width, (const uint8_t*)src + r * spitch, width); }
        e = dpitch == width ? hipMemcpyAsync((void*)(uintptr_t)dst, p, width * rows, hipMemcpyHostToDevice, c->stream)
                            : hipMemcpy2DAsync((void*)(uintptr_t)dst, dpitch, p, width, width, rows,
                                               hipMemcpyHostToDevice, c->stream);
    } else if (e == hipSuccess) {
        e = hipMemcpy2DAsync((void*)(uintptr_t)dst, dpitch, src, spitch, width, rows, hipMemcpyHostToDevice,
                             c->stream);
    }
    return e == hipSuccess ? 0 : fail("hipMemcpy2D H2D", e);
}

int32_t tcg_d2h_2d(tcg_ctx* c, void* dst, size_t dpitch, uint64_t src, size_t spitch, size_t width, size_t rows)
{
    if (width == 0u || rows == 0u) { return 0; }
    if (spitch == width && dpitch == width) {
        hipError_t e1 = hipMemcpyAsync(dst, (const void*)(uintptr_t)src, width * rows, hipMemcpyDeviceToHost,
                                       c->stream);
        if (e1 == hipSuccess) { e1 = hipStreamSynchronize(c->stream); }
        return e1 == hipSuccess ? 0 : fail("hipMemcpy D2H", e1);
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && pin_need(c, width * rows)) {
        uint8_t* p = (uint8_t*)c->pin;
        e = spitch == width ? hipMemcpyAsync(p, (const void*)(uintptr_t)src, width * rows, hipMemcpyDeviceToHost,
                                             c->stream)
                            : hipMemcpy2DAsync(p, width, (const void*)(uintptr_t)src, spitch, width, rows,
                                               hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) { e = hipStreamSynchronize(c->stream); }
        if (e == hipSuccess) {
            if (dpitch == width) {
                memcpy(dst, p, width * rows);
            } else {
                for (size_t r = 0; r < rows; ++r) { memcpy((uint8_t*)dst + r * dpitch, p + r * width, width); }
            }
        }
    } else if (e == hipSuccess) {
        e = hipMemcpy2DAsync(dst, dpitch, (const void*)(uintptr_t)src, spitch, width, rows, hipMemcpyDeviceToHost,
                             c->stream);
        if (e == hipSuccess) { e = hipStreamSynchronize(c->stream); }
    }
    return e == hipSuccess ? 0 : fail("hipMemcpy2D D2H", e);
}

int32_t tcg_sync(tcg_ctx* c)
{
    const hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? 0 : fail("hipStreamSynchronize", e);
}

int32_t tcg_query(tcg_ctx* c)
{
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) { return 0; }
    if (e == hipErrorNotReady) { return 1; }
    return fail("hipStreamQuery", e);
}

int32_t tcg_set_tables(const double* cost4097, const uint32_t* crc256, const uint32_t* x2n32)
{
    TcgConsts h;
    memcpy(h.cost, cost4097, sizeof(h.cost));
    memcpy(h.crc, crc256, sizeof(h.crc));
    memcpy(h.x2n, x2n32, sizeof(h.x2n));
    for (uint32_t j = 0; j < 256u; ++j) { h.x16[j] = tc_gf2_xpow8n(x2n32, 16u * (uint64_t)j); }
    const hipError_t e = hipMemcpy(g_consts, &h, sizeof(h), hipMemcpyHostToDevice);
    return e == hipSuccess ? 0 : fail("upload tables", e);
}

int32_t tcg_set_vlc(const tcg_vlc_books* books)
{
    const hipError_t e = hipMemcpy((char*)g_consts + offsetof(TcgConsts, vlc), books, sizeof(*books),
                                   hipMemcpyHostToDevice);
    return e == hipSuccess ? 0 : fail("upload VLC books", e);
}

void tcg_set_timing(int on) { g_timing = on; }

void tcg_device_info(uint32_t* cus, uint32_t* clock_khz)
{
    *cus = g_ready ? g_cus : 0u;
    *clock_khz = g_ready ? g_clock_khz : 0u;
}

void tcg_set_graphs(int on) { g_graphs = on; }

void tcg_graph_stats(tcg_ctx* c, uint64_t* hits, uint64_t* builds)
{
    *hits = c->graph_hits;
    *builds = c->graph_builds;
}

int tcg_timing_on(void) { return g_timing; }

void tcg_ctx_timing(tcg_ctx* c, tcg_timing* t)
{
    memset(t, 0, sizeof(*t));
    const int n = c->nmarks;
    if (n < 2) { return; }
    float ms[7] = { 0, 0, 0, 0, 0, 0, 0 };
    for (int i = 0; i + 1 < n && i < 7; ++i) {
        if (hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]) != hipSuccess) { ms[i] = -1.0f; }
    }
    if (c->kind == 1) {
        t->fdct_ms = ms[0]; t->slice_enc_ms = ms[1]; t->scan_ms = ms[2]; t->emit_ms = ms[3];
    } else if (c->kind == 2) {
        t->slice_dec_ms = ms[0]; t->idct_ms = ms[1]; t->deblock_ms = ms[2];
    } else if (c->kind == 3) {
        t->convert_ms = ms[0];
    }
    if (c->pairs_valid) {
        unsigned long long n = 0;
        if (hipMemcpyAsync(&n, c->pairs_dev, sizeof(n), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
            hipStreamSynchronize(c->stream) == hipSuccess) {
            t->pairs = n;
        } else {
            (void)hipGetLastError();
        }
    }
}

}  // extern "C"

extern "C" uint32_t tcg_entropy_capacity(int encode, uint32_t wpg, uint32_t max_cols)
{
    return encode ? tcg_entropy_capacity_enc(wpg, max_cols) : tcg_entropy_capacity_dec(wpg, max_cols);
}
