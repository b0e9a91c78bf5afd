// Device, memory, stream, event and graph entry points of libpgpe_hip.so.
#include <cstdio>
#include <cstring>
#include <string>

#include "common.hpp"

namespace pgpe {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t err, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorName(err) + " (" +
                   hipGetErrorString(err) + ")";
    return PGPE_ERR_HIP;
}

// Byte-wise zero fill as a kernel so it can be recorded into a graph
// (hipMemsetAsync nodes are avoided on purpose).
__global__ void fill_zero_kernel(uint4* __restrict__ p16, size_t n16, unsigned char* __restrict__ tail,
                                 size_t ntail) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = i; j < n16; j += stride) p16[j] = make_uint4(0u, 0u, 0u, 0u);
    if (i < ntail) tail[i] = 0;
}

}  // namespace pgpe

using namespace pgpe;

extern "C" {

const char* pgpe_last_error(void) { return g_last_error.c_str(); }

int pgpe_version(void) { return 10000; }

int pgpe_device_count(int* count) {
    if (!count) return fail(PGPE_ERR_ARG, "pgpe_device_count: null output");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return PGPE_OK;
}

int pgpe_device_init(int device) {
    PGPE_HIP(hipSetDevice(device));
    PGPE_HIP(hipFree(nullptr));
    return reduce_scratch_setup(nullptr);
}

int pgpe_device_name(int device, char* buf, size_t len) {
    if (!buf || len == 0) return fail(PGPE_ERR_ARG, "pgpe_device_name: empty buffer");
    PGPE_HIP(hipDeviceGetName(buf, (int)(len > 0x7fffffff ? 0x7fffffff : len), device));
    buf[len - 1] = '\0';
    return PGPE_OK;
}

int pgpe_device_cu_count(int device, int* cus) {
    if (!cus) return fail(PGPE_ERR_ARG, "pgpe_device_cu_count: null output");
    PGPE_HIP(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, device));
    return PGPE_OK;
}

int pgpe_malloc(void** ptr, size_t nbytes) {
    if (!ptr) return fail(PGPE_ERR_ARG, "pgpe_malloc: null output");
    *ptr = nullptr;
    if (nbytes == 0) nbytes = 256;
    PGPE_HIP(hipMalloc(ptr, nbytes));
    return PGPE_OK;
}

int pgpe_free(void* ptr) {
    if (!ptr) return PGPE_OK;
    PGPE_HIP(hipFree(ptr));
    return PGPE_OK;
}

int pgpe_memcpy_h2d(void* dst, const void* src, size_t nbytes, void* stream) {
    if (nbytes == 0) return PGPE_OK;
    if (!dst || !src) return fail(PGPE_ERR_ARG, "pgpe_memcpy_h2d: null pointer");
    PGPE_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    PGPE_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PGPE_OK;
}

int pgpe_memcpy_d2h(void* dst, const void* src, size_t nbytes, void* stream) {
    if (nbytes == 0) return PGPE_OK;
    if (!dst || !src) return fail(PGPE_ERR_ARG, "pgpe_memcpy_d2h: null pointer");
    PGPE_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    PGPE_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PGPE_OK;
}

int pgpe_memcpy_d2d(void* dst, const void* src, size_t nbytes, void* stream) {
    if (nbytes == 0) return PGPE_OK;
    if (!dst || !src) return fail(PGPE_ERR_ARG, "pgpe_memcpy_d2d: null pointer");
    PGPE_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return PGPE_OK;
}

int pgpe_host_alloc(void** ptr, size_t nbytes) {
    if (!ptr) return fail(PGPE_ERR_ARG, "pgpe_host_alloc: null output");
    *ptr = nullptr;
    PGPE_HIP(hipHostMalloc(ptr, nbytes == 0 ? 64 : nbytes, hipHostMallocDefault));
    return PGPE_OK;
}

int pgpe_host_free(void* ptr) {
    if (!ptr) return PGPE_OK;
    PGPE_HIP(hipHostFree(ptr));
    return PGPE_OK;
}

int pgpe_memcpy_d2h_async(void* dst, const void* src, size_t nbytes, void* stream) {
    if (nbytes == 0) return PGPE_OK;
    if (!dst || !src) return fail(PGPE_ERR_ARG, "pgpe_memcpy_d2h_async: null pointer");
    PGPE_HIP(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    return PGPE_OK;
}

int pgpe_fill_zero(void* ptr, size_t nbytes, void* stream) {
    if (nbytes == 0) return PGPE_OK;
    if (!ptr) return fail(PGPE_ERR_ARG, "pgpe_fill_zero: null pointer");
    if (((uintptr_t)ptr & 15u) != 0) return fail(PGPE_ERR_ARG, "pgpe_fill_zero: pointer not 16-byte aligned");
    const size_t n16 = nbytes / 16;
    const size_t ntail = nbytes - n16 * 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (uint4*)ptr, n16, (unsigned char*)ptr + n16 * 16, ntail);
    PGPE_LAUNCHED("fill_zero_kernel");
    return PGPE_OK;
}

int pgpe_stream_create(void** stream) {
    if (!stream) return fail(PGPE_ERR_ARG, "pgpe_stream_create: null output");
    hipStream_t s;
    // Default priority on purpose; only the non-blocking flag is set.
    PGPE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return PGPE_OK;
}

int pgpe_stream_destroy(void* stream) {
    if (!stream) return PGPE_OK;
    PGPE_HIP(hipStreamDestroy((hipStream_t)stream));
    return PGPE_OK;
}

int pgpe_stream_sync(void* stream) {
    PGPE_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PGPE_OK;
}

int pgpe_device_sync(void) {
    PGPE_HIP(hipDeviceSynchronize());
    return PGPE_OK;
}

int pgpe_event_create(void** event) {
    if (!event) return fail(PGPE_ERR_ARG, "pgpe_event_create: null output");
    hipEvent_t e;
    PGPE_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return PGPE_OK;
}

int pgpe_event_destroy(void* event) {
    if (!event) return PGPE_OK;
    PGPE_HIP(hipEventDestroy((hipEvent_t)event));
    return PGPE_OK;
}

int pgpe_event_record(void* event, void* stream) {
    if (!event) return fail(PGPE_ERR_ARG, "pgpe_event_record: null event");
    PGPE_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return PGPE_OK;
}

int pgpe_stream_wait_event(void* stream, void* event) {
    if (!event) return fail(PGPE_ERR_ARG, "pgpe_stream_wait_event: null event");
    PGPE_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
    return PGPE_OK;
}

int pgpe_event_elapsed_ms(void* start, void* stop, float* ms) {
    if (!start || !stop || !ms) return fail(PGPE_ERR_ARG, "pgpe_event_elapsed_ms: null argument");
    PGPE_HIP(hipEventSynchronize((hipEvent_t)stop));
    PGPE_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return PGPE_OK;
}

int pgpe_event_sync(void* event) {
    if (!event) return fail(PGPE_ERR_ARG, "pgpe_event_sync: null event");
    PGPE_HIP(hipEventSynchronize((hipEvent_t)event));
    return PGPE_OK;
}

int pgpe_graph_begin(void* stream) {
    if (!stream) return fail(PGPE_ERR_ARG, "pgpe_graph_begin: capture needs a non-default stream");
    PGPE_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return PGPE_OK;
}

int pgpe_graph_end(void* stream, void** graph_exec) {
    if (!stream || !graph_exec) return fail(PGPE_ERR_ARG, "pgpe_graph_end: null argument");
    hipGraph_t graph = nullptr;
    PGPE_HIP(hipStreamEndCapture((hipStream_t)stream, &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph_exec = (void*)exec;
    return PGPE_OK;
}

int pgpe_graph_launch(void* graph_exec, void* stream) {
    if (!graph_exec) return fail(PGPE_ERR_ARG, "pgpe_graph_launch: null graph");
    PGPE_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return PGPE_OK;
}

int pgpe_graph_destroy(void* graph_exec) {
    if (!graph_exec) return PGPE_OK;
    PGPE_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return PGPE_OK;
}

}  // extern "C"
