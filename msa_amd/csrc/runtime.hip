// C-ABI runtime layer of libmsa_hip.so: status codes, device selection and
// properties, device/pinned memory, copies, streams, events and the named
// scoring tables.  It replaces the reference's CUDA wrapper on the pairwise
// path (src/cuda.cu:42-143, src/cuda.cuh:146-458) and the table registry of
// src/pairwise/table.cu:254-330.  Unlike the reference, whose production
// build compiles cuda::check out (src/cuda.cuh:217-221), every HIP call here
// is checked and reported as a status code.
#include "msa_internal.hpp"

#include <new>

#include <cstdio>
#include <cstring>
#include <string>

#include "scoring_tables.inc"

namespace msa {

namespace {
thread_local std::string g_last_error;
}

int fail(int status, const char* fmt, ...)
{
    char buf[512];
    va_list args;
    va_start(args, fmt);
    vsnprintf(buf, sizeof(buf), fmt, args);
    va_end(args);
    g_last_error = buf;
    return status;
}

int hip_status(hipError_t err, const char* what)
{
    if (err == hipSuccess) return MSA_OK;
    const int status = (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation)
        ? MSA_ERR_OUT_OF_MEMORY
        : (err == hipErrorNoDevice ? MSA_ERR_NO_DEVICE : MSA_ERR_RUNTIME);
    (void)hipGetLastError();
    return fail(status, "%s: %s (%d)", what, hipGetErrorString(err), (int)err);
}

int table_lookup(const char* name, int32_t* table625, int32_t* penalty)
{
    if (!name) return fail(MSA_ERR_INVALID_ARGUMENT, "table name is null");
    for (int t = 0; t < kTableCount; ++t) {
        if (std::strcmp(kTableNames[t], name) != 0) continue;
        if (table625)
            for (int i = 0; i < 625; ++i) table625[i] = kTableData[t][i / 25][i % 25];
        if (penalty) *penalty = kTablePenalty[t];
        return MSA_OK;
    }
    return fail(MSA_ERR_NOT_FOUND, "unknown pairwise scoring table '%s'", name);
}

}  // namespace msa

using msa::fail;
using msa::hip_status;

extern "C" {

const char* msa_error_string(int status)
{
    switch (status) {
    case MSA_OK: return "no error";
    case MSA_ERR_INVALID_ARGUMENT: return "invalid argument";
    case MSA_ERR_NO_DEVICE: return "no compatible device";
    case MSA_ERR_OUT_OF_MEMORY: return "out of device memory";
    case MSA_ERR_RUNTIME: return "HIP runtime error";
    case MSA_ERR_GAP_CODE: return "sequence holds the gap code, which has no score";
    case MSA_ERR_TABLE: return "scoring table outside the supported range";
    case MSA_ERR_TOO_LARGE: return "input exceeds a supported maximum";
    case MSA_ERR_BUSY: return "server busy";
    case MSA_ERR_NOT_READY: return "job not finished";
    case MSA_ERR_NOT_FOUND: return "not found";
    default: return "unknown status";
    }
}

const char* msa_last_error(void) { return msa::g_last_error.c_str(); }

int msa_version(void) { return 10000; }

int msa_device_count(int* count)
{
    if (!count) return fail(MSA_ERR_INVALID_ARGUMENT, "count is null");
    *count = 0;
    const hipError_t err = hipGetDeviceCount(count);
    if (err == hipErrorNoDevice || err == hipErrorInvalidDevice) { (void)hipGetLastError(); *count = 0; return MSA_OK; }
    return hip_status(err, "hipGetDeviceCount");
}

int msa_device_select(int device) { return hip_status(hipSetDevice(device), "hipSetDevice"); }

int msa_device_current(int* device)
{
    if (!device) return fail(MSA_ERR_INVALID_ARGUMENT, "device is null");
    return hip_status(hipGetDevice(device), "hipGetDevice");
}

int msa_device_props_get(int device, msa_device_props* props)
{
    if (!props) return fail(MSA_ERR_INVALID_ARGUMENT, "props is null");
    hipDeviceProp_t p;
    int rc = hip_status(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
    if (rc) return rc;
    std::memset(props, 0, sizeof(*props));
    std::snprintf(props->name, sizeof(props->name), "%s", p.name);
    std::snprintf(props->arch, sizeof(props->arch), "%s", p.gcnArchName);
    props->compute_units = p.multiProcessorCount;
    props->wavefront_size = p.warpSize;
    props->clock_khz = p.clockRate;
    props->max_threads_per_block = p.maxThreadsPerBlock;
    props->lds_per_block = p.sharedMemPerBlock;
    props->lds_per_cu = p.maxSharedMemoryPerMultiProcessor;
    props->global_mem = p.totalGlobalMem;
    return MSA_OK;
}

int msa_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    size_t f = 0, t = 0;
    const int rc = hip_status(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return rc;
}

int msa_device_sync(void) { return hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int msa_malloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(MSA_ERR_INVALID_ARGUMENT, "ptr is null");
    *ptr = nullptr;
    return hip_status(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
}

int msa_free(void* ptr) { return ptr ? hip_status(hipFree(ptr), "hipFree") : MSA_OK; }

int msa_malloc_host(void** ptr, size_t bytes)
{
    if (!ptr) return fail(MSA_ERR_INVALID_ARGUMENT, "ptr is null");
    *ptr = nullptr;
    return hip_status(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
}

int msa_free_host(void* ptr) { return ptr ? hip_status(hipHostFree(ptr), "hipHostFree") : MSA_OK; }

int msa_memcpy(void* dst, const void* src, size_t bytes)
{
    if (!bytes) return MSA_OK;
    if (!dst || !src) return fail(MSA_ERR_INVALID_ARGUMENT, "null pointer in copy");
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "hipMemcpy");
}

int msa_memcpy_async(void* dst, const void* src, size_t bytes, msa_stream stream)
{
    if (!bytes) return MSA_OK;
    if (!dst || !src) return fail(MSA_ERR_INVALID_ARGUMENT, "null pointer in copy");
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream), "hipMemcpyAsync");
}

int msa_memset(void* ptr, int value, size_t bytes)
{
    if (!bytes) return MSA_OK;
    if (!ptr) return fail(MSA_ERR_INVALID_ARGUMENT, "ptr is null");
    return hip_status(hipMemset(ptr, value, bytes), "hipMemset");
}

int msa_stream_create(msa_stream* stream)
{
    if (!stream) return fail(MSA_ERR_INVALID_ARGUMENT, "stream is null");
    hipStream_t s = nullptr;
    const int rc = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = (msa_stream)s;
    return rc;
}

int msa_stream_destroy(msa_stream stream)
{
    return stream ? hip_status(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy") : MSA_OK;
}

int msa_stream_sync(msa_stream stream)
{
    return hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int msa_event_create(msa_event* event)
{
    if (!event) return fail(MSA_ERR_INVALID_ARGUMENT, "event is null");
    hipEvent_t e = nullptr;
    const int rc = hip_status(hipEventCreate(&e), "hipEventCreate");
    *event = (msa_event)e;
    return rc;
}

int msa_event_destroy(msa_event event)
{
    return event ? hip_status(hipEventDestroy((hipEvent_t)event), "hipEventDestroy") : MSA_OK;
}

int msa_event_record(msa_event event, msa_stream stream)
{
    return hip_status(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}

int msa_event_sync(msa_event event)
{
    return hip_status(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}

int msa_event_elapsed_ms(msa_event start, msa_event stop, float* ms)
{
    if (!ms) return fail(MSA_ERR_INVALID_ARGUMENT, "ms is null");
    return hip_status(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}

struct msa_graph_opaque {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int msa_graph_capture_begin(msa_stream stream)
{
    if (!stream) return fail(MSA_ERR_INVALID_ARGUMENT, "graphs are captured on a created stream, not the null stream");
    return hip_status(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
                      "hipStreamBeginCapture");
}

int msa_graph_capture_end(msa_stream stream, msa_graph* graph)
{
    if (!stream) return fail(MSA_ERR_INVALID_ARGUMENT, "graphs are captured on a created stream, not the null stream");
    if (!graph) return fail(MSA_ERR_INVALID_ARGUMENT, "graph is null");
    *graph = nullptr;
    hipGraph_t g = nullptr;
    MSA_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    if (!g) return fail(MSA_ERR_RUNTIME, "stream capture produced no graph");
    auto* out = new (std::nothrow) msa_graph_opaque();
    if (!out) {
        (void)hipGraphDestroy(g);
        return fail(MSA_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    out->graph = g;
    const int rc = hip_status(hipGraphInstantiate(&out->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc != MSA_OK) {
        (void)hipGraphDestroy(g);
        delete out;
        return rc;
    }
    *graph = out;
    return MSA_OK;
}

int msa_graph_launch(msa_graph graph, msa_stream stream)
{
    if (!graph) return fail(MSA_ERR_INVALID_ARGUMENT, "graph is null");
    return hip_status(hipGraphLaunch(graph->exec, (hipStream_t)stream), "hipGraphLaunch");
}

int msa_graph_destroy(msa_graph graph)
{
    if (!graph) return MSA_OK;
    int rc = MSA_OK;
    if (graph->exec) rc = hip_status(hipGraphExecDestroy(graph->exec), "hipGraphExecDestroy");
    if (graph->graph) {
        const int r2 = hip_status(hipGraphDestroy(graph->graph), "hipGraphDestroy");
        if (rc == MSA_OK) rc = r2;
    }
    delete graph;
    return rc;
}

int msa_table_count(void) { return kTableCount; }

const char* msa_table_name(int index)
{
    return (index >= 0 && index < kTableCount) ? kTableNames[index] : nullptr;
}

int msa_table_get(const char* name, int32_t* table625, int32_t* penalty)
{
    return msa::table_lookup(name, table625, penalty);
}

}  // extern "C"
