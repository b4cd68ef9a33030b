// handle.hip -- the rppHandle_t lifecycle of api/rpp.h (src/modules/handle_api.cpp:29-83).
//
// A handle owns: the stream all work is enqueued on, a grow-only device workspace, a small
// pinned staging buffer and the per-parameter caches of GPU-built tables.  Unlike the
// reference (handle_hip.cpp:170-235, 1.49 GB device + 33 MB pinned up front), nothing is
// allocated until an entry point needs it.
#include "common.hpp"

#include <algorithm>
#include <thread>

namespace rppa {

namespace {
void *grow(rppHandle *h, void *&buf, size_t &cap, size_t bytes)
{
    if (bytes <= cap) return buf;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
        fprintf(stderr, "rpp_audio: workspace must grow to %zu bytes while the stream is being captured; "
                        "run the call once before capture\n", bytes);
        return nullptr;
    }
    if (buf) {
        if (hipStreamSynchronize(h->stream) != hipSuccess) return nullptr;
        (void)hipFree(buf);
        buf = nullptr;
        cap = 0;
    }
    size_t want = bytes + bytes / 4;
    if (hipMalloc(&buf, want) != hipSuccess) return nullptr;
    cap = want;
    return buf;
}
}  // namespace

void *workspace(rppHandle *h, size_t bytes) { return grow(h, h->ws, h->wsBytes, bytes); }

void *front_workspace(rppHandle *h, size_t bytes) { return grow(h, h->feWs, h->feWsBytes, bytes); }

}  // namespace rppa

// RPP_HOST_BACKEND handles serve the rppt_*_host entry points and touch no GPU; their thread count
// follows src/modules/handle_host.cpp:54-62 (at most the hardware threads; 0 means one per clip of
// the batch).  RPP_HIP_BACKEND handles serve rppt_*_gpu on the current device.
RPPA_EXPORT rppStatus_t rppCreate(rppHandle_t *handle, size_t nBatchSize, Rpp32u numThreads, void *stream,
                                  RppBackend backend)
{
    if (!handle) return rppStatusBadParm;
    if (backend != RPP_HIP_BACKEND && backend != RPP_HOST_BACKEND) {
        fprintf(stderr, "rpp_audio: only RPP_HOST_BACKEND and RPP_HIP_BACKEND handles are served by this library\n");
        return rppStatusNotImplemented;
    }
    rppHandle *h = new (std::nothrow) rppHandle();
    if (!h) return rppStatusAllocFailed;
    h->backend = backend;
    h->batchSize = nBatchSize ? nBatchSize : 1;
    const unsigned hw = std::thread::hardware_concurrency();
    unsigned nt = hw ? std::min<unsigned>(numThreads, hw) : numThreads;
    if (nt == 0) nt = static_cast<unsigned>(std::min<size_t>(h->batchSize, 1024));
    h->numThreads = static_cast<int>(nt);
    if (backend == RPP_HIP_BACKEND) {
        if (hipGetDevice(&h->device) != hipSuccess) { delete h; return rppStatusInternalError; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) {
            h->numCU = prop.multiProcessorCount;
            h->maxLdsPerBlock = static_cast<int>(prop.sharedMemPerBlock);
            h->maxLdsOptin = static_cast<int>(std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin));
        }
        h->stream = static_cast<hipStream_t>(stream);
    }
    *handle = h;
    return rppStatusSuccess;
}

RPPA_EXPORT rppStatus_t rppDestroy(rppHandle_t handle, RppBackend backend)
{
    if (!handle) return rppStatusBadParm;
    if (backend != handle->backend) return rppStatusBadParm;
    if (backend == RPP_HOST_BACKEND) {
        delete handle;
        return rppStatusSuccess;
    }
    (void)hipStreamSynchronize(handle->stream);
    if (handle->ws) (void)hipFree(handle->ws);
    if (handle->feWs) (void)hipFree(handle->feWs);
    if (handle->pinned) (void)hipHostFree(handle->pinned);
    for (auto &kv : handle->dft) {
        (void)hipFree(kv.second.basis);
        if (kv.second.foldHi) (void)hipFree(kv.second.foldHi);
        if (kv.second.foldLo) (void)hipFree(kv.second.foldLo);
        if (kv.second.foldLo2) (void)hipFree(kv.second.foldLo2);
        if (kv.second.wHalf) (void)hipFree(kv.second.wHalf);
    }
    for (auto &kv : handle->fft) (void)hipFree(kv.second.mem);
    for (auto &kv : handle->melFrag) {
        (void)hipFree(kv.second.frag);
        (void)hipFree(kv.second.pairOff);
    }
    for (auto &kv : handle->mel) {
        (void)hipFree(kv.second.weightsDown);
        (void)hipFree(kv.second.intervals);
        (void)hipFree(kv.second.normFactors);
        (void)hipFree(kv.second.filterBins);
        (void)hipFree(kv.second.dense);
        (void)hipFree(kv.second.denseHi);
        (void)hipFree(kv.second.denseLo);
    }
    delete handle;
    return rppStatusSuccess;
}

RPPA_EXPORT rppStatus_t rppSetBatchSize(rppHandle_t handle, size_t batchSize)
{
    if (!handle) return rppStatusBadParm;
    handle->batchSize = batchSize;
    return rppStatusSuccess;
}

RPPA_EXPORT rppStatus_t rppGetBatchSize(rppHandle_t handle, size_t *batchSize)
{
    if (!handle || !batchSize) return rppStatusBadParm;
    *batchSize = handle->batchSize;
    return rppStatusSuccess;
}

RPPA_EXPORT rppStatus_t rppSetStream(rppHandle_t handle, void *streamID)
{
    if (!handle) return rppStatusBadParm;
    handle->stream = static_cast<hipStream_t>(streamID);
    return rppStatusSuccess;
}

RPPA_EXPORT rppStatus_t rppGetStream(rppHandle_t handle, void **streamID)
{
    if (!handle || !streamID) return rppStatusBadParm;
    *streamID = handle->stream;
    return rppStatusSuccess;
}
