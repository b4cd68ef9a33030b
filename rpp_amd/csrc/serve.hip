// serve.hip -- continuous-batching request loop for the fused log-mel path.
//
// Callers submit clips from any host thread; a scheduler thread forms batches and runs them on a
// pool of device slots, each slot a stream with its own handle and device-resident buffers:
//
//   submit()  -> pending queue (arrival order, arrival timestamp)
//   scheduler -> when a slot is free: take the oldest pending clips while they fit the slot's
//                memory (admission by memory and slots); close the batch immediately when the
//                device has an idle slot (eager) or when the oldest clip has waited `deadlineUs`
//                or `maxBatch` clips are waiting; stage into pinned memory, copy to the slot's
//                device buffers, launch rppt_logmel_fused_packed_gpu, copy the dB rows back and
//                record an event -- on the slot's own stream, so batches overlap on the device
//   scheduler -> polls the slot events; a finished batch stamps each request's completion time
//                (latency = completion - arrival) and frees the slot
//
// The slot buffers are sized once, at creation, from min(memBudgetBytes, 80% of free HBM); a clip
// that can never fit a slot is refused at submit with RPP_ERROR_NOT_ENOUGH_MEMORY.  A clip's result
// depends on that clip only (work items never span clips), so it is identical alone and in any mix.
#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

RppStatus check_logmel_params(const RpptLogMelParams *p, RpptLogMelPrecision precision);
extern "C" RppStatus rppt_logmel_fused_packed_gpu(const Rpp32f *, const long long *, const Rpp32s *, Rpp32s, Rpp32f *,
                                                  const long long *, const RpptLogMelParams *, RpptLogMelPrecision,
                                                  rppHandle_t);


namespace rppa {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    long long id;
    std::vector<float> samples;
    int length = 0, frames = 0;
    Clock::time_point arrival, done;
    bool finished = false;
    int waiters = 0;          // rppa_server_wait calls blocked on this request (they own its lifetime)
    std::vector<float> out;
};

struct Slot {
    rppHandle_t h = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, done = nullptr;
    float *dIn = nullptr, *dOut = nullptr, *hIn = nullptr, *hOut = nullptr;
    long long *dMeta = nullptr, *hMeta = nullptr;  // [srcOff B | dstOff B | lengths B (as int pairs)]
    size_t inCap = 0, outCap = 0;
    int capB = 0;
    bool busy = false;
    std::vector<Request *> reqs;
};

}  // namespace
}  // namespace rppa

struct rppaServer {
    RpptLogMelParams params;
    RpptLogMelPrecision precision;
    rppaServeConfig cfg;
    int device = 0;
    int M = 0, W = 0, H = 0, center = 1;
    std::vector<rppa::Slot> slots;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<rppa::Request *> pending;
    std::unordered_map<long long, rppa::Request *> live;
    std::vector<double> latUs;
    long long nextId = 1, submitted = 0, completed = 0, refused = 0, batches = 0, batchClips = 0;
    double audioSamples = 0, busy = 0;
    std::atomic<bool> stop{false};
    std::thread worker;
    int errors = 0;
};

namespace rppa {
namespace {

int frames_of(const rppaServer *s, int L)
{
    return std::max(0, num_windows(std::max(L, 0), s->W, s->H, s->center != 0));
}

bool launch_batch(rppaServer *s, Slot &sl, std::vector<Request *> &batch)
{
    const int B = static_cast<int>(batch.size());
    long long *srcOff = sl.hMeta, *dstOff = sl.hMeta + sl.capB;
    int *len = reinterpret_cast<int *>(sl.hMeta + 2 * sl.capB);
    long long so = 0, doff = 0;
    for (int b = 0; b < B; b++) {
        Request *r = batch[b];
        srcOff[b] = so;
        dstOff[b] = doff;
        len[b] = r->length;
        std::memcpy(sl.hIn + so, r->samples.data(), sizeof(float) * r->length);
        so += (r->length + 3) / 4 * 4;           // 16-byte aligned clip starts
        doff += static_cast<long long>(r->frames) * s->M;
    }
    const size_t metaBytes = sizeof(long long) * 2 * sl.capB + sizeof(int) * sl.capB;
    if (hipEventRecord(sl.start, sl.stream) != hipSuccess) return false;
    if (hipMemcpyAsync(sl.dIn, sl.hIn, sizeof(float) * so, hipMemcpyHostToDevice, sl.stream) != hipSuccess) return false;
    if (hipMemcpyAsync(sl.dMeta, sl.hMeta, metaBytes, hipMemcpyHostToDevice, sl.stream) != hipSuccess) return false;
    RppStatus st = rppt_logmel_fused_packed_gpu(sl.dIn, sl.dMeta, reinterpret_cast<int *>(sl.dMeta + 2 * sl.capB), B,
                                                sl.dOut, sl.dMeta + sl.capB, &s->params, s->precision, sl.h);
    if (st != RPP_SUCCESS) return false;
    if (hipMemcpyAsync(sl.hOut, sl.dOut, sizeof(float) * doff, hipMemcpyDeviceToHost, sl.stream) != hipSuccess)
        return false;
    if (hipEventRecord(sl.done, sl.stream) != hipSuccess) return false;
    sl.reqs.swap(batch);
    sl.busy = true;
    return true;
}

void finish_batch(rppaServer *s, Slot &sl)
{
    const auto now = Clock::now();
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, sl.start, sl.done);
    std::lock_guard<std::mutex> lk(s->mu);
    long long doff = 0;
    for (Request *r : sl.reqs) {
        r->done = now;
        r->finished = true;
        const size_t n = static_cast<size_t>(r->frames) * s->M;
        if (s->cfg.keepOutputs) r->out.assign(sl.hOut + doff, sl.hOut + doff + n);
        doff += static_cast<long long>(n);
        s->latUs.push_back(std::chrono::duration<double, std::micro>(r->done - r->arrival).count());
        s->audioSamples += r->length;
        std::vector<float>().swap(r->samples);
        // without kept outputs nothing is left to collect: forget the request now, unless a waiter
        // holds it (the last waiter frees it)
        if (!s->cfg.keepOutputs && r->waiters == 0) {
            s->live.erase(r->id);
            delete r;
        }
    }
    s->completed += static_cast<long long>(sl.reqs.size());
    s->batches += 1;
    s->batchClips += static_cast<long long>(sl.reqs.size());
    s->busy += ms * 1e-3;
    sl.reqs.clear();
    sl.busy = false;
    s->cvDone.notify_all();
}

// Marks the requests of a batch that could not run as finished (with an error count).
void fail_requests(rppaServer *s, std::vector<Request *> &reqs)
{
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request *r : reqs) {
        r->finished = true;
        r->done = Clock::now();
        if (!s->cfg.keepOutputs && r->waiters == 0) {
            s->live.erase(r->id);
            delete r;
        }
    }
    s->errors += static_cast<int>(reqs.size());
    reqs.clear();
    s->cvDone.notify_all();
}

void scheduler(rppaServer *s)
{
    (void)hipSetDevice(s->device);
    const auto deadline = std::chrono::microseconds(s->cfg.deadlineUs);
    while (!s->stop.load()) {
        bool anyBusy = false;
        for (Slot &sl : s->slots) {
            if (!sl.busy) continue;
            hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) finish_batch(s, sl);
            else if (q == hipErrorNotReady) anyBusy = true;
            else { fail_requests(s, sl.reqs); sl.busy = false; }
        }
        std::unique_lock<std::mutex> lk(s->mu);
        for (Slot &sl : s->slots) {
            if (sl.busy || s->pending.empty()) continue;
            const auto age = Clock::now() - s->pending.front()->arrival;
            const bool close = s->cfg.eager || static_cast<int>(s->pending.size()) >= s->cfg.maxBatch || age >= deadline;
            if (!close) break;
            std::vector<Request *> batch;
            size_t in = 0, out = 0;
            while (!s->pending.empty() && static_cast<int>(batch.size()) < std::min(s->cfg.maxBatch, sl.capB)) {
                Request *r = s->pending.front();
                const size_t rin = (r->length + 3) / 4 * 4, rout = static_cast<size_t>(r->frames) * s->M;
                if (in + rin > sl.inCap || out + rout > sl.outCap) break;
                in += rin;
                out += rout;
                batch.push_back(r);
                s->pending.pop_front();
            }
            if (batch.empty()) break;
            lk.unlock();
            const bool ok = launch_batch(s, sl, batch);
            if (!ok) fail_requests(s, batch);
            lk.lock();
            anyBusy = true;
        }
        if (s->stop.load()) break;
        if (!anyBusy && s->pending.empty()) {
            s->cvWork.wait_for(lk, std::chrono::milliseconds(5));
        } else if (!anyBusy && !s->cfg.eager) {
            // a batch is forming: sleep until the oldest clip's deadline or a new arrival
            s->cvWork.wait_for(lk, std::chrono::microseconds(std::max(20, s->cfg.deadlineUs / 8)));
        } else {
            lk.unlock();
            std::this_thread::yield();
        }
    }
}

}  // namespace
}  // namespace rppa

using namespace rppa;

static void destroy_server(rppaServer *s)
{
    s->stop.store(true);
    s->cvWork.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.h) rppDestroy(sl.h, RPP_HIP_BACKEND);
        if (sl.dIn) (void)hipFree(sl.dIn);
        if (sl.dOut) (void)hipFree(sl.dOut);
        if (sl.dMeta) (void)hipFree(sl.dMeta);
        if (sl.hIn) (void)hipHostFree(sl.hIn);
        if (sl.hOut) (void)hipHostFree(sl.hOut);
        if (sl.hMeta) (void)hipHostFree(sl.hMeta);
        if (sl.start) (void)hipEventDestroy(sl.start);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    for (auto &kv : s->live) delete kv.second;  // every pending request is also live
    delete s;
}

RPPA_EXPORT RppStatus rppa_server_create(const RpptLogMelParams *params, RpptLogMelPrecision precision,
                                         const rppaServeConfig *config, rppaServer **out)
{
    if (!params || !config || !out || config->maxBatch <= 0 || config->numSlots <= 0 || config->numSlots > 16 ||
        config->deadlineUs < 0)
        return RPP_ERROR_INVALID_ARGUMENTS;
    rppaServer *s = new (std::nothrow) rppaServer();
    if (!s) return RPP_ERROR_NOT_ENOUGH_MEMORY;
    s->params = *params;
    s->precision = precision;
    s->cfg = *config;
    s->M = params->numFilter;
    s->W = params->windowLength;
    s->H = params->windowStep;
    s->center = params->centerWindows;
    if (hipGetDevice(&s->device) != hipSuccess) { delete s; return RPP_ERROR; }
    if (check_logmel_params(params, precision) != RPP_SUCCESS) { delete s; return RPP_ERROR_INVALID_ARGUMENTS; }
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) { delete s; return RPP_ERROR; }
    long long budget = static_cast<long long>(freeB * 0.8);
    budget = std::min(budget, config->memBudgetBytes > 0 ? config->memBudgetBytes : (8LL << 30));
    const long long perSlot = budget / config->numSlots;
    // per input sample: 4 B in (+ pinned twin) and M/H output floats; per clip: metadata + one frame
    const double bytesPerSample = 4.0 * (1.0 + static_cast<double>(s->M) / std::max(s->H, 1));
    const int capB = std::max(config->maxBatch, 1);
    const long long perSlotData = perSlot - static_cast<long long>(capB) * (20 + 8 * s->M);
    if (perSlotData < 1024) { delete s; return RPP_ERROR_NOT_ENOUGH_MEMORY; }
    const size_t inCap = static_cast<size_t>(perSlotData / bytesPerSample) / 4 * 4;
    const size_t outCap = static_cast<size_t>(inCap / std::max(s->H, 1) + 2 * capB) * s->M;
    s->slots.resize(config->numSlots);
    for (Slot &sl : s->slots) {
        sl.capB = capB;
        sl.inCap = inCap;
        sl.outCap = outCap;
        const size_t metaBytes = sizeof(long long) * 2 * capB + sizeof(int) * capB;
        bool ok = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreate(&sl.start) == hipSuccess && hipEventCreate(&sl.done) == hipSuccess &&
                  hipMalloc(&sl.dIn, sizeof(float) * inCap) == hipSuccess &&
                  hipMalloc(&sl.dOut, sizeof(float) * outCap) == hipSuccess &&
                  hipMalloc(&sl.dMeta, metaBytes) == hipSuccess &&
                  hipHostMalloc(&sl.hIn, sizeof(float) * inCap, hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc(&sl.hOut, sizeof(float) * outCap, hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc(&sl.hMeta, metaBytes, hipHostMallocDefault) == hipSuccess &&
                  rppCreate(&sl.h, capB, 0, sl.stream, RPP_HIP_BACKEND) == rppStatusSuccess;
        if (!ok) {
            fprintf(stderr, "rpp_audio: serving slot allocation failed\n");
            destroy_server(s);
            *out = nullptr;
            return RPP_ERROR_NOT_ENOUGH_MEMORY;
        }
    }
    // build every table, size every workspace and take each slot's copy paths once before traffic
    // arrives: capB zero clips spread over the first 64 MB of the pinned input buffer, copied in,
    // run and copied back (a workspace that grew later would free its old buffer, which waits for
    // the whole device; the first copies through a fresh pinned range are slow)
    const size_t warmIn = std::min<size_t>(inCap, static_cast<size_t>(16) << 20);
    int warmLen = static_cast<int>(std::max<size_t>(4, warmIn / capB / 4 * 4));
    while (warmLen > 4 && static_cast<size_t>(frames_of(s, warmLen)) * s->M * capB > outCap) warmLen = (warmLen / 2 + 3) / 4 * 4;
    const bool canWarm = static_cast<size_t>(warmLen) * capB <= inCap &&
                         static_cast<size_t>(frames_of(s, warmLen)) * s->M * capB <= outCap;
    for (Slot &sl : s->slots) {
        if (!canWarm) break;
        long long *srcOff = sl.hMeta, *dstOff = sl.hMeta + capB;
        int *len = reinterpret_cast<int *>(sl.hMeta + 2 * capB);
        for (int b = 0; b < capB; b++) {
            srcOff[b] = static_cast<long long>(b) * warmLen;
            dstOff[b] = static_cast<long long>(b) * frames_of(s, warmLen) * s->M;
            len[b] = warmLen;
        }
        const size_t nIn = static_cast<size_t>(warmLen) * capB;
        const size_t nOut = static_cast<size_t>(frames_of(s, warmLen)) * s->M * capB;
        std::memset(sl.hIn, 0, sizeof(float) * nIn);
        (void)hipMemcpyAsync(sl.dIn, sl.hIn, sizeof(float) * nIn, hipMemcpyHostToDevice, sl.stream);
        (void)hipMemcpyAsync(sl.dMeta, sl.hMeta, sizeof(long long) * 2 * capB + sizeof(int) * capB,
                             hipMemcpyHostToDevice, sl.stream);
        RppStatus st = rppt_logmel_fused_packed_gpu(sl.dIn, sl.dMeta, reinterpret_cast<int *>(sl.dMeta + 2 * capB),
                                                    capB, sl.dOut, sl.dMeta + capB, params, precision, sl.h);
        (void)hipMemcpyAsync(sl.hOut, sl.dOut, sizeof(float) * nOut, hipMemcpyDeviceToHost, sl.stream);
        (void)hipStreamSynchronize(sl.stream);
        if (st != RPP_SUCCESS) {
            destroy_server(s);
            *out = nullptr;
            return st;
        }
    }
    s->worker = std::thread(scheduler, s);
    *out = s;
    return RPP_SUCCESS;
}

RPPA_EXPORT RppStatus rppa_server_destroy(rppaServer *s)
{
    if (!s) return RPP_ERROR_INVALID_ARGUMENTS;
    destroy_server(s);
    return RPP_SUCCESS;
}

// Copies `length` host samples and queues the clip; RPP_ERROR_NOT_ENOUGH_MEMORY when the clip can
// never fit a slot (admission refused).
RPPA_EXPORT RppStatus rppa_server_submit(rppaServer *s, const float *samples, int length, long long *id)
{
    if (!s || !samples || length <= 0 || !id) return RPP_ERROR_INVALID_ARGUMENTS;
    const auto arrival = Clock::now();
    const int frames = frames_of(s, length);
    const Slot &sl = s->slots[0];
    if (static_cast<size_t>((length + 3) / 4 * 4) > sl.inCap || static_cast<size_t>(frames) * s->M > sl.outCap) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->refused++;
        return RPP_ERROR_NOT_ENOUGH_MEMORY;
    }
    Request *r = new (std::nothrow) Request();
    if (!r) return RPP_ERROR_NOT_ENOUGH_MEMORY;
    r->samples.assign(samples, samples + length);
    r->length = length;
    r->frames = frames;
    r->arrival = arrival;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->nextId++;
        s->live[r->id] = r;
        s->pending.push_back(r);
        s->submitted++;
        *id = r->id;
    }
    s->cvWork.notify_one();
    return RPP_SUCCESS;
}

// Blocks until request `id` is done; copies its [numFilter][frames] dB rows into out (when the
// server keeps outputs and out is non-null, up to outCapacity floats) and forgets the request.
// A request is held alive by its waiters: the scheduler never frees one that a waiter is blocked on,
// and the last waiter to return frees it.  Without kept outputs a request that already finished
// (and was forgotten) reports success with *frames = -1 and no latency.
RPPA_EXPORT RppStatus rppa_server_wait(rppaServer *s, long long id, float *out, long long outCapacity, int *frames,
                                       double *latencyUs)
{
    if (!s) return RPP_ERROR_INVALID_ARGUMENTS;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end()) {
        if (!s->cfg.keepOutputs && id >= 1 && id < s->nextId) {
            if (frames) *frames = -1;
            if (latencyUs) *latencyUs = -1.0;
            return out ? RPP_ERROR_INSUFFICIENT_DST_BUFFER_LENGTH : RPP_SUCCESS;
        }
        return RPP_ERROR_INVALID_ARGUMENTS;
    }
    Request *r = it->second;
    r->waiters++;
    s->cvDone.wait(lk, [&] { return r->finished; });
    r->waiters--;
    if (frames) *frames = r->frames;
    if (latencyUs) *latencyUs = std::chrono::duration<double, std::micro>(r->done - r->arrival).count();
    RppStatus st = RPP_SUCCESS;
    if (out) {
        if (static_cast<long long>(r->out.size()) > outCapacity || r->out.empty()) st = RPP_ERROR_INSUFFICIENT_DST_BUFFER_LENGTH;
        else std::memcpy(out, r->out.data(), sizeof(float) * r->out.size());
    }
    if (r->waiters == 0) {
        s->live.erase(r->id);   // looked up again: the iterator may be stale after the wait
        delete r;
    }
    return st;
}

// Waits until every submitted clip has finished (or timeoutMs passes); RPP_ERROR on timeout or a
// failed batch.
RPPA_EXPORT RppStatus rppa_server_drain(rppaServer *s, int timeoutMs)
{
    if (!s) return RPP_ERROR_INVALID_ARGUMENTS;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cvDone.wait_for(lk, std::chrono::milliseconds(timeoutMs), [&] {
        return s->completed + s->errors >= s->submitted;
    });
    return ok && s->errors == 0 ? RPP_SUCCESS : RPP_ERROR;
}

RPPA_EXPORT RppStatus rppa_server_stats(rppaServer *s, rppaServeStats *st)
{
    if (!s || !st) return RPP_ERROR_INVALID_ARGUMENTS;
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> v = s->latUs;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) { return v.empty() ? 0.0 : v[std::min(v.size() - 1, static_cast<size_t>(q * (v.size() - 1) + 0.5))]; };
    double sum = 0;
    for (double x : v) sum += x;
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->refused = s->refused;
    st->batches = s->batches;
    st->p50Us = pct(0.50);
    st->p90Us = pct(0.90);
    st->p99Us = pct(0.99);
    st->meanUs = v.empty() ? 0.0 : sum / v.size();
    st->maxUs = v.empty() ? 0.0 : v.back();
    st->meanBatchClips = s->batches ? static_cast<double>(s->batchClips) / s->batches : 0.0;
    st->audioSamples = s->audioSamples;
    st->gpuBusySeconds = s->busy;
    st->slotCapacitySamples = s->slots.empty() ? 0 : static_cast<long long>(s->slots[0].inCap);
    return RPP_SUCCESS;
}
