// serve.hip -- continuous-batching request loop for the fused log-mel path.
//
// Callers submit clips from any host thread; a scheduler thread forms batches and runs them on a
// pool of device slots, each slot a stream with its own handle and device-resident buffers:
//
//   submit()  -> the clip is copied once, into a pinned arena (a ring allocated in submission order
//                and freed when the clip's batch has finished; a clip that finds it full keeps its
//                own copy and is staged through the slot's pinned buffer), then queued (arrival
//                order, arrival timestamp)
//   scheduler -> when a slot is free: take the oldest pending clips while they fit the slot's
//                memory (admission by memory and slots); close the batch immediately when the
//                device has an idle slot (eager) or when the oldest clip has waited `deadlineUs`
//                or `maxBatch` clips are waiting; copy each run of arena-contiguous clips to the
//                slot's device buffer in one copy, launch rppt_logmel_fused_packed_gpu, copy the dB
//                rows back and record an event -- on the slot's own stream, so batches overlap on
//                the device
//   scheduler -> polls the slot events; a finished batch stamps each request's completion time
//                (latency = completion - arrival) and frees the slot
//
// The slot buffers are sized at creation from min(memBudgetBytes, 80% of free HBM).  With growSlots,
// a clip larger than a slot holds is admitted when 80% of the free device memory at that moment
// (hipMemGetInfo, refreshed by the scheduler every millisecond) covers a slot grown for it, and the
// next idle slot is regrown before the clip's batch; otherwise it is refused at submit with
// RPP_ERROR_NOT_ENOUGH_MEMORY.  A clip's result depends on that clip only (work items never span
// clips), so it is identical alone and in any mix.
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
    std::vector<float> samples;   // the clip's own copy when the arena was full
    const float *src = nullptr;   // its samples: in the arena or in `samples`
    long long span = -1;          // arena span id (-1: staged)
    size_t arenaOff = 0;          // offset of the samples in the arena (floats)
    int length = 0, frames = 0;
    Clock::time_point arrival, launched, done;
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
    // pinned submission arena: a ring of spans in allocation order (span ids spanFront ..)
    struct Span { size_t off, n; bool freed; };
    float *arena = nullptr;
    size_t arenaCap = 0, arenaHead = 0;
    std::deque<Span> spans;
    long long spanFront = 0, spanNext = 0;
    long long arenaSubmits = 0, stagedSubmits = 0, slotGrowths = 0;
    // growth: the slot capacities the queued clips need, the free device memory last read
    size_t growIn = 0, growOut = 0, maxInCap = 0, maxOutCap = 0;
    std::atomic<long long> freeHbm{0};
    double maxLatUs = 0, maxQueueUs = 0, maxServiceUs = 0, maxGpuUs = 0;
    // scheduler turn and launch_batch host times (written by the scheduler thread only)
    std::atomic<double> maxLoopUs{0.0}, maxLaunchUs{0.0};
    std::atomic<double> maxCopyInUs{0.0}, maxKernelUs{0.0}, maxCopyOutUs{0.0};
    long long maxLatLen = 0;
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

// Reserves n floats of the arena (caller holds s->mu); returns the span id or -1 when it is full.
// The used region is [front span, head) -- wrapped past the end when the head is below the front.
long long arena_alloc(rppaServer *s, size_t n, size_t &off)
{
    if (!s->arena || n > s->arenaCap) return -1;
    if (s->spans.empty()) s->arenaHead = 0;
    const size_t tail = s->spans.empty() ? 0 : s->spans.front().off;
    // wrapped: the head has passed the end and restarted below the front (it never reaches it)
    const bool wrapped = !s->spans.empty() && s->arenaHead < tail;
    if (!wrapped) {
        if (s->arenaHead + n <= s->arenaCap) off = s->arenaHead;
        else if (n < tail) off = 0;          // wrap: the end's gap is left unused until the ring passes it
        else return -1;
    } else {
        if (s->arenaHead + n < tail) off = s->arenaHead;
        else return -1;
    }
    s->arenaHead = off + n;
    s->spans.push_back({off, n, false});
    return s->spanNext++;
}

// Releases span `id` (caller holds s->mu); the ring's front advances over released spans.
void arena_free(rppaServer *s, long long id)
{
    if (id < 0) return;
    s->spans[static_cast<size_t>(id - s->spanFront)].freed = true;
    while (!s->spans.empty() && s->spans.front().freed) {
        s->spans.pop_front();
        s->spanFront++;
    }
}

// Writes a slot's pinned buffers and takes every byte of them through a copy once (host to device
// for the input staging, device to host for the outputs): the first copy through a fresh pinned
// range is slow (milliseconds), which would otherwise land on some request's batch.
void warm_slot(Slot &sl)
{
    std::memset(sl.hIn, 0, sizeof(float) * sl.inCap);
    std::memset(sl.hOut, 0, sizeof(float) * sl.outCap);
    (void)hipMemcpyAsync(sl.dIn, sl.hIn, sizeof(float) * sl.inCap, hipMemcpyHostToDevice, sl.stream);
    (void)hipMemcpyAsync(sl.hOut, sl.dOut, sizeof(float) * sl.outCap, hipMemcpyDeviceToHost, sl.stream);
    (void)hipStreamSynchronize(sl.stream);
}

// (Re)allocates a slot's data buffers for inCap input and outCap output floats.
bool slot_buffers(Slot &sl, size_t inCap, size_t outCap)
{
    if (sl.dIn) (void)hipFree(sl.dIn);
    if (sl.dOut) (void)hipFree(sl.dOut);
    if (sl.hIn) (void)hipHostFree(sl.hIn);
    if (sl.hOut) (void)hipHostFree(sl.hOut);
    sl.dIn = sl.dOut = sl.hIn = sl.hOut = nullptr;
    sl.inCap = sl.outCap = 0;
    if (hipMalloc(&sl.dIn, sizeof(float) * inCap) != hipSuccess || hipMalloc(&sl.dOut, sizeof(float) * outCap) != hipSuccess ||
        hipHostMalloc(&sl.hIn, sizeof(float) * inCap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&sl.hOut, sizeof(float) * outCap, hipHostMallocDefault) != hipSuccess)
        return false;
    sl.inCap = inCap;
    sl.outCap = outCap;
    warm_slot(sl);
    return true;
}

bool launch_batch(rppaServer *s, Slot &sl, std::vector<Request *> &batch)
{
    const int B = static_cast<int>(batch.size());
    long long *srcOff = sl.hMeta, *dstOff = sl.hMeta + sl.capB;
    int *len = reinterpret_cast<int *>(sl.hMeta + 2 * sl.capB);
    long long so = 0, doff = 0;
    const auto now = Clock::now();
    for (int b = 0; b < B; b++) {
        Request *r = batch[b];
        r->launched = now;
        srcOff[b] = so;
        dstOff[b] = doff;
        len[b] = r->length;
        if (r->span < 0) std::memcpy(sl.hIn + so, r->src, sizeof(float) * r->length);   // staged clip
        so += (r->length + 3) / 4 * 4;           // 16-byte aligned clip starts
        doff += static_cast<long long>(r->frames) * s->M;
    }
    const size_t metaBytes = sizeof(long long) * 2 * sl.capB + sizeof(int) * sl.capB;
    auto upd = [](std::atomic<double> &m, Clock::time_point t0) {
        m.store(std::max(m.load(), std::chrono::duration<double, std::micro>(Clock::now() - t0).count()));
    };
    auto t0 = Clock::now();
    if (hipEventRecord(sl.start, sl.stream) != hipSuccess) return false;
    // host-to-device: one copy per run of clips that lie back to back in the arena (or in the
    // slot's staging buffer), straight from where they are
    for (int b = 0; b < B;) {
        const bool arena = batch[b]->span >= 0;
        const float *from = arena ? s->arena + batch[b]->arenaOff : sl.hIn + srcOff[b];
        long long n = 0;
        int e = b;
        while (e < B && (batch[e]->span >= 0) == arena &&
               (!arena || s->arena + batch[e]->arenaOff == from + n)) {
            n += (batch[e]->length + 3) / 4 * 4;
            e++;
        }
        if (hipMemcpyAsync(sl.dIn + srcOff[b], from, sizeof(float) * n, hipMemcpyHostToDevice, sl.stream) != hipSuccess)
            return false;
        b = e;
    }
    if (hipMemcpyAsync(sl.dMeta, sl.hMeta, metaBytes, hipMemcpyHostToDevice, sl.stream) != hipSuccess) return false;
    upd(s->maxCopyInUs, t0);
    t0 = Clock::now();
    RppStatus st = rppt_logmel_fused_packed_gpu(sl.dIn, sl.dMeta, reinterpret_cast<int *>(sl.dMeta + 2 * sl.capB), B,
                                                sl.dOut, sl.dMeta + sl.capB, &s->params, s->precision, sl.h);
    if (st != RPP_SUCCESS) return false;
    upd(s->maxKernelUs, t0);
    t0 = Clock::now();
    if (hipMemcpyAsync(sl.hOut, sl.dOut, sizeof(float) * doff, hipMemcpyDeviceToHost, sl.stream) != hipSuccess)
        return false;
    if (hipEventRecord(sl.done, sl.stream) != hipSuccess) return false;
    upd(s->maxCopyOutUs, t0);
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
        const double lat = std::chrono::duration<double, std::micro>(r->done - r->arrival).count();
        s->latUs.push_back(lat);
        if (lat > s->maxLatUs) {
            s->maxLatUs = lat;
            s->maxQueueUs = std::chrono::duration<double, std::micro>(r->launched - r->arrival).count();
            s->maxServiceUs = std::chrono::duration<double, std::micro>(r->done - r->launched).count();
            s->maxLatLen = r->length;
            s->maxGpuUs = ms * 1e3;
        }
        s->audioSamples += r->length;
        arena_free(s, r->span);
        r->span = -1;
        r->src = nullptr;
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
        arena_free(s, r->span);
        r->span = -1;
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
    auto lastMemInfo = Clock::now() - std::chrono::seconds(1);
    auto turn = Clock::now();
    while (!s->stop.load()) {
        {
            // the longest turn of this loop that did not sleep on purpose (waits below restart it)
            const auto now = Clock::now();
            s->maxLoopUs.store(std::max(s->maxLoopUs.load(), std::chrono::duration<double, std::micro>(now - turn).count()));
            turn = now;
        }
        if (Clock::now() - lastMemInfo >= std::chrono::milliseconds(20)) {
            // the free device memory submit() admits oversized clips against
            size_t freeB = 0, totalB = 0;
            if (hipMemGetInfo(&freeB, &totalB) == hipSuccess) s->freeHbm.store(static_cast<long long>(freeB));
            lastMemInfo = Clock::now();
        }
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
            {
                // a queued clip larger than this idle slot: regrow the slot to the capacity the
                // admitted clips need (checked against the free device memory at submit)
                Request *r = s->pending.front();
                const size_t rin = (r->length + 3) / 4 * 4, rout = static_cast<size_t>(r->frames) * s->M;
                if (rin > sl.inCap || rout > sl.outCap) {
                    const size_t gi = std::max(s->growIn, sl.inCap), go = std::max(s->growOut, sl.outCap);
                    lk.unlock();
                    (void)hipStreamSynchronize(sl.stream);
                    const bool ok = slot_buffers(sl, gi, go);
                    lk.lock();
                    if (!ok) {
                        // the memory is gone: fail the clip rather than stall the queue behind it
                        fprintf(stderr, "rpp_audio: serving slot could not grow to %zu samples\n", gi);
                        (void)slot_buffers(sl, s->slots[0].inCap ? s->slots[0].inCap : 1024,
                                           s->slots[0].outCap ? s->slots[0].outCap : 1024);
                        std::vector<Request *> one{r};
                        s->pending.pop_front();
                        lk.unlock();
                        fail_requests(s, one);
                        lk.lock();
                        continue;
                    }
                    s->slotGrowths++;
                    s->maxInCap = std::max(s->maxInCap, sl.inCap);
                    s->maxOutCap = std::max(s->maxOutCap, sl.outCap);
                    if (s->pending.empty()) continue;
                }
            }
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
            const auto l0 = Clock::now();
            const bool ok = launch_batch(s, sl, batch);
            s->maxLaunchUs.store(
                std::max(s->maxLaunchUs.load(), std::chrono::duration<double, std::micro>(Clock::now() - l0).count()));
            if (!ok) fail_requests(s, batch);
            lk.lock();
            anyBusy = true;
        }
        if (s->stop.load()) break;
        if (!anyBusy && s->pending.empty()) {
            s->cvWork.wait_for(lk, std::chrono::milliseconds(5));
            turn = Clock::now();
        } else if (!anyBusy && !s->cfg.eager) {
            // a batch is forming: sleep until the oldest clip's deadline or a new arrival
            s->cvWork.wait_for(lk, std::chrono::microseconds(std::max(20, s->cfg.deadlineUs / 8)));
            turn = Clock::now();
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
    if (s->arena) (void)hipHostFree(s->arena);
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
    s->maxInCap = inCap;
    s->maxOutCap = outCap;
    s->freeHbm.store(static_cast<long long>(freeB));
    // the pinned submission arena: room for two rounds of full slots, at most 256 MiB; every page is
    // written and taken through a host-to-device copy once here (the first copies from a fresh pinned
    // range are slow), into the first slot's device buffer in pieces it holds
    s->arenaCap = std::min<size_t>(2 * inCap * config->numSlots, static_cast<size_t>(64) << 20);
    if (hipHostMalloc(&s->arena, sizeof(float) * s->arenaCap, hipHostMallocDefault) != hipSuccess) {
        s->arena = nullptr;
        s->arenaCap = 0;
    } else {
        std::memset(s->arena, 0, sizeof(float) * s->arenaCap);
        Slot &s0 = s->slots[0];
        for (size_t o = 0; o < s->arenaCap; o += s0.inCap)
            (void)hipMemcpyAsync(s0.dIn, s->arena + o, sizeof(float) * std::min(s0.inCap, s->arenaCap - o),
                                 hipMemcpyHostToDevice, s0.stream);
        (void)hipStreamSynchronize(s0.stream);
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
    for (Slot &sl : s->slots) warm_slot(sl);
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
    const size_t n4 = static_cast<size_t>((length + 3) / 4 * 4), nOut = static_cast<size_t>(frames) * s->M;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (n4 > s->maxInCap || nOut > s->maxOutCap) {
            // admission by the free device memory now: a slot regrown for this clip (a quarter of
            // headroom) must fit 80% of it
            if (!s->cfg.growSlots) {
                s->refused++;
                return RPP_ERROR_NOT_ENOUGH_MEMORY;
            }
            const size_t gi = std::max(s->growIn, n4 + n4 / 4), go = std::max(s->growOut, nOut + nOut / 4);
            const double extra = 4.0 * (static_cast<double>(gi) - static_cast<double>(s->slots[0].inCap)) +
                                 4.0 * (static_cast<double>(go) - static_cast<double>(s->slots[0].outCap));
            if (extra > 0.8 * static_cast<double>(s->freeHbm.load())) {
                s->refused++;
                return RPP_ERROR_NOT_ENOUGH_MEMORY;
            }
            s->growIn = gi;
            s->growOut = go;
        }
    }
    Request *r = new (std::nothrow) Request();
    if (!r) return RPP_ERROR_NOT_ENOUGH_MEMORY;
    r->length = length;
    r->frames = frames;
    r->arrival = arrival;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->span = arena_alloc(s, n4, r->arenaOff);
    }
    if (r->span >= 0) {
        // the one host copy: into the arena the batch's host-to-device copy reads
        r->src = s->arena + r->arenaOff;
        std::memcpy(s->arena + r->arenaOff, samples, sizeof(float) * length);
    } else {
        r->samples.assign(samples, samples + length);
        r->src = r->samples.data();
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->span >= 0) s->arenaSubmits++;
        else s->stagedSubmits++;
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
    st->slotCapacitySamples = static_cast<long long>(s->maxInCap);
    st->maxQueueUs = s->maxQueueUs;
    st->maxServiceUs = s->maxServiceUs;
    st->maxLatencySamples = s->maxLatLen;
    st->maxBatchGpuUs = s->maxGpuUs;
    st->maxLoopUs = s->maxLoopUs.load();
    st->maxLaunchUs = s->maxLaunchUs.load();
    st->maxCopyInCallUs = s->maxCopyInUs.load();
    st->maxKernelCallUs = s->maxKernelUs.load();
    st->maxCopyOutCallUs = s->maxCopyOutUs.load();
    st->arenaSubmits = s->arenaSubmits;
    st->stagedSubmits = s->stagedSubmits;
    st->slotGrowths = s->slotGrowths;
    return RPP_SUCCESS;
}
