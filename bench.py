"""Benchmark of the batched log-mel path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp16x3] [--workload 1]

Workload (default 1): batch 512 mono float32 16 kHz clips, lengths log-uniform 0.5-20 s,
N(0, 0.1^2) noise plus random chirps; pre-emphasis 0.97 (clamp), centered reflect-padded Hann
frames W=400 hop 160 nfft 512, power 2, 80 Slaney mels 0-8 kHz normalised, dB (x10, ref 1, -80 dB).
Inputs are resident in HBM before timing starts.

A step is one pass of the hot path over the whole batch: the fused persistent launch
(rppt_logmel_fused_gpu: the persistent fused kernel, which plans its own tiles) replayed from a
hipGraph.  Before the warmup steps the
graph is replayed, untimed, for --settle-seconds so the timed steps see the clocks of sustained
load (the first tens of milliseconds on an idle GPU run up to 20% slower).  Printed as one JSON line:
value = audio-seconds of the batch * steps / wall time of the timed steps.  Side figures (same
run): the FP32 fused launch, the plain FP32 drop-in chain per clip (one launch per clip per stage,
the 1x denominator), the error of each path against the FP64 oracle on a sample, the roofline of
the fused kernel, and the CPU baseline (the oracle's RPP-HOST-equivalent FP32 chain, OpenMP).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# revisions of the fused kernels (the factored-DFT kernel is selected by 16 frames per tile, the dense
# one runs at 32/64); recorded traffic (profiles/*.md `traffic:` lines) applies only to the revision it
# was measured on
FACTORED_REVISION = {512: "fft32x16-stream", 1024: "fft2x32x16-radix2-stream"}
DENSE_REVISION = "mirror-il2-selfplan"
# (16-mel tile, 32-bin chunk) pairs the factored kernel's mel occupies per work item: the 80-mel
# Slaney bank at nfft 512 / 16 kHz, and the 128-mel bank at nfft 1024 / 44.1 kHz (14 pairs over bins
# 0-255 and 9 over bins 257-512)
FACTORED_MEL_PAIRS = {512: 12, 1024: 23}


def kernel_revision(frames_per_tile, nfft):
    return FACTORED_REVISION.get(nfft, DENSE_REVISION) if frames_per_tile == 16 else DENSE_REVISION
MI355X_PEAK_TFLOPS = {"fp16x3": 2500.0, "bf16x3": 2500.0, "bf16": 2500.0, "fp32": 157.3}
MI355X_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def recorded_traffic(workload, precision, tiles, nfft):
    """HBM bytes per launch of the fused kernel for this configuration, from the rocprofv3
    FETCH_SIZE/WRITE_SIZE runs summarised under profiles/ (a `traffic:` line per configuration,
    counters already scaled by the copy-kernel calibration), or None when none matches."""
    import glob
    key = {"kernel": kernel_revision(tiles[0], nfft), "workload": str(workload), "precision": precision,
           "frames_per_tile": str(tiles[0]), "blocks_per_cu": str(tiles[1])}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            if not line.startswith("traffic:"):
                continue
            kv = dict(f.split("=", 1) for f in line.split()[1:] if "=" in f)
            if all(kv.get(k) == v for k, v in key.items()):
                return int(kv["bytes_per_launch"]), os.path.basename(path)
    return None


def front_end_measurement(api, synth, torch, precision, batch=64, seconds=60.0, steps=5, settle_seconds=0.2):
    """Workload 4 (64 clips): stereo int16-range 48 kHz clips, 60 s each ->
    down-mix -> Kaiser-windowed sinc resample to 16 kHz (16 zero-crossings, 4096-entry table) ->
    fused log-mel through rppt_logmel_front_end_gpu replayed from a hipGraph: as one launch (the
    log-mel work items compute the resampled samples they read) and, for comparison, as the two
    launches (down-mix + resample into a buffer, then the fused log-mel)."""
    bt = synth.workload4(batch=batch, seconds=seconds)
    lut, sc, ce = synth.kaiser_sinc_lut()
    x = torch.from_numpy(bt.x).cuda()
    lut_d = torch.from_numpy(lut).cuda()
    window = api.make_window(lut_d, sc, ce, 16, 4091)
    dims = torch.stack([torch.from_numpy(bt.lengths), torch.full((batch,), 2, dtype=torch.int32)], 1).reshape(-1)
    dims = dims.contiguous().cuda()
    p = api.logmel_params()
    n_out = int(np.ceil(np.float32(bt.lengths.max()) * np.float32(16000.0) / np.float32(48000.0)))
    out = torch.empty((batch, p.numFilter, n_out // p.windowStep + 1), dtype=torch.float32, device="cuda")
    audio = float(bt.lengths.sum()) / 48000.0
    res = {"batch": batch, "seconds_per_clip": seconds}
    stream = torch.cuda.Stream()
    for one in (True, False):
        h = api.Handle(batch, stream=stream)
        with torch.cuda.stream(stream):
            api.logmel_front_end(h, x, out, dims, 48000.0, 16000.0, window, p, precision, channels=2, one_launch=one)
            stream.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=stream):
                api.logmel_front_end(h, x, out, dims, 48000.0, 16000.0, window, p, precision, channels=2,
                                     one_launch=one)
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < settle_seconds:   # untimed, to the clocks of sustained load
            g.replay()
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.replay()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        del g
        h.close()
        key = "" if one else "two_launch_"
        res[key + "ms_per_step"] = round(wall / steps * 1e3, 3)
        res[key + "audio_s_per_s"] = round(audio * steps / wall, 1)
    return res


def workload_measurement(api, synth, torch, workload, precision, steps=20, settle_seconds=0.1):
    """Fused launch (hipGraph) on another workload's batch: audio-s/s and the
    kernel's MFMA fraction by the same algorithmic count as the headline (events over the steps)."""
    bt = synth.make(workload)
    f = bt.params
    x = torch.from_numpy(bt.x).cuda()
    L = torch.from_numpy(bt.lengths).cuda()
    frames = bt.lengths // f["H"] + 1
    out = torch.empty((len(bt.lengths), f["nmel"], int(frames.max())), dtype=torch.float32, device="cuda")
    p = api.logmel_params(nfft=f["nfft"], window_length=f["W"], window_step=f["H"], sample_rate=f["sample_rate"],
                          min_freq=f["fmin"], max_freq=f["fmax"], num_filter=f["nmel"])
    stream = torch.cuda.Stream()
    h = api.Handle(len(bt.lengths), stream=stream)
    with torch.cuda.stream(stream):
        api.logmel_fused(h, x, out, L, p, precision)
        tiles, _ = api.autotune(h, x, out, L, p, precision, iters=3)
        stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            api.logmel_fused(h, x, out, L, p, precision)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < settle_seconds:
        g.replay()
        torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        ev0.record(stream)
        for _ in range(steps):
            g.replay()
        ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_s = ev0.elapsed_time(ev1) / 1e3 / steps
    K = f["nfft"] // 2 + 1
    flops = float(frames.sum()) * (2 * f["W"] * 2 * K + 2 * K * f["nmel"])
    h.close()
    return {"batch": len(bt.lengths), "audio_seconds": round(bt.audio_seconds, 1), "frames": int(frames.sum()),
            "nfft": f["nfft"], "window": f["W"], "hop": f["H"], "mels": f["nmel"],
            "tiles": {"frames_per_tile": tiles[0], "blocks_per_cu": tiles[1]},
            "ms_per_step": round(wall / steps * 1e3, 4), "audio_s_per_s": round(bt.audio_seconds * steps / wall, 1),
            "mfma_frac": round(flops / step_s / 1e12 / MI355X_PEAK_TFLOPS[precision], 4)}


def trim_measurement(api, synth, torch, params, precision, steps=20, settle_seconds=0.1):
    """Workload 1 trimmed first: non-silent region detection (-60 dB cut-off, 2048-sample window,
    reset interval 8192, reference = each clip's maximum) writing each clip's begin/length on the
    device, then the fused launch over the detected regions (rppt_logmel_fused_region_gpu, no
    copy); both launches replayed from one hipGraph, and the detection alone from another."""
    bt = synth.make(1)
    feat = bt.params
    B = len(bt.lengths)
    x = torch.from_numpy(bt.x).cuda()
    L = torch.from_numpy(bt.lengths).cuda()
    frames = bt.lengths // feat["H"] + 1
    out = torch.empty((B, feat["nmel"], int(frames.max())), dtype=torch.float32, device="cuda")
    begin = torch.zeros(B, dtype=torch.int32, device="cuda")
    length = torch.zeros(B, dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    h = api.Handle(B, stream=stream)

    def nsr():
        api.non_silent_region_detection(h, x, L, begin, length, -60.0, 2048, 0.0, 8192)

    def both():
        nsr()
        api.logmel_fused_region(h, x, out, begin, length, params, precision)

    graphs = {}
    with torch.cuda.stream(stream):
        for name, fn in (("nsr", nsr), ("both", both)):
            fn()
            stream.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=stream):
                fn()
            graphs[name] = g
    res = {"batch": B, "audio_seconds": round(bt.audio_seconds, 1)}
    for name, g in graphs.items():
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < settle_seconds:
            g.replay()
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.replay()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        res[f"{name}_ms_per_step"] = round(wall * 1e3, 4)
        res[f"{name}_audio_s_per_s"] = round(bt.audio_seconds / wall, 1)
    res["kept_fraction"] = round(float(length.sum().item()) / float(bt.lengths.sum()), 4)
    h.close()
    return res


def serve_measurement(api, synth, params, precision, seconds, rate=4000.0):
    """Served stream: Poisson arrivals at `rate` clips/s, lengths log-uniform 0.25-30 s at 16 kHz,
    through the continuous-batching server (2 ms batch deadline, 4 device slots, admission by
    memory).  Latency = arrival -> dB rows in host memory.  Then a saturation burst."""
    from rpp_amd.serve import Server
    t_arr, lengths = synth.arrivals(rate, seconds)
    pool = [synth.clip_for_request(i, 480000) for i in range(16)]
    res = {"offered_clips_per_s": rate, "seconds": seconds, "requests": int(len(t_arr)),
           "offered_audio_s_per_s": round(float(lengths.sum()) / 16000.0 / seconds, 1)}
    with Server(params, precision, max_batch=64, deadline_us=2000, mem_budget=2 << 30, slots=4, eager=True,
                keep_outputs=False) as srv:
        t0 = time.perf_counter()
        for i, (t, n) in enumerate(zip(t_arr, lengths)):
            while time.perf_counter() - t0 < t:
                pass
            srv.submit(pool[i % len(pool)][:int(n)])
        srv.drain(120000)
        wall = time.perf_counter() - t0
        st = srv.stats()
    res.update({"p50_ms": round(st["p50Us"] / 1e3, 3), "p90_ms": round(st["p90Us"] / 1e3, 3),
                "p99_ms": round(st["p99Us"] / 1e3, 3), "max_ms": round(st["maxUs"] / 1e3, 3),
                "audio_s_per_s": round(st["audioSamples"] / 16000.0 / wall, 1), "completed": st["completed"],
                "refused": st["refused"], "batches": st["batches"], "mean_batch_clips": round(st["meanBatchClips"], 2),
                "gpu_busy_fraction": round(st["gpuBusySeconds"] / wall, 4),
                # the slowest request: waiting for its batch vs its batch's copies and launch
                "slowest": {"ms": round(st["maxUs"] / 1e3, 3), "queue_ms": round(st["maxQueueUs"] / 1e3, 3),
                            "service_ms": round(st["maxServiceUs"] / 1e3, 3),
                            "batch_gpu_ms": round(st["maxBatchGpuUs"] / 1e3, 3),
                            "scheduler_longest_turn_ms": round(st["maxLoopUs"] / 1e3, 3),
                            "longest_launch_call_ms": round(st["maxLaunchUs"] / 1e3, 3),
                            "longest_call_parts_ms": {"copies_in": round(st["maxCopyInCallUs"] / 1e3, 3),
                                                      "launch": round(st["maxKernelCallUs"] / 1e3, 3),
                                                      "copy_out": round(st["maxCopyOutCallUs"] / 1e3, 3)},
                            "audio_s": round(st["maxLatencySamples"] / 16000.0, 2)},
                "one_copy_submits": st["arenaSubmits"], "staged_submits": st["stagedSubmits"]})
    # saturation: a burst of 4000 requests submitted at once from 8 host threads (the library call
    # releases the interpreter lock, so the copies into the server run in parallel), drained by
    # continuous batching
    import threading
    n_burst, n_threads = 4000, 8
    with Server(params, precision, max_batch=128, deadline_us=2000, mem_budget=8 << 30, slots=4, eager=True,
                keep_outputs=False) as srv:
        def submitter(k):
            for i in range(k, n_burst, n_threads):
                srv.submit(pool[i % len(pool)][:int(lengths[i % len(lengths)])])
        ths = [threading.Thread(target=submitter, args=(k,)) for k in range(n_threads)]
        t0 = time.perf_counter()
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        t_sub = time.perf_counter() - t0
        srv.drain(120000)
        wall = time.perf_counter() - t0
        st = srv.stats()
    sat = st["audioSamples"] / 16000.0 / wall
    res["saturation"] = {"requests": n_burst, "audio_s_per_s": round(sat, 1),
                         "batches": st["batches"], "mean_batch_clips": round(st["meanBatchClips"], 2),
                         "p50_ms": round(st["p50Us"] / 1e3, 3), "p99_ms": round(st["p99Us"] / 1e3, 3),
                         "one_copy_submits": st["arenaSubmits"], "staged_submits": st["stagedSubmits"],
                         "submit_s": round(t_sub, 4), "submit_threads": n_threads}
    # the host link the served clips cross: pinned host -> device copies of 256 MiB, and the
    # samples' share of it at saturation (64,000 B per audio-second in)
    import torch
    hsrc = torch.empty(64 << 20, dtype=torch.float32, pin_memory=True)
    ddst = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
    ddst.copy_(hsrc, non_blocking=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        ddst.copy_(hsrc, non_blocking=True)
    torch.cuda.synchronize()
    h2d = 5 * hsrc.numel() * 4 / (time.perf_counter() - t0) / 1e9
    del hsrc, ddst
    res["saturation"]["h2d_pinned_gbs"] = round(h2d, 2)
    res["saturation"]["input_gbs"] = round(sat * 64000.0 / 1e9, 3)
    res["saturation"]["input_fraction_of_h2d"] = round(sat * 64000.0 / 1e9 / h2d, 4)
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--settle-seconds", type=float, default=0.3,
                    help="untimed replays before the warmup steps, so the timed steps run at the clocks of "
                         "sustained load (an idle GPU needs tens of milliseconds of work to reach them)")
    ap.add_argument("--workload", type=int, default=1, choices=[0, 1, 2])
    ap.add_argument("--precision", default="fp16x3", choices=["fp16x3", "bf16x6", "bf16x3", "bf16", "fp32"])
    ap.add_argument("--no-side", action="store_true", help="skip side measurements (profiling runs)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (known-size kernel for HBM counters)")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="duration of the served-stream measurement (Poisson arrivals, 4000 clips/s); 0 skips it")
    ap.add_argument("--front-end-only", action="store_true",
                    help="only the workload-4 front end (down-mix + resample, then fused log-mel; profiling runs)")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the served-stream measurement (profiling runs)")
    ap.add_argument("--trim-only", action="store_true",
                    help="only the workload-1 detection + region log-mel side line (profiling runs)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus > 1 is not supported "
                                   "(clips shard across GPUs with no exchange between them)"}))
        return 2

    import torch
    from rpp_amd import api, synth

    torch.cuda.set_device(0)
    if a.copy_calibration_mb:
        n = a.copy_calibration_mb * (1 << 20) // 4
        src_c = torch.rand(n, device="cuda")
        dst_c = torch.empty_like(src_c)
        for _ in range(3):
            dst_c.copy_(src_c)
        torch.cuda.synchronize()
        del src_c, dst_c
    if a.serve_only:
        print(json.dumps({"serving": serve_measurement(api, synth, api.logmel_params(), a.precision,
                                                       a.serve_seconds)}))
        return 0
    if a.front_end_only:
        print(json.dumps({"front_end": front_end_measurement(api, synth, torch, a.precision, steps=a.steps)}))
        return 0
    f1 = synth.make(1).params   # the trim and serving side lines run workload 1's 16 kHz features
    params16k = api.logmel_params(nfft=f1["nfft"], window_length=f1["W"], window_step=f1["H"],
                                  sample_rate=f1["sample_rate"], min_freq=f1["fmin"], max_freq=f1["fmax"],
                                  num_filter=f1["nmel"])
    if a.trim_only:
        print(json.dumps({"trim_workload1": trim_measurement(api, synth, torch, params16k, a.precision, steps=a.steps)}))
        return 0
    bt = synth.make(a.workload)
    feat = bt.params
    B = len(bt.lengths)
    H = feat["H"]
    frames = bt.lengths // H + 1
    maxF = int(frames.max())
    audio_s = bt.audio_seconds
    log(f"workload {a.workload}: B={B} audio={audio_s:.1f}s frames={int(frames.sum())} maxF={maxF}")

    x = torch.from_numpy(bt.x).cuda()
    L = torch.from_numpy(bt.lengths).cuda()
    out = torch.empty((B, feat["nmel"], maxF), dtype=torch.float32, device="cuda")
    params = api.logmel_params(nfft=feat["nfft"], window_length=feat["W"], window_step=H,
                               sample_rate=feat["sample_rate"], min_freq=feat["fmin"], max_freq=feat["fmax"],
                               num_filter=feat["nmel"])
    stream = torch.cuda.Stream()
    h = api.Handle(B, stream=stream)

    tuned = {}

    def fused_graph(precision):
        with torch.cuda.stream(stream):
            api.logmel_fused(h, x, out, L, params, precision)  # builds tables, sizes workspace
            tuned[precision] = api.autotune(h, x, out, L, params, precision, iters=3)
            log(f"autotune {precision}: tiles {tuned[precision][0]} {tuned[precision][1]:.3f} ms")
            stream.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=stream):
                api.logmel_fused(h, x, out, L, params, precision)
        return g

    def settle(g, seconds):
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < seconds:
            g.replay()
            n += 1
            if n % 16 == 0:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        return n

    def time_graph(g, steps, warmup):
        for _ in range(warmup):
            g.replay()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        with torch.cuda.stream(stream):
            ev0.record(stream)
            for _ in range(steps):
                g.replay()
            ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, ev0.elapsed_time(ev1) / 1e3

    # ---------------- headline: mixed-precision fused launch under a graph
    g = fused_graph(a.precision)
    n_settle = settle(g, a.settle_seconds)
    wall, gpu_s = time_graph(g, a.steps, a.warmup)
    ms_per_step = wall / a.steps * 1e3
    value = audio_s * a.steps / wall
    head = out.clone()
    log(f"fused {a.precision}: {ms_per_step:.3f} ms/step, {value:,.0f} audio-s/s (events {gpu_s / a.steps * 1e3:.3f} ms)")
    # the kernel's own launch duration: one graph of CHAIN back-to-back launches of the same step,
    # so the few microseconds between graph replays fall out of the per-launch figure (the launches
    # between two kernel nodes of one graph still count, so this stays an upper bound)
    CHAIN = 10
    with torch.cuda.stream(stream):
        gc = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gc, stream=stream):
            for _ in range(CHAIN):
                api.logmel_fused(h, x, out, L, params, a.precision)
    reps = max(1, a.steps // CHAIN)
    _, chain_s = time_graph(gc, reps, 2)
    kernel_s = min(chain_s / (reps * CHAIN), gpu_s / a.steps)
    del gc
    log(f"fused {a.precision}: {kernel_s * 1e6:.1f} us per launch in a graph of {CHAIN} launches")

    # roofline of the dominant kernel (the fused kernel; the planner is ~us): algorithmic FLOPs
    # (dense DFT 2*W*2K + dense mel 2*K*M per frame; split-product passes are overhead, not work)
    K = feat["nfft"] // 2 + 1
    flop_per_frame = 2 * feat["W"] * 2 * K + 2 * K * feat["nmel"]
    flops = float(frames.sum()) * flop_per_frame
    achieved_tf = flops / kernel_s / 1e12
    rec = recorded_traffic(a.workload, a.precision, tuned[a.precision][0], feat["nfft"])
    roofline = {"bound": "mfma", "achieved": round(achieved_tf, 2), "peak": MI355X_PEAK_TFLOPS[a.precision],
                "unit": "TFLOP/s", "frac": round(achieved_tf / MI355X_PEAK_TFLOPS[a.precision], 4),
                "traffic": rec[0] if rec else None, "flop_per_launch": flops, "kernel_ms": round(kernel_s * 1e3, 4),
                "step_ms_events": round(gpu_s / a.steps * 1e3, 4),
                "hbm_bytes_algorithmic": int(4 * bt.lengths.sum() + 4 * feat["nmel"] * frames.sum())}
    if rec:
        roofline["traffic_source"] = "profiles/" + rec[1]
        roofline["hbm_gbs_at_traffic"] = round(rec[0] / kernel_s / 1e9, 1)
    ft_chosen = tuned[a.precision][0][0]
    roofline["kernel"] = kernel_revision(ft_chosen, feat["nfft"])
    if ft_chosen == 16 and a.precision == "fp16x3" and feat["nfft"] in FACTORED_MEL_PAIRS:
        # the factored DFT executes fewer matrix FLOPs than the dense count above: per work item
        # (16 frames at nfft 512, 8 at nfft 1024) 96 + 96 fp16 products (stage 1, stage 2) and one
        # bf16 triple per (16-mel tile, 32-bin chunk) pair of the bank, each v_mfma_*_16x16x32 =
        # 16,384 FLOP; split passes included
        fi = 16 if feat["nfft"] == 512 else 8
        items = int(((frames + fi - 1) // fi).sum())
        mfma = items * (192 + 3 * FACTORED_MEL_PAIRS[feat["nfft"]]) * 16384.0
        roofline["mfma_flop_executed_per_launch"] = mfma
        roofline["mfma_pipe_frac"] = round(mfma / kernel_s / 1e12 / MI355X_PEAK_TFLOPS[a.precision], 4)

    side = {}
    cpu = None
    if not a.no_side:
        from oracle import oracle

        def max_db_err(o, r, fr):
            return max(float(np.abs(o[b, :, :F].astype(np.float64) - r[b, :, :F]).max()) for b, F in enumerate(fr))
        # FP32 fused
        g32 = fused_graph("fp32")
        w32, _ = time_graph(g32, max(3, a.steps // 4), 1)
        side["fp32_fused_audio_s_per_s"] = round(audio_s * max(3, a.steps // 4) / w32, 1)
        # plain FP32 drop-in chain: one launch per clip per stage (the 1x denominator), and batched
        L_host = torch.from_numpy(bt.lengths).pin_memory()
        out32 = torch.empty_like(out)
        with torch.cuda.stream(stream):
            for per_clip, key, reps in ((True, "fp32_per_clip_audio_s_per_s", 2), (False, "fp32_chain_batched_audio_s_per_s", 3)):
                api.logmel_chain_gpu(h, x, out32, L_host, params, per_clip=per_clip)
                stream.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    api.logmel_chain_gpu(h, x, out32, L_host, params, per_clip=per_clip)
                stream.synchronize()
                side[key] = round(audio_s * reps / (time.perf_counter() - t0), 1)
        side["speedup_vs_fp32_per_clip"] = round(value / side["fp32_per_clip_audio_s_per_s"], 2)
        # accuracy of the headline output against the FP64 oracle on 64 clips (the 32 longest and 32
        # spread over the rest): max |dB error| over every output, and the relative error of the
        # mel power max and p99.9 over the outputs above the -80 dB clamp
        order = np.argsort(bt.lengths)[::-1]
        rest = np.sort(order[32:])
        idx = np.sort(np.concatenate([order[:32], rest[np.linspace(0, len(rest) - 1, 32).astype(int)]])) \
            if B > 64 else np.arange(B)
        kw = dict(nfft=feat["nfft"], W=feat["W"], H=H, nmel=feat["nmel"], sample_rate=feat["sample_rate"],
                  fmax=feat["fmax"])
        ref64, mel64, fr64 = oracle.logmel_f64(bt.x[idx], bt.lengths[idx], nthreads=min(16, os.cpu_count() or 1), **kw)
        hd = head[torch.from_numpy(idx).cuda()].cpu().numpy()
        side["max_db_err_vs_fp64"] = round(max_db_err(hd, ref64, fr64), 6)
        rel = np.concatenate([(np.abs(np.power(10.0, hd[b, :, :F].astype(np.float64) / 10.0) - mel64[b, :, :F]) /
                               mel64[b, :, :F])[mel64[b, :, :F] > 1e-8] for b, F in enumerate(fr64)])
        side["rel_power_err_vs_fp64"] = {"max": float(f"{rel.max():.3e}"),
                                         "p99_9": float(f"{np.percentile(rel, 99.9):.3e}"),
                                         "outputs": int(rel.size)}
        side["accuracy_sample_clips"] = int(len(idx))
        # the three-term bf16 split (six products) on the same batch: throughput of its fused
        # graph, and its max |dB error| on the same clips
        g6 = fused_graph("bf16x6")
        n6 = max(3, a.steps // 4)
        w6, _ = time_graph(g6, n6, 1)
        o6 = out[torch.from_numpy(idx).cuda()].cpu().numpy()
        side["bf16x6_fused"] = {"audio_s_per_s": round(audio_s * n6 / w6, 1),
                                "max_db_err_vs_fp64": round(max_db_err(o6, ref64, fr64), 6),
                                "tiles": {"frames_per_tile": tuned["bf16x6"][0][0],
                                          "blocks_per_cu": tuned["bf16x6"][0][1]}}
        del g6
        # CPU baseline: the oracle's RPP-HOST-equivalent FP32 chain on the host cores, over the
        # whole batch: inputs resident and every output preallocated, only the four C calls timed
        # (pre-emphasis, spectrogram, mel, to_decibels), 2 warm-up runs, then the median of at
        # least 20 runs and about 10 s.  Threads: the CPUs this process may run on, at most the 16
        # a one-GPU share of the host holds.
        affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        nthr = a.cpu_threads or min(16, affinity)
        timer = oracle.LogMelF32Timer(bt.x, bt.lengths, nthreads=nthr, **kw)
        for _ in range(2):
            timer.run()
        reps = []
        t_all = time.perf_counter()
        while len(reps) < 20 or time.perf_counter() - t_all < 10.0:
            t0 = time.perf_counter()
            timer.run()
            reps.append(time.perf_counter() - t0)
        cpu = {"value": round(audio_s / float(np.median(reps)), 1), "unit": "audio-s/s", "cores": nthr, "kind": "port",
               "sample": f"the whole batch ({audio_s:.0f} audio-s), outputs preallocated, the four C calls timed: "
                         f"2 warm-up + {len(reps)} runs, {sum(reps):.1f} s, median run",
               "host_cpus": os.cpu_count(), "affinity_cpus": affinity}
        del timer
        log(f"side: {side}  cpu: {cpu}")

    if not a.no_side:
        side["front_end_workload4"] = front_end_measurement(api, synth, torch, a.precision, steps=20)
        log(f"front end: {side['front_end_workload4']}")
        side["trim_workload1"] = trim_measurement(api, synth, torch, params16k, a.precision)
        log(f"trim: {side['trim_workload1']}")
        for wl in (0, 2):
            if wl != a.workload:
                side[f"workload{wl}"] = workload_measurement(api, synth, torch, wl, a.precision)
                log(f"workload {wl}: {side[f'workload{wl}']}")
    serving = None
    if not a.no_side and a.serve_seconds > 0:
        serving = serve_measurement(api, synth, params16k, a.precision, a.serve_seconds)
        log(f"serving: {serving}")

    line = {
        "metric": "log-mel audio-sec/sec (mixed-precision fused vs FP32, <=0.05 dB vs FP64)",
        "value": round(value, 1),
        "unit": "audio-s/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "settle": {"seconds": a.settle_seconds, "replays": n_settle},
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp16x3": "fp16 hi+lo split, fp32 accumulate", "bf16x3": "bf16 hi+lo split, fp32 accumulate",
                  "bf16x6": "bf16 hi+mid+lo split, fp32 accumulate", "bf16": "bf16, fp32 accumulate",
                  "fp32": "fp32"}[a.precision],
        "data": "synthetic (seeded SplitMix64)",
        "config": {"workload": {0: "batch 8 mono 16 kHz 1-3 s", 1: "batch 512 ragged mono 16 kHz 0.5-20 s",
                                2: "batch 256 mono 44.1 kHz 2-10 s nfft 1024"}[a.workload],
                   "batch": B, "audio_seconds": round(audio_s, 2), "frames": int(frames.sum()),
                   "nfft": feat["nfft"], "window": feat["W"], "hop": H, "mels": feat["nmel"],
                   "precision": a.precision, "launch": (("persistent fused kernel, factored DFT (32 x 16), hipGraph replay" if feat["nfft"] == 512 else
                               "persistent fused kernel, factored DFT (radix 2 over 32 x 16), hipGraph replay")
                              if tuned[a.precision][0][0] == 16
                              else "persistent fused kernel, folded dense DFT, hipGraph replay"),
                   "tiles": {"frames_per_tile": tuned[a.precision][0][0], "blocks_per_cu": tuned[a.precision][0][1]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "side": side,
        "serving": serving,
    }
    print(json.dumps(line))
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
