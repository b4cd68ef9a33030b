"""The continuous-batching server through the C-ABI: every clip's result is the same alone and in
the mix (bit for bit: work items never span clips), within 0.05 dB of the FP64 oracle (fp16x3),
admission refuses a clip that cannot fit (or grows a slot for it against the free device memory),
and a non-eager batch closes at its deadline."""
import time

import numpy as np
import pytest
import torch

from rpp_amd import synth

pytestmark = pytest.mark.gpu


def clips(n, seed=3, lo=4000, hi=160000):
    s = synth.Stream(seed)
    L = synth.log_uniform_lengths(s, n, lo, hi)
    return [synth.clip_for_request(i, int(l), seed) for i, l in enumerate(L)]


def alone(gpu, x):
    """The clip by itself through the 2-D fused entry point."""
    h = gpu.Handle(1)
    F = len(x) // 160 + 1
    out = torch.zeros((1, 80, F), device="cuda")
    gpu.logmel_fused(h, torch.from_numpy(x[None]).cuda(), out, torch.tensor([len(x)], dtype=torch.int32).cuda(),
                     gpu.logmel_params(), "fp16x3")
    torch.cuda.synchronize()
    h.close()
    return out[0].cpu().numpy()


def test_same_alone_and_in_the_mix(gpu, orc):
    from rpp_amd.serve import Server
    xs = clips(40)
    with Server(gpu.logmel_params(), "fp16x3", max_batch=8, deadline_us=2000, slots=3) as srv:
        ids = []
        for i, x in enumerate(xs):          # staggered arrivals: batches overlap on the device
            ids.append(srv.submit(x))
            if i % 5 == 4:
                time.sleep(0.0005)
        outs = [srv.wait(r, 80, len(x) // 160 + 1)[0] for r, x in zip(ids, xs)]
        st = srv.stats()
    assert st["completed"] == len(xs) and st["refused"] == 0
    for x, o in zip(xs, outs):
        assert np.array_equal(o, alone(gpu, x))
    L = np.array([len(x) for x in xs[:8]], np.int32)
    pad = np.zeros((8, int(L.max()) + 8), np.float32)
    for b, x in enumerate(xs[:8]):
        pad[b, :len(x)] = x
    ref, _, fr = orc.logmel_f64(pad, L, nthreads=8)
    err = max(float(np.abs(outs[b] - ref[b, :, :F]).max()) for b, F in enumerate(fr))
    assert err <= 0.05, err


def test_admission_refuses_what_cannot_fit(gpu):
    from rpp_amd.serve import Server
    with Server(gpu.logmel_params(), "fp16x3", max_batch=4, mem_budget=1 << 20, slots=1, grow_slots=False) as srv:
        cap = srv.stats()["slotCapacitySamples"]
        assert 1000 < cap < 1 << 20
        big = np.zeros(cap + 1000, np.float32)
        assert srv.submit(big, check_status=False) == -16      # RPP_ERROR_NOT_ENOUGH_MEMORY
        rid = srv.submit(np.ones(2000, np.float32))
        o, lat = srv.wait(rid, 80, 2000 // 160 + 1)
        assert np.isfinite(o).all()
        assert srv.stats()["refused"] == 1


def test_admission_grows_a_slot_for_a_long_clip(gpu):
    """With growSlots a clip longer than the slots were sized for is admitted against the free device
    memory at submit: an idle slot is regrown for it, its result equals the clip alone, and the
    clips before and after it still run (one host copy each, through the pinned arena)."""
    from rpp_amd.serve import Server
    with Server(gpu.logmel_params(), "fp16x3", max_batch=4, deadline_us=2000, mem_budget=1 << 20, slots=2) as srv:
        cap = srv.stats()["slotCapacitySamples"]
        xs = clips(3, seed=21, lo=2000, hi=8000)
        long_clip = synth.clip_for_request(99, int(cap) + 20000, 21)
        seq = [xs[0], long_clip, xs[1], xs[2]]
        ids = [srv.submit(x) for x in seq]
        outs = [srv.wait(r, 80, len(x) // 160 + 1)[0] for r, x in zip(ids, seq)]
        st = srv.stats()
    assert st["refused"] == 0 and st["completed"] == 4
    assert st["slotGrowths"] >= 1 and st["slotCapacitySamples"] >= len(long_clip)
    assert st["arenaSubmits"] + st["stagedSubmits"] == 4
    for x, o in zip(seq, outs):
        assert np.array_equal(o, alone(gpu, x))


def test_deadline_closes_a_partial_batch(gpu):
    from rpp_amd.serve import Server
    with Server(gpu.logmel_params(), "fp16x3", max_batch=16, deadline_us=3000, slots=2, eager=False) as srv:
        ids = [srv.submit(x) for x in clips(3, seed=9, hi=20000)]
        lats = [srv.wait(r, 80, 20000 // 160 + 1)[1] for r in ids]
        st = srv.stats()
    assert st["batches"] == 1                # three clips, fewer than maxBatch: one batch at the deadline
    assert min(lats) >= 2500, lats


def test_wait_without_kept_outputs(gpu):
    """keepOutputs = 0: several threads may wait on one request, a request that finished before its
    wait reports frames -1, and the server never frees a request a waiter still holds."""
    import threading
    from rpp_amd.serve import Server
    xs = clips(24, seed=11, hi=40000)
    with Server(gpu.logmel_params(), "fp16x3", max_batch=8, deadline_us=2000, slots=2,
                keep_outputs=False) as srv:
        ids = [srv.submit(x) for x in xs]
        got = {}

        def waiter(tag):
            got[tag] = [srv.wait_timing(r) for r in ids]

        ts = [threading.Thread(target=waiter, args=(k,)) for k in range(3)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(60)
        assert not any(t.is_alive() for t in ts)
        srv.drain()
        for k in range(3):
            for (fr, lat), x in zip(got[k], xs):
                # a waiter that arrived after the request was forgotten sees -1
                assert fr in (-1, len(x) // 160 + 1)
        late = srv.wait_timing(ids[0])
        assert late == (-1, -1.0)
        st = srv.stats()
    assert st["completed"] == len(xs)
