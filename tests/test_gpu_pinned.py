"""The reference caller's memory convention: every per-clip array in pinned host memory.

RPP's HIP test driver allocates srcLengthTensor, coefficient arrays, srcDims, rates and the
resampling table with hipHostMalloc and hands those host pointers straight to rppt_*_gpu, whose
kernels read them (utilities/test_suite/HIP/Tensor_audio_hip.cpp:146-176; the resampling table
rpp_test_suite_audio.h:413).  Each entry point here takes the same arrays from pinned host memory
(torch's pin_memory allocation) and must give the bits it gives with device arrays; region
detection also writes its outputs into pinned memory.
"""
import numpy as np
import pytest
import torch

from rpp_amd import synth

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def pinned(a):
    t = torch.empty(a.shape, dtype=torch.from_numpy(np.ascontiguousarray(a)).dtype, pin_memory=True)
    t.copy_(torch.from_numpy(np.ascontiguousarray(a)))
    assert t.is_pinned()
    return t


def both(fn, arrays):
    """fn(*arrays as device tensors) and fn(*arrays as pinned host tensors); returns both results."""
    r_dev = fn(*[dev(a) for a in arrays])
    torch.cuda.synchronize()
    r_pin = fn(*[pinned(a) for a in arrays])
    torch.cuda.synchronize()
    return r_dev, r_pin


def test_pinned_pre_emphasis_spectrogram_mel_db(gpu, handle):
    rng = np.random.default_rng(40)
    L = np.array([16000, 401, 9000, 1], np.int32)
    S = 16008
    x = dev(rng.uniform(-1, 1, (len(L), S)).astype(np.float32))
    coeff = np.full(len(L), 0.97, np.float32)

    def pe(Ld, cd):
        y = torch.full((len(L), S), 3.0, device="cuda")
        gpu.pre_emphasis_filter(handle, x, y, Ld, cd, "clamp")
        return y.cpu()

    a, b = both(pe, [L, coeff])
    assert torch.equal(a, b)
    frames = L // 160 + 1
    F = int(frames.max())

    def spec(Ld):
        y = torch.full((len(L), 257, F), -1.0, device="cuda")
        gpu.spectrogram(handle, x, y, Ld)
        return y

    sa, sb = both(spec, [L])
    assert torch.equal(sa, sb)
    dims = np.stack([np.full(len(L), 257), frames], 1).reshape(-1).astype(np.int32)

    def mel(dd):
        y = torch.full((len(L), 80, F), -1.0, device="cuda")
        gpu.mel_filter_bank(handle, sa, y, dd)
        return y

    ma, mb = both(mel, [dims])
    assert torch.equal(ma, mb)
    ddims = np.stack([frames, np.full(len(L), 80)], 1).reshape(-1).astype(np.uint32)

    def db(dd):
        y = torch.full((len(L), 80, F), -1.0, device="cuda")
        gpu.to_decibels(handle, ma, y, dd, reference_magnitude=0.0)
        return y.cpu()

    da, dbb = both(db, [ddims.view(np.int32)])
    assert torch.equal(da, dbb)


def test_pinned_down_mix_and_resample(gpu, handle):
    rng = np.random.default_rng(41)
    n = 24000
    L = np.array([n, 5000, 3], np.int32)
    x = dev(rng.uniform(-1, 1, (len(L), 2 * n)).astype(np.float32))
    dims = np.stack([L, np.full(len(L), 2, np.int32)], 1).reshape(-1).astype(np.int32)

    def dm(dd):
        y = torch.full((len(L), n), 5.0, device="cuda")
        gpu.down_mixing(handle, x, y, dd, channels=2)
        return y

    ya, yb = both(dm, [dims])
    assert torch.equal(ya, yb)
    lut, sc, ce = synth.kaiser_sinc_lut()
    out_stride = n // 3 + 16
    rin = np.full(len(L), 48000.0, np.float32)
    rout = np.full(len(L), 16000.0, np.float32)
    dims1 = np.stack([L, np.ones(len(L), np.int32)], 1).reshape(-1).astype(np.int32)

    def rs(ri, ro, dd, table):
        w = gpu.make_window(table, sc, ce, 16, 4091)
        y = torch.full((len(L), out_stride), 7.0, device="cuda")
        gpu.resample(handle, ya, y, ri, ro, dd, w)
        return y.cpu()

    ra, rb = both(rs, [rin, rout, dims1, lut])
    assert torch.equal(ra, rb)

    def dmr(ri, ro, dd, table):
        w = gpu.make_window(table, sc, ce, 16, 4091)
        y = torch.full((len(L), out_stride), 7.0, device="cuda")
        gpu.down_mix_resample(handle, x, y, ri, ro, dd, w, channels=2)
        return y.cpu()

    fa, fb = both(dmr, [rin, rout, dims, lut])
    assert torch.equal(fa, fb)
    assert torch.equal(fa, ra)


def test_pinned_non_silent_region(gpu, handle):
    rng = np.random.default_rng(42)
    L = np.array([30000, 5000, 20000], np.int32)
    xs = np.zeros((len(L), 30008), np.float32)
    for i, n in enumerate(L):
        env = np.zeros(n)
        env[n // 4: n // 2] = 1.0
        xs[i, :n] = (rng.normal(size=n) * (0.001 + env)).astype(np.float32)
    x = dev(xs)

    def nsr(Ld, bgn, ln):
        gpu.non_silent_region_detection(handle, x, Ld, bgn, ln, -40.0, 512, 0.0, 1000)
        torch.cuda.synchronize()
        return bgn.cpu().numpy().copy(), ln.cpu().numpy().copy()

    z = np.zeros(len(L), np.int32)
    (ba, la), (bb, lb) = both(nsr, [L, z, z])
    assert np.array_equal(ba, bb) and np.array_equal(la, lb)


def test_pinned_fused_and_region(gpu, handle):
    bt = synth.workload1(batch=12, max_seconds=4.0)
    L = bt.lengths
    x = dev(bt.x)
    frames = L // 160 + 1
    p = gpu.logmel_params()

    def fused(Ld):
        y = torch.full((len(L), 80, int(frames.max())), float("nan"), device="cuda")
        gpu.logmel_fused(handle, x, y, Ld, p, "fp16x3")
        return torch.nan_to_num(y, nan=-7.0).cpu()

    a, b = both(fused, [L])
    assert torch.equal(a, b)
    begin = (L // 5).astype(np.int32)
    reg = (L - begin - 7).astype(np.int32)

    def region(bd, ld):
        y = torch.full((len(L), 80, int(frames.max())), float("nan"), device="cuda")
        gpu.logmel_fused_region(handle, x, y, bd, ld, p, "fp16x3")
        return torch.nan_to_num(y, nan=-7.0).cpu()

    ra, rb = both(region, [begin, reg])
    assert torch.equal(ra, rb)
