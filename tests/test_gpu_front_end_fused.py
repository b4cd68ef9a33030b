"""rppt_logmel_front_end_gpu: interleaved multi-channel clips at k times the output rate -> down-mix
-> windowed-sinc resample -> the fused log-mel, through the C-ABI.

The one-launch form (each work item of the factored-DFT kernel computes the down-mixed, resampled
samples it reads) must give, bit for bit, what the two-launch form gives (the down-mix + resample
kernel into a buffer, then the fused launch over it) and what the public calls give in sequence
(rppt_down_mix_resample_gpu, then rppt_logmel_fused_gpu on the resampled clips with their lengths
ceil(L * outRate / inRate)).  Lengths cover clips of one sample, clips whose resampled length is at
most half a window (the short-clip staging), the edges where the FIR taps are clipped, and clips of
several seconds; rows whose stride is odd take the per-sample down-mix path.  One case is checked
against the FP64 oracle chain (down-mix and resample restated from RPP HOST, then FP64 log-mel) at
0.05 dB.
"""
import numpy as np
import pytest
import torch

from rpp_amd import synth

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 5, 300, 602, 603, 604, 1000, 4801, 48000, 3 * 48000 + 7, 5 * 48000]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def clips(C, odd_stride, seed):
    rng = np.random.default_rng(seed)
    L = np.array(LENGTHS, np.int32)
    n = int(L.max())
    stride = n * C + (1 if odd_stride else 0)
    x = np.zeros((len(L), stride), np.float32)
    for b, m in enumerate(L):
        x[b, :m * C] = (rng.integers(-32768, 32767, m * C) / 32768.0).astype(np.float32)
    dims = np.stack([L, np.full(len(L), C, np.int32)], 1).reshape(-1).astype(np.int32)
    return x, L, dims


def resampled_lengths(L, in_rate, out_rate):
    return np.ceil(L.astype(np.float32) * np.float32(out_rate) / np.float32(in_rate)).astype(np.int32)


def run_entry(gpu, x, dims, C, in_rate, out_rate, window, p, F, normalize, one_launch):
    h = gpu.Handle(x.shape[0])
    out = torch.full((x.shape[0], 80, F), float("nan"), device="cuda")
    gpu.logmel_front_end(h, dev(x), out, dev(dims), in_rate, out_rate, window, p, "fp16x3", channels=C,
                         normalize=normalize, one_launch=one_launch)
    torch.cuda.synchronize()
    h.close()
    return torch.nan_to_num(out, nan=-7.0).cpu().numpy()


def run_public(gpu, x, L, dims, C, in_rate, out_rate, window, p, F, normalize):
    B = x.shape[0]
    rs = int(np.ceil(x.shape[1] // C * out_rate / in_rate)) + 8
    res = torch.zeros((B, rs), device="cuda")
    h = gpu.Handle(B)
    gpu.down_mix_resample(h, dev(x), res, dev(np.full(B, in_rate, np.float32)), dev(np.full(B, out_rate, np.float32)),
                          dev(dims), window, normalize=normalize, channels=C)
    Lr = resampled_lengths(L, in_rate, out_rate)
    out = torch.full((B, 80, F), float("nan"), device="cuda")
    gpu.logmel_fused(h, res, out, dev(Lr), p, "fp16x3")
    torch.cuda.synchronize()
    h.close()
    return torch.nan_to_num(out, nan=-7.0).cpu().numpy(), res.cpu().numpy(), Lr


@pytest.mark.parametrize("in_rate,C,normalize,odd_stride", [(48000.0, 2, False, False), (32000.0, 2, False, False),
                                                            (48000.0, 2, True, True), (48000.0, 1, False, False),
                                                            (48000.0, 3, False, False)])
def test_front_end_one_launch_equals_two(gpu, in_rate, C, normalize, odd_stride):
    out_rate = 16000.0
    x, L, dims = clips(C, odd_stride, seed=int(in_rate) + C)
    lut, sc, ce = synth.kaiser_sinc_lut()
    window = gpu.make_window(dev(lut), sc, ce, 16, 4091)
    p = gpu.logmel_params()
    Lr = resampled_lengths(L, in_rate, out_rate)
    frames = Lr // 160 + 1
    F = int(frames.max())
    print(f"front end {in_rate:.0f} -> {out_rate:.0f} Hz, C {C}, normalize {normalize}, odd stride {odd_stride}: "
          f"resampled lengths {Lr.tolist()}", flush=True)
    one = run_entry(gpu, x, dims, C, in_rate, out_rate, window, p, F, normalize, True)
    two = run_entry(gpu, x, dims, C, in_rate, out_rate, window, p, F, normalize, False)
    pub, _, _ = run_public(gpu, x, L, dims, C, in_rate, out_rate, window, p, F, normalize)
    for b, Fb in enumerate(frames):
        assert np.array_equal(two[b, :, :Fb], pub[b, :, :Fb]), ("two launches vs public calls", b)
        assert np.array_equal(one[b, :, :Fb], two[b, :, :Fb]), ("one launch vs two", b, int(L[b]))
        assert np.all(np.isfinite(one[b, :, :Fb])), b


def test_front_end_against_fp64_oracle(gpu, orc):
    """Stereo 48 kHz int16-range clips through the one-launch front end: within 0.05 dB of the
    oracle's down-mix -> resample (both bit-exact restatements of RPP HOST) -> FP64 log-mel."""
    in_rate, out_rate, C = 48000.0, 16000.0, 2
    x, L, dims = clips(C, False, seed=7)
    lut, sc, ce = synth.kaiser_sinc_lut()
    window = gpu.make_window(dev(lut), sc, ce, 16, 4091)
    p = gpu.logmel_params()
    Lr = resampled_lengths(L, in_rate, out_rate)
    frames = Lr // 160 + 1
    F = int(frames.max())
    one = run_entry(gpu, x, dims, C, in_rate, out_rate, window, p, F, False, True)
    m_ref = orc.down_mix_f32(x, L, [C] * len(L))
    rs = int(Lr.max()) + 8
    r_ref = orc.resample_f32(m_ref, L, in_rate, out_rate, lut, sc, ce, 16, out_stride=rs)
    ref64, _, fr = orc.logmel_f64(r_ref, Lr, nthreads=8)
    err = max(float(np.abs(one[i, :, :Fb] - ref64[i, :, :Fb]).max()) for i, Fb in enumerate(fr))
    print(f"front end one launch: max |dB err| vs FP64 = {err:.3e}")
    assert err <= 0.05


def test_front_end_graph_replay(gpu):
    """The one-launch front end captured in a hipGraph: every replay equals the eager launch."""
    in_rate, out_rate, C = 48000.0, 16000.0, 2
    x, L, dims = clips(C, False, seed=9)
    lut, sc, ce = synth.kaiser_sinc_lut()
    window = gpu.make_window(dev(lut), sc, ce, 16, 4091)
    p = gpu.logmel_params()
    Lr = resampled_lengths(L, in_rate, out_rate)
    F = int((Lr // 160 + 1).max())
    xs, ds = dev(x), dev(dims)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        h = gpu.Handle(len(L), stream=s)
        out = torch.full((len(L), 80, F), float("nan"), device="cuda")
        gpu.logmel_front_end(h, xs, out, ds, in_rate, out_rate, window, p, "fp16x3", channels=C)
        s.synchronize()
        eager = out.clone()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            gpu.logmel_front_end(h, xs, out, ds, in_rate, out_rate, window, p, "fp16x3", channels=C)
    for i in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(torch.nan_to_num(out, nan=-7.0), torch.nan_to_num(eager, nan=-7.0)), ("replay", i)
    h.close()
