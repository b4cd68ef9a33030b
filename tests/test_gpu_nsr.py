"""GPU parity of rppt_non_silent_region_detection_gpu (through the C-ABI) against the oracle's
restatement of the HOST kernel (src/modules/tensor/cpu/kernel/non_silent_region_detection.cpp:93-191)
and the reference's golden outputs (utilities/test_suite/rpp_test_suite_audio.h:64-69).

The outputs are indices, and the moving mean square is evaluated in the HOST order on both
sides, so the bar is exact equality."""
import numpy as np
import pytest
import torch

from rpp_amd import synth

pytestmark = pytest.mark.gpu


def run(gpu, handle, x, L, *args):
    xs = torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
    Ls = torch.from_numpy(np.ascontiguousarray(L, np.int32)).cuda()
    b = torch.full((len(L),), -7, dtype=torch.int32, device="cuda")
    n = torch.full((len(L),), -7, dtype=torch.int32, device="cuda")
    gpu.non_silent_region_detection(handle, xs, Ls, b, n, *args)
    torch.cuda.synchronize()
    return b.cpu().numpy(), n.cpu().numpy()


def test_nsr_golden(gpu, handle, golden):
    """QA parameters of HOST/Tensor_audio_host.cpp:196-206: cutoff -60 dB, window 2048, per-clip
    maximum as reference, reset every 8192 samples."""
    b, n = run(gpu, handle, golden["x"], golden["lengths"], -60.0, 2048, 0.0, 8192)
    assert list(zip(b.tolist(), n.tolist())) == [(0, 35840), (0, 33680), (0, 34160)]


def bursty(rng, L, S):
    x = np.zeros((len(L), S), np.float32)
    for i, l in enumerate(L):
        env = np.full(l, 1e-3)
        if l > 8:
            a, c = sorted(rng.integers(0, l, 2))
            env[a:c + 1] = 1.0
        x[i, :l] = (rng.normal(size=l) * env).astype(np.float32)
    return x


@pytest.mark.parametrize("cutoff,W,ref,reset", [
    (-60.0, 2048, 0.0, 8192), (-40.0, 512, 0.0, 1000), (-30.0, 400, 0.01, 8192),
    (-20.0, 1, 0.0, 64), (-50.0, 300, 0.0, -1), (-45.0, 1024, 0.5, -1), (3.0, 256, 0.0, 4096)])
def test_nsr_vs_oracle_ragged(gpu, handle, orc, cutoff, W, ref, reset):
    rng = np.random.default_rng(abs(hash((cutoff, W, ref, reset))) % 2**32)
    L = np.array([0, 1, 2, 5, 300, 2047, 2048, 2049, 8191, 8192, 8193, 30000, 77777, 160000], np.int32)
    x = bursty(rng, L, 160008)
    b, n = run(gpu, handle, x, L, cutoff, W, ref, reset)
    eb, en = orc.non_silent_region_f32(x, L, cutoff, W, ref, reset, nthreads=8)
    eb[L == 0], en[L == 0] = 0, 0   # HOST reads an empty buffer's maximum there; the GPU reports (0, 0)
    assert np.array_equal(b, eb) and np.array_equal(n, en), (b, eb, n, en)


@pytest.mark.parametrize("stride,offset,cutoff,W,ref,reset", [
    (160003, 0, -40.0, 512, 0.0, 999),      # odd row stride, reset interval not a multiple of 4
    (160008, 1, -45.0, 300, 0.0, -1),       # rows one sample off 16-byte alignment
    (160008, 1, -30.0, 400, 0.01, 8192),    # the same with a fixed reference power
    (160002, 0, -60.0, 2048, 0.0, 8190)])
def test_nsr_vs_oracle_4byte_loads(gpu, handle, orc, stride, offset, cutoff, W, ref, reset):
    """Rows that are not 16-byte aligned, or a reset interval that is not a multiple of 4, take the
    4-byte-load recurrence (and its checkpoints) instead of the 16-byte one."""
    rng = np.random.default_rng(abs(hash((stride, offset, cutoff, W, ref, reset))) % 2**32)
    L = np.array([0, 1, 2, 5, 300, 2047, 2048, 2049, 8191, 8192, 8193, 30000, 77777, 160000], np.int32)
    x = bursty(rng, L, stride)
    flat = torch.zeros(len(L) * stride + offset, dtype=torch.float32, device="cuda")
    xs = flat[offset:].view(len(L), stride)
    xs.copy_(torch.from_numpy(x))
    Ls = torch.from_numpy(L).cuda()
    b = torch.full((len(L),), -7, dtype=torch.int32, device="cuda")
    n = torch.full((len(L),), -7, dtype=torch.int32, device="cuda")
    gpu.non_silent_region_detection(handle, xs, Ls, b, n, cutoff, W, ref, reset)
    torch.cuda.synchronize()
    b, n = b.cpu().numpy(), n.cpu().numpy()
    eb, en = orc.non_silent_region_f32(x, L, cutoff, W, ref, reset, nthreads=8)
    eb[L == 0], en[L == 0] = 0, 0   # HOST reads an empty buffer's maximum there; the GPU reports (0, 0)
    assert np.array_equal(b, eb) and np.array_equal(n, en), (b, eb, n, en)


def test_nsr_silence_and_constant(gpu, handle, orc):
    """All-zero clips report the whole clip (cut-off 0 is met everywhere); a constant burst reports
    the burst widened by W - 1 on both sides."""
    x = np.zeros((3, 20000), np.float32)
    x[0, 8000:9000] = 0.5
    x[2, :] = 0.25
    L = np.array([20000, 20000, 20000], np.int32)
    b, n = run(gpu, handle, x, L, -60.0, 100, 0.0, 4096)
    assert list(zip(b.tolist(), n.tolist())) == [(7901, 1198), (0, 20000), (0, 20000)]


def test_nsr_workload1_full(gpu, handle, orc):
    """Every clip of the 512-clip ragged 16 kHz batch, QA parameters."""
    wl = synth.workload1()
    b, n = run(gpu, handle, wl.x, wl.lengths, -60.0, 2048, 0.0, 8192)
    eb, en = orc.non_silent_region_f32(wl.x, wl.lengths, -60.0, 2048, 0.0, 8192, nthreads=8)
    assert np.array_equal(b, eb) and np.array_equal(n, en)


def test_nsr_workload1_bursts(gpu, handle, orc):
    """The 512-clip batch with every clip made quiet (x 1e-4, -80 dB) outside one random burst, so
    the regions begin and end inside reset blocks and inside the recorded 192-output spans."""
    wl = synth.workload1()
    rng = np.random.default_rng(11)
    x = wl.x.copy()
    for i, l in enumerate(wl.lengths):
        a, c = sorted(rng.integers(0, l, 2))
        x[i, :a] *= np.float32(1e-4)
        x[i, c:l] *= np.float32(1e-4)
    b, n = run(gpu, handle, x, wl.lengths, -60.0, 2048, 0.0, 8192)
    eb, en = orc.non_silent_region_f32(x, wl.lengths, -60.0, 2048, 0.0, 8192, nthreads=8)
    assert np.array_equal(b, eb) and np.array_equal(n, en)
    assert (b > 0).mean() > 0.5 and (b + n < wl.lengths).mean() > 0.5   # interior regions


def test_nsr_status(gpu, handle):
    x = torch.zeros((2, 64), device="cuda")
    L = torch.tensor([64, 64], dtype=torch.int32, device="cuda")
    o = torch.zeros(2, dtype=torch.int32, device="cuda")
    assert gpu.non_silent_region_detection(handle, x, L, o, o, -60.0, 0, 0.0, 16, check_status=False) == -2
    assert gpu.non_silent_region_detection(handle, x, L, o, o, -60.0, 8, 0.0, 0, check_status=False) == -2


@pytest.mark.parametrize("precision", ["fp16x3", "fp32"])
def test_nsr_regions_feed_fused_logmel(gpu, handle, precision):
    """Non-silent regions detected on the GPU go straight into the fused launch
    (rppt_logmel_fused_region_gpu): the output equals the fused launch on copies of the trimmed
    clips, bit for bit; begin/length pairs that leave the row are clamped to it."""
    rng = np.random.default_rng(7)
    L = np.array([40000, 16000, 2049, 90000, 800, 5], np.int32)
    S = int(L.max()) + 16
    x = bursty(rng, L, S)
    xs = torch.from_numpy(x).cuda()
    Ls = torch.from_numpy(L).cuda()
    b = torch.zeros(len(L), dtype=torch.int32, device="cuda")
    n = torch.zeros(len(L), dtype=torch.int32, device="cuda")
    gpu.non_silent_region_detection(handle, xs, Ls, b, n, -30.0, 1024, 0.0, 8192)
    p = gpu.logmel_params()
    nb, nn = b.cpu().numpy(), n.cpu().numpy()
    # add two pairs that reach past the row: the launch clamps them
    nb = np.append(nb, [S - 100, -5]).astype(np.int32)
    nn = np.append(nn, [5000, 300]).astype(np.int32)
    x2 = np.concatenate([x, x[:2]], 0)
    starts = np.clip(nb, 0, S)
    lens = np.maximum(np.minimum(nn, S - starts), 0).astype(np.int32)
    frames = lens // 160 + 1
    cap = int(frames.max())
    out = torch.full((len(nb), 80, cap), float("nan"), device="cuda")
    gpu.logmel_fused_region(handle, torch.from_numpy(x2).cuda(), out, torch.from_numpy(nb).cuda(),
                            torch.from_numpy(nn).cuda(), p, precision)
    trimmed = np.zeros((len(nb), S), np.float32)
    for i in range(len(nb)):
        trimmed[i, :lens[i]] = x2[i, starts[i]:starts[i] + lens[i]]
    ref = torch.full((len(nb), 80, cap), float("nan"), device="cuda")
    gpu.logmel_fused(handle, torch.from_numpy(trimmed).cuda(), ref, torch.from_numpy(lens).cuda(), p, precision)
    torch.cuda.synchronize()
    o, r = out.cpu().numpy(), ref.cpu().numpy()
    for i, F in enumerate(frames):
        assert np.array_equal(o[i, :, :F], r[i, :, :F], equal_nan=True), i
