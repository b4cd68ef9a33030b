"""GPU parity of the front-end entry points (rppt_down_mixing_gpu, rppt_resample_gpu and the fused
rppt_down_mix_resample_gpu) through the C-ABI: bit-exact against the oracle's restatement of RPP
HOST and against the reference's golden resample vectors (16 kHz -> 18.4 kHz, Hann-windowed sinc,
16 lobes, 1025 coefficients).  The resample cases cover each path of the kernel: an integer ratio
(taps looked up once, FIR), non-integer down- and up-sampling from the staged span, equal rates
(copy), more than eight inputs per output and more than 32 lobes (taps read from memory), and
multi-channel input resampled per channel."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("C,normalize", [(1, False), (2, False), (2, True), (3, False), (4, False), (7, True),
                                         (8, False), (9, False), (16, True)])
def test_down_mixing(gpu, handle, orc, C, normalize):
    rng = np.random.default_rng(C)
    n = 3001
    L = np.array([n, 17, 1, n - 5], np.int32)
    x = rng.uniform(-1, 1, (len(L), n * C)).astype(np.float32)
    dims = np.stack([L, np.full(len(L), C, np.int32)], 1).reshape(-1).astype(np.int32)
    y = torch.full((len(L), n), 9.0, device="cuda")
    gpu.down_mixing(handle, dev(x), y, dev(dims), normalize=normalize, channels=C)
    ref = orc.down_mix_f32(x, L, np.full(len(L), C), normalize)
    y = y.cpu().numpy()
    for b, m in enumerate(L):
        assert np.array_equal(y[b, :m], ref[b, :m]), b
        assert np.all(y[b, m:] == 9.0)


@pytest.mark.parametrize("normalize", [False, True])
def test_down_mixing_stereo_vectorised(gpu, handle, orc, normalize):
    """Rows with 16-byte aligned, multiple-of-4 strides take the four-samples-per-thread stereo
    kernel (tails of 1-3 samples included); a mono row in the same batch takes the generic one."""
    rng = np.random.default_rng(11)
    n = 4000
    L = np.array([n, 4001 - 1, 1, 6, 2999, 3000], np.int32)
    C = np.array([2, 2, 2, 2, 1, 2], np.int32)
    x = rng.uniform(-1, 1, (len(L), 2 * n)).astype(np.float32)
    dims = np.stack([L, C], 1).reshape(-1).astype(np.int32)
    y = torch.full((len(L), n), 9.0, device="cuda")
    gpu.down_mixing(handle, dev(x), y, dev(dims), normalize=normalize, channels=2)
    y = y.cpu().numpy()
    for b, (m, c) in enumerate(zip(L, C)):
        ref = orc.down_mix_f32(x[b:b + 1], L[b:b + 1], [c], normalize)
        assert np.array_equal(y[b, :m], ref[0, :m]), b
        assert np.all(y[b, m:] == 9.0), b


def window_for(coeffs, lobes):
    from oracle import oracle as O
    lut, sc, ce = O.windowed_sinc_lut(coeffs, lobes)
    return lut, sc, ce


def run_resample(gpu, handle, x, L, in_rate, out_rate, lut, sc, ce, lobes, coeffs, out_stride, C=1):
    B = x.shape[0]
    lut_d = dev(lut)
    w = gpu.make_window(lut_d, sc, ce, lobes, coeffs)
    dims = np.stack([L, np.full(B, C, np.int32)], 1).reshape(-1).astype(np.int32)
    y = torch.full((B, out_stride), 7.0, device="cuda")
    src_desc = gpu.desc_1d(B, x.shape[1])
    gpu.resample(handle, dev(x), y, dev(np.full(B, in_rate, np.float32)), dev(np.full(B, out_rate, np.float32)),
                 dev(dims), w)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def test_resample_golden(gpu, handle, golden, orc):
    x, L = golden["x"], golden["lengths"]
    lut, sc, ce = window_for(1025, 16)
    out_rate = float(np.float32(16000 * 1.15))
    g = golden["resample"]
    y = run_resample(gpu, handle, x, L, 16000.0, out_rate, lut, sc, ce, 16, 1025, g.shape[1])
    for b, n in enumerate(L):
        k = orc.resample_out_length(int(n), 16000.0, out_rate)
        assert np.array_equal(y[b, :k], g[b, :k]), b
        assert np.all(y[b, k:] == 7.0)


@pytest.mark.parametrize("in_rate,out_rate", [(48000.0, 16000.0), (44100.0, 16000.0), (16000.0, 16000.0),
                                              (8000.0, 16000.0)])
def test_resample_rates(gpu, handle, orc, in_rate, out_rate):
    rng = np.random.default_rng(int(in_rate))
    L = np.array([48000, 4097, 257, 3], np.int32)
    x = np.zeros((len(L), 48008), np.float32)
    for b, n in enumerate(L):
        x[b, :n] = rng.uniform(-1, 1, n)
    lut, sc, ce = window_for(4091, 16)
    out_stride = max(orc.resample_out_length(int(n), in_rate, out_rate) for n in L) + 8
    y = run_resample(gpu, handle, x, L, in_rate, out_rate, lut, sc, ce, 16, 4091, out_stride)
    ref = orc.resample_f32(x, L, in_rate, out_rate, lut, sc, ce, 16, out_stride=out_stride)
    for b, n in enumerate(L):
        k = orc.resample_out_length(int(n), in_rate, out_rate)
        assert np.array_equal(y[b, :k], ref[b, :k]), (b, np.abs(y[b, :k] - ref[b, :k]).max())


def test_resample_stereo(gpu, handle, orc):
    rng = np.random.default_rng(21)
    L = np.array([6000, 1000], np.int32)
    x = rng.uniform(-1, 1, (2, 12000)).astype(np.float32)
    lut, sc, ce = window_for(1025, 16)
    out_stride = 2 * orc.resample_out_length(6000, 48000.0, 16000.0) + 8
    y = run_resample(gpu, handle, x, L, 48000.0, 16000.0, lut, sc, ce, 16, 1025, out_stride, C=2)
    ref = orc.resample_f32(x, L, 48000.0, 16000.0, lut, sc, ce, 16, channels=[2, 2], out_stride=out_stride)
    for b, n in enumerate(L):
        k = 2 * orc.resample_out_length(int(n), 48000.0, 16000.0)
        assert np.array_equal(y[b, :k], ref[b, :k]), b


def test_workload4_front_end(gpu, handle, orc):
    """Stereo int16-range 48 kHz clips -> down-mix + Kaiser-sinc resample to 16 kHz (one launch) ->
    fused log-mel: the resampled signal bit-exact against the oracle's down-mix then resample, the
    dB output within 0.05 dB of FP64."""
    from rpp_amd import synth
    bt = synth.workload4(batch=2, seconds=3.0)
    lut, sc, ce = synth.kaiser_sinc_lut()
    x = dev(bt.x)
    L = dev(bt.lengths)
    p = gpu.logmel_params()
    db, res = gpu.front_end_logmel(handle, x, L, 2, dev(lut), sc, ce, 16, 4091, 48000.0, 16000.0, p)
    torch.cuda.synchronize()
    m_ref = orc.down_mix_f32(bt.x, bt.lengths, [2, 2])
    n_out = orc.resample_out_length(int(bt.lengths[0]), 48000.0, 16000.0)
    r_ref = orc.resample_f32(m_ref, bt.lengths, 48000.0, 16000.0, lut, sc, ce, 16, out_stride=res.shape[1])
    r = res.cpu().numpy()
    assert np.array_equal(r[:, :n_out], r_ref[:, :n_out])
    ref64, _, fr = orc.logmel_f64(r_ref[:, :n_out], np.full(2, n_out, np.int32), nthreads=8)
    d = db.cpu().numpy()
    err = max(float(np.abs(d[b, :, :F] - ref64[b, :, :F]).max()) for b, F in enumerate(fr))
    print(f"workload4 front end: max |dB err| vs FP64 = {err:.3e}")
    assert err <= 0.05


@pytest.mark.parametrize("in_rate,out_rate,lobes", [(96000.0, 8000.0, 16), (48000.0, 16000.0, 40),
                                                    (44100.0, 16000.0, 40)])
def test_resample_unstaged_paths(gpu, handle, orc, in_rate, out_rate, lobes):
    """Ratios above 8 inputs per output and windows above 32 lobes read their taps from memory."""
    rng = np.random.default_rng(int(in_rate) + lobes)
    L = np.array([20000, 4097, 100, 3], np.int32)
    x = np.zeros((len(L), 20008), np.float32)
    for b, n in enumerate(L):
        x[b, :n] = rng.uniform(-1, 1, n)
    lut, sc, ce = window_for(4091, lobes)
    out_stride = max(orc.resample_out_length(int(n), in_rate, out_rate) for n in L) + 8
    y = run_resample(gpu, handle, x, L, in_rate, out_rate, lut, sc, ce, lobes, 4091, out_stride)
    ref = orc.resample_f32(x, L, in_rate, out_rate, lut, sc, ce, lobes, out_stride=out_stride)
    for b, n in enumerate(L):
        k = orc.resample_out_length(int(n), in_rate, out_rate)
        assert np.array_equal(y[b, :k], ref[b, :k]), (b, np.abs(y[b, :k] - ref[b, :k]).max())


@pytest.mark.parametrize("C,normalize,in_rate,out_rate,lobes", [
    (2, False, 48000.0, 16000.0, 16), (2, True, 44100.0, 16000.0, 16), (3, False, 8000.0, 16000.0, 16),
    (2, False, 16000.0, 16000.0, 16), (9, True, 48000.0, 16000.0, 15), (2, False, 96000.0, 8000.0, 16),
    (4, False, 32000.0, 16000.0, 40)])
def test_down_mix_resample_fused(gpu, handle, orc, C, normalize, in_rate, out_rate, lobes):
    """The fused launch equals rppt_down_mixing_gpu then rppt_resample_gpu bit for bit, and both the
    oracle's down-mix then resample (odd lobes leave two taps for the per-tap tail)."""
    rng = np.random.default_rng(C * 1000 + lobes)
    n = 30000
    L = np.array([n, 4099, 200, 5], np.int32)
    x = rng.uniform(-1, 1, (len(L), n * C)).astype(np.float32)
    dims = np.stack([L, np.full(len(L), C, np.int32)], 1).reshape(-1).astype(np.int32)
    lut, sc, ce = window_for(4091, lobes)
    w = gpu.make_window(dev(lut), sc, ce, lobes, 4091)
    out_stride = max(orc.resample_out_length(int(m), in_rate, out_rate) for m in L) + 8
    rin = dev(np.full(len(L), in_rate, np.float32))
    rout = dev(np.full(len(L), out_rate, np.float32))
    fused = torch.full((len(L), out_stride), 7.0, device="cuda")
    gpu.down_mix_resample(handle, dev(x), fused, rin, rout, dev(dims), w, normalize=normalize, channels=C)
    mono = torch.zeros((len(L), n), device="cuda")
    gpu.down_mixing(handle, dev(x), mono, dev(dims), normalize=normalize, channels=C)
    dims1 = dev(np.stack([L, np.ones(len(L), np.int32)], 1).reshape(-1).astype(np.int32))
    two = torch.full((len(L), out_stride), 7.0, device="cuda")
    gpu.resample(handle, mono, two, rin, rout, dims1, w)
    torch.cuda.synchronize()
    f, t = fused.cpu().numpy(), two.cpu().numpy()
    m_ref = orc.down_mix_f32(x, L, np.full(len(L), C), normalize)
    ref = orc.resample_f32(m_ref, L, in_rate, out_rate, lut, sc, ce, lobes, out_stride=out_stride)
    for b, m in enumerate(L):
        k = orc.resample_out_length(int(m), in_rate, out_rate) if in_rate != out_rate else int(m)
        assert np.array_equal(f[b, :k], t[b, :k]), ("fused vs two calls", b)
        assert np.array_equal(f[b, :k], ref[b, :k]), ("fused vs oracle", b)
        assert np.all(f[b, k:] == 7.0), b
