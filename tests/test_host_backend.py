"""The HOST backend of librpp_audio.so (rppt_*_host, RPP_HOST_BACKEND handles), on the CPU through
the C-ABI.  Its code is the library's own (rpp_amd/csrc/host_audio.cpp); the oracle is only the
checker here.

  pre-emphasis, mel, to_decibels, resample   the reference's golden vectors bit for bit
  spectrogram                                 the golden spectrogram within 1e-6 of each frame's
                                              maximum (FFTS vs this library's FFT)
  every function                              the oracle's restatement of RPP HOST: bit for bit,
                                              except the spectrogram (1e-6 of the frame maximum)
  the chain                                   within the FP32 parity bar of the FP64 oracle
  region detection                            the reference's hard-coded golden regions
"""
import numpy as np
import pytest

from oracle import oracle as O
from rpp_amd import host as H


@pytest.fixture(scope="module")
def hh():
    h = H.HostHandle(8, threads=4)
    yield h
    h.close()


def valid_mask(shape, lengths):
    m = np.zeros(shape, bool)
    for b, n in enumerate(lengths):
        m[b, :n] = True
    return m


def test_golden_pre_emphasis(hh, golden):
    x, L = golden["x"], golden["lengths"]
    y = np.zeros_like(x)
    H.pre_emphasis_filter(hh, x, y, L, np.full(3, 0.97, np.float32), "clamp")
    m = valid_mask(x.shape, L)
    assert np.array_equal(y[m], golden["pre_emphasis"][m])


def test_golden_to_decibels(hh, golden):
    """The QA case of HOST/Tensor_audio_host.cpp: 1-D clips (width 1), cut-off log(1e-20) dB,
    multiplier log(10), reference 1."""
    x, L = golden["x"], golden["lengths"]
    y = np.zeros_like(x)
    dims = np.zeros(6, np.uint32)
    dims[0::2], dims[1::2] = 1, L
    H.to_decibels(hh, x, y, dims, cutoff_db=float(np.float32(np.log(1e-20))),
                  multiplier=float(np.float32(np.log(10))), reference_magnitude=1.0)
    m = valid_mask(x.shape, L)
    assert np.array_equal(y[m], golden["to_decibels"][m])


def test_golden_spectrogram(hh, golden):
    x, L, F = golden["x"], golden["lengths"], golden["frames"]
    y = np.zeros((3, 257, int(F.max())), np.float32)
    H.spectrogram(hh, x, y, L, nfft=512, window_length=320, window_step=160)
    g = golden["spectrogram"]
    for b, f in enumerate(F):
        a, r = y[b, :, :f].astype(np.float64), g[b, :, :f].astype(np.float64)
        assert (np.abs(a - r) / r.max(axis=0)).max() <= 1e-6


def test_golden_mel(hh, golden):
    spec, F = golden["spectrogram"], golden["frames"]
    y = np.zeros((3, 80, spec.shape[2]), np.float32)
    dims = np.zeros(6, np.int32)
    dims[0::2], dims[1::2] = 257, F
    H.mel_filter_bank(hh, np.ascontiguousarray(spec), y, dims, max_freq=8000.0, min_freq=0.0, num_filter=80,
                      sample_rate=16000.0)
    for b, f in enumerate(F):
        assert np.array_equal(y[b, :, :f], golden["mel"][b, :, :f])


def test_golden_resample(hh, golden):
    x, L = golden["x"], golden["lengths"]
    lut, sc, ce = O.windowed_sinc_lut(1025, 16)
    out_rate = np.float32(16000 * 1.15)
    y = np.zeros((3, golden["resample"].shape[1]), np.float32)
    dims = np.zeros(6, np.int32)
    dims[0::2], dims[1::2] = L, 1
    w = H.make_window(lut, sc, ce, 16, 1025)
    H.resample(hh, x, y, np.full(3, 16000.0, np.float32), np.full(3, out_rate, np.float32), dims, w)
    for b, n in enumerate(L):
        k = O.resample_out_length(int(n), 16000.0, float(out_rate))
        assert np.array_equal(y[b, :k], golden["resample"][b, :k])


def test_golden_non_silent_region(hh, golden):
    """Golden regions of utilities/test_suite/rpp_test_suite_audio.h:64-69 (QA parameters)."""
    x, L = golden["x"], golden["lengths"]
    b, n = np.full(3, -7, np.int32), np.full(3, -7, np.int32)
    H.non_silent_region_detection(hh, x, L, b, n, -60.0, 2048, 0.0, 8192)
    assert list(zip(b.tolist(), n.tolist())) == [(0, 35840), (0, 33680), (0, 34160)]


# ------------------------------------------------------------------ against the oracle
RNG = np.random.default_rng(77)
LENGTHS = np.array([1, 2, 3, 159, 400, 401, 4000, 16000, 23456], np.int32)


def ragged(L, S=None, scale=1.0):
    S = S or int(L.max()) + 8
    x = np.zeros((len(L), S), np.float32)
    for b, n in enumerate(L):
        x[b, :n] = (RNG.uniform(-1, 1, n) * scale).astype(np.float32)
    return x


@pytest.mark.parametrize("border", ["clamp", "reflect", "zero"])
def test_pre_emphasis_vs_oracle(hh, border):
    x = ragged(LENGTHS)
    c = RNG.uniform(0.5, 1.0, len(LENGTHS)).astype(np.float32)
    y = np.full_like(x, 5.0)
    H.pre_emphasis_filter(hh, x, y, LENGTHS, c, border)
    ref = np.zeros_like(x)
    O.lib().orc_pre_emphasis_f32(x, x.shape[1], ref, x.shape[1], LENGTHS, c, len(LENGTHS), O.BORDER[border], 1)
    m = valid_mask(x.shape, LENGTHS)
    assert np.array_equal(y[m], ref[m]) and np.all(y[~m] == 5.0)


@pytest.mark.parametrize("nfft,W,Hop,center,reflect,power,layout", [
    (512, 400, 160, True, True, 2, "NFT"), (512, 400, 160, True, True, 1, "NTF"),
    (256, 256, 100, False, True, 2, "NFT"), (1024, 1024, 256, True, False, 2, "NFT"),
    (400, 400, 160, True, True, 2, "NFT"), (401, 320, 128, True, True, 3, "NTF")])
def test_spectrogram_vs_oracle(hh, nfft, W, Hop, center, reflect, power, layout):
    """Power 1 gives |X|, any other value |X|^2 (the HOST rule); non-power-of-two nfft takes the
    direct DFT.  Against the FP64 oracle within 1e-5 of each frame's maximum; against the FP32
    restatement within 1e-6 of it (1e-5 for a non-power-of-two nfft, whose restatement sums a
    direct DFT in FP32)."""
    L = LENGTHS[LENGTHS >= (0 if center else W)]
    x = ragged(L)
    K = nfft // 2 + 1
    F = np.array([O.num_windows(int(n), W, Hop, center) for n in L])
    shape = (len(L), K, int(F.max())) if layout == "NFT" else (len(L), int(F.max()), K)
    y = np.full(shape, -3.0, np.float32)
    H.spectrogram(hh, x, y, L, nfft=nfft, window_length=W, window_step=Hop, center=center, reflect=reflect,
                  power=power, layout=5 if layout == "NFT" else 6)
    ref = O.spectrogram_f32(x, L, nfft, W, Hop, center, reflect, power, layout)
    ref64 = O.spectrogram_f64(x.astype(np.float64), L, nfft, W, Hop, center, reflect, power, layout)
    pow2 = nfft & (nfft - 1) == 0
    for b, f in enumerate(F):
        a = y[b, :, :f] if layout == "NFT" else y[b, :f, :].T
        for rr, rel in ((ref, 1e-6 if pow2 else 1e-5), (ref64, 1e-5)):
            r = rr[b, :, :f] if layout == "NFT" else rr[b, :f, :].T
            tol = rel * r.max(axis=0) + 1e-30
            assert (np.abs(a.astype(np.float64) - r) / tol).max() <= 1.0, b
        rest = y[b, :, f:] if layout == "NFT" else y[b, f:, :]
        assert np.all(rest == -3.0)


@pytest.mark.parametrize("formula,normalize,fmin,fmax,M", [
    ("slaney", True, 0.0, 8000.0, 80), ("htk", False, 0.0, 8000.0, 64), ("slaney", False, 300.0, 6000.0, 40)])
def test_mel_vs_oracle(hh, formula, normalize, fmin, fmax, M):
    B, K, F = 4, 257, 37
    spec = RNG.uniform(0, 5, (B, K, F)).astype(np.float32)
    frames = np.array([37, 1, 20, 36], np.int32)
    dims = np.zeros(2 * B, np.int32)
    dims[0::2], dims[1::2] = K, frames
    y = np.full((B, M, F), 9.0, np.float32)
    H.mel_filter_bank(hh, spec, y, dims, max_freq=fmax, min_freq=fmin, formula=formula, num_filter=M,
                      sample_rate=16000.0, normalize=normalize)
    ref = O.mel_f32(spec, frames, M, 16000.0, fmin, fmax, formula, normalize)
    for b, f in enumerate(frames):
        assert np.array_equal(y[b, :, :f], ref[b, :, :f]), b


@pytest.mark.parametrize("ref_mag", [1.0, 0.0, 2.5])
def test_to_db_vs_oracle(hh, ref_mag):
    B, M, F = 3, 80, 50
    x = (10.0 ** RNG.uniform(-12, 2, (B, M, F))).astype(np.float32)
    widths = np.array([50, 17, 1], np.uint32)
    dims = np.zeros(2 * B, np.uint32)
    dims[0::2], dims[1::2] = widths, M
    y = np.zeros_like(x)
    H.to_decibels(hh, x, y, dims, reference_magnitude=ref_mag)
    r = O.to_db_f32(x, widths, np.full(B, M), -80.0, 10.0, ref_mag)
    for b, w in enumerate(widths):
        if w == 1:   # 1-D reading: `height` contiguous values
            assert np.array_equal(y[b].reshape(-1)[:M], r[b].reshape(-1)[:M])
        else:
            assert np.array_equal(y[b, :, :w], r[b, :, :w])


@pytest.mark.parametrize("C", [1, 2, 3, 4, 6, 8, 11])
def test_down_mixing_vs_oracle(hh, C):
    L = np.array([1, 7, 1000, 4097], np.int32)
    S = int(L.max()) * C + 8
    x = RNG.uniform(-1, 1, (len(L), S)).astype(np.float32)
    dims = np.zeros(2 * len(L), np.int32)
    dims[0::2], dims[1::2] = L, C
    y = np.zeros((len(L), int(L.max()) + 8), np.float32)
    H.down_mixing(hh, x, y, dims, normalize=C == 3, channels=C)
    ref = O.down_mix_f32(x, L, [C] * len(L), normalize=C == 3)
    for b, n in enumerate(L):
        assert np.array_equal(y[b, :n], ref[b, :n]), b


@pytest.mark.parametrize("in_rate,out_rate,C", [(48000.0, 16000.0, 1), (44100.0, 16000.0, 1),
                                                (8000.0, 16000.0, 1), (48000.0, 16000.0, 2), (16000.0, 16000.0, 1)])
def test_resample_vs_oracle(hh, in_rate, out_rate, C):
    L = np.array([5, 333, 48000, 12345], np.int32)
    x = RNG.uniform(-1, 1, (len(L), int(L.max()) * C + 8)).astype(np.float32)
    lut, sc, ce = O.windowed_sinc_lut(4091, 16)
    n_out = [O.resample_out_length(int(n), in_rate, out_rate) for n in L]
    y = np.zeros((len(L), max(n_out) * C + 8), np.float32)
    dims = np.zeros(2 * len(L), np.int32)
    dims[0::2], dims[1::2] = L, C
    H.resample(hh, x, y, np.full(len(L), in_rate, np.float32), np.full(len(L), out_rate, np.float32), dims,
               H.make_window(lut, sc, ce, 16, 4091))
    ref = O.resample_f32(x, L, in_rate, out_rate, lut, sc, ce, 16, channels=[C] * len(L), out_stride=y.shape[1])
    for b, k in enumerate(n_out):
        assert np.array_equal(y[b, :k * C], ref[b, :k * C]), b


@pytest.mark.parametrize("cutoff,W,ref,reset", [(-60.0, 2048, 0.0, 8192), (-40.0, 512, 0.0, 1000),
                                                (-30.0, 400, 0.01, -1), (-20.0, 1, 0.0, 64)])
def test_non_silent_region_vs_oracle(hh, cutoff, W, ref, reset):
    L = np.array([1, 2, 5, 2047, 2048, 2049, 8193, 30000], np.int32)
    x = ragged(L)
    for b, n in enumerate(L):   # a loud burst inside a quiet clip
        if n > 8:
            x[b, n // 3: n // 2] *= 1000.0
    x *= np.float32(1e-3)
    b, n = np.full(len(L), -7, np.int32), np.full(len(L), -7, np.int32)
    H.non_silent_region_detection(hh, x, L, b, n, cutoff, W, ref, reset)
    eb, en = O.non_silent_region_f32(x, L, cutoff, W, ref, reset)
    assert np.array_equal(b, eb) and np.array_equal(n, en)


def test_chain_vs_fp64(hh):
    """pre-emphasis -> spectrogram -> mel -> to_decibels through the HOST entry points, against the
    FP64 oracle chain: the FP32 parity bar (1e-4 relative in mel power, floor 1e-6 of the frame's
    loudest band)."""
    from helpers import fp32_bar
    from rpp_amd import synth
    bt = synth.workload0()
    x, L = bt.x, bt.lengths
    B, S = x.shape
    frames = L // 160 + 1
    F = int(frames.max())
    pe = np.zeros_like(x)
    spec = np.zeros((B, 257, F), np.float32)
    mel = np.zeros((B, 80, F), np.float32)
    db = np.zeros((B, 80, F), np.float32)
    H.pre_emphasis_filter(hh, x, pe, L, np.full(B, 0.97, np.float32))
    H.spectrogram(hh, pe, spec, L)
    md = np.zeros(2 * B, np.int32)
    md[0::2], md[1::2] = 257, frames
    H.mel_filter_bank(hh, spec, mel, md)
    dd = np.zeros(2 * B, np.uint32)
    dd[0::2], dd[1::2] = frames, 80
    H.to_decibels(hh, mel, db, dd)
    ref64, _, fr = O.logmel_f64(x, L, nthreads=4)
    assert fp32_bar(db, ref64, fr) <= 1.0


def test_host_status_codes(hh):
    x = np.zeros((2, 64), np.float32)
    L = np.array([64, 64], np.int32)
    y = np.zeros((2, 257, 2), np.float32)
    # window longer than nfft, zero multiplier, wrong layout: rejected before any work
    assert H.spectrogram(hh, x, y, L, nfft=256, window_length=400, check_status=False) == -2
    assert H.to_decibels(hh, x, x.copy(), np.ones(4, np.uint32), multiplier=0.0, check_status=False) == -4
    assert H.spectrogram(hh, x, y, L, layout=4, check_status=False) == -10
