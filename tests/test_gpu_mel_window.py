"""The factored launch's mel phase at the edge of its accumulator window, through the C-ABI.

The factored kernel streams the filter tiles (16 filters each) in ascending frequency through a
window of accumulators: all tiles when the bank has at most five, else six.  A 32-bin chunk whose
filters span more tiles than the window cannot be streamed, and the table builder then sends the bank
to the dense kernel.  128 Slaney filters over nfft 512 put five (48 kHz) and six (192 kHz) tiles on
the first chunk, the window's largest runs; 80 filters at 48 kHz put three of five; 72 and 100 filters
end in a partial tile, whose missing rows are never stored.  Checks, on ragged clips with edge
lengths:

  * the factored launch (framesPerTile 16) is within 0.05 dB of the FP64 oracle, and cells past each
    clip's frames stay untouched
  * it is not bit-identical to the dense launch (framesPerTile 32), so the factored kernel ran
"""
import numpy as np
import pytest
import torch

from helpers import max_db_err

pytestmark = pytest.mark.gpu

DB_TOL = 0.05
FILL = -12345.0
CASES = [
    dict(sr=48000.0, nmel=128),
    dict(sr=192000.0, nmel=128),
    dict(sr=48000.0, nmel=80),
    dict(sr=16000.0, nmel=72),    # the last tile's rows 8-15 do not exist
    dict(sr=48000.0, nmel=100),   # seven tiles, the last one partial
]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def clips(seed):
    rng = np.random.default_rng(seed)
    L = np.array([1, 3, 200, 201, 399, 400, 401, 560, 2563, 2720, 2721, 16000 + 17, 41234], np.int32)
    x = np.zeros((len(L), int(L.max()) + 8), np.float32)
    for b, n in enumerate(L):
        t = np.arange(n) / 48000.0
        x[b, :n] = (0.3 * rng.normal(size=n) + 0.4 * np.sin(2 * np.pi * (150 + 91 * b) * t)).astype(np.float32)
    return x, L


def run(gpu, x, L, c, tiles, hs):
    p = gpu.logmel_params(nfft=512, window_length=400, window_step=160, sample_rate=c["sr"], min_freq=0.0,
                          max_freq=c["sr"] / 2, num_filter=c["nmel"], formula="slaney", normalize=True)
    h = gpu.Handle(len(L))
    gpu.set_tiles(h, *tiles)
    out = torch.full((len(L), c["nmel"], hs), FILL, device="cuda")
    gpu.logmel_fused(h, dev(x), out, dev(L), p, "fp16x3")
    torch.cuda.synchronize()
    h.close()
    return out.cpu().numpy()


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_mel_window_factored_vs_fp64_and_dense(gpu, orc, ci):
    c = CASES[ci]
    x, L = clips(ci)
    frames = L // 160 + 1
    hs = int(frames.max()) + 2
    print(f"nfft 512, {c['nmel']} Slaney filters at {c['sr']:.0f} Hz, lengths {L.tolist()}", flush=True)
    fac = run(gpu, x, L, c, (16, 1), hs)
    den = run(gpu, x, L, c, (32, 2), hs)
    for b, F in enumerate(frames):
        assert np.all(fac[b, :, F:] == FILL), ("cells past the clip's frames", b)
    ref, _, fr = orc.logmel_f64(x, L, nfft=512, W=400, H=160, nmel=c["nmel"], sample_rate=c["sr"],
                                fmax=c["sr"] / 2, nthreads=8)
    err = max_db_err(fac, ref, fr)
    d = max(float(np.abs(fac[b, :, :F] - den[b, :, :F]).max()) for b, F in enumerate(frames))
    same = all(np.array_equal(fac[b, :, :F], den[b, :, :F]) for b, F in enumerate(frames))
    print(f"  factored vs FP64 {err:.3e} dB, vs dense {d:.3e} dB")
    assert err <= DB_TOL and d <= DB_TOL
    assert not same, "the factored launch did not run (its output equals the dense launch's bit for bit)"


def test_mel_window_overflow_runs_dense(gpu, orc):
    """128 filters between 0 and 1 kHz at 16 kHz put all eight tiles on the first 32-bin chunk, more than
    the window holds: the table builder refuses the bank, so framesPerTile 16 runs the dense kernel too
    (bit for bit the framesPerTile 32 launch), within 0.05 dB of FP64."""
    c = dict(sr=16000.0, nmel=128)
    x, L = clips(7)
    frames = L // 160 + 1
    hs = int(frames.max()) + 2

    def launch(tiles):
        p = gpu.logmel_params(nfft=512, window_length=400, window_step=160, sample_rate=c["sr"], min_freq=0.0,
                              max_freq=1000.0, num_filter=c["nmel"], formula="slaney", normalize=True)
        h = gpu.Handle(len(L))
        gpu.set_tiles(h, *tiles)
        out = torch.full((len(L), c["nmel"], hs), FILL, device="cuda")
        gpu.logmel_fused(h, dev(x), out, dev(L), p, "fp16x3")
        torch.cuda.synchronize()
        h.close()
        return out.cpu().numpy()

    a, b_ = launch((16, 1)), launch((32, 1))
    for b, F in enumerate(frames):
        assert np.array_equal(a[b, :, :F], b_[b, :, :F]), b
    ref, _, fr = orc.logmel_f64(x, L, nfft=512, W=400, H=160, nmel=c["nmel"], sample_rate=c["sr"], fmax=1000.0,
                                nthreads=8)
    err = max_db_err(a, ref, fr)
    print(f"128 filters to 1 kHz: dense fallback, {err:.3e} dB vs FP64")
    assert err <= DB_TOL
