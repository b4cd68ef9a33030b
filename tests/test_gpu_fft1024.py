"""The factored-DFT fused launch at nfft 1024 (hop 256) through the C-ABI.

A 1024-point frame is two 512-point DFTs (its even and odd samples) on the same MFMA stages as nfft
512, combined radix-2 in registers: X[k] = E[k] + W^k O[k] and X[512 - k] = conj(E[k] - W^k O[k])
for k < 256, X[256] = E[256] - i O[256].  The mel runs over both halves and bin 256.  Checks:

  * against the FP64 oracle at 0.05 dB on ragged clips with the edge lengths (one sample, half a
    window, a window, 8 frames = one work item and one past it), for W = nfft and W < nfft, 128
    filters (eight 16-filter tiles, bin 256 weighted) and 80 (five tiles), Slaney and HTK banks;
    cells past each clip's frames stay untouched
  * the factored launch (framesPerTile 16) and the dense launch (framesPerTile 32) agree within
    0.05 dB and are not bit-identical, so both kernels really ran
  * clips whose samples are not 16-byte aligned (row stride S + 1; the staging's 16-byte buffer
    loads take any 4-byte alignment) give the same bits as aligned ones
  * the shapes the factored nfft-1024 launch does not take (NTF output, power 1, the clip-maximum
    reference) run the dense launch whatever the tiles ask: tiles 16 and 32 give the same bits
  * a batch too large for eight waves' segments beside its clip plan runs with six waves
  * the region and packed entry points give the strided launch's bits
"""
import numpy as np
import pytest
import torch

from helpers import max_db_err
from rpp_amd import synth

pytestmark = pytest.mark.gpu

DB_TOL = 0.05
FILL = -12345.0

CASES = [
    dict(W=1024, nmel=128, sr=44100.0, formula="slaney", normalize=True),
    dict(W=800, nmel=80, sr=16000.0, formula="slaney", normalize=True),
    dict(W=1024, nmel=96, sr=32000.0, formula="htk", normalize=False),
]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def clips(W, H, seed, extra=0):
    rng = np.random.default_rng(seed)
    L = np.array([1, 2, 3, W // 2, W // 2 + 1, W // 2 + 2, W - 1, W, W + 1, W + H, 7 * H + 3, 8 * H, 8 * H + 1,
                  9 * H - 1, 15 * H, 40 * H + 17, 123 * H + 5, 300 * H + 1], np.int32)
    x = np.zeros((len(L), int(L.max()) + 8 + extra), np.float32)
    for b, n in enumerate(L):
        t = np.arange(n) / 16000.0
        x[b, :n] = (0.3 * rng.normal(size=n) + 0.4 * np.sin(2 * np.pi * (200 + 37 * b) * t)).astype(np.float32)
    return x, L


def params(gpu, c, **kw):
    return gpu.logmel_params(nfft=1024, window_length=c["W"], window_step=256, sample_rate=c["sr"], min_freq=0.0,
                             max_freq=c["sr"] / 2, num_filter=c["nmel"], formula=c["formula"],
                             normalize=c["normalize"], **kw)


def run(gpu, x, L, p, nmel, tiles, hs=None, fill=FILL, layout=None):
    frames = L // 256 + 1
    h = gpu.Handle(len(L))
    gpu.set_tiles(h, *tiles)
    F = int(frames.max()) if hs is None else hs
    kw = {}
    if layout is not None:
        kw["layout"] = layout
        out = torch.full((len(L), F, nmel), fill, device="cuda")
    else:
        out = torch.full((len(L), nmel, F), fill, device="cuda")
    gpu.logmel_fused(h, dev(x), out, dev(L), p, "fp16x3", **kw)
    torch.cuda.synchronize()
    h.close()
    return out.cpu().numpy(), frames


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_fft1024_against_fp64(gpu, orc, ci):
    c = CASES[ci]
    x, L = clips(c["W"], 256, seed=ci)
    frames = L // 256 + 1
    hs = int(frames.max()) + 3
    print(f"nfft 1024 factored: W {c['W']} M {c['nmel']} sr {c['sr']:.0f} {c['formula']}, lengths {L.tolist()}",
          flush=True)
    out, _ = run(gpu, x, L, params(gpu, c), c["nmel"], (16, 1), hs=hs)
    for b, F in enumerate(frames):
        assert np.all(out[b, :, F:] == FILL), ("cells past the clip's frames", b)
    ref, _, fr = orc.logmel_f64(x, L, nfft=1024, W=c["W"], H=256, nmel=c["nmel"], sample_rate=c["sr"],
                                fmax=c["sr"] / 2, formula=c["formula"], normalize=c["normalize"], nthreads=8)
    err = max_db_err(out, ref, fr)
    print(f"  max |dB err| vs FP64 = {err:.3e}")
    assert err <= DB_TOL


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_fft1024_factored_vs_dense(gpu, ci):
    c = CASES[ci]
    x, L = clips(c["W"], 256, seed=10 + ci)
    p = params(gpu, c)
    fac, frames = run(gpu, x, L, p, c["nmel"], (16, 1))
    den, _ = run(gpu, x, L, p, c["nmel"], (32, 2))
    d = max(float(np.abs(fac[b, :, :F] - den[b, :, :F]).max()) for b, F in enumerate(frames))
    same = all(np.array_equal(fac[b, :, :F], den[b, :, :F]) for b, F in enumerate(frames))
    print(f"W {c['W']} M {c['nmel']}: factored vs dense max |dB diff| = {d:.3e}")
    assert d <= DB_TOL
    assert not same, "the factored launch did not run (its output equals the dense launch's bit for bit)"


def test_fft1024_alignment_invariance(gpu):
    c = CASES[0]
    x, L = clips(c["W"], 256, seed=21)
    S = x.shape[1]
    x1 = np.zeros((len(L), S + 1), np.float32)
    x1[:, :S] = x
    p = params(gpu, c)
    a, frames = run(gpu, x, L, p, c["nmel"], (16, 1))
    b_, _ = run(gpu, x1, L, p, c["nmel"], (16, 1))
    for b, F in enumerate(frames):
        assert np.array_equal(a[b, :, :F], b_[b, :, :F]), b


@pytest.mark.parametrize("shape", ["ntf", "power1", "clipmax"])
def test_fft1024_general_shapes_run_dense(gpu, orc, shape):
    c = CASES[0]
    x, L = clips(c["W"], 256, seed=31)
    kw = {"power1": dict(power=1), "clipmax": dict(reference_magnitude=0.0)}.get(shape, {})
    p = params(gpu, c, **kw)
    layout = gpu.NTF if shape == "ntf" else None
    a, frames = run(gpu, x, L, p, c["nmel"], (16, 1), layout=layout)
    b_, _ = run(gpu, x, L, p, c["nmel"], (32, 1), layout=layout)
    for b, F in enumerate(frames):
        if shape == "ntf":
            assert np.array_equal(a[b, :F, :], b_[b, :F, :]), b
        else:
            assert np.array_equal(a[b, :, :F], b_[b, :, :F]), b


def test_fft1024_six_waves_for_large_batches(gpu, orc):
    """A batch whose clip plan leaves no room for eight segments beside the 128-mel table (600
    clips: 165 KB at eight waves) runs the factored launch with six waves per workgroup: within
    0.05 dB of FP64 and of the dense launch, and not the dense launch's bits."""
    c = CASES[0]
    rng = np.random.default_rng(41)
    L = rng.integers(1, 12000, 600).astype(np.int32)
    L[:4] = [1, 512, 1024, 11999]
    x = np.zeros((len(L), int(L.max()) + 8), np.float32)
    for b, n in enumerate(L):
        x[b, :n] = (0.3 * rng.normal(size=n)).astype(np.float32)
    p = params(gpu, c)
    fac, frames = run(gpu, x, L, p, c["nmel"], (16, 1))
    den, _ = run(gpu, x, L, p, c["nmel"], (32, 2))
    d = max(float(np.abs(fac[b, :, :F] - den[b, :, :F]).max()) for b, F in enumerate(frames))
    same = all(np.array_equal(fac[b, :, :F], den[b, :, :F]) for b, F in enumerate(frames))
    idx = np.arange(0, len(L), 10)
    ref, _, fr = orc.logmel_f64(x[idx], L[idx], nfft=1024, W=c["W"], H=256, nmel=c["nmel"], sample_rate=c["sr"],
                                fmax=c["sr"] / 2, nthreads=8)
    err = max_db_err(fac[idx], ref, fr)
    print(f"600 clips: factored vs dense {d:.3e} dB, vs FP64 (60 clips) {err:.3e} dB", flush=True)
    assert err <= DB_TOL and d <= DB_TOL
    assert not same, "the factored launch did not run"


def test_fft1024_region_and_packed_entries(gpu):
    """The region entry (per-clip begin/length into each row) and the packed entry (clips back to
    back, outputs as dense per-clip blocks) take the same launch at nfft 1024: bit for bit the
    strided launch on copies of the clips."""
    c = CASES[0]
    x, L = clips(c["W"], 256, seed=51)
    rng = np.random.default_rng(52)
    S = x.shape[1]
    begin = rng.integers(0, 700, len(L)).astype(np.int32)
    wide = np.zeros((len(L), S + 700), np.float32)
    for b, n in enumerate(L):
        wide[b, begin[b]:begin[b] + n] = x[b, :n]
    p = params(gpu, c)
    ref, frames = run(gpu, x, L, p, c["nmel"], (16, 1))
    h = gpu.Handle(len(L))
    gpu.set_tiles(h, 16, 1)
    out = torch.full((len(L), c["nmel"], int(frames.max())), float(FILL), device="cuda")
    gpu.logmel_fused_region(h, dev(wide), out, dev(begin), dev(L), p, "fp16x3")
    src_off = np.concatenate([[0], np.cumsum(L[:-1].astype(np.int64) + 3)]).astype(np.int64)
    flat = np.zeros(int(src_off[-1] + L[-1] + 8), np.float32)
    for b, n in enumerate(L):
        flat[src_off[b]:src_off[b] + n] = x[b, :n]
    dst_off = np.concatenate([[0], np.cumsum(c["nmel"] * frames[:-1].astype(np.int64))]).astype(np.int64)
    dst = torch.full((int(dst_off[-1] + c["nmel"] * frames[-1]),), float(FILL), device="cuda")
    gpu.logmel_fused_packed(h, dev(flat), dev(src_off), dev(L), dst, dev(dst_off), p, "fp16x3")
    torch.cuda.synchronize()
    h.close()
    o, d = out.cpu().numpy(), dst.cpu().numpy()
    for b, F in enumerate(frames):
        assert np.array_equal(o[b, :, :F], ref[b, :, :F]), ("region", b)
        blk = d[dst_off[b]:dst_off[b] + c["nmel"] * F].reshape(c["nmel"], F)
        assert np.array_equal(blk, ref[b, :, :F]), ("packed", b)
