"""GPU parity of the fused persistent log-mel launch (rppt_logmel_fused_gpu) through the C-ABI.

  FP32 mode   bit-identical to the 4-call drop-in chain (same rounding points); against the FP32
              oracle (the RPP HOST restatement) 1e-4 relative in mel power with a floor at 1e-6 of
              the frame's loudest band (helpers.fp32_bar)
  fp16x3      max |dB - FP64 oracle dB| <= 0.05 dB over every output (values clamp at -80 dB)
  bf16x6      (three-term bf16 split, six products) max |dB - FP64 oracle dB| <= 0.05 dB, like fp16x3
  bf16x3      meets 0.05 dB on workload 0 only (a two-term bf16 split keeps ~16 mantissa bits); on
              workloads 1/2 it is held to its recorded error plus a margin (BF16X3_BOUND_DB), so a
              regression fails
  bf16        one product, reported only
"""
import numpy as np
import pytest
import torch

from helpers import fp32_bar, max_db_err, max_rel_power_err
from rpp_amd import synth

pytestmark = pytest.mark.gpu

DB_TOL_MIXED = 0.05
# bf16x3 against the FP64 oracle: measured 7.9e-2 dB on workload 1[:24] and 1.4e-1 dB on workload
# 2[:6]; the bound is that plus about a third
BF16X3_BOUND_DB = {"workload1[:24]": 0.11, "workload2[:6]": 0.19}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def params_for(gpu, feat):
    return gpu.logmel_params(nfft=feat["nfft"], window_length=feat["W"], window_step=feat["H"],
                             sample_rate=feat["sample_rate"], min_freq=feat["fmin"], max_freq=feat["fmax"],
                             num_filter=feat["nmel"])


def run_fused(gpu, h, x, L, feat, precision, hstride=None, fill=np.nan):
    frames = L // feat["H"] + 1
    hs = int(frames.max()) if hstride is None else hstride
    out = torch.full((len(L), feat["nmel"], hs), float(fill), device="cuda")
    gpu.logmel_fused(h, dev(x), out, dev(L), params_for(gpu, feat), precision)
    torch.cuda.synchronize()
    return out.cpu().numpy(), frames


def test_fused_fp32_equals_chain(gpu, handle):
    bt = synth.workload0()
    L = bt.lengths
    db_chain, fr = gpu.logmel_chain(handle, dev(bt.x), dev(L))
    out, frames = run_fused(gpu, handle, bt.x, L, bt.params, "fp32")
    d = db_chain.cpu().numpy()
    for b, F in enumerate(frames):
        assert np.array_equal(out[b, :, :F], d[b, :, :F]), b


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3", "bf16x6"])
def test_fused_mixed_workload0_vs_fp64(gpu, handle, orc, precision):
    bt = synth.workload0()
    ref, _, frames = orc.logmel_f64(bt.x, bt.lengths, nthreads=8)
    out, fr = run_fused(gpu, handle, bt.x, bt.lengths, bt.params, precision)
    err = max_db_err(out, ref, frames)
    print(f"{precision} workload0 max |dB err| vs FP64 = {err:.3e}")
    assert err <= DB_TOL_MIXED


def check_against_oracles(orc, out, x, L, precision, name, **kw):
    """FP32 mode: fp32_bar against the FP32 oracle; fp16x3 and bf16x6: <= 0.05 dB from the FP64
    oracle; bf16x3: its recorded bound.  All figures are printed."""
    ref64, _, frames = orc.logmel_f64(x, L, nthreads=8, **kw)
    ref32, _ = orc.logmel_f32(x, L, nthreads=8, **kw)
    e64 = max_db_err(out, ref64, frames)
    e32 = max_db_err(out, ref32, frames)
    bar = fp32_bar(out, ref32, frames)
    rel64 = max_rel_power_err(out, ref64, frames)
    print(f"{precision} {name}: max |dB err| vs FP64 = {e64:.3e} (max rel power {rel64:.2e}), "
          f"vs FP32 oracle = {e32:.3e} (fp32 bar {bar:.3f})")
    if precision == "fp32":
        assert bar <= 1.0
    elif precision in ("fp16x3", "bf16x6"):
        assert e64 <= DB_TOL_MIXED
    elif precision == "bf16x3":
        assert e64 <= BF16X3_BOUND_DB[name]
    return e64


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp16x3", "bf16x6"])
def test_fused_workload1_subset(gpu, handle, orc, precision):
    bt = synth.workload1(batch=24)
    out, fr = run_fused(gpu, handle, bt.x, bt.lengths, bt.params, precision)
    check_against_oracles(orc, out, bt.x, bt.lengths, precision, "workload1[:24]")


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp16x3", "bf16x6"])
def test_fused_workload2_subset(gpu, handle, orc, precision):
    bt = synth.workload2(batch=6)
    p = bt.params
    kw = dict(nfft=p["nfft"], W=p["W"], H=p["H"], nmel=p["nmel"], sample_rate=p["sample_rate"], fmax=p["fmax"])
    out, fr = run_fused(gpu, handle, bt.x, bt.lengths, p, precision)
    check_against_oracles(orc, out, bt.x, bt.lengths, precision, "workload2[:6]", **kw)


@pytest.mark.parametrize("ft", [64, 32])
@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3", "bf16x6"])
def test_fused_dense_tile_shapes_workload1(gpu, orc, precision, ft):
    """The dense MFMA kernel at both of its nfft-512 tile shapes (64 frames, each wave two 32-frame
    row tiles of one column tile; 32 frames, one unit per wave), chosen through the handle's tiles
    as the autotuner does: against the oracles as above, and the same bits at 1 and 3 persistent
    blocks per CU."""
    bt = synth.workload1(batch=24)
    outs = []
    for bpc in (1, 3):
        h = gpu.Handle(len(bt.lengths))
        gpu.set_tiles(h, ft, bpc)
        out, frames = run_fused(gpu, h, bt.x, bt.lengths, bt.params, precision)
        h.close()
        outs.append(out)
    for b, F in enumerate(frames):
        assert np.array_equal(outs[0][b, :, :F], outs[1][b, :, :F]), ("blocks per CU", b)
    check_against_oracles(orc, outs[0], bt.x, bt.lengths, precision, "workload1[:24]")


def test_fused_bf16_single_reports(gpu, handle, orc):
    bt = synth.workload0(batch=4)
    ref, _, frames = orc.logmel_f64(bt.x, bt.lengths)
    out, fr = run_fused(gpu, handle, bt.x, bt.lengths, bt.params, "bf16")
    err = max_db_err(out, ref, frames)
    print(f"bf16 (one product) workload0 max |dB err| vs FP64 = {err:.3e}")
    assert np.isfinite(err)


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_fused_ragged_edges(gpu, handle, orc, precision):
    """Tiny and boundary lengths; cells beyond each clip's frames stay untouched."""
    rng = np.random.default_rng(11)
    L = np.array([1, 2, 159, 160, 399, 400, 401, 10239, 10240, 20481, 3], np.int32)
    x = np.zeros((len(L), int(L.max()) + 8), np.float32)
    for b, n in enumerate(L):
        x[b, :n] = rng.uniform(-1, 1, n)
    feat = synth.FEATURES_16K
    hs = int((L // 160 + 1).max()) + 5
    out, frames = run_fused(gpu, handle, x, L, feat, precision, hstride=hs, fill=-12345.0)
    ref32, _ = orc.logmel_f32(x, L)
    ref64, _, _ = orc.logmel_f64(x, L)
    for b, F in enumerate(frames):
        assert np.all(out[b, :, F:] == -12345.0), b
    if precision == "fp32":
        assert fp32_bar(out, ref32, frames) <= 1.0
    else:
        assert max_db_err(out, ref64, frames) <= DB_TOL_MIXED


def test_fused_graph_replay(gpu, orc):
    """The planner + persistent kernel captured in a graph give the eager result on every replay."""
    bt = synth.workload0()
    x, L = dev(bt.x), dev(bt.lengths)
    frames = bt.lengths // 160 + 1
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        h = gpu.Handle(8, stream=s)
        out = torch.zeros((8, 80, int(frames.max())), device="cuda")
        p = params_for(gpu, bt.params)
        gpu.logmel_fused(h, x, out, L, p, "bf16x3")  # warm-up: builds tables, sizes workspace
        s.synchronize()
        eager = out.clone()
        g = torch.cuda.CUDAGraph()
        out.zero_()
        with torch.cuda.graph(g, stream=s):
            gpu.logmel_fused(h, x, out, L, p, "bf16x3")
    for i in range(4):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for b, F in enumerate(frames):
            assert torch.equal(out[b, :, :F], eager[b, :, :F]), (i, b)
    h.close()


def test_fused_status_codes(gpu, handle):
    x = torch.zeros((2, 64), device="cuda")
    L = dev(np.array([64, 64], np.int32))
    out = torch.zeros((2, 80, 1), device="cuda")
    p = gpu.logmel_params(multiplier=0.0)
    assert gpu.logmel_fused(handle, x, out, L, p, "bf16x3", check_status=False) == -4
    p = gpu.logmel_params(reference_magnitude=-1.0)
    assert gpu.logmel_fused(handle, x, out, L, p, "bf16x3", check_status=False) == -2
    p = gpu.logmel_params(max_freq=9000.0)
    assert gpu.logmel_fused(handle, x, out, L, p, "bf16x3", check_status=False) == -2


@pytest.mark.parametrize("precision", ["fp16x3", "fp32"])
def test_fused_grid_and_alignment_invariance(gpu, orc, precision):
    """Every tile is computed the same wherever it runs: the output does not depend on the number of
    persistent blocks, and clips whose samples are not 16-byte aligned (a row stride of S + 1, which
    sends every tile down the generic staging path) give the same bits as aligned ones (the
    prefetched fast path)."""
    bt = synth.workload1(batch=24, max_seconds=6.0)
    L = bt.lengths
    S = bt.x.shape[1]
    x1 = np.zeros((len(L), S + 1), np.float32)
    x1[:, :S] = bt.x
    outs = []
    for bpc, x in ((1, bt.x), (3, bt.x), (1, x1)):
        h = gpu.Handle(len(L))
        gpu.set_tiles(h, 0, bpc)
        out, frames = run_fused(gpu, h, x, L, bt.params, precision)
        h.close()
        outs.append(out)
    for b, F in enumerate(frames):
        assert np.array_equal(outs[0][b, :, :F], outs[1][b, :, :F]), ("blocks per CU", b)
        assert np.array_equal(outs[0][b, :, :F], outs[2][b, :, :F]), ("alignment", b)


@pytest.mark.parametrize("nfft,W,H,nmel,sr", [(256, 256, 128, 40, 8000.0), (768, 768, 192, 96, 24000.0),
                                              (1024, 800, 320, 128, 32000.0), (512, 512, 256, 64, 16000.0)])
@pytest.mark.parametrize("precision", ["fp16x3", "bf16x6"])
def test_fused_mfma_shapes(gpu, handle, orc, nfft, W, H, nmel, sr, precision):
    """Every MFMA kernel instance (1-4 column tiles per wave, 64- and 32-frame tiles) against the
    FP64 oracle at 0.05 dB, on ragged clips including edge lengths, with HTK and unnormalised
    filters on one of them, for both split precisions that meet that bound."""
    rng = np.random.default_rng(nfft + W)
    L = np.array([1, W // 2, W, W + H, 7 * H + 3, 40 * H, 123 * H + 17, 300 * H + 5], np.int32)
    x = np.zeros((len(L), int(L.max()) + 8), np.float32)
    for b, n in enumerate(L):
        t = np.arange(n) / sr
        x[b, :n] = (0.2 * rng.normal(size=n) + 0.5 * np.sin(2 * np.pi * (300 + 40 * b) * t)).astype(np.float32)
    htk = nfft == 768
    feat = dict(nfft=nfft, W=W, H=H, nmel=nmel, sample_rate=sr, fmin=0.0, fmax=sr / 2)
    frames = L // H + 1
    out = torch.full((len(L), nmel, int(frames.max())), float("nan"), device="cuda")
    p = gpu.logmel_params(nfft=nfft, window_length=W, window_step=H, sample_rate=sr, min_freq=0.0, max_freq=sr / 2,
                          num_filter=nmel, formula="htk" if htk else "slaney", normalize=not htk)
    gpu.logmel_fused(handle, dev(x), out, dev(L), p, precision)
    torch.cuda.synchronize()
    ref, _, fr = orc.logmel_f64(x, L, nfft=nfft, W=W, H=H, nmel=nmel, sample_rate=sr, fmax=sr / 2,
                                formula="htk" if htk else "slaney", normalize=not htk, nthreads=8)
    err = max_db_err(out.cpu().numpy(), ref, fr)
    print(f"{precision} nfft {nfft} W {W} H {H} M {nmel}: max |dB err| vs FP64 = {err:.3e}")
    assert err <= DB_TOL_MIXED


def test_fused_clip_max_reference(gpu, handle, orc):
    """referenceMagnitude 0 takes each clip's maximum mel power as the dB reference
    (cpu/kernel/to_decibels.cpp:58-75): FP32 mode equals the chain with the same setting bit for
    bit on every multi-frame clip and meets the FP32 bar against the oracle on all of them; fp16x3
    meets 0.05 dB against the FP64 oracle.  The chain reads a one-frame clip's [numFilter x 1]
    mel output as 1-D data (width 1, cpu/kernel/to_decibels.cpp:63, :77), the fused launch as the 2-D block it
    is, like the oracle's log-mel chain, so that clip is held to the oracle only."""
    bt = synth.workload1(batch=20, max_seconds=5.0)
    L = bt.lengths
    L[3] = 1   # a one-frame clip
    feat = bt.params
    p = gpu.logmel_params(nfft=feat["nfft"], window_length=feat["W"], window_step=feat["H"],
                          sample_rate=feat["sample_rate"], min_freq=feat["fmin"], max_freq=feat["fmax"],
                          num_filter=feat["nmel"], reference_magnitude=0.0)
    frames = L // feat["H"] + 1
    db_chain, _ = gpu.logmel_chain(handle, dev(bt.x), dev(L), ref=0.0)
    d = db_chain.cpu().numpy()
    ref64, _, _ = orc.logmel_f64(bt.x, L, ref=0.0, nthreads=8)
    ref32, _ = orc.logmel_f32(bt.x, L, ref=0.0, nthreads=8)
    for precision in ("fp32", "fp16x3"):
        out = torch.full((len(L), feat["nmel"], int(frames.max())), float("nan"), device="cuda")
        gpu.logmel_fused(handle, dev(bt.x), out, dev(L), p, precision)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        err = max_db_err(o, ref64, frames)
        print(f"{precision} ref=clip max: max |dB err| vs FP64 = {err:.3e}")
        if precision == "fp32":
            for b, F in enumerate(frames):
                if F > 1:
                    assert np.array_equal(o[b, :, :F], d[b, :, :F]), b
            assert fp32_bar(o, ref32, frames) <= 1.0
        else:
            assert err <= DB_TOL_MIXED
        assert np.nanmax(o) <= 0.0   # every clip's loudest output is at its own reference


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_fused_ntf_layout(gpu, handle, precision):
    """NTF output (frames as rows) holds exactly the NFT output transposed, also with a row stride
    above numFilter and fewer rows than the longest clip's frames (those frames are not written)."""
    bt = synth.workload1(batch=16, max_seconds=4.0)
    L = bt.lengths
    feat = bt.params
    M = feat["nmel"]
    p = params_for(gpu, feat)
    frames = L // feat["H"] + 1
    cap = int(frames.max()) - 7
    nft = torch.full((len(L), M, cap), float("nan"), device="cuda")
    ntf = torch.full((len(L), cap, M + 3), float("nan"), device="cuda")
    gpu.logmel_fused(handle, dev(bt.x), nft, dev(L), p, precision)
    gpu.logmel_fused(handle, dev(bt.x), ntf, dev(L), p, precision, layout=gpu.NTF)
    torch.cuda.synchronize()
    a, t = nft.cpu().numpy(), ntf.cpu().numpy()
    for b, F in enumerate(np.minimum(frames, cap)):
        assert np.array_equal(t[b, :F, :M], a[b, :, :F].T), b
        assert np.isnan(t[b, F:, :]).all() and np.isnan(t[b, :, M:]).all(), b


@pytest.mark.parametrize("precision", ["fp16x3", "fp32"])
def test_fused_power_of_two_scale_invariance(gpu, handle, precision):
    """With the clip maximum as reference, scaling a clip by 2^-40 changes no output bit: the fp16
    operands are normalised by a power of two per tile and the power is rescaled exactly (far from
    FP32 underflow at this scale).  A silent clip and one of amplitude ~1e-36 (the operand scale at
    its cap) give finite outputs at the -80 dB floor."""
    bt = synth.workload0(batch=3)
    L = bt.lengths.copy()
    feat = bt.params
    x = bt.x.copy()
    xs = np.ldexp(x, -40).astype(np.float32)
    p = gpu.logmel_params(nfft=feat["nfft"], window_length=feat["W"], window_step=feat["H"],
                          sample_rate=feat["sample_rate"], min_freq=feat["fmin"], max_freq=feat["fmax"],
                          num_filter=feat["nmel"], reference_magnitude=0.0)
    frames = L // feat["H"] + 1
    outs = []
    for src in (x, xs):
        out = torch.full((len(L), feat["nmel"], int(frames.max())), float("nan"), device="cuda")
        gpu.logmel_fused(handle, dev(src), out, dev(L), p, precision)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    for b, F in enumerate(frames):
        assert np.array_equal(outs[0][b, :, :F], outs[1][b, :, :F]), b
    quiet = np.zeros_like(x)
    quiet[1] = (x[1] * np.float32(1e-36)).astype(np.float32)
    p1 = params_for(gpu, feat)
    out = torch.full((len(L), feat["nmel"], int(frames.max())), float("nan"), device="cuda")
    gpu.logmel_fused(handle, dev(quiet), out, dev(L), p1, precision)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    for b in (0, 1):
        F = frames[b]
        assert np.all(np.isfinite(o[b, :, :F])), b
        assert np.all(o[b, :, :F] == o[b, :, :F].min()) and o[b, :, :F].min() <= -79.99, b


def test_fused_tile_records_follow_the_batch(gpu):
    """One handle, two batches back to back whose clip lengths differ (the fused launch plans its
    tiles in each block and keeps their records at the same workspace addresses): the second launch
    equals a fresh handle's launch of the same batch bit for bit, so no record of the first batch
    survives into the second."""
    bt = synth.workload1(batch=24, max_seconds=6.0)
    L1 = bt.lengths
    L2 = np.ascontiguousarray(L1[::-1])
    h = gpu.Handle(len(L1))
    run_fused(gpu, h, bt.x, L1, bt.params, "fp16x3")
    o2, frames = run_fused(gpu, h, bt.x, L2, bt.params, "fp16x3")
    h.close()
    h2 = gpu.Handle(len(L1))
    r2, _ = run_fused(gpu, h2, bt.x, L2, bt.params, "fp16x3")
    h2.close()
    for b, F in enumerate(frames):
        assert np.array_equal(o2[b, :, :F], r2[b, :, :F]), b


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_fused_power_rule(gpu, handle, precision):
    """The fused launch reads `power` as rppt_spectrogram_gpu does (hip/kernel/spectrogram.cpp:231-232):
    2 is the squared magnitude, any other value the magnitude, so power 0, 1 and 3 give the same
    bits and differ from power 2."""
    bt = synth.workload0(batch=3)
    feat = bt.params
    frames = bt.lengths // feat["H"] + 1
    outs = {}
    for pw in (0, 1, 2, 3):
        p = gpu.logmel_params(power=pw)
        out = torch.full((3, 80, int(frames.max())), float("nan"), device="cuda")
        gpu.logmel_fused(handle, dev(bt.x), out, dev(bt.lengths), p, precision)
        torch.cuda.synchronize()
        outs[pw] = out.cpu().numpy()
    for b, F in enumerate(frames):
        assert np.array_equal(outs[0][b, :, :F], outs[1][b, :, :F])
        assert np.array_equal(outs[3][b, :, :F], outs[1][b, :, :F])
        assert not np.array_equal(outs[2][b, :, :F], outs[1][b, :, :F])
