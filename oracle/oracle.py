"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
import this module, and only as the checker / the timed CPU baseline.  The
product library never calls into it.

Two libraries sit behind it:

* ``liboracle.so`` -- the project's C restatement of RPP's HOST audio path
  (``rpp_oracle.c``): FP32 functions that follow
  ``src/modules/tensor/cpu/kernel/{pre_emphasis_filter,spectrogram,mel_filter_bank,
  to_decibels,down_mixing,resample}.cpp`` and an FP64 oracle of the same maths.
* ``_ref/libref_audio.so`` (optional; buildable only where the reference sources are,
  and present wherever the built tree is copied) -- RPP's own HOST kernels compiled from
  the reference sources; used only to check the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
L = C.c_long
I = C.c_int
F = C.c_float
D = C.c_double

BORDER = {"zero": 0, "clamp": 1, "reflect": 2}
MEL = {"slaney": 0, "htk": 1}


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and, when /root/reference exists and ref=True, _ref/)."""
    targets = ["all"]
    if ref and os.path.isdir("/root/reference"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, "-j8", *targets], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        l = C.CDLL(path)
        l.orc_pre_emphasis_f32.argtypes = [f32p, L, f32p, L, i32p, f32p, I, I, I]
        l.orc_pre_emphasis_f64.argtypes = [f32p, L, f64p, L, i32p, f32p, I, I]
        l.orc_hann_f32.argtypes = [I, f32p]
        l.orc_hann_f64.argtypes = [I, f64p]
        l.orc_num_windows.argtypes = [I, I, I, I]
        l.orc_num_windows.restype = I
        l.orc_idx_reflect.argtypes = [I, I, I]
        l.orc_idx_reflect.restype = I
        l.orc_spectrogram_f32.argtypes = [f32p, L, i32p, I, f32p, L, L, I, I, I, C.c_void_p, I, I, I, I, I]
        l.orc_spectrogram_f64.argtypes = [f64p, L, i32p, I, f64p, L, L, I, I, I, C.c_void_p, I, I, I, I, I]
        l.orc_mel_tables_f32.argtypes = [I, F, F, F, I, I, I, f32p, i32p, f32p, i32p]
        l.orc_mel_f32.argtypes = [f32p, L, L, i32p, I, f32p, L, L, F, F, I, I, F, I, I]
        l.orc_mel_f64.argtypes = [f64p, L, L, i32p, I, f64p, L, L, D, D, I, I, D, I, I]
        l.orc_to_db_f32.argtypes = [f32p, L, L, f32p, L, L, u32p, I, F, F, F, I, I]
        l.orc_to_db_f64.argtypes = [f64p, L, L, f64p, L, L, u32p, I, D, D, D, I]
        l.orc_down_mix_f32.argtypes = [f32p, L, f32p, L, i32p, I, I, I]
        l.orc_resample_f32.argtypes = [f32p, L, f32p, L, f32p, f32p, i32p, I, F, F, I, I, f32p, I]
        l.orc_windowed_sinc_lut.argtypes = [I, I, f32p, C.POINTER(F), C.POINTER(F)]
        l.orc_nsr_f32.argtypes = [f32p, L, i32p, I, F, I, F, I, i32p, i32p, I]
        l.orc_num_threads_available.restype = I
        _LIB = l
    return _LIB


def ref_lib():
    """RPP's own HOST kernels (None when they were not built here)."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_audio.so")
        if not os.path.exists(path):
            return None
        r = C.CDLL(path)
        U = C.c_uint
        r.ref_pre_emphasis.argtypes = [f32p, U, f32p, U, i32p, f32p, I, I, I]
        r.ref_mel_filter_bank.argtypes = [f32p, U, U, U, i32p, f32p, U, U, I, F, F, I, I, F, I, I]
        r.ref_to_decibels.argtypes = [f32p, U, U, f32p, U, U, u32p, I, F, F, F, I]
        r.ref_down_mixing.argtypes = [f32p, U, U, f32p, U, i32p, I, I, I]
        r.ref_resample.argtypes = [f32p, U, f32p, U, f32p, f32p, i32p, I, F, F, I, I, I, f32p, I]
        r.ref_non_silent_region_detection.argtypes = [f32p, U, i32p, I, F, I, F, I, i32p, i32p, I]
        for fn in ("ref_pre_emphasis", "ref_mel_filter_bank", "ref_to_decibels", "ref_down_mixing", "ref_resample",
                   "ref_non_silent_region_detection"):
            getattr(r, fn).restype = I
        _REF = r
    return _REF


def threads() -> int:
    return max(1, lib().orc_num_threads_available())


# ----------------------------------------------------------------------------------------------
# FP32 restatement of RPP HOST
# ----------------------------------------------------------------------------------------------

def pre_emphasis_f32(x: np.ndarray, lengths, coeff=0.97, border="clamp", nthreads=1) -> np.ndarray:
    """x: [B, stride] float32.  Returns y of the same shape (only [0, len) written, rest zero)."""
    x = np.ascontiguousarray(x, np.float32)
    B, S = x.shape
    y = np.zeros_like(x)
    lens = np.ascontiguousarray(lengths, np.int32)
    c = np.ascontiguousarray(np.broadcast_to(np.float32(coeff), (B,)), np.float32)
    lib().orc_pre_emphasis_f32(x, S, y, S, lens, c, B, BORDER[border], nthreads)
    return y


def pre_emphasis_f64(x, lengths, coeff=0.97, border="clamp") -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    B, S = x.shape
    y = np.zeros((B, S), np.float64)
    lens = np.ascontiguousarray(lengths, np.int32)
    c = np.ascontiguousarray(np.broadcast_to(np.float32(coeff), (B,)), np.float32)
    lib().orc_pre_emphasis_f64(x, S, y, S, lens, c, B, BORDER[border])
    return y


def hann_f32(W: int) -> np.ndarray:
    w = np.zeros(W, np.float32)
    lib().orc_hann_f32(W, w)
    return w


def hann_f64(W: int) -> np.ndarray:
    w = np.zeros(W, np.float64)
    lib().orc_hann_f64(W, w)
    return w


def num_windows(L_: int, W: int, H: int, center=True) -> int:
    return lib().orc_num_windows(int(L_), W, H, int(center))


def idx_reflect(loc: int, lo: int, hi: int) -> int:
    return lib().orc_idx_reflect(loc, lo, hi)


def spectrogram_f32(x, lengths, nfft=512, W=400, H=160, center=True, reflect=True, power=2,
                    layout="NFT", window=None, hstride=None, nthreads=1) -> np.ndarray:
    """Returns [B, K, hStride] (NFT) or [B, maxF, hStride] (NTF) float32; unwritten cells are 0."""
    x = np.ascontiguousarray(x, np.float32)
    B, S = x.shape
    lens = np.ascontiguousarray(lengths, np.int32)
    n = nfft or W
    K = n // 2 + 1
    maxF = max(num_windows(int(l), W, H, center) for l in lens) if B else 0
    nft = layout == "NFT"
    if hstride is None:
        hstride = maxF if nft else K
    rows = K if nft else maxF
    y = np.zeros((B, rows, hstride), np.float32)
    win = None
    if window is not None:
        win = np.ascontiguousarray(window, np.float32)
    lib().orc_spectrogram_f32(x, S, lens, B, y, rows * hstride, hstride, int(nft), int(center), int(reflect),
                              win.ctypes.data if win is not None else None, nfft, power, W, H, nthreads)
    return y


def spectrogram_f64(xd, lengths, nfft=512, W=400, H=160, center=True, reflect=True, power=2,
                    layout="NFT", window=None, hstride=None, nthreads=1) -> np.ndarray:
    xd = np.ascontiguousarray(xd, np.float64)
    B, S = xd.shape
    lens = np.ascontiguousarray(lengths, np.int32)
    n = nfft or W
    K = n // 2 + 1
    maxF = max(num_windows(int(l), W, H, center) for l in lens) if B else 0
    nft = layout == "NFT"
    if hstride is None:
        hstride = maxF if nft else K
    rows = K if nft else maxF
    y = np.zeros((B, rows, hstride), np.float64)
    win = None
    if window is not None:
        win = np.ascontiguousarray(window, np.float64)
    lib().orc_spectrogram_f64(xd, S, lens, B, y, rows * hstride, hstride, int(nft), int(center), int(reflect),
                              win.ctypes.data if win is not None else None, nfft, power, W, H, nthreads)
    return y


def mel_tables_f32(numBins, sampleRate, maxFreq, minFreq, nmel, formula="slaney", normalize=True):
    wd = np.zeros(numBins, np.float32)
    iv = np.zeros(numBins, np.int32)
    nf = np.zeros(nmel, np.float32)
    rng = np.zeros(2, np.int32)
    lib().orc_mel_tables_f32(numBins, sampleRate, maxFreq, minFreq, MEL[formula], nmel, int(normalize), wd, iv, nf, rng)
    return wd, iv, nf, (int(rng[0]), int(rng[1]))


def mel_f32(spec, frames, nmel=80, sample_rate=16000.0, fmin=0.0, fmax=0.0, formula="slaney",
            normalize=True, hstride=None, nthreads=1) -> np.ndarray:
    """spec: [B, K, srcHStride] NFT float32; frames: per-clip frame counts.  Returns [B, nmel, hstride]."""
    spec = np.ascontiguousarray(spec, np.float32)
    B, K, sh = spec.shape
    if hstride is None:
        hstride = sh
    dims = np.zeros(2 * B, np.int32)
    dims[0::2] = K
    dims[1::2] = np.asarray(frames, np.int32)
    y = np.zeros((B, nmel, hstride), np.float32)
    lib().orc_mel_f32(spec, K * sh, sh, dims, B, y, nmel * hstride, hstride, fmax, fmin, MEL[formula], nmel,
                      sample_rate, int(normalize), nthreads)
    return y


def mel_f64(spec, frames, nmel=80, sample_rate=16000.0, fmin=0.0, fmax=0.0, formula="slaney",
            normalize=True, nthreads=1) -> np.ndarray:
    spec = np.ascontiguousarray(spec, np.float64)
    B, K, sh = spec.shape
    dims = np.zeros(2 * B, np.int32)
    dims[0::2] = K
    dims[1::2] = np.asarray(frames, np.int32)
    y = np.zeros((B, nmel, sh), np.float64)
    lib().orc_mel_f64(spec, K * sh, sh, dims, B, y, nmel * sh, sh, fmax, fmin, MEL[formula], nmel,
                      sample_rate, int(normalize), nthreads)
    return y


def to_db_f32(x, widths, heights, cutoff_db=-80.0, multiplier=10.0, ref=1.0, nthreads=1,
              force2d=False) -> np.ndarray:
    """x: [B, rows, hStride] (or [B, stride] for 1-D data with widths == 1).  force2d keeps
    one-column clips 2-D instead of the reference's 1-D reading of them."""
    x = np.ascontiguousarray(x, np.float32)
    B = x.shape[0]
    if x.ndim == 2:
        hs, ns = 1, x.shape[1]
    else:
        hs, ns = x.shape[2], x.shape[1] * x.shape[2]
    dims = np.zeros(2 * B, np.uint32)
    dims[0::2] = np.asarray(widths, np.uint32)
    dims[1::2] = np.asarray(heights, np.uint32)
    y = np.zeros_like(x)
    lib().orc_to_db_f32(x, ns, hs, y, ns, hs, dims, B, cutoff_db, multiplier, ref, nthreads, int(force2d))
    return y


def to_db_f64(x, widths, heights, cutoff_db=-80.0, multiplier=10.0, ref=1.0, force2d=False) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    B = x.shape[0]
    if x.ndim == 2:
        hs, ns = 1, x.shape[1]
    else:
        hs, ns = x.shape[2], x.shape[1] * x.shape[2]
    dims = np.zeros(2 * B, np.uint32)
    dims[0::2] = np.asarray(widths, np.uint32)
    dims[1::2] = np.asarray(heights, np.uint32)
    y = np.zeros_like(x)
    lib().orc_to_db_f64(x, ns, hs, y, ns, hs, dims, B, cutoff_db, multiplier, ref, int(force2d))
    return y


def down_mix_f32(x, lengths, channels, normalize=False, nthreads=1) -> np.ndarray:
    """x: [B, stride] interleaved; returns [B, stride // channels]."""
    x = np.ascontiguousarray(x, np.float32)
    B, S = x.shape
    dims = np.zeros(2 * B, np.int32)
    dims[0::2] = np.asarray(lengths, np.int32)
    dims[1::2] = np.asarray(channels, np.int32)
    out_stride = S // int(np.max(channels)) if B else 0
    y = np.zeros((B, out_stride), np.float32)
    lib().orc_down_mix_f32(x, S, y, out_stride, dims, B, int(normalize), nthreads)
    return y


def windowed_sinc_lut(coeffs: int, lobes: int):
    lut = np.zeros(coeffs + 5, np.float32)
    sc, ce = F(), F()
    lib().orc_windowed_sinc_lut(coeffs, lobes, lut, C.byref(sc), C.byref(ce))
    return lut, sc.value, ce.value


def resample_out_length(L_: int, in_rate: float, out_rate: float) -> int:
    """ceil(L*out/in) in FP32, as the reference kernel computes it (cpu/kernel/resample.cpp:57)."""
    if np.float32(in_rate) == np.float32(out_rate):
        return int(L_)
    v = np.float32(np.float32(L_) * np.float32(out_rate)) / np.float32(in_rate)
    return int(np.ceil(np.float32(v)))


def resample_f32(x, lengths, in_rate, out_rate, lut, scale, center, lobes, channels=None, out_stride=None,
                 nthreads=1) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    B, S = x.shape
    ir = np.ascontiguousarray(np.broadcast_to(np.float32(in_rate), (B,)), np.float32)
    orr = np.ascontiguousarray(np.broadcast_to(np.float32(out_rate), (B,)), np.float32)
    ch = np.ones(B, np.int32) if channels is None else np.asarray(channels, np.int32)
    dims = np.zeros(2 * B, np.int32)
    dims[0::2] = np.asarray(lengths, np.int32)
    dims[1::2] = ch
    if out_stride is None:
        out_stride = max(resample_out_length(int(l), float(a), float(b)) * int(c)
                         for l, a, b, c in zip(dims[0::2], ir, orr, ch)) if B else 0
    y = np.zeros((B, out_stride), np.float32)
    lut = np.ascontiguousarray(lut, np.float32)
    lib().orc_resample_f32(x, S, y, out_stride, ir, orr, dims, B, scale, center, lobes, len(lut), lut, nthreads)
    return y


def non_silent_region_f32(x, lengths, cutoff_db=-60.0, window_length=2048, reference_power=0.0,
                          reset_interval=8192, nthreads=1):
    """Returns (begin[B], length[B]) int32 (src/modules/tensor/cpu/kernel/non_silent_region_detection.cpp:93-191)."""
    x = np.ascontiguousarray(x, np.float32)
    B, S = x.shape
    lens = np.ascontiguousarray(lengths, np.int32)
    b = np.zeros(B, np.int32)
    n = np.zeros(B, np.int32)
    lib().orc_nsr_f32(x, S, lens, B, cutoff_db, window_length, reference_power, reset_interval, b, n, nthreads)
    return b, n


# ----------------------------------------------------------------------------------------------
# whole-chain helpers
# ----------------------------------------------------------------------------------------------

def logmel_f32(x, lengths, *, nfft=512, W=400, H=160, nmel=80, sample_rate=16000.0, fmin=0.0, fmax=0.0,
               coeff=0.97, cutoff_db=-80.0, multiplier=10.0, ref=1.0, formula="slaney", normalize=True,
               nthreads=1):
    """RPP-HOST-equivalent chain: pre-emphasis -> spectrogram (NFT) -> mel -> to_decibels.
    Single-frame clips are converted as 2-D (see to_db_f32's force2d).

    Returns (db [B, nmel, maxF], frames[B])."""
    lens = np.asarray(lengths, np.int32)
    pe = pre_emphasis_f32(x, lens, coeff, "clamp", nthreads)
    spec = spectrogram_f32(pe, lens, nfft, W, H, True, True, 2, "NFT", nthreads=nthreads)
    frames = np.array([num_windows(int(l), W, H, True) for l in lens], np.int32)
    mel = mel_f32(spec, frames, nmel, sample_rate, fmin, fmax, formula, normalize, nthreads=nthreads)
    db = to_db_f32(mel, frames, np.full(len(lens), nmel), cutoff_db, multiplier, ref, nthreads, force2d=True)
    return db, frames


class LogMelF32Timer:
    """The FP32 chain of logmel_f32 with every buffer allocated (and touched) up front, so that
    run() makes only the four C calls -- pre-emphasis, spectrogram, mel, to_decibels -- as a
    caller of RPP HOST's four rppt_*_host functions would time them (inputs resident, outputs
    preallocated).  run() returns the dB array [B, nmel, maxF]."""

    def __init__(self, x, lengths, *, nfft=512, W=400, H=160, nmel=80, sample_rate=16000.0, fmin=0.0, fmax=0.0,
                 coeff=0.97, cutoff_db=-80.0, multiplier=10.0, ref=1.0, formula="slaney", normalize=True,
                 nthreads=1):
        self.x = np.ascontiguousarray(x, np.float32)
        B, S = self.x.shape
        self.B, self.S = B, S
        self.lens = np.ascontiguousarray(lengths, np.int32)
        self.coeffs = np.full(B, coeff, np.float32)
        self.frames = np.array([num_windows(int(n), W, H, True) for n in self.lens], np.int32)
        self.maxF = int(self.frames.max()) if B else 0
        self.K = (nfft or W) // 2 + 1
        self.nfft, self.W, self.H, self.nmel = nfft, W, H, nmel
        self.sample_rate, self.fmin, self.fmax = sample_rate, fmin, fmax
        self.formula, self.normalize = MEL[formula], int(normalize)
        self.cutoff_db, self.multiplier, self.ref = cutoff_db, multiplier, ref
        self.nthreads = nthreads
        # np.ones, not np.zeros: every page is written now, so no first-touch fault lands in run()
        self.pe = np.ones((B, S), np.float32)
        self.spec = np.ones((B, self.K, self.maxF), np.float32)
        self.mel = np.ones((B, nmel, self.maxF), np.float32)
        self.db = np.ones((B, nmel, self.maxF), np.float32)
        self.mel_dims = np.zeros(2 * B, np.int32)
        self.mel_dims[0::2] = self.K
        self.mel_dims[1::2] = self.frames
        self.db_dims = np.zeros(2 * B, np.uint32)
        self.db_dims[0::2] = self.frames
        self.db_dims[1::2] = nmel

    def run(self) -> np.ndarray:
        l = lib()
        F = self.maxF
        l.orc_pre_emphasis_f32(self.x, self.S, self.pe, self.S, self.lens, self.coeffs, self.B, BORDER["clamp"],
                               self.nthreads)
        l.orc_spectrogram_f32(self.pe, self.S, self.lens, self.B, self.spec, self.K * F, F, 1, 1, 1, None, self.nfft, 2,
                              self.W, self.H, self.nthreads)
        l.orc_mel_f32(self.spec, self.K * F, F, self.mel_dims, self.B, self.mel, self.nmel * F, F, self.fmax,
                      self.fmin, self.formula, self.nmel, self.sample_rate, self.normalize, self.nthreads)
        l.orc_to_db_f32(self.mel, self.nmel * F, F, self.db, self.nmel * F, F, self.db_dims, self.B, self.cutoff_db,
                        self.multiplier, self.ref, self.nthreads, 1)
        return self.db


def logmel_f64(x, lengths, *, nfft=512, W=400, H=160, nmel=80, sample_rate=16000.0, fmin=0.0, fmax=0.0,
               coeff=0.97, cutoff_db=-80.0, multiplier=10.0, ref=1.0, formula="slaney", normalize=True,
               nthreads=1):
    """FP64 oracle of the same chain.  Returns (db, mel_power, frames)."""
    lens = np.asarray(lengths, np.int32)
    pe = pre_emphasis_f64(x, lens, coeff, "clamp")
    spec = spectrogram_f64(pe, lens, nfft, W, H, True, True, 2, "NFT", nthreads=nthreads)
    frames = np.array([num_windows(int(l), W, H, True) for l in lens], np.int32)
    mel = mel_f64(spec, frames, nmel, sample_rate, fmin, fmax, formula, normalize, nthreads=nthreads)
    db = to_db_f64(mel, frames, np.full(len(lens), nmel), cutoff_db, multiplier, ref, force2d=True)
    return db, mel, frames
