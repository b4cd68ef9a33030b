"""ctypes front end of librpp_audio.so (the C-ABI of include/rpp_audio.h).

Mirrors the reference's audio operator interface (api/rppt_tensor_audio_augmentations.h):
same entry-point names, same descriptor structs (api/rppdefs.h:570-649), same status codes.
torch supplies device memory and the stream; every compute call goes through the C-ABI into
the HIP kernels.  There is no CPU path: when the library is missing this module raises.

torch must be imported before the library is loaded so that one HIP runtime serves both.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

from . import LIBRARY

# ---------------------------------------------------------------- enums (api/rppdefs.h)
RPP_SUCCESS = 0
STATUS = {
    0: "RPP_SUCCESS", -1: "RPP_ERROR", -2: "RPP_ERROR_INVALID_ARGUMENTS", -4: "RPP_ERROR_ZERO_DIVISION",
    -6: "RPP_ERROR_NOT_IMPLEMENTED", -9: "RPP_ERROR_INVALID_SRC_LAYOUT", -10: "RPP_ERROR_INVALID_DST_LAYOUT",
    -13: "RPP_ERROR_INVALID_SRC_OR_DST_DATATYPE", -16: "RPP_ERROR_NOT_ENOUGH_MEMORY",
    -22: "RPP_ERROR_OUT_OF_BOUND_SCRATCH_MEMORY_SIZE", -23: "RPP_ERROR_INVALID_SRC_DIMS",
    -24: "RPP_ERROR_INVALID_DST_DIMS",
}
RPP_HOST_BACKEND, RPP_HIP_BACKEND = 0, 1
U8, F32, F16, I8, I16 = range(5)
NCHW, NHWC, NCDHW, NDHWC, NHW, NFT, NTF = range(7)
BORDER = {"zero": 0, "clamp": 1, "reflect": 2}
MEL = {"slaney": 0, "htk": 1}
PRECISION = {"fp32": 0, "bf16x3": 1, "fp16x3": 2, "bf16": 3, "bf16x6": 4}


class RppError(RuntimeError):
    def __init__(self, fn, status):
        super().__init__(f"{fn} returned {status} ({STATUS.get(status, '?')})")
        self.status = status


class RpptStrides(C.Structure):
    _fields_ = [("nStride", C.c_uint32), ("cStride", C.c_uint32), ("hStride", C.c_uint32), ("wStride", C.c_uint32)]


class RpptDesc(C.Structure):
    _fields_ = [("numDims", C.c_size_t), ("offsetInBytes", C.c_uint32), ("dataType", C.c_int),
                ("n", C.c_uint32), ("c", C.c_uint32), ("h", C.c_uint32), ("w", C.c_uint32),
                ("strides", RpptStrides), ("layout", C.c_int)]


class RpptImagePatch(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32)]


class RpptResamplingWindow(C.Structure):
    """Layout twin of api/rppdefs.h:751-790 (96 bytes; the std::vector is left empty)."""
    _fields_ = [("scale", C.c_float), ("center", C.c_float), ("lobes", C.c_int32), ("coeffs", C.c_int32),
                ("lookupSize", C.c_int32), ("_pad", C.c_int32), ("lookupPinned", C.c_void_p),
                ("lookup_vec", C.c_void_p * 3), ("_pad2", C.c_uint8 * 8),
                ("pCenter", C.c_float * 4), ("pScale", C.c_float * 4)]


class RpptLogMelParams(C.Structure):
    _fields_ = [("preEmphasisCoeff", C.c_float), ("borderType", C.c_int),
                ("nfft", C.c_int32), ("windowLength", C.c_int32), ("windowStep", C.c_int32),
                ("centerWindows", C.c_int32), ("reflectPadding", C.c_int32), ("power", C.c_int32),
                ("sampleRate", C.c_float), ("minFreq", C.c_float), ("maxFreq", C.c_float),
                ("numFilter", C.c_int32), ("melFormula", C.c_int), ("normalize", C.c_int32),
                ("cutOffDB", C.c_float), ("multiplier", C.c_float), ("referenceMagnitude", C.c_float)]


class RpptLogMelTiles(C.Structure):
    _fields_ = [("framesPerTile", C.c_int32), ("blocksPerCU", C.c_int32)]


assert C.sizeof(RpptDesc) == 56 and C.sizeof(RpptImagePatch) == 8 and C.sizeof(RpptResamplingWindow) == 96

_LIB = None
P = C.c_void_p
DP = C.POINTER(RpptDesc)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(f"{LIBRARY} is missing: build it with `python -m rpp_amd.build` "
                               "(there is no CPU fallback)")
        l = C.CDLL(LIBRARY)
        l.rppCreate.argtypes = [C.POINTER(P), C.c_size_t, C.c_uint32, P, C.c_int]
        l.rppDestroy.argtypes = [P, C.c_int]
        l.rppSetStream.argtypes = [P, P]
        l.rppGetStream.argtypes = [P, C.POINTER(P)]
        l.rppSetBatchSize.argtypes = [P, C.c_size_t]
        l.rppGetBatchSize.argtypes = [P, C.POINTER(C.c_size_t)]
        l.rppt_pre_emphasis_filter_gpu.argtypes = [P, DP, P, DP, P, P, C.c_int, P]
        l.rppt_spectrogram_gpu.argtypes = [P, DP, P, DP, P, C.c_bool, C.c_bool, P, C.c_int32, C.c_int32,
                                           C.c_int32, C.c_int32, P]
        l.rppt_mel_filter_bank_gpu.argtypes = [P, DP, P, DP, P, C.c_float, C.c_float, C.c_int, C.c_int32,
                                               C.c_float, C.c_bool, P]
        l.rppt_to_decibels_gpu.argtypes = [P, DP, P, DP, P, C.c_float, C.c_float, C.c_float, P]
        l.rppt_non_silent_region_detection_gpu.argtypes = [P, DP, P, P, P, C.c_float, C.c_int32, C.c_float,
                                                           C.c_int32, P]
        l.rppt_down_mixing_gpu.argtypes = [P, DP, P, DP, P, C.c_bool, P]
        l.rppt_resample_gpu.argtypes = [P, DP, P, DP, P, P, P, C.POINTER(RpptResamplingWindow), P]
        l.rppt_down_mix_resample_gpu.argtypes = [P, DP, P, DP, P, P, P, C.c_bool, C.POINTER(RpptResamplingWindow), P]
        l.rppt_logmel_front_end_gpu.argtypes = [P, DP, P, DP, P, C.c_float, C.c_float, C.c_bool,
                                                C.POINTER(RpptResamplingWindow), C.POINTER(RpptLogMelParams), C.c_int,
                                                C.c_int32, P]
        l.rppt_logmel_chain_gpu.argtypes = [P, DP, P, DP, P, C.POINTER(RpptLogMelParams), C.c_int, P]
        l.rppt_logmel_fused_gpu.argtypes = [P, DP, P, DP, P, C.POINTER(RpptLogMelParams), C.c_int, P]
        l.rppt_logmel_fused_noalloc_gpu.argtypes = [P, DP, P, DP, P, C.POINTER(RpptLogMelParams), C.c_int, P]
        l.rppt_logmel_fused_region_gpu.argtypes = [P, DP, P, DP, P, P, C.POINTER(RpptLogMelParams), C.c_int, P]
        l.rppt_logmel_fused_packed_gpu.argtypes = [P, P, P, C.c_int32, P, P, C.POINTER(RpptLogMelParams), C.c_int, P]
        l.rppt_logmel_set_tiles.argtypes = [P, C.POINTER(RpptLogMelTiles)]
        l.rppt_logmel_get_tiles.argtypes = [P, C.POINTER(RpptLogMelTiles)]
        l.rppt_logmel_autotune.argtypes = [P, DP, P, DP, P, C.POINTER(RpptLogMelParams), C.c_int, C.c_int32, P,
                                           C.POINTER(RpptLogMelTiles), C.POINTER(C.c_float)]
        l.rppa_debug_mel_tables.argtypes = [P, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                            C.c_int, P, P, P, P]
        _LIB = l
    return _LIB


def check(fn: str, status: int) -> None:
    if status != RPP_SUCCESS:
        raise RppError(fn, status)


def ptr(t) -> int:
    return t.data_ptr() if t is not None else None


def desc_1d(n: int, stride: int, num_dims: int = 2, channels: int = 1, layout: int = NHW) -> RpptDesc:
    d = RpptDesc()
    d.numDims = num_dims
    d.dataType = F32
    d.n, d.c, d.h, d.w = n, channels, 1, stride // max(channels, 1)
    d.strides = RpptStrides(stride, 1, stride, channels)
    d.layout = layout
    return d


def desc_2d(n: int, h: int, w: int, h_stride: int | None = None, layout: int = NFT) -> RpptDesc:
    hs = w if h_stride is None else h_stride
    d = RpptDesc()
    d.numDims = 3
    d.dataType = F32
    d.n, d.c, d.h, d.w = n, 1, h, w
    d.strides = RpptStrides(h * hs, 1, hs, 1)
    d.layout = layout
    return d


class Handle:
    """rppCreate(..., RPP_HIP_BACKEND) on the given (or current) torch stream."""

    def __init__(self, batch_size: int = 1, stream: "torch.cuda.Stream | None" = None):
        self.h = P()
        s = stream if stream is not None else torch.cuda.current_stream()
        self.stream = s
        st = lib().rppCreate(C.byref(self.h), batch_size, 0, P(s.cuda_stream), RPP_HIP_BACKEND)
        if st != 0:
            raise RppError("rppCreate", st)

    def close(self):
        if self.h:
            lib().rppDestroy(self.h, RPP_HIP_BACKEND)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# --------------------------------------------------------------------- rppt_* mirrors
def pre_emphasis_filter(h: Handle, src, dst, lengths, coeffs, border="clamp", check_status=True) -> int:
    """src/dst [B, stride] float32 cuda; lengths int32 cuda [B]; coeffs float32 cuda [B]."""
    B, S = src.shape
    sd, dd = desc_1d(B, S), desc_1d(B, dst.shape[1])
    st = lib().rppt_pre_emphasis_filter_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(lengths), ptr(coeffs),
                                            BORDER[border], h.h)
    if check_status:
        check("rppt_pre_emphasis_filter_gpu", st)
    return st


def spectrogram(h: Handle, src, dst, lengths, *, nfft=512, window_length=400, window_step=160, center=True,
                reflect=True, power=2, layout=NFT, window=None, dst_desc: RpptDesc | None = None,
                check_status=True) -> int:
    """src [B, stride]; dst [B, K, maxF] (NFT) or [B, maxF, K] (NTF) float32 cuda."""
    B, S = src.shape
    sd = desc_1d(B, S)
    if dst_desc is None:
        dst_desc = desc_2d(B, dst.shape[1], dst.shape[2], layout=layout)
    win = None
    if window is not None:
        import numpy as np
        win = np.ascontiguousarray(window, np.float32)
    st = lib().rppt_spectrogram_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dst_desc), ptr(lengths), center, reflect,
                                    win.ctypes.data if win is not None else None, nfft, power, window_length,
                                    window_step, h.h)
    if check_status:
        check("rppt_spectrogram_gpu", st)
    return st


def mel_filter_bank(h: Handle, src, dst, src_dims, *, max_freq=8000.0, min_freq=0.0, formula="slaney", num_filter=80,
                    sample_rate=16000.0, normalize=True, src_desc=None, dst_desc=None, check_status=True) -> int:
    """src [B, K, hStride] NFT; dst [B, M, hStride] NFT; src_dims int32 cuda [B*2] = (K, frames)."""
    B = src.shape[0]
    if src_desc is None:
        src_desc = desc_2d(B, src.shape[1], src.shape[2])
    if dst_desc is None:
        dst_desc = desc_2d(B, dst.shape[1], dst.shape[2])
    st = lib().rppt_mel_filter_bank_gpu(ptr(src), C.byref(src_desc), ptr(dst), C.byref(dst_desc), ptr(src_dims),
                                        max_freq, min_freq, MEL[formula], num_filter, sample_rate, normalize, h.h)
    if check_status:
        check("rppt_mel_filter_bank_gpu", st)
    return st


def to_decibels(h: Handle, src, dst, src_dims, *, cutoff_db=-80.0, multiplier=10.0, reference_magnitude=1.0,
                src_desc=None, dst_desc=None, check_status=True) -> int:
    """src/dst [B, rows, hStride] (2-D) or [B, stride] (1-D); src_dims: RpptImagePatch array as uint32
    cuda tensor [B*2] = (width, height)."""
    B = src.shape[0]
    if src_desc is None:
        src_desc = desc_1d(B, src.shape[1]) if src.dim() == 2 else desc_2d(B, src.shape[1], src.shape[2])
    if dst_desc is None:
        dst_desc = desc_1d(B, dst.shape[1]) if dst.dim() == 2 else desc_2d(B, dst.shape[1], dst.shape[2])
    st = lib().rppt_to_decibels_gpu(ptr(src), C.byref(src_desc), ptr(dst), C.byref(dst_desc), ptr(src_dims),
                                    cutoff_db, multiplier, reference_magnitude, h.h)
    if check_status:
        check("rppt_to_decibels_gpu", st)
    return st


def down_mixing(h: Handle, src, dst, src_dims, normalize=False, channels=2, check_status=True) -> int:
    B, S = src.shape
    sd = desc_1d(B, S, num_dims=3, channels=channels)
    dd = desc_1d(B, dst.shape[1])
    st = lib().rppt_down_mixing_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(src_dims), normalize, h.h)
    if check_status:
        check("rppt_down_mixing_gpu", st)
    return st


def non_silent_region_detection(h: Handle, src, lengths, begin, length, cutoff_db=-60.0, window_length=2048,
                                reference_power=0.0, reset_interval=8192, check_status=True) -> int:
    """rppt_non_silent_region_detection_gpu: begin[b], length[b] (int32 device tensors) of clip b."""
    B, S = src.shape
    sd = desc_1d(B, S)
    st = lib().rppt_non_silent_region_detection_gpu(ptr(src), C.byref(sd), ptr(lengths), ptr(begin), ptr(length),
                                                    cutoff_db, window_length, reference_power, reset_interval, h.h)
    if check_status:
        check("rppt_non_silent_region_detection_gpu", st)
    return st


def make_window(lut_dev, scale: float, center: float, lobes: int, coeffs: int) -> RpptResamplingWindow:
    w = RpptResamplingWindow()
    w.scale, w.center, w.lobes, w.coeffs = scale, center, lobes, coeffs
    w.lookupSize = lut_dev.numel()
    w.lookupPinned = lut_dev.data_ptr()
    w._table = lut_dev   # the struct holds a raw pointer: keep the table alive with it
    return w


def resample(h: Handle, src, dst, in_rate, out_rate, src_dims, window: RpptResamplingWindow,
             check_status=True) -> int:
    B, S = src.shape
    sd, dd = desc_1d(B, S), desc_1d(B, dst.shape[1])
    st = lib().rppt_resample_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(in_rate), ptr(out_rate),
                                 ptr(src_dims), C.byref(window), h.h)
    if check_status:
        check("rppt_resample_gpu", st)
    return st


def down_mix_resample(h: Handle, src, dst, in_rate, out_rate, src_dims, window: RpptResamplingWindow, normalize=False,
                      channels=2, check_status=True) -> int:
    """rppt_down_mix_resample_gpu: src [B, L*C] interleaved, src_dims int32 [B*2] = (L, C); dst mono [B, n]."""
    B, S = src.shape
    sd = desc_1d(B, S, num_dims=3, channels=channels)
    dd = desc_1d(B, dst.shape[1])
    st = lib().rppt_down_mix_resample_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(in_rate), ptr(out_rate),
                                          ptr(src_dims), normalize, C.byref(window), h.h)
    if check_status:
        check("rppt_down_mix_resample_gpu", st)
    return st


def logmel_params(*, coeff=0.97, border="clamp", nfft=512, window_length=400, window_step=160, center=True,
                  reflect=True, power=2, sample_rate=16000.0, min_freq=0.0, max_freq=8000.0, num_filter=80,
                  formula="slaney", normalize=True, cutoff_db=-80.0, multiplier=10.0,
                  reference_magnitude=1.0) -> RpptLogMelParams:
    p = RpptLogMelParams()
    p.preEmphasisCoeff = coeff
    p.borderType = BORDER[border]
    p.nfft, p.windowLength, p.windowStep = nfft, window_length, window_step
    p.centerWindows, p.reflectPadding, p.power = int(center), int(reflect), power
    p.sampleRate, p.minFreq, p.maxFreq = sample_rate, min_freq, max_freq
    p.numFilter, p.melFormula, p.normalize = num_filter, MEL[formula], int(normalize)
    p.cutOffDB, p.multiplier, p.referenceMagnitude = cutoff_db, multiplier, reference_magnitude
    return p


def logmel_fused(h: Handle, src, dst, lengths, params: RpptLogMelParams, precision="bf16x3",
                 check_status=True, layout=NFT) -> int:
    """src [B, stride]; lengths int32 cuda [B]; dst [B, numFilter, hStride] (NFT) or
    [B, maxFrames, hStride >= numFilter] (NTF)."""
    B, S = src.shape
    sd = desc_1d(B, S)
    dd = desc_2d(B, dst.shape[1], dst.shape[2], layout=layout)
    st = lib().rppt_logmel_fused_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(lengths), C.byref(params),
                                     PRECISION[precision], h.h)
    if check_status:
        check("rppt_logmel_fused_gpu", st)
    return st


def logmel_front_end(h: Handle, src, dst, src_dims, in_rate: float, out_rate: float, window: RpptResamplingWindow,
                     params: RpptLogMelParams, precision="fp16x3", channels=2, normalize=False, one_launch=True,
                     check_status=True, layout=NFT) -> int:
    """rppt_logmel_front_end_gpu: src [B, L*C] interleaved at in_rate, src_dims int32 [B*2] = (L, C);
    dst as logmel_fused at out_rate.  one_launch=False forces the two-launch form."""
    B, S = src.shape
    sd = desc_1d(B, S, num_dims=3, channels=channels)
    dd = desc_2d(B, dst.shape[1], dst.shape[2], layout=layout)
    st = lib().rppt_logmel_front_end_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(src_dims), in_rate,
                                         out_rate, normalize, C.byref(window), C.byref(params), PRECISION[precision],
                                         1 if one_launch else 0, h.h)
    if check_status:
        check("rppt_logmel_front_end_gpu", st)
    return st


def logmel_fused_noalloc(h: Handle, src, dst, lengths, params: RpptLogMelParams, precision="fp16x3",
                         check_status=True, layout=NFT, batch=None) -> int:
    """rppt_logmel_fused_noalloc_gpu: rppt_logmel_fused_gpu with the library's own device array for
    every buffer beyond src/dst (no allocation).  `batch` overrides the descriptor's clip count."""
    B, S = src.shape
    sd = desc_1d(B if batch is None else batch, S)
    dd = desc_2d(B if batch is None else batch, dst.shape[1], dst.shape[2], layout=layout)
    st = lib().rppt_logmel_fused_noalloc_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(lengths),
                                             C.byref(params), PRECISION[precision], h.h)
    if check_status:
        check("rppt_logmel_fused_noalloc_gpu", st)
    return st


def logmel_fused_packed(h: Handle, src, src_offsets, lengths, dst, dst_offsets, params: RpptLogMelParams,
                        precision="fp16x3", check_status=True) -> int:
    """rppt_logmel_fused_packed_gpu: clip b is src[src_offsets[b] : + lengths[b]] (1-D f32 cuda), its
    dB a dense [numFilter][frames_b] block at dst[dst_offsets[b]:]; offsets int64, lengths int32 (cuda)."""
    st = lib().rppt_logmel_fused_packed_gpu(ptr(src), ptr(src_offsets), ptr(lengths), int(lengths.shape[0]), ptr(dst),
                                            ptr(dst_offsets), C.byref(params), PRECISION[precision], h.h)
    if check_status:
        check("rppt_logmel_fused_packed_gpu", st)
    return st


def logmel_fused_region(h: Handle, src, dst, begin, lengths, params: RpptLogMelParams, precision="fp16x3",
                        check_status=True, layout=NFT) -> int:
    """rppt_logmel_fused_region_gpu: clip b is src[b, begin[b] : begin[b] + lengths[b]] (int32 cuda)."""
    B, S = src.shape
    sd = desc_1d(B, S)
    dd = desc_2d(B, dst.shape[1], dst.shape[2], layout=layout)
    st = lib().rppt_logmel_fused_region_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(begin), ptr(lengths),
                                            C.byref(params), PRECISION[precision], h.h)
    if check_status:
        check("rppt_logmel_fused_region_gpu", st)
    return st


def logmel_chain_gpu(h: Handle, src, dst, lengths_host, params: RpptLogMelParams, per_clip=False,
                     check_status=True) -> int:
    """The 4-call chain through rppt_logmel_chain_gpu; lengths_host: int32 CPU tensor (pinned or not)."""
    B, S = src.shape
    sd = desc_1d(B, S)
    dd = desc_2d(B, dst.shape[1], dst.shape[2])
    st = lib().rppt_logmel_chain_gpu(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), lengths_host.data_ptr(),
                                     C.byref(params), int(per_clip), h.h)
    if check_status:
        check("rppt_logmel_chain_gpu", st)
    return st


def set_tiles(h: Handle, frames_per_tile=0, blocks_per_cu=0) -> None:
    t = RpptLogMelTiles(frames_per_tile, blocks_per_cu)
    check("rppt_logmel_set_tiles", lib().rppt_logmel_set_tiles(h.h, C.byref(t)))


def autotune(h: Handle, src, dst, lengths, params: RpptLogMelParams, precision="fp16x3", iters=5):
    """rppt_logmel_autotune: returns ((framesPerTile, blocksPerCU), best ms per launch)."""
    B, S = src.shape
    sd = desc_1d(B, S)
    dd = desc_2d(B, dst.shape[1], dst.shape[2])
    t = RpptLogMelTiles()
    ms = C.c_float()
    check("rppt_logmel_autotune", lib().rppt_logmel_autotune(ptr(src), C.byref(sd), ptr(dst), C.byref(dd), ptr(lengths),
                                                           C.byref(params), PRECISION[precision], iters, h.h,
                                                           C.byref(t), C.byref(ms)))
    return (t.framesPerTile, t.blocksPerCU), ms.value


def get_tiles(h: Handle):
    t = RpptLogMelTiles()
    check("rppt_logmel_get_tiles", lib().rppt_logmel_get_tiles(h.h, C.byref(t)))
    return t.framesPerTile, t.blocksPerCU


def front_end_logmel(h: Handle, stereo, lengths, channels, lut_dev, scale, center, lobes, coeffs, in_rate, out_rate,
                     params: RpptLogMelParams, precision="fp16x3", workspace=None):
    """Workload-4 front end: down-mix (interleaved channels -> mono) and resample in_rate ->
    out_rate with the caller's table in one launch (rppt_down_mix_resample_gpu, bit-identical to
    rppt_down_mixing_gpu + rppt_resample_gpu), then the fused log-mel launch.
    stereo [B, L*C] cuda; lengths int32 cuda [B] (frames per channel).  Returns (db, resampled)."""
    B, S = stereo.shape
    dev = stereo.device
    n_mono = S // channels
    ws = workspace if workspace is not None else {}
    if "res" not in ws:
        n_out = int(-(-n_mono * out_rate // in_rate)) + 8
        ws["res"] = torch.empty((B, n_out), dtype=torch.float32, device=dev)
        ws["dims_dm"] = torch.stack([lengths, torch.full_like(lengths, channels)], 1).reshape(-1).contiguous()
        ws["in_rate"] = torch.full((B,), in_rate, dtype=torch.float32, device=dev)
        ws["out_rate"] = torch.full((B,), out_rate, dtype=torch.float32, device=dev)
        out_len = torch.ceil(lengths.to(torch.float32) * torch.tensor(out_rate, dtype=torch.float32, device=dev)
                             / torch.tensor(in_rate, dtype=torch.float32, device=dev)).to(torch.int32)
        ws["res_len"] = out_len
        F = int(out_len.max().item()) // params.windowStep + 1
        ws["db"] = torch.empty((B, params.numFilter, F), dtype=torch.float32, device=dev)
        ws["window"] = make_window(lut_dev, scale, center, lobes, coeffs)
    down_mix_resample(h, stereo, ws["res"], ws["in_rate"], ws["out_rate"], ws["dims_dm"], ws["window"],
                      normalize=False, channels=channels)
    logmel_fused(h, ws["res"], ws["db"], ws["res_len"], params, precision)
    return ws["db"], ws["res"]


def logmel_chain(h: Handle, x, lengths_dev, *, nfft=512, W=400, H=160, nmel=80, sample_rate=16000.0, fmin=0.0,
                 fmax=None, coeff=0.97, cutoff_db=-80.0, multiplier=10.0, ref=1.0, frames=None, workspace=None):
    """The reference's 4-call chain through the drop-in entry points.  Returns (db, frames)."""
    B, S = x.shape
    if fmax is None:
        fmax = sample_rate / 2
    if frames is None:
        frames = (lengths_dev.cpu().long() // H + 1).to(torch.int32)
    maxF = int(frames.max()) if B else 0
    K = nfft // 2 + 1
    dev = x.device
    ws = workspace if workspace is not None else {}
    pe = ws.get("pe") if ws.get("pe") is not None and ws["pe"].shape == x.shape else torch.empty_like(x)
    spec = torch.empty((B, K, maxF), dtype=torch.float32, device=dev)
    mel = torch.empty((B, nmel, maxF), dtype=torch.float32, device=dev)
    db = torch.empty((B, nmel, maxF), dtype=torch.float32, device=dev)
    coeffs = torch.full((B,), coeff, dtype=torch.float32, device=dev)
    pre_emphasis_filter(h, x, pe, lengths_dev, coeffs, "clamp")
    spectrogram(h, pe, spec, lengths_dev, nfft=nfft, window_length=W, window_step=H)
    sdims = torch.stack([torch.full_like(frames, K), frames], 1).reshape(-1).to(torch.int32).to(dev)
    mel_filter_bank(h, spec, mel, sdims, max_freq=fmax, min_freq=fmin, num_filter=nmel, sample_rate=sample_rate)
    ddims = torch.stack([frames, torch.full_like(frames, nmel)], 1).reshape(-1).to(torch.int32).to(dev)
    to_decibels(h, mel, db, ddims, cutoff_db=cutoff_db, multiplier=multiplier, reference_magnitude=ref)
    return db, frames
