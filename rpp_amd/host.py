"""ctypes front end of the HOST backend of librpp_audio.so: the rppt_*_host entry points of
include/rpp_audio.h (api/rppt_tensor_audio_augmentations.h) on host buffers.

Arrays are numpy arrays (C-contiguous, host memory); descriptors, enum values and status codes are
those of rpp_amd.api.  A HostHandle is rppCreate(..., RPP_HOST_BACKEND): it touches no GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .api import (BORDER, MEL, NFT, RPP_HOST_BACKEND, RpptResamplingWindow, RppError, check, desc_1d, desc_2d,
                  lib)

P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data


def _bind():
    l = lib()
    if getattr(l, "_host_bound", False):
        return l
    DP = C.c_void_p
    l.rppt_pre_emphasis_filter_host.argtypes = [P, DP, P, DP, P, P, C.c_int, P]
    l.rppt_spectrogram_host.argtypes = [P, DP, P, DP, P, C.c_bool, C.c_bool, P, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_int32, P]
    l.rppt_mel_filter_bank_host.argtypes = [P, DP, P, DP, P, C.c_float, C.c_float, C.c_int, C.c_int32, C.c_float,
                                            C.c_bool, P]
    l.rppt_to_decibels_host.argtypes = [P, DP, P, DP, P, C.c_float, C.c_float, C.c_float, P]
    l.rppt_down_mixing_host.argtypes = [P, DP, P, DP, P, C.c_bool, P]
    l.rppt_resample_host.argtypes = [P, DP, P, DP, P, P, P, C.POINTER(RpptResamplingWindow), P]
    l.rppt_non_silent_region_detection_host.argtypes = [P, DP, P, P, P, C.c_float, C.c_int32, C.c_float, C.c_int32,
                                                        P]
    l._host_bound = True
    return l


def _ref(d):
    return C.cast(C.pointer(d), C.c_void_p)


class HostHandle:
    """rppCreate(..., RPP_HOST_BACKEND): `threads` host threads (0 = one per clip, as RPP)."""

    def __init__(self, batch_size: int = 1, threads: int = 0):
        self.h = P()
        st = _bind().rppCreate(C.byref(self.h), batch_size, threads, None, RPP_HOST_BACKEND)
        if st != 0:
            raise RppError("rppCreate", st)

    def close(self):
        if self.h:
            lib().rppDestroy(self.h, RPP_HOST_BACKEND)
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


def pre_emphasis_filter(h: HostHandle, src, dst, lengths, coeffs, border="clamp", check_status=True) -> int:
    B, S = src.shape
    sd, dd = desc_1d(B, S), desc_1d(B, dst.shape[1])
    st = _bind().rppt_pre_emphasis_filter_host(_p(src), _ref(sd), _p(dst), _ref(dd), _p(lengths), _p(coeffs),
                                               BORDER[border], h.h)
    if check_status:
        check("rppt_pre_emphasis_filter_host", st)
    return st


def spectrogram(h: HostHandle, src, dst, lengths, *, nfft=512, window_length=400, window_step=160, center=True,
                reflect=True, power=2, layout=NFT, window=None, check_status=True) -> int:
    """dst [B, K, maxF] (NFT) or [B, maxF, K] (NTF)."""
    B, S = src.shape
    sd = desc_1d(B, S)
    dd = desc_2d(B, dst.shape[1], dst.shape[2], layout=layout)
    win = None if window is None else np.ascontiguousarray(window, np.float32)
    st = _bind().rppt_spectrogram_host(_p(src), _ref(sd), _p(dst), _ref(dd), _p(lengths), center, reflect, _p(win),
                                       nfft, power, window_length, window_step, h.h)
    if check_status:
        check("rppt_spectrogram_host", st)
    return st


def mel_filter_bank(h: HostHandle, src, dst, src_dims, *, max_freq=8000.0, min_freq=0.0, formula="slaney",
                    num_filter=80, sample_rate=16000.0, normalize=True, check_status=True) -> int:
    """src [B, K, hStride] NFT, dst [B, M, hStride] NFT, src_dims int32 [B*2] = (K, frames)."""
    B = src.shape[0]
    sd, dd = desc_2d(B, src.shape[1], src.shape[2]), desc_2d(B, dst.shape[1], dst.shape[2])
    st = _bind().rppt_mel_filter_bank_host(_p(src), _ref(sd), _p(dst), _ref(dd), _p(src_dims), max_freq, min_freq,
                                           MEL[formula], num_filter, sample_rate, normalize, h.h)
    if check_status:
        check("rppt_mel_filter_bank_host", st)
    return st


def to_decibels(h: HostHandle, src, dst, src_dims, *, cutoff_db=-80.0, multiplier=10.0, reference_magnitude=1.0,
                check_status=True) -> int:
    """src/dst [B, rows, hStride] or [B, stride]; src_dims uint32 [B*2] = (width, height) pairs."""
    B = src.shape[0]
    sd = desc_1d(B, src.shape[1]) if src.ndim == 2 else desc_2d(B, src.shape[1], src.shape[2])
    dd = desc_1d(B, dst.shape[1]) if dst.ndim == 2 else desc_2d(B, dst.shape[1], dst.shape[2])
    st = _bind().rppt_to_decibels_host(_p(src), _ref(sd), _p(dst), _ref(dd), _p(src_dims), cutoff_db, multiplier,
                                       reference_magnitude, h.h)
    if check_status:
        check("rppt_to_decibels_host", st)
    return st


def down_mixing(h: HostHandle, src, dst, src_dims, normalize=False, channels=2, check_status=True) -> int:
    B, S = src.shape
    sd = desc_1d(B, S, num_dims=3, channels=channels)
    dd = desc_1d(B, dst.shape[1])
    st = _bind().rppt_down_mixing_host(_p(src), _ref(sd), _p(dst), _ref(dd), _p(src_dims), normalize, h.h)
    if check_status:
        check("rppt_down_mixing_host", st)
    return st


def make_window(lut, scale: float, center: float, lobes: int, coeffs: int) -> RpptResamplingWindow:
    """RpptResamplingWindow whose std::vector `lookup` views `lut` (begin, end, end of storage)."""
    lut = np.ascontiguousarray(lut, np.float32)
    w = RpptResamplingWindow()
    w.scale, w.center, w.lobes, w.coeffs = scale, center, lobes, coeffs
    w.lookupSize = lut.size
    base = lut.ctypes.data
    w.lookup_vec[0], w.lookup_vec[1], w.lookup_vec[2] = base, base + 4 * lut.size, base + 4 * lut.size
    w._table = lut   # the struct holds raw pointers: keep the table alive with it
    return w


def resample(h: HostHandle, src, dst, in_rate, out_rate, src_dims, window: RpptResamplingWindow,
             check_status=True) -> int:
    B, S = src.shape
    sd, dd = desc_1d(B, S), desc_1d(B, dst.shape[1])
    st = _bind().rppt_resample_host(_p(src), _ref(sd), _p(dst), _ref(dd), _p(in_rate), _p(out_rate), _p(src_dims),
                                    C.byref(window), h.h)
    if check_status:
        check("rppt_resample_host", st)
    return st


def non_silent_region_detection(h: HostHandle, src, lengths, begin, length, cutoff_db=-60.0, window_length=2048,
                                reference_power=0.0, reset_interval=8192, check_status=True) -> int:
    B, S = src.shape
    sd = desc_1d(B, S)
    st = _bind().rppt_non_silent_region_detection_host(_p(src), _ref(sd), _p(lengths), _p(begin), _p(length),
                                                       cutoff_db, window_length, reference_power, reset_interval, h.h)
    if check_status:
        check("rppt_non_silent_region_detection_host", st)
    return st
