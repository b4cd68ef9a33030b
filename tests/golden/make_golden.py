"""Writes tests/golden/rpp_qa_audio.npz from the reference's own QA data files.

Run in the build container (it reads /root/reference, which the GPU box does not
have):  python tests/golden/make_golden.py

Inputs:   utilities/test_suite/TEST_AUDIO_FILES/three_samples_single_channel_src1/sample{1,2,3}.wav
Outputs:  utilities/test_suite/REFERENCE_OUTPUTS_AUDIO/<case>/<case>.bin (raw f32, batch 3)

The arrays are stored in the output-descriptor layout the reference QA driver
uses (utilities/test_suite/rpp_test_suite_audio.h:84-115, HOST/Tensor_audio_host.cpp:187-375):
  1-D cases   [3, w]  with w = (maxLen // 8) * 8 + 8
  spectrogram [3, 257, maxFrames]   (NFT, hStride = maxFrames)
  mel         [3, 80, maxFrames]    (NFT, hStride = maxFrames)
  resample    [3, ceil(1.15 * maxLen)]
"""
from __future__ import annotations

import os
import struct

import numpy as np

REF = "/root/reference/utilities/test_suite"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rpp_qa_audio.npz")


def read_wav_f32(path: str) -> np.ndarray:
    """Minimal RIFF/WAVE PCM16 reader; scales like libsndfile's sf_read_float (x / 32768)."""
    with open(path, "rb") as f:
        data = f.read()
    assert data[:4] == b"RIFF" and data[8:12] == b"WAVE", path
    pos, fmt, pcm = 12, None, None
    while pos + 8 <= len(data):
        cid, size = data[pos:pos + 4], struct.unpack("<I", data[pos + 4:pos + 8])[0]
        body = data[pos + 8:pos + 8 + size]
        if cid == b"fmt ":
            fmt = struct.unpack("<HHIIHH", body[:16])
        elif cid == b"data":
            pcm = body
        pos += 8 + size + (size & 1)
    audio_format, channels, rate, _, _, bits = fmt
    assert audio_format == 1 and bits == 16, (path, fmt)
    x = np.frombuffer(pcm, dtype="<i2").astype(np.float32) / np.float32(32768.0)
    return x.reshape(-1, channels), rate


def main() -> None:
    clips = []
    for i in (1, 2, 3):
        x, rate = read_wav_f32(f"{REF}/TEST_AUDIO_FILES/three_samples_single_channel_src1/sample{i}.wav")
        assert rate == 16000 and x.shape[1] == 1
        clips.append(x[:, 0])
    lengths = np.array([len(c) for c in clips], np.int32)
    w = (int(lengths.max()) // 8) * 8 + 8
    x = np.zeros((3, w), np.float32)
    for i, c in enumerate(clips):
        x[i, :len(c)] = c

    def load(case, shape):
        a = np.fromfile(f"{REF}/REFERENCE_OUTPUTS_AUDIO/{case}/{case}.bin", dtype="<f4")
        assert a.size == int(np.prod(shape)), (case, a.size, shape)
        return a.reshape(shape)

    frames = lengths // 160 + 1
    maxF = int(frames.max())
    out_len = int(np.ceil(np.float32(1.15) * np.float32(lengths.max())))
    np.savez_compressed(
        OUT,
        x=x,
        lengths=lengths,
        frames=frames.astype(np.int32),
        pre_emphasis=load("pre_emphasis_filter", (3, w)),
        to_decibels=load("to_decibels", (3, w)),
        spectrogram=load("spectrogram", (3, 257, maxF)),
        mel=load("mel_filter_bank", (3, 80, maxF)),
        resample=load("resample", (3, out_len)),
    )
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
