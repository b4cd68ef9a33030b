"""The drop-in boundary, CPU only: librpp_audio.so loads, exports every entry point that
include/rpp_audio.h declares (unmangled), and the ctypes mirrors of the descriptor structs have
the reference's layout (api/rppdefs.h:570-649, :751-790).  No compute call is made here."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rpp_audio.h")
LIB = os.path.join(ROOT, "rpp_amd", "librpp_audio.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^(?:RppStatus|rppStatus_t)\s+(\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for n in ("rppCreate", "rppDestroy", "rppSetStream", "rppGetStream", "rppSetBatchSize", "rppGetBatchSize",
              "rppt_pre_emphasis_filter_gpu", "rppt_spectrogram_gpu", "rppt_mel_filter_bank_gpu",
              "rppt_to_decibels_gpu", "rppt_down_mixing_gpu", "rppt_resample_gpu",
              "rppt_non_silent_region_detection_gpu"):
        assert n in names, n


def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  -- one HIP runtime for the process
    assert os.path.exists(LIB), "build with python -m rpp_amd.build"
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for n in declared_functions():
        assert n in exported, f"{n} is not an unmangled global text symbol"


def test_struct_layouts():
    from rpp_amd import api
    assert C.sizeof(api.RpptDesc) == 56
    assert api.RpptDesc.offsetInBytes.offset == 8 and api.RpptDesc.dataType.offset == 12
    assert api.RpptDesc.n.offset == 16 and api.RpptDesc.strides.offset == 32 and api.RpptDesc.layout.offset == 48
    assert C.sizeof(api.RpptImagePatch) == 8
    assert C.sizeof(api.RpptResamplingWindow) == 96
    assert api.RpptResamplingWindow.lookupSize.offset == 16
    assert api.RpptResamplingWindow.lookupPinned.offset == 24
    assert api.RpptResamplingWindow.pCenter.offset == 64 and api.RpptResamplingWindow.pScale.offset == 80
    assert (api.NFT, api.NTF, api.F32) == (5, 6, 1)
    assert api.BORDER == {"zero": 0, "clamp": 1, "reflect": 2} and api.MEL == {"slaney": 0, "htk": 1}


def test_create_host_and_hip_backends_only():
    """RPP_HOST_BACKEND handles are made without touching a device (their thread count follows
    handle_host.cpp:54-62) and destroyed with the same backend; RPP_OCL_BACKEND is refused with
    rppStatusNotImplemented; destroying with the wrong backend is rppStatusBadParm."""
    import torch  # noqa: F401
    lib = C.CDLL(LIB)
    lib.rppCreate.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint32, C.c_void_p, C.c_int]
    lib.rppDestroy.argtypes = [C.c_void_p, C.c_int]
    h = C.c_void_p()
    assert lib.rppCreate(C.byref(h), 4, 0, None, 0) == 0 and h.value
    assert lib.rppDestroy(h, 1) == -1
    assert lib.rppDestroy(h, 0) == 0
    assert lib.rppCreate(C.byref(h), 4, 0, None, 2) == -7


def test_cxx_caller_builds_and_links(tmp_path):
    """A C++ caller (tests/cxx/integration_example.cpp: INTEGRATION.md's call sequence) compiles
    against include/rpp_audio.h -- whose static_asserts pin sizeof/offsetof of RpptDesc,
    RpptImagePatch and RpptResamplingWindow and the enum values to the reference's -- and links
        against librpp_audio.so, resolving every entry point.  The binary runs the log-mel chain
    through the HOST entry points on the CPU (a 1 kHz tone peaks at 0 dB in its filter) and
    touches no GPU."""
    src = os.path.join(ROOT, "tests", "cxx", "integration_example.cpp")
    exe = str(tmp_path / "integration_example")
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
           "-L", os.path.join(ROOT, "rpp_amd"), "-lrpp_audio", "-L/opt/rocm/lib",
           "-Wl,-rpath," + os.path.join(ROOT, "rpp_amd"), "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "36 entry points resolved" in r.stdout and "host chain" in r.stdout


def test_cxx_header_is_plain_c_up_to_the_cpp_only_parts(tmp_path):
    """The header also compiles as C (the reference's rppt_resample_gpu takes a C++ reference and
    the resampling window holds a std::vector, so those two stay C++-only)."""
    src = tmp_path / "c_caller.c"
    src.write_text('#include "rpp_audio.h"\nint main(void) { RpptDesc d; d.n = 1; return (int)d.n - 1; }\n')
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                        "-o", str(tmp_path / "c_caller.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
