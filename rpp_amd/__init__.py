"""rpp_amd -- MI355X (gfx950) batched log-mel audio path behind RPP's rppt_* C API.

The product is the shared library ``librpp_audio.so`` (HIP kernels + C-ABI, include/rpp_audio.h).
This package only locates/builds it and offers a thin ctypes front end (``rpp_amd.api``) that
the tests and the benchmark use; torch is used for device memory and streams.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "librpp_audio.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include", "rpp_audio.h")

__all__ = ["HERE", "LIBRARY", "INCLUDE"]
