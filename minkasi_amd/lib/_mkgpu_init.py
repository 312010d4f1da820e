"""Loads libmkgpu, the gfx950 HIP library, with ctypes.

Mirrors the reference's loader (minkasi/lib/_minkasi_init.py:9-16): the
package-local build (``_libmkgpu.so`` next to this file, produced by
``python -m minkasi_amd.build``) is tried first, then ``libmkgpu.so`` on the
loader path.  Unlike the reference, there is no CPU fallback: if neither
loads, importing ``minkasi_amd.lib.mkgpu`` raises OSError.
"""
import ctypes
import os
from pathlib import Path

__all__ = ["libmkgpu", "libmkgpu_path"]

_local = Path(__file__).parent.joinpath("_libmkgpu.so")
try:
    libmkgpu = ctypes.CDLL(_local.as_posix())
    libmkgpu_path = _local.as_posix()
except OSError as e1:
    try:
        libmkgpu = ctypes.CDLL("libmkgpu.so")
        libmkgpu_path = "libmkgpu.so"
    except OSError as e2:
        raise OSError(
            "libmkgpu not found (%s; %s). Build it with `python -m minkasi_amd.build`." % (e1, e2)
        )

if os.environ.get("MK_VERBOSE"):
    print("loaded", libmkgpu_path)
