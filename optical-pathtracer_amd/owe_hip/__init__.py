"""owe_hip — MI355X (gfx950) transport backend of the Optical World Engine.

The compute path is the C ABI of libowe_hip.so (include/owe_hip.h) over hand-written HIP kernels;
this package is its Python face. Importing it does not touch the GPU; the first call that needs the
library loads it and fails loudly when it has not been built.
"""
from __future__ import annotations

from . import _lib
from ._lib import OweHipError, available, device_info

BACKEND_NAME = "hip"

__all__ = ["BACKEND_NAME", "OweHipError", "available", "device_info", "_lib"]
