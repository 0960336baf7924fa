"""GPUBackend -- the device selection surface of src/KalnajsLogSpiral/GPUBackend.jl.

This build has exactly one backend (HIP on gfx950).  ``GPUType`` keeps the
reference's three names so signatures stay compatible; ``NVIDIA`` is accepted
but never detected, and ``NONE`` means "no device" -- there is no CPU path
behind it (the compute entry points raise instead).
"""
from __future__ import annotations

import enum

from . import _ffi

__all__ = ["GPUType", "detect_gpu_type", "gpu_device", "current_device", "gpu_ndevices",
           "gpu_synchronize", "gpu_device_name", "gpu_total_memory", "NVIDIA", "AMD", "NONE"]


class GPUType(enum.Enum):
    NVIDIA = 0
    AMD = 1
    NONE = 2


NVIDIA, AMD, NONE = GPUType.NVIDIA, GPUType.AMD, GPUType.NONE

_current = 0


def detect_gpu_type() -> GPUType:
    """GPUBackend.jl:50: AMD when a HIP device is visible, else NONE."""
    try:
        return GPUType.AMD if _ffi.device_count() > 0 else GPUType.NONE
    except RuntimeError:
        return GPUType.NONE


def gpu_ndevices() -> int:
    return _ffi.device_count()


def gpu_device(dev_id: int) -> None:
    """GPUBackend.jl:147 (gpu_device!): device for contexts created afterwards."""
    global _current
    n = gpu_ndevices()
    if dev_id < 0 or dev_id >= n:
        raise ValueError(f"GPU {dev_id} not available (have {n} GPUs)")
    _current = dev_id


def current_device() -> int:
    return _current


def gpu_synchronize() -> None:
    """Every libklsp call is blocking; nothing to wait for."""
    return None


def gpu_device_name(dev_id: int = -1) -> str:
    """GPUBackend.jl:246: name of device dev_id (-1: the current device)."""
    return _ffi.device_info(_current if dev_id < 0 else dev_id)[0]


def gpu_total_memory(dev_id: int = -1) -> int:
    """Device memory in bytes (run_kalnajs_gpu.jl:73 gpu_total_memory)."""
    return _ffi.device_info(_current if dev_id < 0 else dev_id)[1]
