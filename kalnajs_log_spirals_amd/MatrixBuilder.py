"""MatrixBuilder -- mirror of src/KalnajsLogSpiral/MatrixBuilder.jl.

Every entry point runs on the device through libklsp: the weights, the
M(omega) contraction, the G*N epilogue and the FP64 complex LU determinant.
``Tcpu`` selects the dtype of the returned determinant (the LU itself is
always FP64).  The reference's CPU route (devprecomp === nothing or gpu_type
NONE) does not exist here and raises.
"""
from __future__ import annotations

import numpy as np

from . import _ffi
from . import GPUBackend
from .BasisFunctions import DevicePrecomputed, PrecomputedData

__all__ = ["compute_determinant", "compute_determinant_batched", "build_M_matrix",
           "build_M_matrix_gpu", "compute_determinant_gpu"]


def _ctype(Tcpu):
    return np.complex128 if np.dtype(Tcpu) == np.float64 else np.complex64


def _device(devprecomp, gpu_type, who):
    if devprecomp is None or gpu_type == GPUBackend.GPUType.NONE:
        raise RuntimeError(f"{who}: the reference's CPU route is not part of this build; "
                           "pass a DevicePrecomputed from to_device()")
    return devprecomp


def build_M_matrix_gpu(omega: complex, devprecomp: DevicePrecomputed, gpu_type=None) -> np.ndarray:
    """MatrixBuilder.jl:100-152: M [N, N] complex (returned on the host)."""
    return devprecomp.ctx.build_M(complex(omega), devprecomp.precision)


def compute_determinant_gpu(omega: complex, precomputed: PrecomputedData,
                            devprecomp: DevicePrecomputed, gpu_type, Tcpu=np.float64):
    """MatrixBuilder.jl:162-185: det(I - M) with the LU on the device."""
    d = devprecomp.ctx.det([complex(omega)], devprecomp.precision)[0]
    return _ctype(Tcpu)(d)


def compute_determinant(omega: complex, precomputed: PrecomputedData,
                        devprecomp: DevicePrecomputed, gpu_type, Tcpu=np.float64):
    """MatrixBuilder.jl:197-207."""
    dev = _device(devprecomp, gpu_type, "compute_determinant")
    return compute_determinant_gpu(omega, precomputed, dev, gpu_type, Tcpu)


def compute_determinant_batched(omegas, precomputed: PrecomputedData,
                                devprecomp: DevicePrecomputed, gpu_type, Tcpu=np.float64,
                                batch_size: int = 100) -> np.ndarray:
    """MatrixBuilder.jl:215-239: one device call for all omegas (omegas that
    share a pass over W are batched inside the library)."""
    dev = _device(devprecomp, gpu_type, "compute_determinant_batched")
    om = np.asarray(omegas, dtype=np.complex128).ravel()
    return dev.ctx.det(om, dev.precision).astype(_ctype(Tcpu))


def build_M_matrix(omega: complex, precomputed: PrecomputedData) -> np.ndarray:
    """MatrixBuilder.jl:242-244 (debug entry): FP64 M(omega) from the device W."""
    if precomputed.ctx is None:
        raise RuntimeError("build_M_matrix: precomputed data has no device context")
    return precomputed.ctx.build_M(complex(omega), _ffi.PREC_FP64)
