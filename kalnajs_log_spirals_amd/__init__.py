"""kalnajs_log_spirals_amd -- MI355X-native hot path of the Kalnajs log-spiral
matrix method (det(I - M(omega)) = 0 for unstable modes of hot stellar disks).

Mirrors the package facade src/KalnajsLogSpiral.jl:13-43: the same module
and function names, backed by libklsp.so (hand-written gfx950 HIP kernels
behind a C ABI, include/klsp.h).  Build the library with
``python -m kalnajs_log_spirals_amd.build_lib``.
"""
from . import Configuration, GPUBackend, Models, OrbitIntegration  # noqa: F401
from . import BasisFunctions, MatrixBuilder, NewtonSolver, GridScan  # noqa: F401
from .Configuration import (KalnajsConfig, load_config, load_model, save_config,  # noqa: F401
                            get_float_type, get_complex_type, get_n_l, get_NPh)
from .GPUBackend import GPUType, detect_gpu_type  # noqa: F401
from .Models import ToomreZangModel, create_toomre_zang_model  # noqa: F401
from .OrbitIntegration import GridData, OrbitData, compute_grids, compute_orbits  # noqa: F401
from .BasisFunctions import (PrecomputedData, DevicePrecomputed, precompute_all,  # noqa: F401
                             precompute_device_orbits, compute_W_l, compute_N_kernel,
                             gamma_complex, to_device)
from .MatrixBuilder import (compute_determinant, compute_determinant_batched,  # noqa: F401
                            build_M_matrix)
from .NewtonSolver import NewtonResult, newton_search  # noqa: F401
from .GridScan import grid_scan, find_minimum  # noqa: F401
from . import UnevenBatch, LinearForm  # noqa: F401
from .LinearForm import pencil_blocks, assemble_pencil, solve_pencil, linear_form_modes  # noqa: F401

VERSION = "0.1.0"
