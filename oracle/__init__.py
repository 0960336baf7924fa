"""FP64 CPU oracle for the Kalnajs log-spiral hot path -- TEST INFRASTRUCTURE.

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import
it.  ``klsp_oracle.c`` restates the reference Julia code
(``src/KalnajsLogSpiral/*.jl``) function by function; ``blas_form`` restates the
BLAS form of the det (``NL_julia_direct.jl:299-310``) on numpy/OpenBLAS.

Parity pinning of each function is done by ``tests/test_oracle_*.py``.
"""
from .core import *  # noqa: F401,F403
from .core import __all__  # noqa: F401
