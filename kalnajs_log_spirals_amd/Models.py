"""Models -- the Toomre-Zang disk of src/KalnajsLogSpiral/Models.jl.

Scalar host functions (the orbit integration itself runs in libklsp's host
code, which restates the same formulas in C++).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from .Configuration import KalnajsConfig

__all__ = ["ToomreZangModel", "create_toomre_zang_model", "potential", "potential_derivative",
           "rotation_frequency", "epicyclic_frequency", "surface_density", "taper",
           "taper_derivative", "distribution_function", "DF_E", "DF_L"]


@dataclass(frozen=True)
class ToomreZangModel:
    """Models.jl:24-32."""
    L0: float
    n_zang: int
    q1: int
    q_zang: int
    sigma_r0: float
    G: float
    DF_const: float


def create_toomre_zang_model(config: KalnajsConfig) -> ToomreZangModel:
    """Models.jl:40-56."""
    L0 = float(config.model.L0)
    n_zang = int(config.model.n_zang)
    q1 = int(config.model.q1)
    G = float(config.physics.G)
    sigma_r0 = 1.0 / math.sqrt(q1)
    q = q1 - 1
    DF_const = 1.0 / (2.0 * math.pi * 2.0 ** (q / 2) * math.sqrt(math.pi) *
                      math.gamma((q + 1) / 2) * sigma_r0 ** (q + 2))
    return ToomreZangModel(L0, n_zang, q1, q, sigma_r0, G, DF_const)


def potential(model, r):            # Models.jl:64
    return math.log(r)


def potential_derivative(model, r):  # Models.jl:67
    return 1.0 / r


def rotation_frequency(model, r):    # Models.jl:70
    return 1.0 / r


def epicyclic_frequency(model, r):   # Models.jl:73
    return math.sqrt(2.0) / r


def surface_density(model, r):       # Models.jl:76
    return 1.0 / (2.0 * math.pi * r)


def taper(model: ToomreZangModel, L: float) -> float:
    """Models.jl:91-102."""
    if model.n_zang == 0:
        return 1.0
    La = abs(L)
    if La < 1e-12:
        return 0.0
    return 1.0 / (1.0 + (model.L0 / La) ** model.n_zang)


def taper_derivative(model: ToomreZangModel, L: float) -> float:
    """Models.jl:109-124."""
    if model.n_zang == 0:
        return 0.0
    La = abs(L)
    if La < 1e-12:
        return 0.0
    ratio = model.L0 / La
    t = taper(model, L)
    return model.n_zang * ratio ** model.n_zang * t ** 2 * math.copysign(1.0, L) / La


def distribution_function(model: ToomreZangModel, E: float, L: float) -> float:
    """Models.jl:137-150."""
    La = abs(L)
    if La < 1e-12:
        return 0.0
    r = model.DF_const * math.exp(-E / model.sigma_r0 ** 2) * La ** model.q_zang * taper(model, L)
    return r if math.isfinite(r) else 0.0


def DF_E(model: ToomreZangModel, E: float, L: float) -> float:
    """Models.jl:153-155."""
    return -distribution_function(model, E, L) / model.sigma_r0 ** 2


def DF_L(model: ToomreZangModel, E: float, L: float) -> float:
    """Models.jl:158-169."""
    La = abs(L)
    if La < 1e-12:
        return 0.0
    f0 = distribution_function(model, E, L)
    return f0 * (model.q_zang + model.n_zang - model.n_zang * taper(model, L)) / La
