"""Configuration -- mirror of src/KalnajsLogSpiral/Configuration.jl.

Same struct names, field names, defaults (Configuration.jl:31-149) and TOML
keys (load_config :160-261, load_model :269-300, save_config :307-371).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List

import numpy as np

__all__ = [
    "PhysicsConfig", "ModelConfig", "GridConfig", "ScanConfig", "NewtonConfig", "GPUConfig",
    "CPUConfig", "IOConfig", "ReferenceConfig", "KalnajsConfig", "load_config", "load_model",
    "save_config", "get_float_type", "get_complex_type", "get_n_l", "get_NPh",
]


@dataclass
class PhysicsConfig:
    m: int = 2
    G: float = 1.0


@dataclass
class ModelConfig:
    L0: float = 1.0
    n_zang: int = 4
    q1: int = 7


@dataclass
class GridConfig:
    NR: int = 201
    Ne: int = 21
    Rc_min: float = 1e-3
    Rc_max: float = 100.0
    v_min: float = 1e-2
    N_alpha: int = 101
    alpha_max: float = 10.0
    l_min: int = -25
    l_max: int = 25
    nw: int = 10001
    nwa: int = 201


@dataclass
class ScanConfig:
    n_Op: int = 41
    n_gamma: int = 20
    Omega_p_min: float = 0.2
    Omega_p_max: float = 0.6
    gamma_min: float = 0.01
    gamma_max: float = 0.2


@dataclass
class NewtonConfig:
    max_iter: int = 50
    tol: float = 1e-10
    delta: float = 1e-6
    use_line_search: bool = True
    max_line_search: int = 10
    use_complex_step: bool = True
    complex_step_h: float = 1e-20
    enable_convergence_study: bool = False


@dataclass
class GPUConfig:
    backend: str = "auto"
    precision_double: bool = False
    devices: List[int] = field(default_factory=list)


@dataclass
class CPUConfig:
    precision_double: bool = False
    max_threads: int = 32


@dataclass
class IOConfig:
    output_path: str = "results"
    save_intermediate: bool = False
    verbose: bool = True


@dataclass
class ReferenceConfig:
    Omega_p: float = 0.44
    gamma: float = 0.13


@dataclass
class KalnajsConfig:
    physics: PhysicsConfig = field(default_factory=PhysicsConfig)
    model: ModelConfig = field(default_factory=ModelConfig)
    grid: GridConfig = field(default_factory=GridConfig)
    scan: ScanConfig = field(default_factory=ScanConfig)
    newton: NewtonConfig = field(default_factory=NewtonConfig)
    gpu: GPUConfig = field(default_factory=GPUConfig)
    cpu: CPUConfig = field(default_factory=CPUConfig)
    io: IOConfig = field(default_factory=IOConfig)
    reference: ReferenceConfig = field(default_factory=ReferenceConfig)


def _toml(path: str) -> dict:
    try:
        import tomllib  # Python >= 3.11
    except ImportError:  # pragma: no cover - 3.10 image
        import tomli as tomllib
    with open(path, "rb") as f:
        return tomllib.load(f)


def _apply(obj, sec: dict, keys, cast=None):
    for k in keys:
        if k in sec:
            setattr(obj, k, sec[k])


def load_config(filepath: str) -> KalnajsConfig:
    """Configuration.jl:160-261 (the [precision] section is read twice there;
    once is equivalent)."""
    if not os.path.isfile(filepath):
        raise FileNotFoundError(f"Configuration file not found: {filepath}")
    data = _toml(filepath)
    c = KalnajsConfig()
    _apply(c.physics, data.get("physics", {}), ("m", "G"))
    _apply(c.model, data.get("model", {}), ("L0", "n_zang", "q1"))
    _apply(c.grid, data.get("grid", {}), ("NR", "Ne", "Rc_min", "Rc_max", "v_min", "N_alpha",
                                          "alpha_max", "l_min", "l_max", "nw", "nwa"))
    _apply(c.scan, data.get("scan", {}), ("n_Op", "n_gamma", "Omega_p_min", "Omega_p_max",
                                          "gamma_min", "gamma_max"))
    _apply(c.newton, data.get("newton", {}), ("max_iter", "tol", "delta", "use_line_search",
                                              "max_line_search", "use_complex_step",
                                              "complex_step_h", "enable_convergence_study"))
    prec = data.get("precision", {})
    if "gpu_double_precision" in prec:
        c.gpu.precision_double = prec["gpu_double_precision"]
    if "cpu_double_precision" in prec:
        c.cpu.precision_double = prec["cpu_double_precision"]
    gpu = data.get("gpu", {})
    if "backend" in gpu:
        c.gpu.backend = str(gpu["backend"])
    _apply(c.gpu, gpu, ("precision_double", "devices"))
    _apply(c.cpu, data.get("cpu", {}), ("precision_double", "max_threads"))
    _apply(c.io, data.get("io", {}), ("output_path", "save_intermediate", "verbose"))
    return c


def load_model(config: KalnajsConfig, filepath: str) -> KalnajsConfig:
    """Configuration.jl:269-300: [physics], [model] (G may sit there), [reference]."""
    if not os.path.isfile(filepath):
        raise FileNotFoundError(f"Model file not found: {filepath}")
    data = _toml(filepath)
    _apply(config.physics, data.get("physics", {}), ("m", "G"))
    m = data.get("model", {})
    _apply(config.model, m, ("L0", "n_zang", "q1"))
    if "G" in m:
        config.physics.G = m["G"]
    _apply(config.reference, data.get("reference", {}), ("Omega_p", "gamma"))
    return config


def _b(x: bool) -> str:
    return "true" if x else "false"


def save_config(config: KalnajsConfig, filepath: str) -> None:
    """Configuration.jl:307-371 (same sections and keys)."""
    c = config
    lines = ["# Kalnajs Log-Spiral Configuration", "",
             "[physics]", f"m = {c.physics.m}", f"G = {float(c.physics.G)}", "",
             "[model]", f"L0 = {float(c.model.L0)}", f"n_zang = {c.model.n_zang}",
             f"q1 = {c.model.q1}", "",
             "[grid]"]
    g = c.grid
    lines += [f"NR = {g.NR}", f"Ne = {g.Ne}", f"Rc_min = {float(g.Rc_min)}",
              f"Rc_max = {float(g.Rc_max)}", f"v_min = {float(g.v_min)}", f"N_alpha = {g.N_alpha}",
              f"alpha_max = {float(g.alpha_max)}", f"l_min = {g.l_min}", f"l_max = {g.l_max}",
              f"nw = {g.nw}", f"nwa = {g.nwa}", ""]
    s = c.scan
    lines += ["[scan]", f"n_Op = {s.n_Op}", f"n_gamma = {s.n_gamma}",
              f"Omega_p_min = {float(s.Omega_p_min)}", f"Omega_p_max = {float(s.Omega_p_max)}",
              f"gamma_min = {float(s.gamma_min)}", f"gamma_max = {float(s.gamma_max)}", ""]
    n = c.newton
    lines += ["[newton]", f"max_iter = {n.max_iter}", f"tol = {float(n.tol)}",
              f"delta = {float(n.delta)}", f"use_line_search = {_b(n.use_line_search)}",
              f"use_complex_step = {_b(n.use_complex_step)}",
              f"enable_convergence_study = {_b(n.enable_convergence_study)}", ""]
    lines += ["[gpu]", f'backend = "{c.gpu.backend}"',
              f"precision_double = {_b(c.gpu.precision_double)}", f"devices = {list(c.gpu.devices)}",
              "", "[cpu]", f"precision_double = {_b(c.cpu.precision_double)}",
              f"max_threads = {c.cpu.max_threads}", "",
              "[io]", f'output_path = "{c.io.output_path}"',
              f"save_intermediate = {_b(c.io.save_intermediate)}", f"verbose = {_b(c.io.verbose)}"]
    with open(filepath, "w") as f:
        f.write("\n".join(lines) + "\n")


def get_float_type(config: KalnajsConfig, for_gpu: bool = True):
    """Configuration.jl:380-386: numpy.float64 or numpy.float32."""
    dbl = config.gpu.precision_double if for_gpu else config.cpu.precision_double
    return np.float64 if dbl else np.float32


def get_complex_type(config: KalnajsConfig, for_gpu: bool = True):
    dbl = config.gpu.precision_double if for_gpu else config.cpu.precision_double
    return np.complex128 if dbl else np.complex64


def get_n_l(config: KalnajsConfig) -> int:
    return config.grid.l_max - config.grid.l_min + 1


def get_NPh(config: KalnajsConfig) -> int:
    return config.grid.NR * config.grid.Ne
