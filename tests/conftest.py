import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CONFIGS = os.path.join(ROOT, "kalnajs_log_spirals_amd", "configs")
MODELS = os.path.join(ROOT, "kalnajs_log_spirals_amd", "models")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs")


@pytest.fixture(scope="session")
def small_cfg_paths():
    return os.path.join(CONFIGS, "small.toml"), os.path.join(MODELS, "kalnajs_zang4.toml")


@pytest.fixture(scope="session")
def oracle_small(small_cfg_paths):
    """Oracle precompute of the published small grid (convergence_table.txt:13)."""
    import oracle
    p = oracle.load_toml_params(*small_cfg_paths)
    return oracle.precompute(p)


TINY = {
    # general orbits plus the circular v = 1 column
    "general": dict(NR=9, Ne=7, Rc_min=0.05, Rc_max=20.0, v_min=0.01, N_alpha=21,
                    alpha_max=8.0, l_min=-3, l_max=3, nw=401, nwa=41),
    # v close to 1: the epicyclic branch (e < 0.01) and the circular column
    "epicyclic": dict(NR=5, Ne=7, Rc_min=0.1, Rc_max=10.0, v_min=0.97, N_alpha=19,
                      alpha_max=6.0, l_min=-2, l_max=4, nw=201, nwa=21),
}


@pytest.fixture(scope="session", params=sorted(TINY))
def oracle_tiny(request):
    """Tiny grids that exercise every orbit branch of compute_orbits."""
    import oracle
    return oracle.precompute(oracle.Params(**TINY[request.param]))
