"""Configuration mirror (Configuration.jl): defaults, TOML keys, round trip."""
import os

import numpy as np

import kalnajs_log_spirals_amd as K
from kalnajs_log_spirals_amd import Configuration as C


def test_struct_defaults_match_reference():
    c = K.KalnajsConfig()
    assert (c.physics.m, c.physics.G) == (2, 1.0)
    assert (c.model.L0, c.model.n_zang, c.model.q1) == (1.0, 4, 7)
    g = c.grid
    assert (g.NR, g.Ne, g.Rc_min, g.Rc_max, g.v_min) == (201, 21, 1e-3, 100.0, 1e-2)
    assert (g.N_alpha, g.alpha_max, g.l_min, g.l_max, g.nw, g.nwa) == (101, 10.0, -25, 25, 10001, 201)
    assert (c.newton.max_iter, c.newton.tol, c.newton.delta, c.newton.use_line_search) == \
        (50, 1e-10, 1e-6, True)
    assert (c.scan.n_Op, c.scan.n_gamma) == (41, 20)


def test_load_default_and_model(small_cfg_paths):
    cfgdir = os.path.dirname(small_cfg_paths[0])
    c = K.load_model(K.load_config(os.path.join(cfgdir, "default.toml")), small_cfg_paths[1])
    assert (c.grid.NR, c.grid.Ne, c.grid.N_alpha, c.grid.l_min, c.grid.l_max) == (501, 51, 301, -100, 100)
    assert (c.grid.nw, c.grid.nwa, c.grid.alpha_max) == (100001, 2001, 30.0)
    assert c.newton.tol == 1e-12 and c.newton.use_line_search is False
    assert c.reference.Omega_p == 0.4397142267 and c.reference.gamma == 0.1230663944
    assert c.physics.m == 2 and c.model.n_zang == 4
    assert C.get_float_type(c, True) == np.float64
    assert C.get_n_l(c) == 201 and C.get_NPh(c) == 501 * 51


def test_save_load_round_trip(tmp_path, small_cfg_paths):
    c = K.load_model(K.load_config(small_cfg_paths[0]), small_cfg_paths[1])
    c.scan.n_Op = 7
    f = tmp_path / "x.toml"
    K.save_config(c, str(f))
    d = K.load_config(str(f))
    assert d.grid == c.grid and d.scan == c.scan and d.physics == c.physics
    assert d.model == c.model


def test_uneven_batch_draws_match_fixture():
    """The 48 draws of the uneven batch (SplitMix64 seed 0) are the recorded
    ones, and every draw lies in its stated ranges."""
    import json
    from dataclasses import asdict
    from kalnajs_log_spirals_amd.UnevenBatch import draw_batch, config_for
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "uneven_batch_draws.json")))
    draws = draw_batch(48, 0)
    assert [asdict(d) for d in draws] == fx["draws"]
    for d in draws:
        assert d.n_zang in (2, 3, 4, 5, 6) and d.q1 in (4, 6, 7, 9, 12) and d.m in (1, 2, 3)
        assert 201 <= d.NR <= 801 and 21 <= d.Ne <= 101 and 101 <= d.N_alpha <= 401
        assert 50 <= d.l_max <= 150 and 0.3 <= d.Omega_p0 < 0.6 and 0.05 <= d.gamma0 < 0.2
    c = config_for(draws[0])
    assert (c.grid.NR, c.grid.Ne, c.grid.l_min, c.grid.nwa) == (draws[0].NR, draws[0].Ne,
                                                               -draws[0].l_max, 2001)
    cs = config_for(draws[0], scale=0.1)
    assert cs.grid.NR == max(11, round(draws[0].NR * 0.1)) and cs.grid.nwa == 201
