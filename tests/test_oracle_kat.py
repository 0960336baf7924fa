"""End-to-end known answers: the published convergence-table roots.

convergence_table.txt:13-24 lists ten Newton roots of det(I - M) = 0 at ten
grids, printed to 10 decimals (MATLAB FP64 runs of the same algorithm).  The
oracle's whole chain -- grids, orbits, Jacobian, DF, W_l, N_m, M, LU det and
the Newton driver -- must land within 1e-10 of each (print rounding 5e-11 plus
the runs' own stopping tolerance).
"""
import json
import os

import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "convergence_table.json")))


@pytest.mark.parametrize("row", range(len(GOLD["rows"])))
def test_published_root(row):
    Na, am, lm, vm, r0, r1, nw, nwa, NR, Ne, Op, g = GOLD["rows"][row]
    p = oracle.Params(N_alpha=Na, alpha_max=am, l_min=-lm, l_max=lm, v_min=vm, Rc_min=r0,
                      Rc_max=r1, nw=nw, nwa=nwa, NR=NR, Ne=Ne, **GOLD["model"])
    pre = oracle.precompute(p)
    r = oracle.newton(pre, Op + 1e-4, g - 1e-4, max_iter=50, tol=1e-12, delta=1e-6)
    assert r["converged"]
    assert abs(r["Omega_p"] - Op) <= 1e-10
    assert abs(r["gamma"] - g) <= 1e-10
