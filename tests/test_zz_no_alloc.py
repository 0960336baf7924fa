"""klsp_det_noalloc (the FP64 det with its workspaces in the library's fixed
per-GPU scratch array) against klsp_det through the C ABI: two calls in a
row on one context, the second on a smaller problem; the set-up it needs;
and a problem too large for the scratch, which it refuses."""
import numpy as np
import pytest

from kalnajs_log_spirals_amd import _ffi
from kalnajs_log_spirals_amd.BasisFunctions import from_arrays

pytestmark = pytest.mark.gpu


def det_noalloc(ctx, omegas):
    om = np.ascontiguousarray(np.atleast_1d(np.asarray(omegas, dtype=np.complex128)))
    out = np.full(om.size, complex(np.nan, np.nan))
    st = _ffi.lib().klsp_det_noalloc(ctx.handle, _ffi.dptr(om.view(np.float64)), om.size,
                                     _ffi.dptr(out.view(np.float64)))
    return st, out


def det_c(ctx, omegas):
    om = np.ascontiguousarray(np.atleast_1d(np.asarray(omegas, dtype=np.complex128)))
    out = np.empty(om.size, dtype=np.complex128)
    st = _ffi.lib().klsp_det(ctx.handle, _ffi.dptr(om.view(np.float64)), om.size,
                             _ffi.PREC_FP64, _ffi.dptr(out.view(np.float64)))
    assert st == _ffi.KLSP_OK
    return out


def problem(rng, NPh, N, n_l, symmetric):
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l)))
    W *= 1.0 / np.sqrt(NPh * n_l)
    if symmetric:                  # W[:, N-1-a] = conj W[:, a], real centre column
        h = N // 2
        for a in range(h):
            W[:, N - 1 - a, :] = np.conj(W[:, a, :])
        if N % 2:
            W[:, h, :] = W[:, h, :].real
    return dict(W=W, DJ=rng.uniform(0.2, 1.0, NPh), F0l=rng.standard_normal((NPh, n_l)),
                Ores=rng.uniform(-2, 2, (NPh, n_l)), Nk=rng.uniform(0.5, 1.5, N))


def test_det_noalloc_matches_det_on_two_calls():
    rng = np.random.default_rng(77)
    big = problem(rng, 3001, 130, 4, symmetric=True)        # real-basis kernel
    small = problem(rng, 509, 37, 2, symmetric=False)      # complex kernel
    pre = from_arrays(2, 0.9, big["W"], big["DJ"], big["F0l"], big["Ores"], big["Nk"])
    ctx = pre.ctx
    try:
        oms1 = [0.30 + 0.40j, 0.9 + 0.1j, 0.55 + 0.2j, 1.2 - 0.3j, 0.7 + 0.05j]
        # the work table and the real basis are made by klsp_prepare, not here
        st, _ = det_noalloc(ctx, oms1[:1])
        assert st == _ffi.KLSP_ERR_NOT_READY
        ctx.prepare(_ffi.PREC_FP64)
        assert ctx.symmetric_basis()
        ref1 = det_c(ctx, oms1)
        st, got1 = det_noalloc(ctx, oms1)
        print("call 1:", st, got1, ref1, flush=True)
        assert st == _ffi.KLSP_OK
        assert np.array_equal(got1, ref1)              # same kernels, same data
        # second call: a smaller problem on the same context (the scratch keeps
        # the first call's contents, which must not leak into this one)
        ctx.upload_problem(509, 37, 2, 2, 0.9, small["DJ"], small["F0l"], small["Ores"],
                           small["Nk"])
        ctx.upload_W(small["W"])
        ctx.prepare(_ffi.PREC_FP64)
        assert not ctx.symmetric_basis()
        oms2 = [0.45 + 0.15j, 0.8 - 0.2j]
        ref2 = det_c(ctx, oms2)
        st, got2 = det_noalloc(ctx, oms2)
        print("call 2:", st, got2, ref2, flush=True)
        assert st == _ffi.KLSP_OK
        assert np.array_equal(got2, ref2)
        # and against the oracle's BLAS-form det (parity bar of the FP64 path)
        from oracle import blas_form
        for k, om in enumerate(oms2):
            d = blas_form.det(small["W"], small["DJ"], small["F0l"], small["Ores"], small["Nk"],
                              0.9, om)
            assert abs(got2[k] - d) <= 1e-11 * max(abs(d), 1.0)
    finally:
        ctx.close()


def test_det_noalloc_refuses_a_problem_beyond_the_scratch():
    """K = 17.6M rows: the weights alone need 282 MB, more than the 256 MiB
    scratch; klsp_det (which allocates) still evaluates it."""
    rng = np.random.default_rng(78)
    NPh, N, n_l = 2_200_000, 2, 8
    W = (rng.standard_normal((NPh, N, n_l)) + 1j * rng.standard_normal((NPh, N, n_l))) * 1e-4
    pre = from_arrays(2, 1.0, W, rng.uniform(0.2, 1.0, NPh), rng.standard_normal((NPh, n_l)),
                      rng.uniform(-2, 2, (NPh, n_l)), rng.uniform(0.5, 1.5, N))
    ctx = pre.ctx
    try:
        ctx.prepare(_ffi.PREC_FP64)
        st, _ = det_noalloc(ctx, [0.5 + 0.2j])
        assert st == _ffi.KLSP_ERR_INVALID_ARG
        assert "scratch" in _ffi.lib().klsp_last_error(ctx.handle).decode()
        assert np.isfinite(det_c(ctx, [0.5 + 0.2j])[0])
    finally:
        ctx.close()
