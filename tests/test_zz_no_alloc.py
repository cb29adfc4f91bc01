"""pfem_element_batch_noalloc (no device allocation; its buffers live in one static device
array) against pfem_element_batch, through the C-ABI: two calls in a row, the second
smaller and of another element kind with optional inputs left out (so stale contents of
the shared array from the first call would show), bit for bit equal; an inverted element
is reported the same way; sizes above the limit are refused."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import plasticity_amd._lib as L

pytestmark = pytest.mark.gpu
MAT = dict(E=200e9, nu=0.3, sy0=250e6, Hiso=1e9, Hkin=0.0, ssat=250e6, delta=0.0)


def _call(name, kind, mat, X, u, epsp, beta, abar, cpinv, commit):
    n = X.shape[0]
    Fe, Ke, sig = np.empty((n, 24)), np.empty((n, 576)), np.empty((n, 8, 6))
    f = L.Fail()
    m = L.Material(**mat)
    p = lambda a: None if a is None else L.ptr(a)  # noqa: E731
    rc = getattr(L.lib(), name)(kind, C.byref(m), n, L.ptr(X), L.ptr(u), p(epsp), p(beta), p(abar), p(cpinv),
                                int(commit), L.ptr(Fe), L.ptr(Ke), L.ptr(sig), C.byref(f))
    return rc, Fe, Ke, sig, (f.e, f.g, f.J)


def _elements(n, seed, scale):
    rng = np.random.default_rng(seed)
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    X = np.stack([(nodes[conn[0]] * (1 + 0.1 * rng.random(3)) + 0.05 * rng.standard_normal((8, 3))).ravel()
                  for _ in range(n)])
    u = scale * rng.standard_normal((n, 24))
    return np.ascontiguousarray(X), np.ascontiguousarray(u)


def _history(n, seed):
    rng = np.random.default_rng(seed)
    epsp = 1e-3 * rng.standard_normal((n, 8, 6))
    abar = 1e-3 * rng.random((n, 8))
    cpinv = np.tile([1.0, 1, 1, 0, 0, 0], (n, 8, 1)) + 1e-3 * rng.standard_normal((n, 8, 6))
    beta = 1e6 * rng.standard_normal((n, 8, 6))
    return epsp, beta, abar, cpinv


def _compare(kind, mat, n, seed, scale, with_beta, with_cpinv, commit):
    X, u = _elements(n, seed, scale)
    outs = []
    for name in ("pfem_element_batch", "pfem_element_batch_noalloc"):
        epsp, beta, abar, cpinv = _history(n, seed + 1)
        r = _call(name, kind, mat, X, u, epsp, beta if with_beta else None, abar, cpinv if with_cpinv else None,
                  commit)
        outs.append(r + (epsp, beta, abar, cpinv))
    (rc0, *a), (rc1, *b) = outs
    assert rc0 == rc1 == L.PFEM_OK, (rc0, rc1)
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y)
    return a


def test_noalloc_equals_the_allocating_batch_twice_in_a_row(need_gpu):
    print("[noalloc] call 1: 1500 F-bar elements, plastic, committed history", flush=True)
    Fe, Ke, sig, *_ = _compare(L.HEX8_FINITE_FBAR, MAT, 1500, 1, 2e-3, True, True, True)
    assert np.abs(Fe).max() > 0 and np.abs(Ke).max() > 0
    print("[noalloc] call 2: 37 small-strain elements, kinematic hardening, no beta / cpinv", flush=True)
    _compare(L.HEX8_SMALL, dict(MAT, Hkin=5e8), 37, 7, 1e-3, False, False, True)
    print("[noalloc] call 3: 2048 finite-strain elements, kinematic hardening (the maximum)", flush=True)
    _compare(L.HEX8_FINITE, dict(MAT, Hkin=5e8), L.BATCH_NOALLOC_MAX, 3, 2e-3, True, True, False)


def test_noalloc_reports_inversion_and_refuses_oversize(need_gpu):
    X, u = _elements(5, 11, 1e-3)
    u[3] = -2.0 * X[3]  # element 3 turned inside out
    res = []
    for name in ("pfem_element_batch", "pfem_element_batch_noalloc"):
        epsp, beta, abar, cpinv = _history(5, 12)
        res.append(_call(name, L.HEX8_FINITE_FBAR, MAT, X, u, epsp, None, abar, cpinv, False))
    assert res[0][0] == res[1][0] == L.PFEM_ERR_INVERTED
    assert res[0][4] == res[1][4] and res[0][4][0] == 3
    X, u = _elements(L.BATCH_NOALLOC_MAX + 1, 13, 1e-3)
    epsp, beta, abar, cpinv = _history(L.BATCH_NOALLOC_MAX + 1, 14)
    rc, *_ = _call("pfem_element_batch_noalloc", L.HEX8_FINITE_FBAR, MAT, X, u, epsp, None, abar, cpinv, False)
    assert rc == L.PFEM_ERR_BAD_ARG
