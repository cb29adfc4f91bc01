"""Coarsest-level AMG solve: the device Gauss-Jordan inverse (blocked panels of 16/8/4
pivots while a panel fits in LDS, one launch pair per pivot beyond that) against LAPACK,
including a matrix that needs pivoting in every column and an all-zero row (decoupled
with a unit diagonal)."""
import numpy as np
import pytest

import plasticity_amd._lib as L

pytestmark = pytest.mark.gpu


def dense_inverse(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    inv = np.empty_like(A)
    L.check(L.lib().pfem_dense_inverse(A.shape[0], L.ptr(A), L.ptr(inv)), "pfem_dense_inverse")
    return inv


@pytest.mark.parametrize("n", [1, 5, 37, 378, 700, 1278, 3200])
def test_dense_inverse_matches_lapack(n, need_gpu):
    print(f"[dense inverse] n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 2.0 * np.eye(n)
    inv = dense_inverse(A)
    ref = np.linalg.inv(A)
    err = np.linalg.norm(inv - ref) / np.linalg.norm(ref)
    print(f"  relative difference to LAPACK {err:.2e}", flush=True)
    assert err < 1e-10


def test_dense_inverse_pivoting_and_zero_row(need_gpu):
    n = 200
    rng = np.random.default_rng(1)
    perm = rng.permutation(n)
    A = np.zeros((n, n))
    A[np.arange(n), perm] = 1.0 + rng.random(n)  # zero diagonal: every column needs a row exchange
    A += 1e-3 * rng.normal(size=(n, n))
    inv = dense_inverse(A)
    assert np.linalg.norm(inv @ A - np.eye(n)) < 1e-10 * n
    B = rng.normal(size=(n, n)) + 2.0 * np.eye(n)
    B[17] = 0.0
    Bm = B.copy()
    Bm[17, 17] = 1.0
    assert np.linalg.norm(dense_inverse(B) - np.linalg.inv(Bm)) < 1e-10 * np.linalg.norm(np.linalg.inv(Bm))
