"""Host eigen-solvers behind the recycled PCG deflation space
(maplab_amd/csrc/ba_ritz.cc) against numpy/scipy on random problems and
matrices with known spectra; the library code is compiled here with g++."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def lib():
    out = os.path.join(tempfile.mkdtemp(), "libritz_test.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out,
                    os.path.join(ROOT, "maplab_amd", "csrc", "ba_ritz.cc"),
                    os.path.join(HERE, "harness", "ritz_harness.cc")], check=True)
    L = ctypes.CDLL(out)
    P = ctypes.POINTER(ctypes.c_double)
    for f in (L.ritz_tridiag, L.ritz_generalized):
        f.restype = ctypes.c_int
    L.ritz_tridiag.argtypes = [ctypes.c_int, P, P, ctypes.c_int, P, P]
    L.ritz_generalized.argtypes = [ctypes.c_int, P, P, ctypes.c_int, P, P]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def tridiag(lib, d, e, k):
    m = len(d)
    vals, vecs = np.zeros(k), np.zeros((m, k))
    assert lib.ritz_tridiag(m, _p(d), _p(e), k, _p(vals), _p(vecs)) == 0
    return vals, vecs


@pytest.mark.parametrize("m,k,seed", [(8, 3, 0), (40, 12, 1), (64, 12, 2), (64, 20, 3), (5, 5, 4)])
def test_tridiag_matches_dense_eigh(lib, m, k, seed):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.5, 3.0, m)
    e = rng.uniform(-1.0, 1.0, m - 1)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    vals, vecs = tridiag(lib, d, e, k)
    ref = np.linalg.eigvalsh(T)[:k]
    np.testing.assert_allclose(vals, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # Eigen-residual and orthonormality (vectors are defined up to sign).
    np.testing.assert_allclose(T @ vecs, vecs * vals, atol=1e-9)
    np.testing.assert_allclose(vecs.T @ vecs, np.eye(k), atol=1e-9)


def test_tridiag_known_spectrum(lib):
    # The (-1, 2, -1) matrix of order m has eigenvalues 2 - 2 cos(j pi / (m + 1)).
    m, k = 50, 6
    vals, vecs = tridiag(lib, np.full(m, 2.0), np.full(m - 1, -1.0), k)
    j = np.arange(1, k + 1)
    np.testing.assert_allclose(vals, 2 - 2 * np.cos(j * np.pi / (m + 1)), atol=1e-13)
    # Eigenvector j is sin(i j pi / (m + 1)), up to sign and scale.
    i = np.arange(1, m + 1)
    for c in range(k):
        ref = np.sin(i * (c + 1) * np.pi / (m + 1))
        ref /= np.linalg.norm(ref)
        assert abs(abs(ref @ vecs[:, c]) - 1.0) < 1e-9


def test_tridiag_clustered_pair_stays_orthogonal(lib):
    # A direct sum of two identical blocks: every eigenvalue is double.
    b = 10
    d = np.concatenate([np.linspace(1, 2, b), np.linspace(1, 2, b)])
    e = np.concatenate([np.full(b - 1, 0.3), [0.0], np.full(b - 1, 0.3)])
    vals, vecs = tridiag(lib, d, e, 4)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    np.testing.assert_allclose(vals, np.linalg.eigvalsh(T)[:4], atol=1e-12)
    np.testing.assert_allclose(vecs.T @ vecs, np.eye(4), atol=1e-8)
    np.testing.assert_allclose(T @ vecs, vecs * vals, atol=1e-8)


@pytest.mark.parametrize("n,k,seed", [(6, 2, 0), (24, 12, 1), (32, 16, 2)])
def test_generalized_matches_scipy(lib, n, k, seed):
    import scipy.linalg as sla
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    A = G @ G.T + n * np.eye(n)
    H = rng.standard_normal((n, n))
    B = H @ H.T + 0.5 * np.eye(n)
    vals, C = np.zeros(k), np.zeros((n, k))
    assert lib.ritz_generalized(n, _p(A), _p(B), k, _p(vals), _p(C)) == 0
    ref = sla.eigh(A, B, eigvals_only=True)[:k]
    np.testing.assert_allclose(vals, ref, rtol=1e-10)
    np.testing.assert_allclose(A @ C, B @ C * vals, atol=1e-8 * np.abs(A).max())
    np.testing.assert_allclose(C.T @ B @ C, np.eye(k), atol=1e-9)


def test_generalized_rejects_indefinite_b(lib):
    n = 4
    A = np.eye(n)
    B = np.diag([1.0, 1.0, -1.0, 1.0])
    vals, C = np.zeros(2), np.zeros((n, 2))
    assert lib.ritz_generalized(n, _p(A), _p(B), 2, _p(vals), _p(C)) == -1
