"""The library's dense device product against numpy, through vsvb_debug_gemm:
small products (vector-FMA tiles) and large ones (FP64 matrix-core tiles), edge
shapes smaller than, equal to and straddling the 16 x 16 tile, every transpose
combination, and beta != 0.
Tolerance: 1e-13 of K * max|A| * max|B| (+ |beta| max|C0|) per element."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# small shapes take the vector-FMA tiles, the last three (>= 8e6 multiply-adds) the matrix-core tiles
SHAPES = [(1, 1, 1), (16, 16, 4), (63, 65, 17), (64, 64, 64), (130, 7, 300), (5, 200, 33), (257, 129, 70),
          (300, 200, 150), (1031, 77, 129), (200, 333, 257)]


@pytest.fixture(scope="module")
def lib():
    from valence_amd import lib as L
    L.load()
    return L


@pytest.mark.parametrize("shape", SHAPES)
def test_gemm_matches_numpy(lib, shape):
    M, N, K = shape
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    for ta, tb in itertools.product([False, True], [False, True]):
        a = rng.normal(size=(M, K)); b = rng.normal(size=(K, N))
        A = a.T.copy() if ta else a
        B = b.T.copy() if tb else b
        got = lib.device_gemm(A, B, ta, tb)
        ref = a @ b
        tol = 1e-13 * K * np.abs(a).max() * np.abs(b).max()
        err = np.abs(got - ref).max()
        print(f"M={M} N={N} K={K} ta={ta} tb={tb} max err {err:.2e} tol {tol:.2e}")
        assert err <= tol


@pytest.mark.parametrize("shape", [(70, 90, 45), (300, 250, 200)])
def test_gemm_alpha_beta(lib, shape):
    M, N, K = shape
    rng = np.random.default_rng(5)
    a = rng.normal(size=(M, K)); b = rng.normal(size=(K, N)); c0 = rng.normal(size=(M, N))
    got = lib.device_gemm(a, b, alpha=-0.5, beta=2.0, C0=c0)
    ref = -0.5 * (a @ b) + 2.0 * c0
    assert np.abs(got - ref).max() <= 1e-13 * (K * np.abs(a).max() * np.abs(b).max() + 2 * np.abs(c0).max())
