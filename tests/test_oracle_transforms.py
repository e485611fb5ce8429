"""Pins the oracle's padded-FFT transforms and eigen-tables to library
routines and to linear algebra (CPU only)."""

import numpy as np
import pytest
import scipy.fft


@pytest.fixture(params=[5, 9, 17, 33])
def n(request):
    return request.param


def test_dst2d_is_scipy_dstn_type1(oracle, n):
    a = np.random.default_rng(n).standard_normal((n, n))
    ref = scipy.fft.dstn(a, type=1)
    assert np.allclose(oracle.dst2d(a), ref, rtol=0, atol=1e-12 * abs(ref).max())


def test_dct2d_is_minus_scipy_dctn_type1(oracle, n):
    a = np.random.default_rng(n + 1).standard_normal((n, n))
    ref = -scipy.fft.dctn(a, type=1)
    assert np.allclose(oracle.dct2d(a), ref, rtol=0, atol=1e-12 * abs(ref).max())


@pytest.mark.parametrize("m,nn", [(3, 5), (7, 9), (15, 17), (31, 33)])
def test_mix2d_is_dst_then_dct(oracle, m, nn):
    a = np.random.default_rng(m).standard_normal((m, nn))
    out = oracle.mix2d(a)
    assert out.shape == (m + 2, nn)
    assert np.all(out[0] == 0) or np.abs(out[0]).max() < 1e-12
    assert np.abs(out[-1]).max() < 1e-12 * np.abs(out).max()
    ref = scipy.fft.dct(scipy.fft.dst(a, type=1, axis=0), type=1, axis=1)
    assert np.allclose(out[1:-1], ref, rtol=0, atol=1e-12 * abs(ref).max())


def test_transforms_are_self_inverse(oracle, n):
    rng = np.random.default_rng(7)
    a = rng.standard_normal((n, n))
    # DST-I of length L: applied twice gives (2(L+1))^2 in 2D.
    assert np.allclose(oracle.dst2d(oracle.dst2d(a)), (2 * (n + 1))**2 * a)
    assert np.allclose(oracle.dct2d(oracle.dct2d(a)), (2 * (n - 1))**2 * a)
    b = rng.standard_normal((n - 2, n))
    twice = oracle.mix2d(oracle.mix2d(b)[1:-1])[1:-1]
    assert np.allclose(twice, (2 * (n - 1))**2 * b)


def _second_difference(n, h, kind):
    """1D -d2/dx2 matrix: 'dirichlet' on n interior points, 'neumann' with
    even-reflection ghosts on n points."""
    a = (2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)) / h**2
    if kind == 'neumann':
        a[0, 1] = a[-1, -2] = -2 / h**2
    return a


@pytest.mark.parametrize("N,h", [(9, .1), (17, .05), (33, .025)])
def test_eigen_tables_match_brute_force_eigenvalues(oracle, N, h):
    lam_d = np.sort(np.linalg.eigvals(_second_difference(N - 2, h, 'dirichlet')).real)
    lam_n = np.sort(np.linalg.eigvals(_second_difference(N, h, 'neumann')).real)
    scale = (2 * (N - 1))**2
    md = oracle.dirichlet_matrix(N, h) * scale
    assert np.allclose(np.sort(np.diag(1 / md)) / 2, lam_d)
    mm = oracle.mixed_matrix(N, h, 1) * scale
    # 1/mm[i, j] - 1 = lam_d[i] + lam_n[j]; take j = 0 (lam_n[0] = 0).
    assert np.allclose(np.sort(1 / mm[:, 0] - 1), lam_d)
    assert np.allclose(np.sort(1 / mm[0, :] - 1 - 1 / mm[0, 0] + 1), lam_n)
    mm0 = oracle.mixed_matrix(N, h, 0) * scale
    assert np.allclose(1 / mm0, 1 / mm - 1)
    mn = oracle.neumann_matrix(N, h) * scale
    assert mn[0, 0] == 0
    assert np.allclose(np.sort(1 / mn[1:, 0]), lam_n[1:])
