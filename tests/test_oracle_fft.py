"""Pins of the oracle's FFT (the stand-in for FFTW 3, which the reference links
at g2evolutionpert.cpp:183-195, g2spectra.cpp:47-61, g2init.cpp:55-63).

Checked against the textbook DFT by brute force, against numpy's pocketfft
(an independent library), and by the c2r/r2c inverse relation.  The c2r of a
non-Hermitian half spectrum (the reference's IC input, Q8) is checked against
pocketfft's irfftn, which like FFTW's c2r reads only the Hermitian part of the
kz = 0 and kz = N/2 planes.
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as po


def brute_r2c(x):
    N = x.shape[0]
    out = np.zeros((N, N, N // 2 + 1), dtype=complex)
    n = np.arange(N)
    for kx, ky, kz in itertools.product(range(N), range(N), range(N // 2 + 1)):
        ph = np.exp(-2j * np.pi * (kx * n[:, None, None] + ky * n[None, :, None] + kz * n[None, None, :]) / N)
        out[kx, ky, kz] = np.sum(x * ph)
    return out


@pytest.mark.parametrize("N", [4, 6])
def test_r2c_equals_brute_force_dft(N):
    x = np.random.default_rng(N).standard_normal((N, N, N))
    got = po.r2c3d(x)
    ref = brute_r2c(x)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("N", [8, 10, 12, 16, 24, 32, 48])
def test_r2c_matches_pocketfft(N):
    x = np.random.default_rng(N).standard_normal((N, N, N))
    got = po.r2c3d(x)
    ref = np.fft.rfftn(x, axes=(0, 1, 2))
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("N", [8, 12, 16])
def test_c2r_inverts_r2c_unnormalised(N):
    x = np.random.default_rng(N + 1).standard_normal((N, N, N))
    y = po.c2r3d(po.r2c3d(x))
    assert np.abs(y / N ** 3 - x).max() < 1e-13


@pytest.mark.parametrize("N", [8, 12])
def test_c2r_of_non_hermitian_planes(N):
    rng = np.random.default_rng(N + 2)
    X = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    got = po.c2r3d(X)
    ref = np.fft.irfftn(X, s=(N, N, N), axes=(0, 1, 2)) * N ** 3
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    # and it equals the transform of the explicitly symmetrised planes
    S = X.copy()
    for kz in (0, N // 2):
        P = X[:, :, kz]
        S[:, :, kz] = 0.5 * (P + np.conj(np.roll(P[::-1, ::-1], 1, axis=(0, 1))))
    assert np.abs(po.c2r3d(S) - got).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 9, 12, 30])
def test_fft1d_mixed_radix(n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 9).standard_normal(n)
    assert np.abs(po.fft1d(x, -1) - np.fft.fft(x)).max() < 1e-13 * max(1, np.abs(x).sum())
    assert np.abs(po.fft1d(x, +1) - np.fft.ifft(x) * n).max() < 1e-13 * max(1, np.abs(x).sum())
