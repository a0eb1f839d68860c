"""Pins the oracle's own FFT: against an O(N^2) DFT written from the
definition (forward exp(-i), unnormalised, as cufftExecC2C), against numpy's
FP64 FFT, and against the shift / Parseval theorems."""
import numpy as np
import pytest

import orc

SIZES = [1, 2, 3, 4, 5, 7, 8, 11, 12, 13, 16, 17, 20, 27, 30, 61, 64, 107, 220, 229, 459]


def dft(x, inverse=False):
    n = x.size
    k = np.arange(n)
    s = 1.0 if inverse else -1.0
    w = np.exp(s * 2j * np.pi * np.outer(k, k) / n)
    return w @ x


@pytest.mark.parametrize("n", SIZES)
def test_fft1d_vs_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for inv in (False, True):
        ref = dft(x, inv)
        got = orc.fft1d(x, inv)
        assert np.abs(got - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max())
        assert np.allclose(got, np.fft.ifft(x) * n if inv else np.fft.fft(x), atol=1e-10)


@pytest.mark.parametrize("shape", [(6, 10), (17, 12), (459, 220), (33, 45)])
def test_fft2d_vs_numpy(shape):
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = orc.fft2d(a)
    assert np.allclose(got, np.fft.fft2(a), atol=1e-9 * np.sqrt(a.size))
    back = orc.fft2d(got, inverse=True) / a.size
    assert np.allclose(back, a, atol=1e-12)


def test_fft_theorems():
    rng = np.random.default_rng(1)
    n = 459
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = orc.fft1d(x)
    # Parseval
    assert np.isclose(np.sum(np.abs(X) ** 2), n * np.sum(np.abs(x) ** 2), rtol=1e-12)
    # shift theorem: x[j-1] <-> X[k] exp(-2 pi i k/n)
    Xs = orc.fft1d(np.roll(x, 1))
    k = np.arange(n)
    assert np.allclose(Xs, X * np.exp(-2j * np.pi * k / n), atol=1e-9)
    # impulse -> constant
    e = np.zeros(n, complex)
    e[0] = 1
    assert np.allclose(orc.fft1d(e), np.ones(n))
