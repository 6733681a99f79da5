"""Pinning of the FP64 oracle (CPU only).

Each oracle function is checked against something independent of its own
formula: brute-force DFT matrices, numpy.fft / numpy.fft.rfft (a separate
implementation), closed forms and invariants, and brute-force correlation sums.
A plausible mistake (wrong sign, missing 1/N, wrong axis, transposed operand,
dropped conjugate, off-by-one Hermitian index) fails at least one test here.
"""
import numpy as np
import pytest

import oracle as O


RNG = np.random.default_rng(1234)


def crandn(*shape):
    return RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)


@pytest.mark.parametrize("n", list(range(1, 65)) + [81, 97, 121, 128, 162, 169, 210, 343, 720, 1001])
def test_dft_matches_bruteforce_matrix(n):
    x = crandn(2, n)
    ref = x @ O.dft_matrix(n, -1).T
    assert O.rel_l2(O.dft(x), ref) < 1e-12
    refi = x @ O.dft_matrix(n, +1).T / n
    assert O.rel_l2(O.idft(x), refi) < 1e-12


@pytest.mark.parametrize("shape,axis", [((8, 12), 0), ((8, 12), 1), ((3, 5, 7), 0), ((3, 5, 7), 1), ((3, 5, 7), 2),
                                        ((16, 64, 3), 1), ((4, 4096), 1)])
def test_dft_matches_numpy_per_axis(shape, axis):
    x = crandn(*shape)
    assert O.rel_l2(O.fft_axis(x, axis), np.fft.fft(x, axis=axis)) < 1e-13
    assert O.rel_l2(O.fft_axis(x, axis, inverse=True), np.fft.ifft(x, axis=axis)) < 1e-13
    assert O.rel_l2(O.fft_axis(x, axis, inverse=True, normalize=False), np.fft.ifft(x, axis=axis) * shape[axis]) < 1e-13


@pytest.mark.parametrize("n", [67, 101, 257, 1031, 65521])
def test_bluestein_large_primes(n):
    x = crandn(n)
    assert O.rel_l2(O.bluestein_dft(x), np.fft.fft(x)) < 1e-12
    assert O.rel_l2(O.dft(x), np.fft.fft(x)) < 1e-12


@pytest.mark.parametrize("n", [5, 7, 13, 17, 31])
def test_bluestein_small_primes_bruteforce(n):
    x = crandn(3, n)
    assert O.rel_l2(O.bluestein_dft(x), x @ O.dft_matrix(n).T) < 1e-12
    assert O.rel_l2(O.bluestein_dft(x, inverse=True), x @ O.dft_matrix(n, +1).T) < 1e-12


def test_closed_forms():
    n = 48
    imp = np.zeros(n, complex)
    imp[0] = 1
    assert np.allclose(O.dft(imp), np.ones(n))
    const = np.ones(n, complex)
    e = np.zeros(n, complex)
    e[0] = n
    assert np.allclose(O.dft(const), e)
    # shift theorem: x[n - s] <-> X[k] exp(-2 pi i k s / n)
    x = crandn(n)
    s = 5
    k = np.arange(n)
    assert np.allclose(O.dft(np.roll(x, s)), O.dft(x) * np.exp(-2j * np.pi * k * s / n))
    # Parseval and linearity
    X = O.dft(x)
    assert np.isclose(np.sum(np.abs(x) ** 2), np.sum(np.abs(X) ** 2) / n)
    y = crandn(n)
    assert np.allclose(O.dft(2 * x - 3j * y), 2 * O.dft(x) - 3j * O.dft(y))
    # a single tone lands in one bin with the forward sign convention
    tone = np.exp(2j * np.pi * 3 * k / n)
    T = O.dft(tone)
    assert np.isclose(T[3], n) and np.allclose(np.delete(T, 3), 0)


def test_signal_range():
    x = crandn(4, 30)
    z = x.copy()
    z[:, :7] = 0
    z[:, 19:] = 0
    assert np.allclose(O.fft_axis(x, 1, signal_range=(7, 19)), np.fft.fft(z, axis=1))
    z2 = x.copy()
    z2[:2] = 0
    assert np.allclose(O.fft_axis(x, 0, signal_range=(2, None)), np.fft.fft(z2, axis=0))


@pytest.mark.parametrize("n", [1, 2, 3, 8, 13, 162, 255, 512])
def test_rfft_irfft_rows(n):
    x = RNG.standard_normal((5, n))
    X = O.rfft_rows(x)
    assert X.shape == (5, n // 2 + 1)
    assert np.allclose(X, np.fft.rfft(x, axis=-1))
    assert np.allclose(O.irfft_rows(X, n), x)
    # c2r ignores the imaginary part of DC (and Nyquist for even n), like numpy.irfft
    X2 = X.copy()
    X2[:, 0] += 0.5j
    if n % 2 == 0:
        X2[:, -1] += 0.25j
    assert np.allclose(O.irfft_rows(X2, n), np.fft.irfft(X2, n, axis=-1))


def test_pack_real_layout():
    x = RNG.standard_normal((3, 7))
    p = O.pack_real(x)
    assert p.shape == (3, 8)
    assert np.all(p[:, 7:] == 0) and np.all(p[:, :7] == x)
    assert np.array_equal(O.unpack_real(p, 7), x)


def test_convolve_axis_is_circular_cross_correlation():
    x = crandn(3, 20)
    k = crandn(20)
    K = np.fft.fft(k)
    y = O.convolve_axis(x, np.broadcast_to(K, x.shape), axis=1)
    assert np.allclose(y, O.correlate_bruteforce(x, k))
    # identity kernel (K = 1) returns the input, impulse at m returns a shift
    assert np.allclose(O.convolve_axis(x, np.ones_like(x), axis=1), x)
    d = np.zeros(20, complex)
    d[4] = 1
    assert np.allclose(O.convolve_axis(x, np.broadcast_to(np.fft.fft(d), x.shape), axis=1), np.roll(x, -4, axis=1))


def test_convolve_inner_only_broadcast():
    x = crandn(4, 6, 5)
    k = crandn(6, 5)
    K = np.fft.fft(k, axis=0)
    y = O.convolve_axis(x, K, axis=1, kernel_inner_only=True)
    ref = np.stack([O.convolve_axis(x[b], K, axis=0) for b in range(4)])
    assert np.allclose(y, ref)


def test_convolve2D_and_2DR():
    x = crandn(2, 12, 10)
    k = crandn(12, 10)
    K = np.fft.fft2(k)
    y = O.convolve2D(x, K, kernel_inner_only=True)
    assert np.allclose(y, O.correlate_bruteforce(x, k))
    xr = RNG.standard_normal((3, 9, 8))
    kr = RNG.standard_normal((9, 8))
    Kh = np.fft.rfft2(kr)
    yr = O.convolve2DR(xr, Kh, kernel_inner_only=True)
    assert np.allclose(yr, O.correlate_bruteforce(xr, kr).real)
    # reference test formula (tests/test_conv.py:16-21): ifft2(fft2(x) * conj(fft2(k)))
    assert np.allclose(yr, np.fft.ifft2(np.fft.fft2(xr) * np.conj(np.fft.fft2(kr))).real)


def test_vkfft_convolve2D_is_circular_convolution():
    """vd.vkfft.convolve2D semantics: ifft2(fft2(x) * fft2(k)) is circular convolution
    (the reference test's numpy formula, tests/test_vkfft_conv.py:10-15); the
    conjugated kernel gives cross-correlation; no 1/(HW) without normalize."""
    x = crandn(3, 5, 6)
    k = crandn(3, 5, 6)
    K = np.fft.fft2(k)
    y = O.vkfft_convolve2D(x, K, normalize=True)
    assert np.allclose(y, O.circular_convolve_bruteforce(x, k))
    assert np.allclose(y, np.fft.ifft2(np.fft.fft2(x) * K))
    assert np.allclose(O.vkfft_convolve2D(x, K, normalize=False), 30 * y)
    k1 = crandn(5, 6)
    y1 = O.vkfft_convolve2D(x, np.fft.fft2(k1), normalize=True)          # one kernel, broadcast
    assert np.allclose(y1, O.circular_convolve_bruteforce(x, k1))
    yc = O.vkfft_convolve2D(x, np.fft.fft2(k1), normalize=True, conjugate_kernel=True)
    assert np.allclose(yc, O.correlate_bruteforce(x, k1))
    d = np.zeros((5, 6), complex)
    d[2, 3] = 1                                                            # impulse -> shift
    assert np.allclose(O.vkfft_convolve2D(x, np.fft.fft2(d), normalize=True), np.roll(x, (2, 3), axis=(1, 2)))


def test_plausible_mistakes_are_caught():
    """The checks above distinguish the oracle from its likeliest wrong variants."""
    n = 12
    x = crandn(n)
    good = O.dft(x)
    wrong_sign = O.idft(x, normalize=False)
    assert not np.allclose(wrong_sign, x @ O.dft_matrix(n).T)
    assert np.allclose(good, x @ O.dft_matrix(n).T)
    K = np.fft.fft(crandn(n))
    no_conj = O.idft(O.dft(x) * K)
    assert not np.allclose(no_conj, O.convolve_axis(x, K))
