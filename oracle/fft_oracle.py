"""FP64 restatement of the reference vd.fft semantics (test infrastructure only).

Reference behaviour restated here:
  * forward sign exp(-2 pi i k n / N), inverse +, inverse normalised by 1/N per
    axis when ``normalize`` (vkdispatch/fft/cooley_tukey.py:8-9,
    vkdispatch/fft/config.py:276-277, vkdispatch/fft/shader_factories.py:52-53);
  * axis geometry: stride = prod(shape[axis+1:]) (vkdispatch/fft/config.py:217-233);
  * ``input_signal_range=(s, e)`` zeroes inputs outside [s, e) along the axis
    (vkdispatch/fft/global_memory_iterators.py:167-195,231-244);
  * r2c: full N-point complex FFT of (x, 0), keep k < N/2+1
    (global_memory_iterators.py:85-88,214-219); c2r: bins k >= N/2+1 are
    conj(X[N-k]), full complex inverse, real part (:90-92,221-227);
  * RFFTBuffer rows: N/2+1 complex = 2*(N/2+1) reals, the first N the signal
    (vkdispatch/base/buffer.py:377-448);
  * convolve: X <- X * conj(K[io_index]) between forward and normalised
    inverse FFT along the axis (vkdispatch/fft/shader_factories.py:107-195);
    ``kernel_inner_only`` broadcasts K over the outer batch
    (global_memory_iterators.py:26-27);
  * convolve2D / convolve2DR = fft(last) -> convolve(ndim-2) -> ifft(last)
    (vkdispatch/fft/functions.py:549-646);
  * vd.vkfft.convolve2D(x, K) = ifft2(fft2(x) * K) with K = fft2(k) prepared by
    transpose_kernel2D, no conjugate unless ``conjugate_kernel``, unnormalised
    unless ``normalize`` (vkdispatch/vkfft/vkfft_dispatcher.py:109-167, checked by
    tests/test_vkfft_conv.py:10-15,62-66).

The DFT core is a mixed-radix Stockham decimation-in-time in float64, written
for this oracle (it does not call numpy.fft); numpy.fft is used only by the
tests, as an independent check.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np

__all__ = [
    "prime_factors", "dft", "idft", "dft_matrix", "bluestein_dft", "fft_axis", "fftn_axes", "rfft_rows",
    "irfft_rows", "pack_real", "unpack_real", "convolve_axis", "convolve2D", "convolve2DR",
    "correlate_bruteforce", "circular_convolve_bruteforce", "vkfft_convolve2D", "rel_l2",
]


def prime_factors(n: int) -> List[int]:
    out, f = [], 2
    while n % 2 == 0:
        out.append(2)
        n //= 2
    f = 3
    while f * f <= n:
        while n % f == 0:
            out.append(f)
            n //= f
        f += 2
    if n > 1:
        out.append(n)
    return out


def _radices(n: int) -> List[int]:
    """Group prime factors into radices <= 16 (largest prime factors kept alone)."""
    out: List[int] = []
    for p in sorted(prime_factors(n), reverse=True):
        for i, r in enumerate(out):
            if r * p <= 16:
                out[i] *= p
                break
        else:
            out.append(p)
    return out


def dft_matrix(n: int, sign: int = -1) -> np.ndarray:
    k = np.arange(n)
    m = np.outer(k, k) % n
    return np.exp(sign * 2j * np.pi * m / n)


_BLUESTEIN_MIN_PRIME = 64


def _stockham_last(x: np.ndarray, sign: int) -> np.ndarray:
    n = x.shape[-1]
    lead = x.shape[:-1]
    y = np.array(x, dtype=np.complex128, copy=True)
    ns = 1
    for R in _radices(n):
        W = n // R
        v = y.reshape(lead + (R, W))
        j = np.arange(W)
        k = j % ns
        if ns > 1:
            r = np.arange(R)[:, None]
            tw = np.exp(sign * 2j * np.pi * ((r * k[None, :]) % (ns * R)) / (ns * R))
            v = v * tw
        if R >= _BLUESTEIN_MIN_PRIME:
            u = np.moveaxis(_bluestein_last(np.moveaxis(v, -2, -1), sign), -1, -2)
        else:
            u = np.einsum("qr,...rj->...qj", dft_matrix(R, sign), v)
        base = (j // ns) * ns * R + k
        pos = base[None, :] + (np.arange(R) * ns)[:, None]
        out = np.empty_like(y)
        out[..., pos.reshape(-1)] = u.reshape(lead + (R * W,))
        y = out
        ns *= R
    return y


def _next_pow2(n: int) -> int:
    m = 1
    while m < n:
        m *= 2
    return m


def _chirp(n: int, count: int, sign: int) -> np.ndarray:
    m = np.arange(count, dtype=np.int64)
    ph = (m * m) % (2 * n)          # exact integer phase reduction
    return np.exp(sign * 1j * np.pi * ph / n)


def _bluestein_last(x: np.ndarray, sign: int) -> np.ndarray:
    n = x.shape[-1]
    M = _next_pow2(2 * n - 1)
    w = _chirp(n, n, sign)                      # exp(sign i pi m^2 / n)
    a = np.zeros(x.shape[:-1] + (M,), dtype=np.complex128)
    a[..., :n] = x * w
    b = np.zeros(M, dtype=np.complex128)
    wc = np.conj(_chirp(n, n, sign))
    b[:n] = wc
    b[M - n + 1:] = wc[1:][::-1]
    A = _stockham_last(a, -1)
    Bf = _stockham_last(b, -1)
    c = _stockham_last(A * Bf, +1) / M
    return c[..., :n] * w


def bluestein_dft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """DFT of the last axis of any length by the chirp-z transform (unnormalised)."""
    return _bluestein_last(np.asarray(x, dtype=np.complex128), +1 if inverse else -1)


def dft(x: np.ndarray, axis: int = -1) -> np.ndarray:
    """Forward DFT (exp(-2 pi i k n / N)) along ``axis`` in float64."""
    x = np.moveaxis(np.asarray(x, dtype=np.complex128), axis, -1)
    return np.moveaxis(_stockham_last(x, -1), -1, axis)


def idft(x: np.ndarray, axis: int = -1, normalize: bool = True) -> np.ndarray:
    x = np.moveaxis(np.asarray(x, dtype=np.complex128), axis, -1)
    y = _stockham_last(x, +1)
    if normalize:
        y = y / x.shape[-1]
    return np.moveaxis(y, -1, axis)


def fft_axis(x: np.ndarray, axis: Optional[int] = None, inverse: bool = False, normalize: bool = True,
             signal_range: Optional[Tuple[Optional[int], Optional[int]]] = None) -> np.ndarray:
    """vd.fft.fft(buf, axis=..., inverse=..., normalize_inverse=..., input_signal_range=...)."""
    x = np.asarray(x, dtype=np.complex128)
    if axis is None:
        axis = x.ndim - 1
    if signal_range is not None:
        s, e = signal_range
        s = 0 if s is None else s
        e = x.shape[axis] if e is None else e
        x = x.copy()
        sl = [slice(None)] * x.ndim
        sl[axis] = slice(0, s)
        x[tuple(sl)] = 0
        sl[axis] = slice(e, None)
        x[tuple(sl)] = 0
    return idft(x, axis, normalize) if inverse else dft(x, axis)


def fftn_axes(x: np.ndarray, axes: Sequence[int], inverse: bool = False, normalize: bool = True) -> np.ndarray:
    y = np.asarray(x, dtype=np.complex128)
    for a in axes:
        y = fft_axis(y, a, inverse=inverse, normalize=normalize)
    return y


def rfft_rows(x: np.ndarray) -> np.ndarray:
    """r2c along the last axis: full complex FFT of (x, 0), first N/2+1 bins."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    return dft(x.astype(np.complex128), -1)[..., : n // 2 + 1]


def irfft_rows(X: np.ndarray, n: int, normalize: bool = True) -> np.ndarray:
    """c2r along the last axis: bins k >= N/2+1 are conj(X[N-k]); complex inverse; real part."""
    X = np.asarray(X, dtype=np.complex128)
    half = n // 2 + 1
    full = np.empty(X.shape[:-1] + (n,), dtype=np.complex128)
    full[..., :half] = X[..., :half]
    k = np.arange(half, n)
    full[..., half:] = np.conj(X[..., n - k])
    return idft(full, -1, normalize).real


def pack_real(x: np.ndarray) -> np.ndarray:
    """RFFTBuffer packed rows: (..., N) reals -> (..., 2*(N/2+1)) with zero tail."""
    x = np.asarray(x)
    n = x.shape[-1]
    out = np.zeros(x.shape[:-1] + (2 * (n // 2 + 1),), dtype=x.dtype)
    out[..., :n] = x
    return out


def unpack_real(p: np.ndarray, n: int) -> np.ndarray:
    return np.asarray(p)[..., :n]


def convolve_axis(x: np.ndarray, K: np.ndarray, axis: Optional[int] = None, normalize: bool = True,
                  kernel_inner_only: bool = False) -> np.ndarray:
    """vd.fft.convolve(x, K): ifft(fft(x) * conj(K)) along ``axis`` with K already in
    the frequency domain.  With ``kernel_inner_only`` K has the shape of one outer
    batch (shape[axis:]) and is broadcast over the outer axes."""
    x = np.asarray(x, dtype=np.complex128)
    if axis is None:
        axis = x.ndim - 1
    K = np.asarray(K, dtype=np.complex128)
    if kernel_inner_only:
        K = np.broadcast_to(K.reshape((1,) * axis + x.shape[axis:]), x.shape)
    X = dft(x, axis)
    return idft(X * np.conj(K), axis, normalize)


def convolve2D(x: np.ndarray, K: np.ndarray, normalize: bool = True, kernel_inner_only: bool = False):
    """vd.fft.convolve2D(buf, K) with K = fft2(k) (complex data)."""
    x = np.asarray(x, dtype=np.complex128)
    y = dft(x, x.ndim - 1)
    y = convolve_axis(y, K, x.ndim - 2, normalize, kernel_inner_only)
    return idft(y, x.ndim - 1, normalize)


def convolve2DR(x: np.ndarray, K_half: np.ndarray, normalize: bool = True, kernel_inner_only: bool = False):
    """vd.fft.convolve2DR(buf, K) with K = rfft2(k): real result of
    irfft2(rfft2(x) * conj(K)) (vkdispatch/fft/functions.py:614-646)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    X = rfft_rows(x)
    X = convolve_axis(X, K_half, x.ndim - 2, normalize, kernel_inner_only)
    return irfft_rows(X, n, normalize)


def vkfft_convolve2D(x: np.ndarray, K: np.ndarray, normalize: bool = False, conjugate_kernel: bool = False):
    """vd.vkfft.convolve2D(buf, K) over the last two axes of x; K is one spectrum per
    image or one broadcast over the leading axis (vkfft_dispatcher.py:137-167)."""
    X = fftn_axes(x, (x.ndim - 1, x.ndim - 2))
    Kc = np.conj(K) if conjugate_kernel else np.asarray(K, dtype=np.complex128)
    return fftn_axes(X * Kc, (x.ndim - 1, x.ndim - 2), inverse=True, normalize=normalize)


def circular_convolve_bruteforce(x: np.ndarray, k: np.ndarray) -> np.ndarray:
    """y[n] = sum_m x[m] k[(n - m) mod N] over the last two axes (independent pin of
    ifft2(fft2(x) * fft2(k)), normalised); x may carry one leading batch axis."""
    x = np.asarray(x, dtype=np.complex128)
    k = np.asarray(k, dtype=np.complex128)
    H, W = x.shape[-2:]
    y = np.zeros(np.broadcast_shapes(x.shape, k.shape), dtype=np.complex128)
    for m0 in range(H):
        for m1 in range(W):
            y += x[..., m0:m0 + 1, m1:m1 + 1] * np.roll(np.roll(k, m0, axis=-2), m1, axis=-1)
    return y


def correlate_bruteforce(x: np.ndarray, k: np.ndarray) -> np.ndarray:
    """y[n] = sum_m x[(n + m) mod N] * conj(k[m]) over every axis of k (independent pin
    of convolve with K = DFT(k)); x may carry extra leading batch axes."""
    x = np.asarray(x, dtype=np.complex128)
    k = np.asarray(k, dtype=np.complex128)
    nd = k.ndim
    y = np.zeros_like(x)
    for m in np.ndindex(*k.shape):
        shifted = x
        for d, md in enumerate(m):
            shifted = np.roll(shifted, -md, axis=x.ndim - nd + d)
        y = y + shifted * np.conj(k[m])
    return y


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)
