"""FP64 CPU oracle of the vd.fft path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only to check results or as the CPU baseline.  The product
(``vkdispatch_amd``) never imports it and has no CPU fallback.

Contents (see ``fft_oracle.py`` for per-function reference citations):
  dft / idft           mixed-radix Stockham DFT in float64 (own core, not numpy.fft)
  bluestein_dft        chirp-z DFT for any length (exact-integer chirp phase)
  fft_axis             vd.fft.fft / ifft semantics on a buffer (axis, 1/N, signal range)
  rfft_rows/irfft_rows r2c = full complex FFT of (x, 0); c2r via conj(X[N-k])
  pack_real/unpack_real  RFFTBuffer packed-real layout
  convolve_axis        ifft(fft(x) * conj(K)) along an axis (default kernel map)
  convolve2D/2DR       the 2-D compositions
  correlate_bruteforce O(N^2) circular cross-correlation (independent pin)

Pinning (tests/test_oracle.py): brute-force DFT matrices for every N <= 64 and
assorted composites/primes, numpy.fft (an independent implementation) on
random multi-axis inputs, closed forms (impulse, constant, shift theorem,
Parseval, linearity), brute-force correlation for the convolutions, and the
committed golden fixtures in tests/golden/.  Every oracle function has such a
check, so none is "parity unpinned".
"""
from .fft_oracle import *  # noqa: F401,F403
