"""Fused 2-D real FFT-convolution (convolve2DR fast path).

``try_convolve2DR`` returns False when the call is not eligible, and the
caller falls back to the reference's three-dispatch composition
(rfft -> convolve -> irfft, vkdispatch/fft/functions.py:614-646).
"""
from __future__ import annotations


def try_convolve2DR(buffer, kernel, **kw) -> bool:
    return False
