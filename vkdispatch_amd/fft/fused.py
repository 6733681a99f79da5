"""2-D real FFT-convolution (convolve2DR) execution paths and their profiling.

``try_convolve2DR`` returns False when the call is not eligible for a fused
path, and the caller falls back to the reference's three-dispatch composition
(rfft -> convolve(axis ndim-2) -> irfft, vkdispatch/fft/functions.py:614-646).
"""
from __future__ import annotations

from typing import Dict, List

from .. import dtypes as dt
from .. import runtime as rt
from .shader_factories import make_convolution_shader, make_fft_shader


def try_convolve2DR(buffer, kernel, **kw) -> bool:
    return False


def composed_kernels(real_shape, kernel_inner_only=True, compute_type=dt.complex64):
    """The three kernels convolve2DR dispatches for ``real_shape`` (..., H, W)."""
    real_shape = tuple(real_shape)
    half = real_shape[:-1] + (real_shape[-1] // 2 + 1,)
    ct = dt.complex64
    r = make_fft_shader(real_shape, None, r2c=True, output_type=ct, input_type=ct, compute_type=compute_type)
    c = make_convolution_shader(half, None, 1, len(half) - 2, kernel_inner_only=kernel_inner_only, normalize=True,
                                input_type=ct, output_type=ct, kernel_type=ct, compute_type=compute_type)
    i = make_fft_shader(real_shape, None, inverse=True, normalize_inverse=True, r2c=True, output_type=ct,
                        input_type=ct, compute_type=compute_type)
    return r, c, i


def standard_kernels() -> List:
    return list(composed_kernels((64, 512, 512)))


def profile_convolve2DR(buffer, kernel, reps: int = 20) -> List[Dict]:
    """Average duration (HIP events on the launch stream) and algorithmic bytes of each
    kernel convolve2DR runs for ``buffer`` (broadcast kernel)."""
    shape = buffer.real_shape
    batch = 1
    for s in shape[:-2]:
        batch *= s
    H, W = shape[-2], shape[-1]
    Wh = W // 2 + 1
    r, c, i = composed_kernels(shape)
    real_b = batch * H * W * 4
    spec_b = batch * H * Wh * 8
    kb = H * Wh * 8
    out = []
    for name, k, vals, nbytes in (("rfft_rows", r, [buffer], real_b + spec_b),
                                  ("convolve_cols", c, [buffer, kernel], 2 * spec_b + kb),
                                  ("irfft_rows", i, [buffer], spec_b + real_b)):
        out.append({"name": name, "ms": _event_ms(k, vals, reps), "bytes": nbytes})
    return out


def _event_ms(kernel, values, reps: int) -> float:
    e0, e1 = rt.Event(), rt.Event()
    kernel.launch(values)
    e0.record()
    for _ in range(reps):
        kernel.launch(values)
    e1.record()
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps
