"""``*_src`` / ``*_print_src``: the HIP source of the kernels a call would launch
(the reference's vkdispatch/fft/src_functions.py:7-351 returns generated
GLSL/CUDA/OpenCL; here it is the HIP instantiation compiled with hiprtc)."""
from __future__ import annotations

from .. import dtypes as dt
from .shader_factories import make_convolution_shader, make_fft_shader, make_transpose_shader


def fft_src(buffer_shape, axis=None, inverse=False, normalize_inverse=True, r2c=False, input_map=None,
            output_map=None, input_type=None, output_type=None, compute_type=None, input_signal_range=None,
            **_ignored) -> str:
    return make_fft_shader(tuple(buffer_shape), axis, inverse=inverse, normalize_inverse=normalize_inverse, r2c=r2c,
                           input_map=input_map, output_map=output_map, input_type=input_type,
                           output_type=output_type, compute_type=compute_type,
                           input_signal_range=input_signal_range).source


def ifft_src(buffer_shape, axis=None, normalize=True, **kw) -> str:
    return fft_src(buffer_shape, axis=axis, inverse=True, normalize_inverse=normalize, **kw)


def _multi(buffer_shape, axes, **kw) -> str:
    return "\n\n".join(fft_src(buffer_shape, axis=a, **kw) for a in axes)


def fft2_src(buffer_shape, **kw) -> str:
    n = len(buffer_shape)
    return _multi(buffer_shape, (n - 2, n - 1), **kw)


def fft3_src(buffer_shape, **kw) -> str:
    return _multi(buffer_shape, (0, 1, 2), **kw)


def ifft2_src(buffer_shape, normalize=True, **kw) -> str:
    n = len(buffer_shape)
    return _multi(buffer_shape, (n - 2, n - 1), inverse=True, normalize_inverse=normalize, **kw)


def ifft3_src(buffer_shape, normalize=True, **kw) -> str:
    return _multi(buffer_shape, (0, 1, 2), inverse=True, normalize_inverse=normalize, **kw)


def _half(shape):
    return tuple(shape[:-1]) + (shape[-1] // 2 + 1,)


def rfft_src(buffer_shape, compute_type=None, **_ignored) -> str:
    return fft_src(buffer_shape, r2c=True, output_type=dt.complex64, input_type=dt.complex64,
                   compute_type=compute_type)


def irfft_src(buffer_shape, normalize=True, compute_type=None, **_ignored) -> str:
    return fft_src(buffer_shape, r2c=True, inverse=True, normalize_inverse=normalize, output_type=dt.complex64,
                   input_type=dt.complex64, compute_type=compute_type)


def rfft2_src(buffer_shape, **kw) -> str:
    return rfft_src(buffer_shape) + "\n\n" + fft_src(_half(buffer_shape), axis=len(buffer_shape) - 2)


def rfft3_src(buffer_shape, **kw) -> str:
    h = _half(buffer_shape)
    return "\n\n".join([rfft_src(buffer_shape), fft_src(h, axis=1), fft_src(h, axis=0)])


def irfft2_src(buffer_shape, normalize=True, **kw) -> str:
    return ifft_src(_half(buffer_shape), axis=len(buffer_shape) - 2, normalize=normalize) + "\n\n" + \
        irfft_src(buffer_shape, normalize=normalize)


def irfft3_src(buffer_shape, normalize=True, **kw) -> str:
    h = _half(buffer_shape)
    return "\n\n".join([ifft_src(h, axis=0, normalize=normalize), ifft_src(h, axis=1, normalize=normalize),
                        irfft_src(buffer_shape, normalize=normalize)])


def convolve_src(buffer_shape, kernel_map=None, kernel_num=1, axis=None, normalize=True, transposed_kernel=False,
                 kernel_inner_only=False, input_map=None, output_map=None, input_type=None, output_type=None,
                 kernel_type=None, compute_type=None, input_signal_range=None, **_ignored) -> str:
    return make_convolution_shader(tuple(buffer_shape), kernel_map, kernel_num, axis, normalize=normalize,
                                   transposed_kernel=transposed_kernel, kernel_inner_only=kernel_inner_only,
                                   input_map=input_map, output_map=output_map, input_type=input_type,
                                   output_type=output_type, kernel_type=kernel_type, compute_type=compute_type,
                                   input_signal_range=input_signal_range).source


def convolve2D_src(buffer_shape, **kw) -> str:
    n = len(buffer_shape)
    return "\n\n".join([fft_src(buffer_shape), convolve_src(buffer_shape, axis=n - 2, **kw),
                        ifft_src(buffer_shape)])


def convolve2DR_src(buffer_shape, **kw) -> str:
    n = len(buffer_shape)
    return "\n\n".join([rfft_src(buffer_shape), convolve_src(_half(buffer_shape), axis=n - 2, **kw),
                        irfft_src(buffer_shape)])


def transpose_src(buffer_shape, axis=None, kernel_inner_only=False, **_ignored) -> str:
    return make_transpose_shader(tuple(buffer_shape), axis=axis, kernel_inner_only=kernel_inner_only).source


def _printer(fn):
    def p(*a, **k):
        print(fn(*a, **k))
    p.__name__ = fn.__name__.replace("_src", "_print_src")
    return p


fft_print_src = _printer(fft_src)
fft2_print_src = _printer(fft2_src)
fft3_print_src = _printer(fft3_src)
ifft_print_src = _printer(ifft_src)
ifft2_print_src = _printer(ifft2_src)
ifft3_print_src = _printer(ifft3_src)
rfft_print_src = _printer(rfft_src)
rfft2_print_src = _printer(rfft2_src)
rfft3_print_src = _printer(rfft3_src)
irfft_print_src = _printer(irfft_src)
irfft2_print_src = _printer(irfft2_src)
irfft3_print_src = _printer(irfft3_src)
convolve_print_src = _printer(convolve_src)
convolve2D_print_src = _printer(convolve2D_src)
convolve2DR_print_src = _printer(convolve2DR_src)
