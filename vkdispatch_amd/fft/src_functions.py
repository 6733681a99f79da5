"""``*_src`` / ``*_print_src``: the HIP sources of exactly the kernels a call launches
(the reference's vkdispatch/fft/src_functions.py:7-351 returns its generated
GLSL/CUDA/OpenCL).  Single-kernel calls return one source string; calls that launch
several kernels return a tuple of sources in launch order, as the reference does.
Every source is a complete hiprtc translation unit whose first line describes its
launch (grid, block, twiddle radix lists), so it can be compiled and bound on its own
(vkdispatch_amd/native_table.py does exactly that)."""
from __future__ import annotations

from typing import Tuple

from .. import dtypes as dt
from .shader_factories import make_convolution_shader, make_fft_shader, make_transpose_shader


def _check_dims(buffer_shape, dims, what="Buffer Shape"):
    if len(buffer_shape) not in dims:
        raise ValueError(f"{what} must have {' or '.join(str(d) for d in dims)} dimensions")


def fft_src(buffer_shape, axis=None, inverse=False, normalize_inverse=True, r2c=False, input_map=None,
            output_map=None, input_signal_range=None, line_numbers=False, input_type=None, output_type=None,
            compute_type=None) -> str:
    src = make_fft_shader(tuple(buffer_shape), axis, inverse=inverse, normalize_inverse=normalize_inverse, r2c=r2c,
                          input_map=input_map, output_map=output_map, input_type=input_type,
                          output_type=output_type, compute_type=compute_type,
                          input_signal_range=input_signal_range).source
    return _numbered(src) if line_numbers else src


def _numbered(src: str) -> str:
    return "\n".join(f"{i + 1:4d}: {line}" for i, line in enumerate(src.split("\n")))


def fft2_src(buffer_shape, input_map=None, output_map=None) -> Tuple[str, str]:
    _check_dims(buffer_shape, (2, 3))
    n = len(buffer_shape)
    return (fft_src(buffer_shape, axis=n - 2, input_map=input_map),
            fft_src(buffer_shape, axis=n - 1, output_map=output_map))


def fft3_src(buffer_shape, input_map=None, output_map=None) -> Tuple[str, str, str]:
    _check_dims(buffer_shape, (3,))
    return (fft_src(buffer_shape, axis=0, input_map=input_map), fft_src(buffer_shape, axis=1),
            fft_src(buffer_shape, axis=2, output_map=output_map))


def ifft_src(buffer_shape, axis=None, normalize=True, input_map=None, output_map=None) -> str:
    return fft_src(buffer_shape, axis=axis, inverse=True, normalize_inverse=normalize, input_map=input_map,
                   output_map=output_map)


def ifft2_src(buffer_shape, normalize=True, input_map=None, output_map=None) -> Tuple[str, str]:
    _check_dims(buffer_shape, (2, 3))
    n = len(buffer_shape)
    return (ifft_src(buffer_shape, axis=n - 2, normalize=normalize, input_map=input_map),
            ifft_src(buffer_shape, axis=n - 1, normalize=normalize, output_map=output_map))


def ifft3_src(buffer_shape, normalize=True, input_map=None, output_map=None) -> Tuple[str, str, str]:
    _check_dims(buffer_shape, (3,))
    return (ifft_src(buffer_shape, axis=0, normalize=normalize, input_map=input_map),
            ifft_src(buffer_shape, axis=1, normalize=normalize),
            ifft_src(buffer_shape, axis=2, normalize=normalize, output_map=output_map))


def _half(shape):
    return tuple(shape[:-1]) + (shape[-1] // 2 + 1,)


def rfft_src(buffer_shape) -> str:
    return fft_src(buffer_shape, r2c=True, output_type=dt.complex64, input_type=dt.complex64)


def irfft_src(buffer_shape, normalize=True) -> str:
    return fft_src(buffer_shape, r2c=True, inverse=True, normalize_inverse=normalize, output_type=dt.complex64,
                   input_type=dt.complex64)


def rfft2_src(buffer_shape) -> Tuple[str, str]:
    _check_dims(buffer_shape, (2, 3))
    return rfft_src(buffer_shape), fft_src(_half(buffer_shape), axis=len(buffer_shape) - 2)


def rfft3_src(buffer_shape) -> Tuple[str, str, str]:
    _check_dims(buffer_shape, (3,))
    h = _half(buffer_shape)
    return rfft_src(buffer_shape), fft_src(h, axis=1), fft_src(h, axis=0)


def irfft2_src(buffer_shape, normalize=True) -> Tuple[str, str]:
    _check_dims(buffer_shape, (2, 3))
    return (ifft_src(_half(buffer_shape), axis=len(buffer_shape) - 2, normalize=normalize),
            irfft_src(buffer_shape, normalize=normalize))


def irfft3_src(buffer_shape, normalize=True) -> Tuple[str, str, str]:
    _check_dims(buffer_shape, (3,))
    h = _half(buffer_shape)
    return (ifft_src(h, axis=0, normalize=normalize), ifft_src(h, axis=1, normalize=normalize),
            irfft_src(buffer_shape, normalize=normalize))


def convolve_src(buffer_shape, kernel_map=None, kernel_num=1, axis=None, normalize=True, transposed_kernel=False,
                 kernel_inner_only=False, input_map=None, output_map=None, input_signal_range=None,
                 line_numbers=False, input_type=None, output_type=None, kernel_type=None, compute_type=None) -> str:
    src = make_convolution_shader(tuple(buffer_shape), kernel_map, kernel_num, axis, normalize=normalize,
                                  transposed_kernel=transposed_kernel, kernel_inner_only=kernel_inner_only,
                                  input_map=input_map, output_map=output_map, input_type=input_type,
                                  output_type=output_type, kernel_type=kernel_type, compute_type=compute_type,
                                  input_signal_range=input_signal_range).source
    return _numbered(src) if line_numbers else src


def convolve2D_src(buffer_shape, kernel_map=None, normalize=True, transposed_kernel=False, kernel_inner_only=False,
                   input_map=None, output_map=None) -> Tuple[str, str, str]:
    _check_dims(buffer_shape, (2, 3), "Buffer")
    n = len(buffer_shape)
    return (fft_src(buffer_shape, input_map=input_map),
            convolve_src(buffer_shape, kernel_map=kernel_map, axis=n - 2, normalize=normalize,
                         transposed_kernel=transposed_kernel, kernel_inner_only=kernel_inner_only),
            ifft_src(buffer_shape, normalize=normalize, output_map=output_map))


def convolve2DR_src(buffer_shape, kernel_map=None, transposed_kernel=False, kernel_inner_only=False,
                    normalize=True) -> Tuple[str, ...]:
    """The kernels convolve2DR launches for complex64 RFFTBuffers of ``buffer_shape`` (the
    real shape) with one kernel spectrum: the pair-packed kernels when the call takes
    that path (one source for the single launch, three for the phase kernels), else the
    reference's rfft -> convolve -> irfft composition."""
    _check_dims(buffer_shape, (2, 3), "Buffer")
    from . import fused
    plan = fused.plan_for_shape(tuple(buffer_shape), kernel_map, transposed_kernel, kernel_inner_only, normalize)
    if plan is not None:
        return tuple(k.source for k in plan.kernels())
    n = len(buffer_shape)
    return (rfft_src(buffer_shape),
            convolve_src(_half(buffer_shape), kernel_map=kernel_map, axis=n - 2, normalize=normalize,
                         transposed_kernel=transposed_kernel, kernel_inner_only=kernel_inner_only,
                         input_type=dt.complex64, output_type=dt.complex64, kernel_type=dt.complex64),
            irfft_src(buffer_shape, normalize=normalize))


def transpose_src(buffer_shape, axis=None, kernel_inner_only=False, line_numbers=False) -> str:
    src = make_transpose_shader(tuple(buffer_shape), axis=axis, kernel_inner_only=kernel_inner_only).source
    return _numbered(src) if line_numbers else src


# ---- printers (vkdispatch/fft/src_functions.py:212-351) ----
def _print_stages(labels, srcs):
    print("\n".join(f"// {label}:\n{src}" for label, src in zip(labels, srcs)))


def fft_print_src(buffer_shape, axis=None, inverse=False, normalize_inverse=True, r2c=False, input_map=None,
                  output_map=None, input_signal_range=None, line_numbers=False):
    print(fft_src(buffer_shape, axis, inverse, normalize_inverse, r2c, input_map, output_map, input_signal_range,
                  line_numbers))


def fft2_print_src(buffer_shape, input_map=None, output_map=None):
    n = len(buffer_shape)
    _print_stages([f"FFT Stage (axis {n - 2})", f"FFT Stage (axis {n - 1})"],
                  fft2_src(buffer_shape, input_map, output_map))


def fft3_print_src(buffer_shape, input_map=None, output_map=None):
    _print_stages([f"FFT Stage (axis {a})" for a in range(3)], fft3_src(buffer_shape, input_map, output_map))


def ifft_print_src(buffer_shape, axis=None, normalize=True, input_map=None, output_map=None):
    print(ifft_src(buffer_shape, axis, normalize, input_map, output_map))


def ifft2_print_src(buffer_shape, normalize=True, input_map=None, output_map=None):
    n = len(buffer_shape)
    _print_stages([f"IFFT Stage (axis {n - 2})", f"IFFT Stage (axis {n - 1})"],
                  ifft2_src(buffer_shape, normalize, input_map, output_map))


def ifft3_print_src(buffer_shape, normalize=True, input_map=None, output_map=None):
    _print_stages([f"IFFT Stage (axis {a})" for a in range(3)],
                  ifft3_src(buffer_shape, normalize, input_map, output_map))


def rfft_print_src(buffer_shape):
    print(rfft_src(buffer_shape))


def rfft2_print_src(buffer_shape):
    _print_stages(["RFFT Stage", f"FFT Stage (axis {len(buffer_shape) - 2})"], rfft2_src(buffer_shape))


def rfft3_print_src(buffer_shape):
    _print_stages(["RFFT Stage", "FFT Stage (axis 1)", "FFT Stage (axis 0)"], rfft3_src(buffer_shape))


def irfft_print_src(buffer_shape, normalize=True):
    print(irfft_src(buffer_shape, normalize))


def irfft2_print_src(buffer_shape, normalize=True):
    _print_stages([f"IFFT Stage (axis {len(buffer_shape) - 2})", "IRFFT Stage"], irfft2_src(buffer_shape, normalize))


def irfft3_print_src(buffer_shape, normalize=True):
    _print_stages(["IFFT Stage (axis 0)", "IFFT Stage (axis 1)", "IRFFT Stage"], irfft3_src(buffer_shape, normalize))


def convolve_print_src(buffer_shape, kernel_map=None, kernel_num=1, axis=None, normalize=True,
                       transposed_kernel=False, kernel_inner_only=False, input_map=None, output_map=None,
                       input_signal_range=None, line_numbers=False):
    print(convolve_src(buffer_shape, kernel_map, kernel_num, axis, normalize, transposed_kernel, kernel_inner_only,
                       input_map, output_map, input_signal_range, line_numbers))


def convolve2D_print_src(buffer_shape, kernel_map=None, normalize=True, transposed_kernel=False,
                         kernel_inner_only=False, input_map=None, output_map=None):
    n = len(buffer_shape)
    _print_stages([f"FFT Stage (axis {n - 2})", f"Convolution Stage (axis {n - 2})", "IFFT Stage"],
                  convolve2D_src(buffer_shape, kernel_map, normalize, transposed_kernel, kernel_inner_only,
                                 input_map, output_map))


def convolve2DR_print_src(buffer_shape, kernel_map=None, transposed_kernel=False, kernel_inner_only=False,
                          normalize=True):
    srcs = convolve2DR_src(buffer_shape, kernel_map, transposed_kernel, kernel_inner_only, normalize)
    n = len(buffer_shape)
    if len(srcs) == 3 and "conv2d_phase_body" not in srcs[0]:
        labels = ["RFFT Stage", f"Convolution Stage (axis {n - 2})", "IRFFT Stage"]
    elif len(srcs) == 3:
        labels = ["Pair rows FFT", "Pair columns FFT * Kt * IFFT", "Pair rows IFFT"]
    else:
        labels = ["Pair-packed single launch"]
    _print_stages(labels, srcs)


def transpose_print_src(buffer_shape, axis=None, kernel_inner_only=False, line_numbers=False):
    print(transpose_src(buffer_shape, axis, kernel_inner_only, line_numbers))
