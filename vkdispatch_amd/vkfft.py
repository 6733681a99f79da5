"""``vd.vkfft``: the reference's VkFFT-style plan API on the HIP kernels of ``vd.fft``.

Same names, signatures and defaults as vkdispatch/vkfft/vkfft_dispatcher.py:74-277
(and the package exports of vkdispatch/vkfft/__init__.py).  There is no VkFFT library
underneath: every call lowers to the ``vd.fft`` plans (one kernel per transformed
axis, plans cached by ``vd.fft``), so the semantics are the ones the reference tests
check (tests/test_vkfft.py, tests/test_vkfft_conv.py):

- ``fft``/``ifft`` over a list of axes (default: the last); ``ifft`` normalises by the
  transformed lengths unless ``normalize=False``;
- ``rfft*``/``irfft*`` on an ``RFFTBuffer`` (half spectrum along the last axis);
- ``transpose_kernel2D`` turns a (batch of) 2-D kernel(s) into the layout
  ``convolve2D`` consumes, which here is simply its 2-D spectrum in natural order;
- ``convolve2D(buffer, kernel)`` = ifft2(fft2(x) * K) per image, K = conj(K) with
  ``conjugate_kernel=True``, unnormalised unless ``normalize=True`` (VkFFT's default),
  with either one kernel per image or one kernel broadcast over the batch.

``padding=((l0, r0), (l1, r1))`` follows VkFFT's zero-padding ranges: along each of the
two image axes the samples in [l, r) are read as zero.  Only ranges that touch an end
of the axis are supported (they map onto ``input_signal_range``); the output is written
over the whole image.  (No reference test exercises padding.)
"""
from __future__ import annotations

from functools import lru_cache
from typing import Optional, Sequence, Tuple, Union

from . import codegen as vc
from . import dtypes as dt
from . import runtime as rt
from .buffer import Buffer, RFFTBuffer
from .fft import functions as F
from .fft.shader_factories import ZeroBlock, cache_clear as _fft_cache_clear
from .fft.hooks import read_op
from .graph import CommandGraph
from .shader_map import map as vd_map

__all__ = ["fft", "fft2", "fft3", "ifft", "ifft2", "ifft3", "rfft", "rfft2", "rfft3", "irfft", "irfft2", "irfft3",
           "clear_plan_cache", "convolve2D", "transpose_kernel2D"]


def clear_plan_cache() -> None:
    """Drop the cached plans (vkfft_dispatcher.py:74-75)."""
    _fft_cache_clear()
    _product_map.cache_clear()


def _axes(buffer, axis) -> Tuple[int, ...]:
    nd = len(buffer.shape)
    if axis is None:
        axes = (nd - 1,)
    elif isinstance(axis, int):
        axes = (axis,)
    else:
        axes = tuple(int(a) for a in axis)
    for a in axes:
        if not 0 <= a < nd:
            raise ValueError(f"axis {a} out of range for a {nd}-D buffer")
    if len(set(axes)) != len(axes):
        raise ValueError("repeated axis")
    return axes


def _copy_in(buffer: Buffer, input_buffer: Buffer, graph: Optional[CommandGraph]) -> None:
    """Out-of-place transform: the input buffer is copied into the output first."""
    if input_buffer.mem_size != buffer.mem_size:
        raise ValueError("input_buffer must have the byte size of the output buffer")
    dst, src, n = buffer.ptr, input_buffer.ptr, buffer.mem_size
    if graph is None:
        rt.dtod(dst, src, n)
    else:
        graph.record_callable(lambda stream: rt.dtod(dst, src, n, stream))


def fft(buffer: Buffer, input_buffer: Buffer = None, buffer_shape: Tuple = None, graph: CommandGraph = None,
        print_shader: bool = False, axis: Union[int, Sequence[int]] = None, inverse: bool = False,
        normalize_inverse: bool = True, r2c: bool = False):
    """Transform ``buffer`` along ``axis`` (an int or a list; default the last axis),
    in place, or from ``input_buffer`` (vkfft_dispatcher.py:169-208)."""
    axes = _axes(buffer, axis)
    if input_buffer is not None:
        _copy_in(buffer, input_buffer, graph)
    if not r2c:
        for a in axes:
            F.fft(buffer, buffer_shape=buffer_shape, graph=graph, print_shader=print_shader, axis=a,
                  inverse=inverse, normalize_inverse=normalize_inverse)
        return
    if not isinstance(buffer, RFFTBuffer):
        raise ValueError("r2c transforms take an RFFTBuffer")
    last = len(buffer.shape) - 1
    if last not in axes:
        raise ValueError("an r2c transform includes the last (contiguous) axis")
    others = [a for a in axes if a != last]
    real_shape = buffer.real_shape if buffer_shape is None else tuple(buffer_shape)
    t = buffer.var_type
    if not inverse:
        F.fft(buffer, buffer_shape=real_shape, graph=graph, print_shader=print_shader, r2c=True, output_type=t,
              input_type=t)
        for a in others:
            F.fft(buffer, graph=graph, print_shader=print_shader, axis=a, output_type=t, input_type=t)
    else:
        for a in others:
            F.fft(buffer, graph=graph, print_shader=print_shader, axis=a, inverse=True,
                  normalize_inverse=normalize_inverse, output_type=t, input_type=t)
        F.fft(buffer, buffer_shape=real_shape, graph=graph, print_shader=print_shader, inverse=True,
              normalize_inverse=normalize_inverse, r2c=True, output_type=t, input_type=t)


def fft2(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False):
    if len(buffer.shape) < 2:
        raise ValueError("Buffer must have at least 2 dimensions")
    fft(buffer, graph=graph, print_shader=print_shader, axis=(len(buffer.shape) - 2, len(buffer.shape) - 1))


def fft3(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False):
    if len(buffer.shape) < 3:
        raise ValueError("Buffer must have at least 3 dimensions")
    fft(buffer, graph=graph, print_shader=print_shader, axis=(0, 1, 2))


def ifft(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False, axis: Union[int, Sequence[int]] = None,
         normalize: bool = True):
    fft(buffer, graph=graph, print_shader=print_shader, axis=axis, inverse=True, normalize_inverse=normalize)


def ifft2(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False):
    if len(buffer.shape) < 2:
        raise ValueError("Buffer must have at least 2 dimensions")
    ifft(buffer, graph=graph, print_shader=print_shader, axis=(len(buffer.shape) - 2, len(buffer.shape) - 1))


def ifft3(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False):
    if len(buffer.shape) != 3:
        raise ValueError("Buffer must have 3 dimensions")
    ifft(buffer, graph=graph, print_shader=print_shader, axis=(0, 1, 2))


def rfft(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False, axis: int = None):
    fft(buffer, buffer_shape=buffer.real_shape, graph=graph, print_shader=print_shader, r2c=True, axis=axis)


def rfft2(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False):
    if len(buffer.real_shape) < 2:
        raise ValueError("Buffer must have at least 2 dimensions")
    rfft(buffer, graph=graph, print_shader=print_shader, axis=(len(buffer.shape) - 2, len(buffer.shape) - 1))


def rfft3(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False):
    if len(buffer.real_shape) != 3:
        raise ValueError("Buffer must have 3 dimensions")
    rfft(buffer, graph=graph, print_shader=print_shader, axis=(0, 1, 2))


def irfft(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False, normalize: bool = True,
          axis: int = None):
    fft(buffer, buffer_shape=buffer.real_shape, graph=graph, print_shader=print_shader, inverse=True,
        normalize_inverse=normalize, r2c=True, axis=axis)


def irfft2(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False):
    if len(buffer.real_shape) < 2:
        raise ValueError("Buffer must have at least 2 dimensions")
    irfft(buffer, graph=graph, print_shader=print_shader, axis=(len(buffer.shape) - 2, len(buffer.shape) - 1))


def irfft3(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False):
    if len(buffer.real_shape) != 3:
        raise ValueError("Buffer must have 3 dimensions")
    irfft(buffer, graph=graph, print_shader=print_shader, axis=(0, 1, 2))


def _image_shape(shape) -> Tuple[int, int, int]:
    shape = tuple(int(s) for s in shape)
    if len(shape) == 2:
        shape = (1,) + shape
    if len(shape) != 3:
        raise ValueError("Kernel shape must be 2D or 3D!")
    return shape


def transpose_kernel2D(kernel: Buffer, shape: Tuple[int, ...] = None, graph: Optional[CommandGraph] = None,
                       keep_shader_code: bool = False):
    """Prepare a (batch of) 2-D kernel(s) for ``convolve2D``: its forward 2-D spectrum,
    in place (vkfft_dispatcher.py:109-135)."""
    shp = _image_shape(kernel.shape if shape is None else shape)
    if shp[0] * shp[1] * shp[2] != kernel.size:
        raise ValueError("shape does not match the kernel buffer")
    for a in (2, 1):
        F.fft(kernel, buffer_shape=shp, graph=graph, print_shader=keep_shader_code, axis=a)


@lru_cache(maxsize=None)
def _product_map(inner: int, conjugate: bool):
    """Kernel map x * K (or x * conj(K)); K read at the element's index, modulo ``inner``
    when one kernel is broadcast over the batch."""
    @vd_map
    def kernel_product(kbuf: vc.Buffer[dt.complex64]):
        op = read_op()
        kv = vc.new_register(dt.complex64)
        op.read_from_buffer(kbuf, register=kv, io_index=(op.io_index % inner) if inner else op.io_index)
        op.register[:] = vc.mult_complex(op.register, kv.conjugate() if conjugate else kv)
    return kernel_product


def _signal_range(pad, n: int):
    """VkFFT zero range [l, r) along an axis of length n (fft_zeropad_left / right,
    vkdispatch_native/stages/stage_fft.cpp:194-204) -> the input range the forward
    transform reads: the kept range when the block touches an end of the axis, a
    ZeroBlock for an interior block (None when nothing is zero)."""
    if pad is None:
        return None
    lo, hi = int(pad[0]), int(pad[1])
    if not 0 <= lo <= hi <= n:
        raise ValueError(f"padding range {pad} outside an axis of {n}")
    if lo == hi:
        return None
    if lo == 0:
        return (hi, n)
    if hi == n:
        return (0, lo)
    return ZeroBlock((lo, hi))


def convolve2D(buffer: Buffer, kernel: Buffer, normalize: bool = False, conjugate_kernel: bool = False,
               graph: Optional[CommandGraph] = None, keep_shader_code: bool = False,
               padding: Tuple[Tuple[int, int]] = None):
    """y = ifft2(fft2(x) * K) per 2-D image of ``buffer`` (complex, (H, W) or (B, H, W)),
    with K the output of ``transpose_kernel2D``: one per image or one for all
    (vkfft_dispatcher.py:137-167)."""
    if buffer.var_type is not dt.complex64 or kernel.var_type is not dt.complex64:
        raise ValueError("convolve2D takes complex64 buffers")
    B, H, W = _image_shape(buffer.shape)
    if kernel.size == B * H * W:
        inner = 0
    elif kernel.size == H * W:
        inner = H * W if B > 1 else 0
    else:
        raise ValueError(f"kernel has {kernel.size} elements; expected {H * W} or {B * H * W}")
    shp = (B, H, W)
    pads = tuple(padding) if padding is not None else (None, None)
    if len(pads) != 2:
        raise ValueError("padding gives one (start, end) range per image axis")
    rng_h, rng_w = _signal_range(pads[0], H), _signal_range(pads[1], W)
    F.fft(buffer, buffer_shape=shp, graph=graph, print_shader=keep_shader_code, axis=2, input_signal_range=rng_w)
    F.convolve(buffer, kernel, kernel_map=_product_map(inner, bool(conjugate_kernel)), buffer_shape=shp, graph=graph,
               print_shader=keep_shader_code, axis=1, normalize=normalize, input_signal_range=rng_h)
    F.fft(buffer, buffer_shape=shp, graph=graph, print_shader=keep_shader_code, axis=2, inverse=True,
          normalize_inverse=normalize)
