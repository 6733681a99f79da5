"""Kernel factories of vd.fft (counterparts of vkdispatch/fft/shader_factories.py).

``make_fft_shader`` / ``make_convolution_shader`` / ``make_transpose_shader``
are ``lru_cache``d on all their arguments, like the reference's
(shader_factories.py:11, 72, 106); mapping functions hash by instance.  Each
returns a :class:`CompiledKernel` whose source is the HIP instantiation.
"""
from __future__ import annotations

from functools import lru_cache
from typing import Optional, Tuple

from .. import dtypes as dt
from . import hooks
from .kernels import CompiledKernel, cfg_struct, kernel_symbol, translation_unit
from .planner import AxisGeometry, KernelPlan, make_plan, transposed_size
from .precision import kernel_real_type



def _unaligned(*pitch_bytes: int) -> bool:
    """Whether rows at these byte pitches start off 128-byte cache lines."""
    return any(int(p) % 128 for p in pitch_bytes)


def _compute_dtype(compute_type):
    return dt.complex128 if compute_type is dt.complex128 else dt.complex64


class ZeroBlock(tuple):
    """An input range read as zeros, ``(l, r)`` with 0 < l < r < N: the kept samples are
    [0, l) and [r, N) (VkFFT's interior zero padding, vd.vkfft's ``padding``).  The
    kernel configuration encodes it as the wrapped kept range SIG_LO = r > SIG_HI = l."""


def _resolve_signal_range(signal_range, N: int) -> Tuple[int, int]:
    if signal_range is None:
        return 0, N
    if isinstance(signal_range, ZeroBlock):
        lo, hi = int(signal_range[0]), int(signal_range[1])
        if not 0 < lo < hi < N:
            raise ValueError(f"zero block {tuple(signal_range)} must lie strictly inside [0, {N}]")
        return hi, lo
    s, e = signal_range
    s = 0 if s is None else int(s)
    e = N if e is None else int(e)
    if s < 0 or e > N or s > e:
        raise ValueError(f"input_signal_range {signal_range} is not within [0, {N}]")
    return s, e


def conv_plan(buffer_shape, axis, compute_type) -> KernelPlan:
    geom = AxisGeometry.of(buffer_shape, axis)
    return make_plan(geom, kernel_real_type(compute_type), kind="conv")


def _io_functors(args, output_map, input_map, output_type, input_type, compute, N, knum):
    out_args = [a for a in args if a.role == "output"]
    in_args = [a for a in args if a.role == "input"]
    functors = []
    if output_map is not None:
        functors.append(hooks.functor_source("HookOut", output_map, out_args, compute, N, writing=True,
                                             per_kernel=True, kernel_num=knum))
    else:
        functors.append(f"using HookOut = vkx::DefaultOut<{out_args[0].slot}, {output_type.child_type.cpp}>;")
    if input_map is not None:
        functors.append(hooks.functor_source("HookIn", input_map, in_args, compute, N, writing=False,
                                             per_kernel=False))
    elif output_map is not None:
        functors.append(f"using HookIn = vkx::DefaultIn<{in_args[0].slot}, {input_type.child_type.cpp}>;")
    else:
        functors.append(f"using HookIn = vkx::DefaultIn<{out_args[0].slot}, {input_type.child_type.cpp}>;")
    return functors


@lru_cache(maxsize=None)
def make_fft_shader(buffer_shape: Tuple, axis: int = None, inverse: bool = False, normalize_inverse: bool = True,
                    r2c: bool = False, input_map=None, output_map=None, input_type=None, output_type=None,
                    compute_type=None, input_signal_range: Optional[Tuple] = None) -> CompiledKernel:
    if output_type is None:
        output_type = dt.complex64
    if input_type is None and input_map is None:
        input_type = output_type
    if compute_type is None:
        compute_type = dt.complex64
    compute = _compute_dtype(compute_type)
    geom = AxisGeometry.of(buffer_shape, axis)
    N = geom.N
    if r2c and geom.S != 1:
        raise ValueError("R2C io only supported for contiguous data")
    if (r2c and N % 2 == 0 and input_map is None and output_map is None and output_type is input_type
            and (not inverse or input_signal_range is None)):
        return _make_real_row_shader(geom, inverse, normalize_inverse, output_type, compute_type,
                                     input_signal_range)
    plan = make_plan(geom, kernel_real_type(compute_type))
    args = hooks.declare_args(output_map, input_map, None, output_type, input_type, None, conv=False)
    if r2c:
        half = N // 2 + 1
        in_os, out_os = (half, 2 * half) if inverse else (2 * half, half)
    else:
        in_os = out_os = N * geom.S
    sig = _resolve_signal_range(input_signal_range, N)
    normalize = bool(inverse and normalize_inverse)
    # rows at unaligned offsets (hooked gathers and scatters, strides off the cache line)
    # share lines between neighbouring tiles: those tiles go to one XCD (one fetch)
    xcd = bool(input_map or output_map) or _unaligned((geom.S if geom.S > 1 else N) * output_type.item_size,
                                                      (geom.S if geom.S > 1 else N) * (input_type or output_type).item_size)
    cfg = cfg_struct(plan, in_os=in_os, out_os=out_os, k_os=N * geom.S, sig=sig, r2c=r2c, inverse=inverse,
                     normalize=normalize, scale=1.0 / N, xcd_map=xcd)
    functors = _io_functors(args, output_map, input_map, output_type, input_type, compute, N, 1)
    kind = ("irfft" if inverse else "rfft") if r2c else "fft"
    tag = ("inv" if inverse else "fwd") + ("_hook" if (input_map or output_map) else "")
    src = translation_unit(cfg, functors, "vkx::fft_body<Cfg, HookIn, HookOut>(prm);", plan.BLOCK,
                           kernel_symbol(kind, plan, tag))
    name = f"fft_shader_{tuple(buffer_shape)}_{axis}_{inverse}_{normalize_inverse}_{r2c}"
    return CompiledKernel(src, plan, args, list(plan.radices), None, name=name)


def _make_real_row_shader(geom: AxisGeometry, inverse: bool, normalize_inverse: bool, io_type, compute_type,
                          input_signal_range, oop: bool = False, real_pitch: Optional[int] = None,
                          spec_pitch: Optional[int] = None) -> CompiledKernel:
    """r2c / c2r of even-length rows through an N/2-point complex FFT (rfft_body / irfft_body).

    In place on an RFFTBuffer by default.  ``oop``: the real rows live in a separate
    buffer of row pitch ``real_pitch`` reals (slot 1 for r2c input, slot 0 for c2r
    output) and the spectrum in one of row pitch ``spec_pitch`` complex values."""
    N = geom.N
    M = N // 2
    half_geom = AxisGeometry((geom.OUTER, M), 1, M, 1, geom.OUTER)
    plan = make_plan(half_geom, kernel_real_type(compute_type), need_lds=not inverse,
                     kind="irfft" if inverse else "rfft")
    rp = 2 * M + 2 if real_pitch is None else int(real_pitch)
    sp = M + 1 if spec_pitch is None else int(spec_pitch)
    in_os, out_os = (sp, rp) if inverse else (rp, sp)
    sig = _resolve_signal_range(input_signal_range, N)
    normalize = bool(inverse and normalize_inverse)
    # plain tile order: with the XCD-contiguous order the 1024-point rows of 513-value
    # spectra measured slower (67.1 / 57.2 against 61.6 / 54.2 us)
    cfg = cfg_struct(plan, in_os=in_os, out_os=out_os, k_os=0, sig=sig, r2c=True, inverse=inverse,
                     normalize=normalize, scale=1.0 / N)
    from .hooks import Arg
    if oop:
        real_t = io_type.child_type
        args = ([Arg("output", "buffer", real_t, 0), Arg("input", "buffer", io_type, 1)] if inverse else
                [Arg("output", "buffer", io_type, 0), Arg("input", "buffer", real_t, 1)])
    else:
        args = [Arg("output", "buffer", io_type, 0)]
    kind = "irfft" if inverse else "rfft"
    body = f"vkx::{kind}_body<Cfg, {io_type.child_type.cpp}, {'true' if oop else 'false'}>(prm);"
    tag = ("inv" if inverse else "fwd") + ("_oop" if oop else "")
    src = translation_unit(cfg, [], body, plan.BLOCK, kernel_symbol(kind + "h", plan, tag))
    return CompiledKernel(src, plan, args, list(plan.radices), None, name=f"{kind}_shader_{geom.shape}",
                          roots_n=N)


@lru_cache(maxsize=None)
def make_real_rows_shader(real_shape: Tuple, inverse: bool, normalize: bool = True, io_type=dt.complex64,
                          compute_type=dt.complex64, oop: bool = False, real_pitch: Optional[int] = None,
                          spec_pitch: Optional[int] = None) -> CompiledKernel:
    """Row r2c / c2r kernel for even row length (internal building block of fused paths)."""
    geom = AxisGeometry.of(tuple(real_shape), None)
    if geom.N % 2:
        raise ValueError("real-row kernels need an even row length")
    return _make_real_row_shader(geom, inverse, normalize, io_type, compute_type, None, oop, real_pitch,
                                 spec_pitch)


@lru_cache(maxsize=None)
def get_transposed_size(buffer_shape: Tuple, axis: int = None, compute_type=dt.complex64) -> int:
    return transposed_size(conv_plan(tuple(buffer_shape), axis, compute_type))


@lru_cache(maxsize=None)
def make_transpose_shader(buffer_shape: Tuple, axis: int = None, kernel_inner_only: bool = False,
                          input_type=dt.complex64, output_type=dt.complex64,
                          compute_type=dt.complex64) -> CompiledKernel:
    geom = AxisGeometry.of(buffer_shape, axis)
    plan = conv_plan(tuple(buffer_shape), axis, compute_type)
    from .hooks import Arg
    args = [Arg("output", "buffer", output_type, 0), Arg("input", "buffer", input_type, 1)]
    cfg = cfg_struct(plan, in_os=geom.N * geom.S, out_os=geom.N * geom.S, k_os=geom.N * geom.S, sig=(0, geom.N),
                     r2c=False, inverse=False, normalize=False, scale=1.0, k_inner_only=kernel_inner_only,
                     k_transposed=True)
    body = (f"vkx::transpose_body<Cfg, {input_type.child_type.cpp}, {output_type.child_type.cpp}>(prm);")
    src = translation_unit(cfg, [], body, plan.BLOCK, kernel_symbol("transpose", plan))
    return CompiledKernel(src, plan, args, None, None, name=f"transpose_shader_{tuple(buffer_shape)}_{axis}")


_kernel_index_doc = "index of the kernel being traced in a kernel_num > 1 convolution (None otherwise)"


def mapped_kernel_index():
    return hooks.mapped_kernel_index()


@lru_cache(maxsize=None)
def make_convolution_shader(buffer_shape: Tuple, kernel_map=None, kernel_num: int = 1, axis: int = None,
                            normalize: bool = True, transposed_kernel: bool = False,
                            kernel_inner_only: bool = False, input_map=None, output_map=None, input_type=None,
                            output_type=None, kernel_type=None, compute_type=None,
                            input_signal_range: Optional[Tuple] = None) -> CompiledKernel:
    if output_type is None:
        output_type = dt.complex64
    if input_type is None and input_map is None:
        input_type = output_type
    if kernel_type is None:
        kernel_type = dt.complex64
    if compute_type is None:
        compute_type = dt.complex64
    if kernel_num < 1:
        raise ValueError("kernel_num must be >= 1")
    compute = _compute_dtype(compute_type)
    geom = AxisGeometry.of(buffer_shape, axis)
    N = geom.N
    plan = conv_plan(tuple(buffer_shape), axis, compute_type)
    args = hooks.declare_args(output_map, input_map, kernel_map, output_type, input_type, kernel_type, conv=True)
    k_args = [a for a in args if a.role == "kernel"]
    sig = _resolve_signal_range(input_signal_range, N)
    cfg = cfg_struct(plan, in_os=N * geom.S, out_os=N * geom.S, k_os=N * geom.S, sig=sig, r2c=False,
                     inverse=False, normalize=bool(normalize), scale=1.0 / N, knum=kernel_num,
                     k_inner_only=kernel_inner_only, k_transposed=transposed_kernel,
                     xcd_map=bool(input_map or output_map) or _unaligned(
                         (geom.S if geom.S > 1 else N) * (output_type or compute_type).item_size))
    functors = _io_functors(args, output_map, input_map, output_type, input_type, compute, N, kernel_num)
    if kernel_map is not None:
        functors.append(hooks.functor_source("HookKernel", kernel_map, k_args, compute, N, writing=False,
                                             per_kernel=True, kernel_num=kernel_num))
    else:
        functors.append(f"using HookKernel = vkx::DefaultKernel<{k_args[0].slot}, {kernel_type.child_type.cpp}>;")
    tag = ("inner" if kernel_inner_only else "") + ("_t" if transposed_kernel else "") + \
        ("_hook" if (input_map or output_map or kernel_map) else "") + (f"_k{kernel_num}" if kernel_num > 1 else "")
    src = translation_unit(cfg, functors, "vkx::conv_body<Cfg, HookIn, HookKernel, HookOut>(prm);", plan.BLOCK,
                           kernel_symbol("conv", plan, tag))
    return CompiledKernel(src, plan, args, list(plan.radices), list(reversed(plan.radices)),
                          name=f"convolution_shader_{tuple(buffer_shape)}_{axis}")


def get_cache_info():
    return make_fft_shader.cache_info()


def get_convoliution_cache_info():  # reference spelling
    return make_convolution_shader.cache_info()


get_convolution_cache_info = get_convoliution_cache_info


def print_cache_info():
    print(get_cache_info())
    print(get_convoliution_cache_info())


# Caches of plans that own device memory (pair-convolution scratch, mixed-precision
# intermediates, ragged-batch queues): cache_clear() empties them too, so their
# buffers are released once no CommandGraph that recorded them is alive.
_PLAN_CACHE_CLEARERS = []


def register_plan_cache(clear) -> None:
    """Register a callable that empties one module-level cache of plans."""
    _PLAN_CACHE_CLEARERS.append(clear)


def cache_clear():
    make_convolution_shader.cache_clear()
    make_fft_shader.cache_clear()
    make_transpose_shader.cache_clear()
    make_real_rows_shader.cache_clear()
    get_transposed_size.cache_clear()
    for clear in _PLAN_CACHE_CLEARERS:
        clear()
