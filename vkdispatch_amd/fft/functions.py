"""Public vd.fft functions.

Signatures, defaults, axis orchestration and argument validation are those of
the reference (vkdispatch/fft/functions.py:13-713) so this module is a drop-in:
``fft2`` transforms axis ndim-2 then ndim-1, ``fft3`` axes 0,1,2, the inverse
variants normalise every axis, ``rfft2`` is ``rfft`` on the last axis then
``fft`` on axis ndim-2, ``convolve2DR`` is rfft -> convolve(axis ndim-2) ->
irfft, and mapping hooks go on the first (input) and last (output) pass.

Two deliberate differences, both documented in the package README:
* ``convolve2D(..., output_map=...)`` works (the reference passes two buffers
  to ``ifft``, which takes one, and raises TypeError, functions.py:603-611);
* ``convolve2DR`` with the default kernel map runs a fused 2-D path when the
  shape allows it (same result; see ``fft/fused.py``).
"""
from __future__ import annotations

from typing import List, Optional, Tuple, Union

from .. import dtypes as dt
from ..buffer import Buffer, RFFTBuffer
from ..graph import CommandGraph, dispatch
from ..shader_map import MappingFunction
from . import precision as prec
from .shader_factories import (get_transposed_size, make_convolution_shader, make_fft_shader,
                               make_transpose_shader, register_plan_cache)


def _validate_map_argument_annotations(map_fn: MappingFunction, map_name: str) -> None:
    for t in map_fn.buffer_types:
        if not hasattr(t, "__args__") or len(t.__args__) != 1:
            raise ValueError(f"{map_name} contains an annotation without exactly one type argument: {t}")


def _resolve_output_precision(buffers, output_map, output_type):
    if output_map is not None:
        if output_type is not None:
            raise ValueError("output_type cannot be provided when output_map is used")
        return None
    resolved = buffers[0].var_type if output_type is None else output_type
    prec.validate_complex_precision(resolved, arg_name="output_type")
    prec.ensure_supported_complex_precision(resolved, role="Output")
    return resolved


def _resolve_input_precision(buffers, input_map, output_map, input_type, output_precision):
    if input_map is not None:
        if input_type is not None:
            raise ValueError("input_type cannot be provided when input_map is used")
        return None
    if output_map is not None:
        n_out = len(output_map.buffer_types)
        if len(buffers) <= n_out:
            raise ValueError("When output_map is used without input_map, an input buffer argument must be "
                             "provided after output_map arguments")
        resolved = input_type
        if resolved is None:
            inferred = buffers[n_out]
            if not hasattr(inferred, "var_type"):
                raise ValueError("When output_map is used without input_map, the argument after output_map "
                                 "arguments must be a buffer")
            resolved = inferred.var_type
        prec.validate_complex_precision(resolved, arg_name="input_type")
        prec.ensure_supported_complex_precision(resolved, role="Input")
        return resolved
    if output_precision is None:
        raise ValueError("output_precision must be provided when output_map is not used")
    resolved = output_precision if input_type is None else input_type
    prec.validate_complex_precision(resolved, arg_name="input_type")
    prec.ensure_supported_complex_precision(resolved, role="Input")
    if resolved != output_precision:
        raise ValueError("input_type must match output_type when input_map is None (default FFT path is in-place)")
    return resolved


def _resolve_kernel_precision(buffers, kernel_map, kernel_type):
    if kernel_map is not None:
        if kernel_type is not None:
            raise ValueError("kernel_type cannot be provided when kernel_map is used")
        return None
    if len(buffers) < 2:
        raise ValueError("Kernel precision inference requires a kernel buffer argument")
    resolved = buffers[1].var_type if kernel_type is None else kernel_type
    prec.validate_complex_precision(resolved, arg_name="kernel_type")
    prec.ensure_supported_complex_precision(resolved, role="Kernel")
    return resolved


def fft(*buffers: Buffer, buffer_shape: Tuple = None, graph: CommandGraph = None, print_shader: bool = False,
        axis: int = None, name: str = None, inverse: bool = False, normalize_inverse: bool = True,
        r2c: bool = False, input_map: MappingFunction = None, output_map: MappingFunction = None,
        output_type=None, input_type=None, compute_type=None,
        input_signal_range: Union[Tuple[Optional[int], Optional[int]], None] = None):
    if len(buffers) < 1:
        raise ValueError("At least one buffer must be provided")
    if input_map is None and output_map is None and len(buffers) != 1:
        raise ValueError("fft() expects exactly one buffer unless input_map/output_map are used")
    if buffer_shape is None:
        buffer_shape = buffers[0].shape
    out_t = _resolve_output_precision(buffers, output_map, output_type)
    in_t = _resolve_input_precision(buffers, input_map, output_map, input_type, out_t)
    io: List = []
    if output_map is None:
        io.append(out_t)
    else:
        _validate_map_argument_annotations(output_map, "output_map")
    if input_map is None:
        if in_t is not None:
            io.append(in_t)
    else:
        _validate_map_argument_annotations(input_map, "input_map")
    comp_t = prec.resolve_compute_precision(io, compute_type)
    try:
        kernel = make_fft_shader(tuple(buffer_shape), axis, inverse=inverse, normalize_inverse=normalize_inverse,
                                 r2c=r2c, input_map=input_map, output_map=output_map, input_type=in_t,
                                 output_type=out_t, compute_type=comp_t, input_signal_range=input_signal_range)
    except ValueError:
        if _large_or_prime_fallback(buffers, buffer_shape, axis, inverse, normalize_inverse, r2c, input_map,
                                    output_map, out_t, comp_t, input_signal_range, graph):
            return
        raise
    if print_shader:
        print(kernel)
    if input_map is None and output_map is None and not r2c:
        n = 1
        for s in buffer_shape:
            n *= s
        if buffers[0].size < n:
            raise ValueError(f"buffer of {buffers[0].size} elements is smaller than buffer_shape {tuple(buffer_shape)}")
    dispatch(kernel, buffers, graph)


def _large_or_prime_fallback(buffers, buffer_shape, axis, inverse, normalize, r2c, input_map, output_map, out_t,
                             comp_t, signal_range, graph) -> bool:
    """Lengths outside the single-workgroup kernels (a prime factor above 31, or more
    than 16384 points) run on the ragged executor (mixed radix through global scratch,
    Bluestein for large primes): contiguous axis, complex64, no hooks, in place."""
    from .planner import AxisGeometry
    geom = AxisGeometry.of(tuple(buffer_shape), axis)
    if (geom.S != 1 or r2c or input_map is not None or output_map is not None or signal_range is not None
            or out_t is not dt.complex64 or comp_t not in (dt.complex64, dt.complex32)):
        return False
    from .ragged import RaggedFFT
    plan = _uniform_ragged(geom.N, geom.OUTER, bool(inverse), bool(normalize))
    buf = buffers[0]
    if graph is not None:
        graph.record_callable(plan.bind(buf))
    else:
        plan(buf)
    return True


def _uniform_ragged(n, count, inverse, normalize):
    from .ragged import RaggedFFT
    key = (n, count, inverse, normalize)
    if key not in _UNIFORM_RAGGED:
        plan = RaggedFFT([n] * count, inverse=inverse, normalize=normalize, block=512)
        if plan.n_items > count and count > 128:
            # many tiled transforms: 256-thread workgroups (65536 x 256 points: 298 vs
            # 371 us); whole transforms and small tiled batches run faster at 512
            plan.block = plan.block_first = 256
        _UNIFORM_RAGGED[key] = plan
    return _UNIFORM_RAGGED[key]


_UNIFORM_RAGGED: dict = {}
register_plan_cache(_UNIFORM_RAGGED.clear)


def fft2(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False, input_map: MappingFunction = None,
         output_map: MappingFunction = None, output_type=None, input_type=None, compute_type=None):
    if len(buffer.shape) != 2 and len(buffer.shape) != 3:
        raise ValueError("Buffer must have 2 or 3 dimensions")
    fft(buffer, graph=graph, print_shader=print_shader, axis=len(buffer.shape) - 2, input_map=input_map,
        output_type=output_type, input_type=input_type, compute_type=compute_type)
    fft(buffer, graph=graph, print_shader=print_shader, axis=len(buffer.shape) - 1, output_map=output_map,
        output_type=output_type if output_map is None else None,
        input_type=input_type if output_map is None else None, compute_type=compute_type)


def fft3(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False, input_map: MappingFunction = None,
         output_map: MappingFunction = None, output_type=None, input_type=None, compute_type=None):
    if len(buffer.shape) != 3:
        raise ValueError("Buffer must have 3 dimensions")
    fft(buffer, graph=graph, print_shader=print_shader, axis=0, input_map=input_map, output_type=output_type,
        input_type=input_type, compute_type=compute_type)
    fft(buffer, graph=graph, print_shader=print_shader, axis=1, output_type=output_type, input_type=input_type,
        compute_type=compute_type)
    fft(buffer, graph=graph, print_shader=print_shader, axis=2, output_map=output_map,
        output_type=output_type if output_map is None else None,
        input_type=input_type if output_map is None else None, compute_type=compute_type)


def ifft(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False, axis: int = None, name: str = None,
         normalize: bool = True, input_map: MappingFunction = None, output_map: MappingFunction = None,
         output_type=None, input_type=None, compute_type=None):
    fft(buffer, graph=graph, print_shader=print_shader, axis=axis, name=name, inverse=True,
        normalize_inverse=normalize, input_map=input_map, output_map=output_map, output_type=output_type,
        input_type=input_type, compute_type=compute_type)


def ifft2(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False, normalize: bool = True,
          input_map: MappingFunction = None, output_map: MappingFunction = None, output_type=None, input_type=None,
          compute_type=None):
    if len(buffer.shape) != 2 and len(buffer.shape) != 3:
        raise ValueError("Buffer must have 2 or 3 dimensions")
    ifft(buffer, graph=graph, print_shader=print_shader, axis=len(buffer.shape) - 2, normalize=normalize,
         input_map=input_map, output_type=output_type, input_type=input_type, compute_type=compute_type)
    ifft(buffer, graph=graph, print_shader=print_shader, axis=len(buffer.shape) - 1, normalize=normalize,
         output_map=output_map, output_type=output_type if output_map is None else None,
         input_type=input_type if output_map is None else None, compute_type=compute_type)


def ifft3(buffer: Buffer, graph: CommandGraph = None, print_shader: bool = False, normalize: bool = True,
          input_map: MappingFunction = None, output_map: MappingFunction = None, output_type=None, input_type=None,
          compute_type=None):
    if len(buffer.shape) != 3:
        raise ValueError("Buffer must have 3 dimensions")
    ifft(buffer, graph=graph, print_shader=print_shader, axis=0, normalize=normalize, input_map=input_map,
         output_type=output_type, input_type=input_type, compute_type=compute_type)
    ifft(buffer, graph=graph, print_shader=print_shader, axis=1, normalize=normalize, output_type=output_type,
         input_type=input_type, compute_type=compute_type)
    ifft(buffer, graph=graph, print_shader=print_shader, axis=2, normalize=normalize, output_map=output_map,
         output_type=output_type if output_map is None else None,
         input_type=input_type if output_map is None else None, compute_type=compute_type)


def rfft(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False, name: str = None,
         compute_type=None):
    fft(buffer, buffer_shape=buffer.real_shape, graph=graph, print_shader=print_shader, name=name, r2c=True,
        output_type=buffer.var_type, input_type=buffer.var_type, compute_type=compute_type)


def rfft2(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False, compute_type=None):
    if len(buffer.real_shape) != 2 and len(buffer.real_shape) != 3:
        raise ValueError("Buffer must have 2 or 3 dimensions")
    rfft(buffer, graph=graph, print_shader=print_shader, compute_type=compute_type)
    fft(buffer, graph=graph, print_shader=print_shader, axis=len(buffer.real_shape) - 2,
        output_type=buffer.var_type, input_type=buffer.var_type, compute_type=compute_type)


def rfft3(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False, compute_type=None):
    if len(buffer.real_shape) != 3:
        raise ValueError("Buffer must have 3 dimensions")
    rfft(buffer, graph=graph, print_shader=print_shader, compute_type=compute_type)
    fft(buffer, graph=graph, print_shader=print_shader, axis=1, output_type=buffer.var_type,
        input_type=buffer.var_type, compute_type=compute_type)
    fft(buffer, graph=graph, print_shader=print_shader, axis=0, output_type=buffer.var_type,
        input_type=buffer.var_type, compute_type=compute_type)


def irfft(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False, name: str = None,
          normalize: bool = True, compute_type=None):
    fft(buffer, buffer_shape=buffer.real_shape, graph=graph, print_shader=print_shader, name=name, inverse=True,
        normalize_inverse=normalize, r2c=True, output_type=buffer.var_type, input_type=buffer.var_type,
        compute_type=compute_type)


def irfft2(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False, normalize: bool = True,
           compute_type=None):
    if len(buffer.real_shape) != 2 and len(buffer.real_shape) != 3:
        raise ValueError("Buffer must have 2 or 3 dimensions")
    ifft(buffer, graph=graph, print_shader=print_shader, axis=len(buffer.real_shape) - 2, normalize=normalize,
         output_type=buffer.var_type, input_type=buffer.var_type, compute_type=compute_type)
    irfft(buffer, graph=graph, print_shader=print_shader, normalize=normalize, compute_type=compute_type)


def irfft3(buffer: RFFTBuffer, graph: CommandGraph = None, print_shader: bool = False, normalize: bool = True,
           compute_type=None):
    if len(buffer.real_shape) != 3:
        raise ValueError("Buffer must have 3 dimensions")
    ifft(buffer, graph=graph, print_shader=print_shader, axis=0, normalize=normalize, output_type=buffer.var_type,
         input_type=buffer.var_type, compute_type=compute_type)
    ifft(buffer, graph=graph, print_shader=print_shader, axis=1, normalize=normalize, output_type=buffer.var_type,
         input_type=buffer.var_type, compute_type=compute_type)
    irfft(buffer, graph=graph, print_shader=print_shader, normalize=normalize, compute_type=compute_type)


def convolve(*buffers: Buffer, kernel_map: MappingFunction = None, kernel_num: int = 1, buffer_shape: Tuple = None,
             graph: CommandGraph = None, print_shader: bool = False, axis: int = None, normalize: bool = True,
             name: str = None, transposed_kernel: bool = False, kernel_inner_only: bool = False,
             input_map: MappingFunction = None, output_map: MappingFunction = None, output_type=None,
             input_type=None, kernel_type=None, compute_type=None,
             input_signal_range: Union[Tuple[Optional[int], Optional[int]], None] = None):
    if len(buffers) < 1:
        raise ValueError("At least one buffer must be provided")
    if kernel_map is None and len(buffers) < 2:
        raise ValueError("convolve() requires at least an output buffer and kernel buffer")
    if buffer_shape is None:
        buffer_shape = buffers[0].shape
    out_t = _resolve_output_precision(buffers, output_map, output_type)
    in_t = _resolve_input_precision(buffers, input_map, output_map, input_type, out_t)
    k_t = _resolve_kernel_precision(buffers, kernel_map, kernel_type)
    io: List = []
    if output_map is None:
        io.append(out_t)
    else:
        _validate_map_argument_annotations(output_map, "output_map")
    if input_map is None:
        if in_t is not None:
            io.append(in_t)
    else:
        _validate_map_argument_annotations(input_map, "input_map")
    if kernel_map is None:
        io.append(k_t)
    else:
        _validate_map_argument_annotations(kernel_map, "kernel_map")
    comp_t = prec.resolve_compute_precision(io, compute_type)
    kernel = make_convolution_shader(tuple(buffer_shape), kernel_map, kernel_num, axis,
                                     transposed_kernel=transposed_kernel, kernel_inner_only=kernel_inner_only,
                                     normalize=normalize, input_map=input_map, output_map=output_map,
                                     input_type=in_t, output_type=out_t, kernel_type=k_t, compute_type=comp_t,
                                     input_signal_range=input_signal_range)
    if print_shader:
        print(kernel)
    if kernel_map is None and len(buffers) >= 2 and hasattr(buffers[1], "size"):
        _check_kernel_extent(kernel, buffers, tuple(buffer_shape), axis, transposed_kernel, kernel_inner_only)
    dispatch(kernel, buffers, graph)


def _check_kernel_extent(kernel, buffers, buffer_shape, axis, transposed, inner_only) -> None:
    """The default kernel map reads K at the data's io_index (or the transposed layout);
    refuse a kernel buffer too small for that instead of reading past it."""
    from .planner import AxisGeometry
    geom = AxisGeometry.of(buffer_shape, axis)
    kbuf = [b for a, b in zip(kernel.args, buffers) if a.role == "kernel" and a.kind == "buffer"]
    if not kbuf:
        return
    kb = kbuf[0]
    if transposed:
        need = kernel.plan.NWG_INNER * kernel.plan.KSLOTS * kernel.plan.BLOCK if inner_only else \
            kernel.plan.NWG * kernel.plan.KSLOTS * kernel.plan.BLOCK
        if inner_only and geom.S == 1:
            need = kernel.plan.KSLOTS * kernel.plan.TT
    elif inner_only:
        need = geom.N * geom.S
    else:
        need = geom.N * geom.S * geom.OUTER
    if kb.size < need:
        raise ValueError(f"kernel buffer has {kb.size} elements but the convolution reads {need} "
                         f"(pass kernel_inner_only=True to broadcast one kernel over the outer batch)")


def convolve2D(buffer: Buffer, kernel: Buffer, kernel_map: MappingFunction = None, buffer_shape: Tuple = None,
               graph: CommandGraph = None, print_shader: bool = False, normalize: bool = True,
               transposed_kernel: bool = False, kernel_inner_only: bool = False, input_map: MappingFunction = None,
               output_map: MappingFunction = None, output_type=None, input_type=None, kernel_type=None,
               compute_type=None):
    if len(buffer.shape) != 2 and len(buffer.shape) != 3:
        raise ValueError("Buffer must have 2 or 3 dimensions")
    input_buffers = [buffer]
    if input_map is not None:
        input_buffers.append(buffer)
    fft(*input_buffers, graph=graph, print_shader=print_shader, input_map=input_map, output_type=output_type,
        input_type=input_type, compute_type=compute_type)
    convolve(buffer, kernel, kernel_map=kernel_map, buffer_shape=buffer_shape, graph=graph,
             transposed_kernel=transposed_kernel, kernel_inner_only=kernel_inner_only, print_shader=print_shader,
             axis=len(buffer.shape) - 2, normalize=normalize, output_type=output_type, input_type=input_type,
             kernel_type=kernel_type, compute_type=compute_type)
    if output_map is not None:
        fft(*([buffer] * (len(output_map.buffer_types) + 1)), graph=graph, print_shader=print_shader,
            inverse=True, normalize_inverse=normalize, output_map=output_map, compute_type=compute_type)
    else:
        ifft(buffer, graph=graph, print_shader=print_shader, normalize=normalize, output_type=output_type,
             input_type=input_type, compute_type=compute_type)


def convolve2DR(buffer: RFFTBuffer, kernel: RFFTBuffer, kernel_map: MappingFunction = None,
                buffer_shape: Tuple = None, transposed_kernel: bool = False, kernel_inner_only: bool = False,
                graph: CommandGraph = None, print_shader: bool = False, normalize: bool = True, compute_type=None):
    if len(buffer.real_shape) != 2 and len(buffer.real_shape) != 3:
        raise ValueError("Buffer must have 2 or 3 dimensions")
    from . import fused
    if fused.try_convolve2DR(buffer, kernel, kernel_map=kernel_map, buffer_shape=buffer_shape,
                             transposed_kernel=transposed_kernel, kernel_inner_only=kernel_inner_only, graph=graph,
                             normalize=normalize, compute_type=compute_type, print_shader=print_shader):
        return
    rfft(buffer, graph=graph, print_shader=print_shader, compute_type=compute_type)
    convolve(buffer, kernel, kernel_map=kernel_map, buffer_shape=buffer_shape, graph=graph,
             transposed_kernel=transposed_kernel, kernel_inner_only=kernel_inner_only, print_shader=print_shader,
             axis=len(buffer.shape) - 2, normalize=normalize, output_type=buffer.var_type,
             input_type=buffer.var_type, kernel_type=kernel.var_type, compute_type=compute_type)
    irfft(buffer, graph=graph, print_shader=print_shader, normalize=normalize, compute_type=compute_type)


def transpose(in_buffer: Buffer, conv_shape: Tuple = None, axis: int = None, out_buffer: Buffer = None,
              graph: CommandGraph = None, kernel_inner_only: bool = False, print_shader: bool = False,
              input_type=None, output_type=None, compute_type=None) -> Buffer:
    in_t = in_buffer.var_type if input_type is None else input_type
    prec.validate_complex_precision(in_t, arg_name="input_type")
    prec.ensure_supported_complex_precision(in_t, role="Input")
    if out_buffer is not None and output_type is None:
        out_t = out_buffer.var_type
    elif output_type is None:
        out_t = in_buffer.var_type
    else:
        out_t = output_type
    prec.validate_complex_precision(out_t, arg_name="output_type")
    prec.ensure_supported_complex_precision(out_t, role="Output")
    comp_t = prec.resolve_compute_precision([in_t, out_t], compute_type)
    if conv_shape is None:
        conv_shape = in_buffer.shape
    size = get_transposed_size(tuple(conv_shape), axis=axis, compute_type=comp_t)
    if out_buffer is None:
        out_buffer = Buffer((size,), var_type=out_t)
    elif out_buffer.var_type != out_t:
        raise ValueError(f"out_buffer type ({out_buffer.var_type.name}) does not match output_type ({out_t.name})")
    if out_buffer.size < size:
        raise ValueError(f"Output buffer size {out_buffer.size} is smaller than expected transposed size {size}")
    kern = make_transpose_shader(tuple(conv_shape), axis=axis, kernel_inner_only=kernel_inner_only,
                                 input_type=in_t, output_type=out_t, compute_type=comp_t)
    if print_shader:
        print(kern)
    dispatch(kern, [out_buffer, in_buffer], graph)
    return out_buffer
