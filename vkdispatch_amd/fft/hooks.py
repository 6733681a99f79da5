"""Mapping hooks -> compile-time functors.

Kernel arguments are declared in the reference's order
(vkdispatch/fft/io_manager.py:58-105): output-map arguments (or the output
buffer), then input-map arguments (or, when only an output map is given, an
explicit input buffer), then kernel-map arguments.  With neither an input map
nor an explicit input, the input aliases the output (in-place transform).

Each hook is traced once per kernel index with ``vd.fft.read_op()`` /
``vd.fft.write_op()`` returning symbolic operations whose ``register``,
``io_index`` and ``fft_index`` are the per-element values of the kernel's
load / kernel-multiply / store loop (the reference runs the hook once per
register instead, io_manager.py:107-155; the generated code is equivalent
because the reference's per-register values are the same expressions).
"""
from __future__ import annotations

import dataclasses
import threading
from typing import List, Optional, Sequence, Tuple

from .. import codegen as vc
from .. import dtypes as dt
from ..shader_map import MappingFunction

_state = threading.local()


def _get(name):
    return getattr(_state, name, None)


def read_op():
    op = _get("read_op")
    if op is None:
        raise ValueError("No global read operation is set for the current thread!")
    return op


def write_op():
    op = _get("write_op")
    if op is None:
        raise ValueError("No global write operation is set for the current thread!")
    return op


def mapped_kernel_index() -> Optional[int]:
    return _get("kernel_index")


class _Op:
    """Symbolic per-element operation handed to a mapping hook."""

    def __init__(self, compute_type, fft_size: int, writing: bool):
        self.compute_type = compute_type
        self.register = vc.ShaderVariable(compute_type, "op.reg", lvalue=True)
        self.io_index = vc.ShaderVariable(dt.uint32, "op.io_index")
        self.fft_index = vc.ShaderVariable(dt.uint32, "op.fft_index")
        self.fft_size = fft_size
        self._writing = writing

    def read_from_buffer(self, buffer, register=None, io_index=None):
        if self._writing:
            raise ValueError("read_from_buffer is not available on a write operation")
        if not isinstance(buffer, vc.BufferVariable):
            raise ValueError("read_from_buffer expects a buffer argument of the mapping function")
        reg = self.register if register is None else register
        io = self.io_index if io_index is None else io_index
        io_expr = vc._convert(io, dt.uint32)
        b = vc._builder()
        if not isinstance(reg, vc.ShaderVariable) or not reg.lvalue:
            raise ValueError("register must be a register variable (vc.new_register / op.register)")
        if reg.var_type is self.compute_type:
            b.emit(f"op.read_from_buffer({buffer.ptr}, {reg.expr}, {io_expr});")
        else:
            tmp = vc.new_register(self.compute_type)
            b.emit(f"op.read_from_buffer({buffer.ptr}, {tmp.expr}, {io_expr});")
            reg[:] = tmp

    def write_to_buffer(self, buffer, register=None, io_index=None):
        if not self._writing:
            raise ValueError("write_to_buffer is not available on a read operation")
        if not isinstance(buffer, vc.BufferVariable):
            raise ValueError("write_to_buffer expects a buffer argument of the mapping function")
        reg = self.register if register is None else register
        io = self.io_index if io_index is None else io_index
        vc._builder().emit(
            f"op.write_to_buffer({buffer.ptr}, {vc._convert(reg, self.compute_type)}, {vc._convert(io, dt.uint32)});")


@dataclasses.dataclass(frozen=True)
class Arg:
    role: str          # "output" | "input" | "kernel"
    kind: str          # "buffer" | "scalar"
    dtype: object
    slot: int          # pointer slot (buffers) or byte offset (scalars)


def declare_args(output_map, input_map, kernel_map, output_type, input_type, kernel_type,
                 conv: bool) -> List[Arg]:
    roles: List[Tuple[str, object]] = []
    if output_map is not None:
        roles += [("output", a) for a in output_map.buffer_types]
    else:
        roles.append(("output", vc.Buffer[output_type]))
    if input_map is not None:
        roles += [("input", a) for a in input_map.buffer_types]
    elif output_map is not None:
        if input_type is None:
            raise ValueError("input_type must be provided when output_map is used without input_map")
        roles.append(("input", vc.Buffer[input_type]))
    if conv:
        if kernel_map is not None:
            roles += [("kernel", a) for a in kernel_map.buffer_types]
        else:
            roles.append(("kernel", vc.Buffer[kernel_type]))
    args: List[Arg] = []
    nptr, off = 0, 0
    for role, ann in roles:
        kind = vc.annotation_kind(ann)
        t = vc.annotation_type(ann)
        if not dt.is_dtype(t):
            raise ValueError(f"annotation type {t} is not a vkdispatch_amd dtype")
        if kind == "buffer":
            if nptr >= 12:
                raise ValueError("at most 12 buffer arguments per FFT kernel")
            args.append(Arg(role, kind, t, nptr))
            nptr += 1
        else:
            align = max(1, t.item_size)
            off = -(-off // align) * align
            if off + t.item_size > 256:
                raise ValueError("Var/Const arguments exceed 256 bytes")
            args.append(Arg(role, kind, t, off))
            off += t.item_size
    return args


def _vars_for(args: Sequence[Arg]):
    out = []
    for a in args:
        if a.kind == "buffer":
            out.append(vc.BufferVariable(a.dtype, f"vkx::bufp<{a.dtype.cpp}>(prm, {a.slot})"))
        else:
            out.append(vc.scalar_arg(a.dtype, a.slot))
    return out


def _trace_body(map_fn: MappingFunction, args: Sequence[Arg], op: _Op, writing: bool,
                kernel_index: Optional[int]) -> List[str]:
    with vc.trace_scope(base_indent=3) as b:
        key = "write_op" if writing else "read_op"
        setattr(_state, key, op)
        _state.kernel_index = kernel_index
        try:
            map_fn.callback(*_vars_for(args))
        finally:
            setattr(_state, key, None)
            _state.kernel_index = None
    return b.lines


def functor_source(name: str, map_fn: MappingFunction, args: Sequence[Arg], compute_type, fft_size: int,
                   writing: bool, per_kernel: bool, kernel_num: int = 1) -> str:
    """C++ struct ``name`` whose ``apply`` runs the traced hook for one element."""
    lines = [f"struct {name} {{"]
    if per_kernel:
        lines.append("  template <int K, class Op>")
    else:
        lines.append("  template <class Op>")
    lines.append("  __device__ __forceinline__ static void apply(Op& op, const vkx::KParams& prm) {")
    count = kernel_num if per_kernel else 1
    for k in range(count):
        body = _trace_body(map_fn, args, _Op(compute_type, fft_size, writing), writing,
                           k if per_kernel else None)
        if per_kernel and count > 1:
            lines.append(f"    if constexpr (K == {k}) {{")
            lines += ["  " + ln for ln in body]
            lines.append("    }")
        else:
            lines += body
    lines.append("  }")
    lines.append("};")
    return "\n".join(lines)
