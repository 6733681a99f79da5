"""Procedural FFT kernel construction (the reference's ``vd.fft.fft_context``,
vkdispatch/fft/context.py:17-204, used as in docs/tutorials/reductions_and_fft.rst:225-247).

The reference iterates ``global_reads_iterator`` / ``global_writes_iterator``
once per register and emits straight-line code for each.  In this build the
kernel applies one functor to every element, so each iterator yields exactly
one symbolic operation: the loop body the user writes is traced once and
becomes the load (or store) functor.  ``ctx.execute(inverse=...)`` selects the
direction; ``ctx.get_callable()`` builds the kernel.
"""
from __future__ import annotations

import contextlib
from typing import List, Optional

from .. import codegen as vc
from .. import dtypes as dt
from ..graph import dispatch
from . import hooks
from .kernels import CompiledKernel, cfg_struct, translation_unit
from .planner import AxisGeometry, make_plan
from .precision import kernel_real_type


class _Registers:
    def __init__(self, ctx):
        self.ctx = ctx


class FFTContext:
    def __init__(self, buffer_shape, axis=None, compute_type=dt.complex64, name=None):
        self.geom = AxisGeometry.of(tuple(buffer_shape), axis)
        self.buffer_shape = tuple(buffer_shape)
        self.axis = axis
        self.compute_type = dt.complex128 if compute_type is dt.complex128 else dt.complex64
        self.plan = make_plan(self.geom, kernel_real_type(compute_type))
        self.name = name or f"fft_context_{self.buffer_shape}_{axis}"
        self.registers = _Registers(self)
        self.args: List[hooks.Arg] = []
        self._vars = []
        self._in_lines: Optional[List[str]] = None
        self._out_lines: Optional[List[str]] = None
        self._inverse = None
        self._normalize = False
        self._kernel: Optional[CompiledKernel] = None

    def declare_shader_args(self, types):
        nptr, off = 0, 0
        for ann in types:
            kind = vc.annotation_kind(ann)
            t = vc.annotation_type(ann)
            if kind == "buffer":
                self.args.append(hooks.Arg("output", "buffer", t, nptr))
                self._vars.append(vc.BufferVariable(t, f"vkx::bufp<{t.cpp}>(prm, {nptr})"))
                nptr += 1
            else:
                align = max(1, t.item_size)
                off = -(-off // align) * align
                self.args.append(hooks.Arg("output", "scalar", t, off))
                self._vars.append(vc.scalar_arg(t, off))
                off += t.item_size
        return list(self._vars)

    def execute(self, inverse: bool = False):
        self._inverse = bool(inverse)

    def normalize(self):
        self._normalize = True

    def get_callable(self) -> CompiledKernel:
        if self._kernel is not None:
            return self._kernel
        if self._in_lines is None or self._out_lines is None or self._inverse is None:
            raise ValueError("fft_context needs a read phase, execute() and a write phase")
        N = self.geom.N
        inv = self._inverse
        cfg = cfg_struct(self.plan, in_os=N * self.geom.S, out_os=N * self.geom.S, k_os=N * self.geom.S,
                         sig=(0, N), r2c=False, inverse=inv, normalize=inv and self._normalize, scale=1.0 / N)
        fin = ["struct HookIn {", "  template <class Op>",
               "  __device__ __forceinline__ static void apply(Op& op, const vkx::KParams& prm) {",
               *self._in_lines, "  }", "};"]
        fout = ["struct HookOut {", "  template <int K, class Op>",
                "  __device__ __forceinline__ static void apply(Op& op, const vkx::KParams& prm) {",
                *self._out_lines, "  }", "};"]
        src = translation_unit(cfg, ["\n".join(fin), "\n".join(fout)], "vkx::fft_body<Cfg, HookIn, HookOut>(prm);",
                               self.plan.BLOCK, "vkx_fft_context")
        self._kernel = _ContextKernel(src, self.plan, self.args, list(self.plan.radices), None, name=self.name)
        return self._kernel


class _ContextKernel(CompiledKernel):
    def __call__(self, *args, graph=None):
        dispatch(self, args, graph)


@contextlib.contextmanager
def fft_context(buffer_shape, axis=None, compute_type=dt.complex64, name=None):
    ctx = FFTContext(buffer_shape, axis=axis, compute_type=compute_type, name=name)
    yield ctx


def _trace_phase(ctx: FFTContext, writing: bool):
    op = hooks._Op(ctx.compute_type, ctx.geom.N, writing)
    key = "write_op" if writing else "read_op"
    with vc.trace_scope(base_indent=2) as b:
        setattr(hooks._state, key, op)
        try:
            yield op
        finally:
            setattr(hooks._state, key, None)
    if writing:
        ctx._out_lines = b.lines
    else:
        ctx._in_lines = b.lines


def global_reads_iterator(registers, r2c: bool = False, inverse: bool = None, format_transposed: bool = False,
                          inner_only: bool = False, signal_range=None):
    """Yields the (single) symbolic read operation of the load phase."""
    if r2c or format_transposed or inner_only or signal_range is not None:
        raise ValueError("fft_context supports plain C2C reads; use vd.fft.fft(r2c=..., input_signal_range=...)")
    yield from _trace_phase(registers.ctx, writing=False)


def global_writes_iterator(registers, r2c: bool = False, inverse: bool = None):
    """Yields the (single) symbolic write operation of the store phase."""
    if r2c:
        raise ValueError("fft_context supports plain C2C writes; use vd.fft.fft(r2c=True)")
    yield from _trace_phase(registers.ctx, writing=True)
