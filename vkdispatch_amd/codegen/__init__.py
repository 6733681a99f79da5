"""``vc``: the hook language of vd.fft mapping functions, traced to HIP C++.

In the reference a mapping hook (``vd.map``) is Python executed at codegen time
against symbolic ``ShaderVariable``s that append GLSL/CUDA/OpenCL source
(vkdispatch/shader/map.py:29-40, vkdispatch/fft/io_manager.py:107-155).  Here
the same Python runs once against the symbolic values below and the emitted
statements become the body of a compile-time functor that the FFT kernel
template calls for every element of its load, kernel-multiply or store phase.

Supported subset (what the reference's hooks, tests and docs use):
annotations ``Buffer[T]``, ``Var[T]``, ``Const[T]``; ``+ - * / // %`` and
comparisons; ``[:] =`` / ``.real`` / ``.imag`` / ``[0]`` / ``[1]`` assignment;
``.conjugate()``, ``.to_register()``; ``mult_complex``, ``to_dtype``,
``to_complex``, ``new_register`` (and typed variants), ``ravel_index`` /
``unravel_index`` (reference semantics: ravel = flat -> tuple,
vkdispatch/codegen/functions/index_raveling.py:55-95), scalar math,
``if_block``/``else_block`` and ``comment``.  As in GLSL, ``*`` between two
complex values is component-wise; complex multiplication is ``mult_complex``.
"""
from __future__ import annotations

import contextlib
import math
import threading
import typing

import numpy as np

from .. import dtypes as _dt
from ..dtypes import (c32, c64, c128, cbf16, f16, bf16, f32, f64, i16, u16, i32, u32, i64, u64,  # noqa: F401
                     complex32, complex64, complex128, complex_bf16, float16, bfloat16, float32, float64,
                     int16, uint16, int32, uint32, int64, uint64)

_T = typing.TypeVar("_T")


# ---------------------------------------------------------------------------
# argument annotations
# ---------------------------------------------------------------------------
class Buffer(typing.Generic[_T]):
    """``vc.Buffer[T]``: a device buffer argument of element type T."""


class Var(typing.Generic[_T]):
    """``vc.Var[T]``: a scalar argument passed at call time (push constant in the reference)."""


class Const(typing.Generic[_T]):
    """``vc.Const[T]``: a scalar argument passed at call time (uniform in the reference)."""


Variable = Var
Constant = Const


def annotation_kind(ann):
    origin = getattr(ann, "__origin__", None)
    if origin is Buffer:
        return "buffer"
    if origin in (Var, Const):
        return "scalar"
    raise ValueError(f"unsupported mapping-function annotation {ann!r}")


def annotation_type(ann):
    args = getattr(ann, "__args__", None)
    if not args or len(args) != 1:
        raise ValueError(f"annotation without exactly one type argument: {ann}")
    return args[0]


# ---------------------------------------------------------------------------
# builder
# ---------------------------------------------------------------------------
class _Builder:
    def __init__(self, base_indent: int = 2):
        self.lines: typing.List[str] = []
        self.indent = base_indent
        self.counter = 0

    def emit(self, line: str) -> None:
        self.lines.append("  " * self.indent + line)

    def tmp(self, prefix: str = "t") -> str:
        self.counter += 1
        return f"{prefix}{self.counter}"


_tls = threading.local()


def _builder() -> _Builder:
    stack = getattr(_tls, "stack", None)
    if not stack:
        raise RuntimeError("vc operations are only valid while a vd.fft mapping function is being traced")
    return stack[-1]


@contextlib.contextmanager
def trace_scope(base_indent: int = 2):
    stack = getattr(_tls, "stack", None)
    if stack is None:
        stack = _tls.stack = []
    b = _Builder(base_indent)
    stack.append(b)
    try:
        yield b
    finally:
        stack.pop()


def is_tracing() -> bool:
    return bool(getattr(_tls, "stack", None))


# ---------------------------------------------------------------------------
# values
# ---------------------------------------------------------------------------
_BOOL = "bool"


def _scalar_cpp(t) -> str:
    return t.cpp


def _rank_float(t) -> int:
    return {float16: 1, bfloat16: 1, float32: 2, float64: 3}.get(t, 0)


def _promote_scalar(a, b):
    if a is _BOOL:
        a = int32
    if b is _BOOL:
        b = int32
    if a.is_float or b.is_float:
        fa, fb = _rank_float(a), _rank_float(b)
        return a if fa >= fb else b
    if a.item_size != b.item_size:
        return a if a.item_size > b.item_size else b
    if a.name.startswith("u") or b.name.startswith("u"):
        return a if a.name.startswith("u") else b
    return a


def _promote_complex(a, b):
    order = {complex32: 1, complex_bf16: 1, complex64: 2, complex128: 3}
    return a if order[a] >= order[b] else b


def _fmt_float(v: float, t) -> str:
    if math.isinf(v) or math.isnan(v):
        raise ValueError("non-finite literal in mapping function")
    r = repr(float(v))
    if "e" not in r and "." not in r:
        r += ".0"
    if t is float64:
        return r
    if t is float32:
        return r + "f"
    return f"(({t.cpp}){r}f)"


class ShaderVariable:
    """A symbolic value: a C++ expression plus its vd dtype (or 'bool')."""

    def __init__(self, var_type, expr: str, lvalue: bool = False):
        self.var_type = var_type
        self.expr = expr
        self.lvalue = lvalue

    def __repr__(self):
        return f"ShaderVariable({getattr(self.var_type, 'name', self.var_type)}: {self.expr})"

    def __str__(self):
        return self.expr

    # -- conversions ----------------------------------------------------------
    def __bool__(self):
        raise TypeError("a traced ShaderVariable has no Python truth value; use vc.if_block(...)")

    def _require_lvalue(self):
        if not self.lvalue:
            raise ValueError(f"cannot assign to non-register expression {self.expr}")

    def to_register(self, var_name: str = None) -> "ShaderVariable":
        b = _builder()
        name = var_name if var_name else b.tmp("r")
        if self.var_type is _BOOL:
            b.emit(f"bool {name} = {self.expr};")
        else:
            b.emit(f"{self.var_type.cpp} {name} = {self.expr};")
        return ShaderVariable(self.var_type, name, lvalue=True)

    def conjugate(self) -> "ShaderVariable":
        if not _dt.is_complex(self.var_type):
            return self
        return ShaderVariable(self.var_type, f"vkx::conj({self.expr})")

    # -- components ---------------------------------------------------------------
    @property
    def real(self):
        if not _dt.is_complex(self.var_type):
            return self
        return ShaderVariable(self.var_type.child_type, f"({self.expr}).x", lvalue=self.lvalue)

    @real.setter
    def real(self, value):
        self._require_lvalue()
        _builder().emit(f"{self.expr}.x = {_convert(value, self.var_type.child_type)};")

    @property
    def imag(self):
        if not _dt.is_complex(self.var_type):
            return ShaderVariable(self.var_type, _literal(0, self.var_type).expr)
        return ShaderVariable(self.var_type.child_type, f"({self.expr}).y", lvalue=self.lvalue)

    @imag.setter
    def imag(self, value):
        self._require_lvalue()
        _builder().emit(f"{self.expr}.y = {_convert(value, self.var_type.child_type)};")

    x = real
    y = imag

    def __getitem__(self, key):
        if isinstance(key, slice):
            return self
        if _dt.is_complex(self.var_type):
            if isinstance(key, (int, np.integer)):
                comp = "x" if int(key) == 0 else "y"
                return ShaderVariable(self.var_type.child_type, f"({self.expr}).{comp}", lvalue=self.lvalue)
            idx = _as_var(key)
            return ShaderVariable(self.var_type.child_type, f"(&({self.expr}).x)[{idx.expr}]", lvalue=self.lvalue)
        raise TypeError(f"cannot index a {self.var_type.name} value")

    def __setitem__(self, key, value):
        self._require_lvalue()
        b = _builder()
        if isinstance(key, slice):
            b.emit(f"{self.expr} = {_convert(value, self.var_type)};")
            return
        target = self[key]
        b.emit(f"{target.expr} = {_convert(value, target.var_type)};")

    # -- arithmetic -----------------------------------------------------------------
    def _bin(self, other, op, reverse=False):
        a, c = (_as_var(other, self.var_type), self) if reverse else (self, _as_var(other, self.var_type))
        ta, tc = a.var_type, c.var_type
        if _dt.is_complex(ta) or _dt.is_complex(tc):
            if _dt.is_complex(ta) and _dt.is_complex(tc):
                t = _promote_complex(ta, tc)
                return ShaderVariable(t, f"({_convert(a, t)} {op} {_convert(c, t)})")
            ct = ta if _dt.is_complex(ta) else tc
            st = ct.child_type
            if op in ("*", "/") and _dt.is_complex(ta):
                return ShaderVariable(ct, f"({a.expr} {op} {_convert(c, st)})")
            if op == "*":
                return ShaderVariable(ct, f"({_convert(a, st)} {op} {c.expr})")
            # GLSL vec2 (op) scalar: component-wise with the scalar splatted
            sa = a.expr if _dt.is_complex(ta) else f"vkx::mk<{st.cpp}>({_convert(a, st)}, {_convert(a, st)})"
            sc_ = c.expr if _dt.is_complex(tc) else f"vkx::mk<{st.cpp}>({_convert(c, st)}, {_convert(c, st)})"
            return ShaderVariable(ct, f"({sa} {op} {sc_})")
        t = _promote_scalar(ta, tc)
        if op == "//":
            if t.is_float:
                return ShaderVariable(t, f"floor({_convert(a, t)} / {_convert(c, t)})")
            op = "/"
        if op == "%" and t.is_float:
            return ShaderVariable(t, f"fmod({_convert(a, t)}, {_convert(c, t)})")
        return ShaderVariable(t, f"({_convert(a, t)} {op} {_convert(c, t)})")

    def __add__(self, o): return self._bin(o, "+")
    def __radd__(self, o): return self._bin(o, "+", True)
    def __sub__(self, o): return self._bin(o, "-")
    def __rsub__(self, o): return self._bin(o, "-", True)
    def __mul__(self, o): return self._bin(o, "*")
    def __rmul__(self, o): return self._bin(o, "*", True)
    def __truediv__(self, o): return self._bin(o, "/")
    def __rtruediv__(self, o): return self._bin(o, "/", True)
    def __floordiv__(self, o): return self._bin(o, "//")
    def __rfloordiv__(self, o): return self._bin(o, "//", True)
    def __mod__(self, o): return self._bin(o, "%")
    def __rmod__(self, o): return self._bin(o, "%", True)

    def __lshift__(self, o): return self._bin(o, "<<")
    def __rshift__(self, o): return self._bin(o, ">>")
    def __and__(self, o): return self._bin(o, "&")
    def __or__(self, o): return self._bin(o, "|")
    def __xor__(self, o): return self._bin(o, "^")

    def _inplace(self, other, op):
        self._require_lvalue()
        res = self._bin(other, op)
        _builder().emit(f"{self.expr} = {_convert(res, self.var_type)};")
        return self

    def __iadd__(self, o): return self._inplace(o, "+")
    def __isub__(self, o): return self._inplace(o, "-")
    def __imul__(self, o): return self._inplace(o, "*")
    def __itruediv__(self, o): return self._inplace(o, "/")

    def __neg__(self):
        return ShaderVariable(self.var_type, f"(-{self.expr})")

    def __pos__(self):
        return self

    def __abs__(self):
        return abs_(self)

    def _cmp(self, other, op):
        o = _as_var(other, self.var_type)
        return ShaderVariable(_BOOL, f"({self.expr} {op} {o.expr})")

    def __lt__(self, o): return self._cmp(o, "<")
    def __le__(self, o): return self._cmp(o, "<=")
    def __gt__(self, o): return self._cmp(o, ">")
    def __ge__(self, o): return self._cmp(o, ">=")

    def __eq__(self, o):  # noqa: D105  (symbolic comparison, like the reference)
        return self._cmp(o, "==")

    def __ne__(self, o):
        return self._cmp(o, "!=")

    __hash__ = object.__hash__


class VectorVariable:
    """A small tuple of integer values (the reference's uvec2/uvec3 from ravel_index)."""

    def __init__(self, comps: typing.List[ShaderVariable]):
        self.comps = list(comps)
        self.var_type = uint32

    def __len__(self):
        return len(self.comps)

    def __getitem__(self, i):
        return self.comps[i]

    def __iter__(self):
        return iter(self.comps)

    @property
    def x(self): return self.comps[0]

    @property
    def y(self): return self.comps[1]

    @property
    def z(self): return self.comps[2]

    def to_register(self, var_name: str = None) -> "VectorVariable":
        return VectorVariable([c.to_register(None if var_name is None else f"{var_name}_{i}")
                               for i, c in enumerate(self.comps)])


class BufferVariable:
    """A buffer argument of a traced hook (``buf[i]`` loads and stores)."""

    def __init__(self, var_type, ptr: str):
        self.var_type = var_type
        self.ptr = ptr

    def __getitem__(self, idx) -> ShaderVariable:
        i = _as_var(idx, uint32)
        return ShaderVariable(self.var_type, f"{self.ptr}[{i.expr}]", lvalue=True)

    def __setitem__(self, idx, value) -> None:
        i = _as_var(idx, uint32)
        _builder().emit(f"{self.ptr}[{i.expr}] = {_convert(value, self.var_type)};")

    def __repr__(self):
        return f"BufferVariable({self.var_type.name}: {self.ptr})"


def _literal(value, like=None) -> ShaderVariable:
    if isinstance(value, (bool, np.bool_)):
        return ShaderVariable(_BOOL, "true" if value else "false")
    if isinstance(value, (int, np.integer)):
        if like is not None and like is not _BOOL and _dt.is_complex(like):
            like = like.child_type
        if like is not None and like is not _BOOL and like.is_float:
            return ShaderVariable(like, _fmt_float(float(value), like))
        v = int(value)
        if like is not None and like is not _BOOL and like.is_integer:
            suffix = "u" if like.name.startswith("u") else ""
            if like.item_size == 8:
                suffix += "ll"
            return ShaderVariable(like, f"{v}{suffix}" if v >= 0 else f"({v}{suffix})")
        if v < 0:
            return ShaderVariable(int32, f"({v})")
        return ShaderVariable(int32 if v < 2 ** 31 else int64, str(v))
    if isinstance(value, (float, np.floating)):
        t = float32
        if like is not None and like is not _BOOL:
            if _dt.is_complex(like):
                t = like.child_type
            elif like.is_float:
                t = like
        if isinstance(value, np.float64) and like is None:
            t = float64
        return ShaderVariable(t, _fmt_float(float(value), t))
    if isinstance(value, (complex, np.complexfloating)):
        t = complex64
        if like is not None and like is not _BOOL and _dt.is_complex(like):
            t = like
        st = t.child_type
        return ShaderVariable(t, f"vkx::mk<{st.cpp}>({_fmt_float(value.real, st)}, {_fmt_float(value.imag, st)})")
    raise TypeError(f"cannot use {value!r} in a traced expression")


def _as_var(value, like=None) -> ShaderVariable:
    if isinstance(value, ShaderVariable):
        return value
    return _literal(value, like)


def _convert(value, target) -> str:
    """C++ expression of ``value`` converted to dtype ``target``."""
    v = _as_var(value, target)
    src = v.var_type
    if src is target:
        return v.expr
    if target is _BOOL:
        return f"(bool)({v.expr})"
    if src is _BOOL:
        return f"(({target.cpp})({v.expr}))"
    if _dt.is_complex(target):
        st = target.child_type.cpp
        if _dt.is_complex(src):
            return f"vkx::ccvt<{st}>({v.expr})"
        return f"vkx::rcvt<{st}>({v.expr})"
    if _dt.is_complex(src):
        raise ValueError(f"cannot convert complex value {v.expr} to {target.name}; take .real or .imag")
    return f"vkx::sc<{target.cpp}>({v.expr})"


# ---------------------------------------------------------------------------
# functions
# ---------------------------------------------------------------------------
def new_register(var_type, *args, var_name: str = None) -> ShaderVariable:
    b = _builder()
    name = var_name if var_name else b.tmp("r")
    if len(args) == 0:
        init = _convert(0, var_type)
    elif len(args) == 1:
        init = _convert(args[0], var_type)
    elif len(args) == 2 and _dt.is_complex(var_type):
        st = var_type.child_type
        init = f"vkx::mk<{st.cpp}>({_convert(args[0], st)}, {_convert(args[1], st)})"
    else:
        raise ValueError("new_register takes at most one initial value (two for complex)")
    b.emit(f"{var_type.cpp} {name} = {init};")
    return ShaderVariable(var_type, name, lvalue=True)


def new(var_type, *args, var_name: str = None):
    return new_register(var_type, *args, var_name=var_name)


def new_float_register(*args, var_name=None):
    return new_register(float32, *args, var_name=var_name)


def new_int_register(*args, var_name=None):
    return new_register(int32, *args, var_name=var_name)


def new_uint_register(*args, var_name=None):
    return new_register(uint32, *args, var_name=var_name)


def new_complex_register(*args, var_name=None):
    return new_register(complex64, *args, var_name=var_name)


def mult_complex(a, b) -> ShaderVariable:
    a, b = _as_var(a), _as_var(b)
    if not (_dt.is_complex(a.var_type) and _dt.is_complex(b.var_type)):
        raise ValueError("mult_complex takes two complex values")
    t = _promote_complex(a.var_type, b.var_type)
    return ShaderVariable(t, f"vkx::cmul({_convert(a, t)}, {_convert(b, t)})")


def to_dtype(var_type, value) -> ShaderVariable:
    return ShaderVariable(var_type, _convert(value, var_type))


def to_complex(real, imag=None) -> ShaderVariable:
    r = _as_var(real)
    if _dt.is_complex(r.var_type):
        return r
    st = r.var_type if (r.var_type is not _BOOL and r.var_type.is_float) else float32
    ct = _dt.complex_from_float(st)
    im = "0" if imag is None else _convert(imag, st)
    if imag is None:
        im = _convert(0, st)
    return ShaderVariable(ct, f"vkx::mk<{st.cpp}>({_convert(r, st)}, {im})")


def to_float(value):
    return to_dtype(float32, value)


def to_int(value):
    return to_dtype(int32, value)


def to_uint(value):
    return to_dtype(uint32, value)


def _sanitize(x):
    if isinstance(x, VectorVariable):
        return list(x.comps)
    if isinstance(x, (tuple, list)):
        return list(x)
    return [x]


def ravel_index(index, shape) -> VectorVariable:
    """Flat index -> (x, y[, z]) for a 2-D/3-D shape (reference naming)."""
    shp = _sanitize(shape)
    idx = _sanitize(index)
    if len(idx) != 1:
        raise ValueError(f"Index must be a single integer value, not '{index}'!")
    i = _as_var(idx[0], uint32)
    if len(shp) == 2:
        return VectorVariable([i // shp[1], i % shp[1]])
    if len(shp) == 3:
        return VectorVariable([i // (shp[1] * shp[2]), (i // shp[2]) % shp[1], i % shp[2]])
    raise ValueError(f"Shape must have 2 or 3 elements, not '{shape}'!")


def unravel_index(index, shape):
    """(x, y[, z]) -> flat index for a 2-D/3-D shape (reference naming)."""
    shp = _sanitize(shape)
    idx = _sanitize(index)
    if len(idx) > len(shp):
        raise ValueError(f"Index ({index}) must have the same number of elements as shape ({shp})!")
    if len(idx) == 1:
        return index
    if len(idx) == 2:
        return idx[0] * shp[1] + idx[1]
    if len(idx) == 3:
        return idx[0] * (shp[1] * shp[2]) + idx[1] * shp[2] + idx[2]
    raise ValueError("unravel_index supports shapes with 2 or 3 elements")


def _unary(name, x):
    """Scalar math in the argument's float precision (float32 for integers and halves)."""
    v = _as_var(x, float32)
    t = v.var_type if (v.var_type is not _BOOL and v.var_type.is_float) else float32
    if t in (float16, bfloat16):
        t = float32
    fn = name + "f" if t is float32 else name
    return ShaderVariable(t, f"{fn}({_convert(v, t)})")


def sin(x): return _unary("sin", x)
def cos(x): return _unary("cos", x)
def tan(x): return _unary("tan", x)
def exp(x): return _unary("exp", x)
def log(x): return _unary("log", x)
def sqrt(x): return _unary("sqrt", x)
def floor(x): return _unary("floor", x)
def ceil(x): return _unary("ceil", x)


def abs_(x):
    v = _as_var(x)
    if _dt.is_complex(v.var_type):
        raise ValueError("abs of a complex value: use vc.sqrt(z.real*z.real + z.imag*z.imag)")
    if v.var_type.is_float:
        return ShaderVariable(v.var_type, f"fabs({v.expr})")
    return ShaderVariable(v.var_type, f"abs({v.expr})")


abs = abs_  # noqa: A001


def min(a, b):  # noqa: A001
    a, b = _as_var(a), _as_var(b)
    t = _promote_scalar(a.var_type, b.var_type)
    return ShaderVariable(t, f"(({_convert(a, t)}) < ({_convert(b, t)}) ? ({_convert(a, t)}) : ({_convert(b, t)}))")


def max(a, b):  # noqa: A001
    a, b = _as_var(a), _as_var(b)
    t = _promote_scalar(a.var_type, b.var_type)
    return ShaderVariable(t, f"(({_convert(a, t)}) > ({_convert(b, t)}) ? ({_convert(a, t)}) : ({_convert(b, t)}))")


def atan2(y, x):
    y, x = _as_var(y, float32), _as_var(x, float32)
    t = _promote_scalar(y.var_type, x.var_type) if y.var_type.is_float else float32
    return ShaderVariable(t, f"atan2({_convert(y, t)}, {_convert(x, t)})")


def complex_from_euler_angle(angle) -> ShaderVariable:
    a = _as_var(angle, float32)
    t = a.var_type if a.var_type.is_float else float32
    ct = _dt.complex_from_float(t)
    return ShaderVariable(ct, f"vkx::mk<{t.cpp}>(cos({_convert(a, t)}), sin({_convert(a, t)}))")


def all(*conds):  # noqa: A001
    return ShaderVariable(_BOOL, "(" + " && ".join(_as_var(c).expr for c in conds) + ")")


def any(*conds):  # noqa: A001
    return ShaderVariable(_BOOL, "(" + " || ".join(_as_var(c).expr for c in conds) + ")")


def comment(text: str, preceding_new_line: bool = True) -> None:
    b = _builder()
    for line in str(text).splitlines() or [""]:
        b.emit(f"// {line}")


# -- control flow ---------------------------------------------------------------
def if_statement(cond) -> None:
    b = _builder()
    b.emit(f"if ({_as_var(cond).expr}) {{")
    b.indent += 1


def else_if_statement(cond) -> None:
    b = _builder()
    b.indent -= 1
    b.emit(f"}} else if ({_as_var(cond).expr}) {{")
    b.indent += 1


def else_statement() -> None:
    b = _builder()
    b.indent -= 1
    b.emit("} else {")
    b.indent += 1


def end() -> None:
    b = _builder()
    b.indent -= 1
    b.emit("}")


@contextlib.contextmanager
def if_block(cond):
    if_statement(cond)
    try:
        yield
    finally:
        end()


@contextlib.contextmanager
def else_block():
    b = _builder()
    # reopen the block closed by the preceding if_block
    if not b.lines or b.lines[-1].strip() != "}":
        raise ValueError("else_block must directly follow an if_block")
    b.lines.pop()
    b.emit("} else {")
    b.indent += 1
    try:
        yield
    finally:
        end()


@contextlib.contextmanager
def scope_block():
    b = _builder()
    b.emit("{")
    b.indent += 1
    try:
        yield
    finally:
        end()


def scalar_arg(var_type, offset: int) -> ShaderVariable:
    """Value of a Var/Const argument stored at byte ``offset`` of the kernel constants."""
    return ShaderVariable(var_type, f"vkx::cval<{var_type.cpp}>(prm, {offset})")
