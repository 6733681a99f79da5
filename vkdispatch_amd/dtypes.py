"""Data types of the vd.fft path.

Mirrors the reference's dtype objects (vkdispatch/base/dtype.py:63-130): each
dtype is a class with ``name``, ``item_size``, ``child_type`` (the scalar of a
complex type) and a numpy mapping.  ``complex32`` is a pair of float16,
``complex64`` a pair of float32 and ``complex128`` a pair of float64, as in the
reference.  ``bfloat16`` and ``complex_bf16`` are additions of this build
(BF16-stored transforms); numpy has no bf16, so those map to uint16 storage.

Each dtype also carries ``cpp`` (the C++ type used in generated kernels).
"""
from __future__ import annotations

import numpy as np


class dtype:
    name = "dtype"
    item_size = 0
    cpp = "void"
    child_type = None
    is_complex = False
    is_float = False
    is_integer = False
    numpy = None
    # shape of one element in numpy terms (reference: true_numpy_shape)
    true_numpy_shape = ()


class _Scalar(dtype):
    pass


class int16(_Scalar):
    name = "int16"; item_size = 2; cpp = "short"; is_integer = True; numpy = np.int16


class uint16(_Scalar):
    name = "uint16"; item_size = 2; cpp = "unsigned short"; is_integer = True; numpy = np.uint16


class int32(_Scalar):
    name = "int32"; item_size = 4; cpp = "int"; is_integer = True; numpy = np.int32


class uint32(_Scalar):
    name = "uint32"; item_size = 4; cpp = "unsigned int"; is_integer = True; numpy = np.uint32


class int64(_Scalar):
    name = "int64"; item_size = 8; cpp = "long long"; is_integer = True; numpy = np.int64


class uint64(_Scalar):
    name = "uint64"; item_size = 8; cpp = "unsigned long long"; is_integer = True; numpy = np.uint64


class float16(_Scalar):
    name = "float16"; item_size = 2; cpp = "_Float16"; is_float = True; numpy = np.float16


class bfloat16(_Scalar):
    name = "bfloat16"; item_size = 2; cpp = "vkx::bf16"; is_float = True; numpy = np.uint16


class float32(_Scalar):
    name = "float32"; item_size = 4; cpp = "float"; is_float = True; numpy = np.float32


class float64(_Scalar):
    name = "float64"; item_size = 8; cpp = "double"; is_float = True; numpy = np.float64


class _Complex(dtype):
    is_complex = True


class complex32(_Complex):
    name = "complex32"; item_size = 4; child_type = float16; cpp = "vkx::cpx<_Float16>"; numpy = np.complex64
    true_numpy_shape = (2,)


class complex_bf16(_Complex):
    name = "complex_bf16"; item_size = 4; child_type = bfloat16; cpp = "vkx::cpx<vkx::bf16>"; numpy = np.complex64
    true_numpy_shape = (2,)


class complex64(_Complex):
    name = "complex64"; item_size = 8; child_type = float32; cpp = "vkx::cpx<float>"; numpy = np.complex64


class complex128(_Complex):
    name = "complex128"; item_size = 16; child_type = float64; cpp = "vkx::cpx<double>"; numpy = np.complex128


# short aliases used by hook code (vc.c64, vc.f32, ...)
c32 = complex32
c64 = complex64
c128 = complex128
cbf16 = complex_bf16
f16 = float16
bf16 = bfloat16
f32 = float32
f64 = float64
i16 = int16
u16 = uint16
i32 = int32
u32 = uint32
i64 = int64
u64 = uint64

ALL = (int16, uint16, int32, uint32, int64, uint64, float16, bfloat16, float32, float64,
       complex32, complex_bf16, complex64, complex128)

_COMPLEX_OF = {float16: complex32, bfloat16: complex_bf16, float32: complex64, float64: complex128}


def is_complex(t) -> bool:
    return isinstance(t, type) and issubclass(t, dtype) and t.is_complex


def is_dtype(t) -> bool:
    return isinstance(t, type) and issubclass(t, dtype)


def complex_from_float(t):
    """complex dtype whose components are the float dtype ``t``."""
    if t not in _COMPLEX_OF:
        raise ValueError(f"no complex dtype with {getattr(t, 'name', t)} components")
    return _COMPLEX_OF[t]


def make_floating_dtype(t):
    """Map an integer scalar dtype to float32 (the reference's helper of the same name)."""
    if t.is_float:
        return t
    return float32


def from_numpy_dtype(np_dtype):
    np_dtype = np.dtype(np_dtype)
    table = {
        np.dtype(np.int16): int16, np.dtype(np.uint16): uint16,
        np.dtype(np.int32): int32, np.dtype(np.uint32): uint32,
        np.dtype(np.int64): int64, np.dtype(np.uint64): uint64,
        np.dtype(np.float16): float16, np.dtype(np.float32): float32, np.dtype(np.float64): float64,
        np.dtype(np.complex64): complex64, np.dtype(np.complex128): complex128,
    }
    if np_dtype not in table:
        raise ValueError(f"unsupported numpy dtype {np_dtype}")
    return table[np_dtype]


def to_numpy_dtype(t):
    """numpy dtype of the host representation returned by Buffer.read()."""
    if t is complex32 or t is complex_bf16:
        return np.dtype(np.complex64)
    if t is bfloat16:
        return np.dtype(np.float32)
    return np.dtype(t.numpy)


def storage_scalar_numpy(t):
    """numpy dtype of one stored scalar component (float16 for complex32, uint16 for bf16)."""
    if t is complex32:
        return np.dtype(np.float16)
    if t is complex_bf16 or t is bfloat16:
        return np.dtype(np.uint16)
    if t.is_complex:
        return np.dtype(t.child_type.numpy)
    return np.dtype(t.numpy)


def bf16_bits_from_float32(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even) and return the raw uint16 bits."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    quiet = ((u >> 16) | 0x40).astype(np.uint16)
    return np.where(nan, quiet, rounded)


def float32_from_bf16_bits(b: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(b, dtype=np.uint16)
    return (b.astype(np.uint32) << 16).view(np.float32)


def quantize_bf16(a: np.ndarray) -> np.ndarray:
    """Values of ``a`` (float32) after a round trip through bfloat16."""
    return float32_from_bf16_bits(bf16_bits_from_float32(a))
