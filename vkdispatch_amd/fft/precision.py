"""Precision policy of vd.fft (the reference's vkdispatch/fft/precision.py:6-99).

Complex precisions are ordered complex32 < complex64 < complex128 (this build
adds complex_bf16 at the rank of complex32).  The default compute precision is
the highest I/O precision but at least complex64; an explicit ``compute_type``
wins.  gfx950 supports fp16 and fp64, so nothing falls back.  A requested
complex32 / complex_bf16 *compute* type runs in fp32 (the kernels do not do
half-precision arithmetic); storage stays half precision.
"""
from __future__ import annotations

from typing import Iterable, List, Optional

from .. import dtypes as dt
from ..context import get_context

_ORDER = (dt.complex32, dt.complex64, dt.complex128)
_RANK = {dt.complex32: 0, dt.complex_bf16: 0, dt.complex64: 1, dt.complex128: 2}


def is_complex_precision(t) -> bool:
    return t in _RANK


def validate_complex_precision(t, *, arg_name: str) -> None:
    if not is_complex_precision(t):
        raise ValueError(f"{arg_name} must be one of complex32, complex_bf16, complex64, or complex128 (got {t})")


def promote_complex_precisions(dtypes: Iterable):
    c = list(dtypes)
    if not c:
        raise ValueError("At least one complex dtype is required for promotion")
    for t in c:
        validate_complex_precision(t, arg_name="dtype")
    return max(c, key=lambda t: _RANK[t])


def default_compute_precision(io_precisions: Iterable):
    p = promote_complex_precisions(io_precisions)
    return dt.complex64 if _RANK[p] < _RANK[dt.complex64] else p


def supports_complex_precision(t) -> bool:
    validate_complex_precision(t, arg_name="dtype")
    st = t.child_type
    for dev in get_context().device_infos:
        if st in (dt.float16, dt.bfloat16):
            if dev.float_16_support != 1:
                return False
            if dev.storage_buffer_16_bit_access != 1 and dev.uniform_and_storage_buffer_16_bit_access != 1:
                return False
        if st is dt.float64 and dev.float_64_support != 1:
            return False
    return True


def ensure_supported_complex_precision(t, *, role: str) -> None:
    if not supports_complex_precision(t):
        raise ValueError(f"{role} precision '{t.name}' is not supported on the active device set")


def resolve_compute_precision(io_precisions: List, compute_precision: Optional[object]):
    if compute_precision is not None:
        validate_complex_precision(compute_precision, arg_name="compute_type")
        ensure_supported_complex_precision(compute_precision, role="Compute")
        return compute_precision
    for p in io_precisions:
        validate_complex_precision(p, arg_name="io_precision")
    if len(io_precisions) == 0:
        for cand in (dt.complex64, dt.complex32):
            if supports_complex_precision(cand):
                return cand
        raise ValueError("Unable to resolve a default compute precision supported by all active devices")
    target = default_compute_precision(io_precisions)
    if supports_complex_precision(target):
        return target
    for cand in (dt.complex64, dt.complex32):
        if _RANK[cand] <= _RANK[target] and supports_complex_precision(cand):
            return cand
    raise ValueError("Unable to resolve an auto compute precision supported by all active devices")


def kernel_real_type(compute_type) -> str:
    """C++ arithmetic type of the kernel for a resolved compute precision."""
    return "double" if compute_type is dt.complex128 else "float"
