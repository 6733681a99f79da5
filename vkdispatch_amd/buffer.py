"""Device buffers of the vd.fft path.

``Buffer`` and ``RFFTBuffer`` keep the reference's contract
(vkdispatch/base/buffer.py:40-466): C-order device memory of 1-3 dims and at
most 2^30 elements, ``write``/``read`` through the host, and the packed
real<->Hermitian ``RFFTBuffer`` layout whose rows hold N/2+1 complex values
(equivalently 2*(N/2+1) reals, the first N of them the real signal).

Memory comes from the native runtime (hipMalloc).  ``from_cuda_array`` /
``asbuffer(tensor)`` alias external device memory (``__cuda_array_interface__``,
e.g. a PyTorch-ROCm tensor) without copying, like the reference's CAI path
(vkdispatch/base/buffer.py:309-375).
"""
from __future__ import annotations

import ctypes
import typing

import numpy as np

from . import dtypes
from . import runtime as rt
from .dtypes import complex64


class Buffer:
    """A typed device buffer of 1, 2 or 3 dimensions."""

    def __init__(self, shape, var_type=complex64, external=None) -> None:
        if isinstance(shape, (int, np.integer)):
            shape = (int(shape),)
        shape = tuple(int(s) for s in shape)
        if len(shape) > 3 or len(shape) == 0:
            raise ValueError("Buffer shape must be 1, 2, or 3 dimensions!")
        if any(s <= 0 for s in shape):
            raise ValueError("Buffer dimensions must be positive")
        if not dtypes.is_dtype(var_type):
            raise ValueError(f"var_type must be a vkdispatch_amd dtype, got {var_type}")
        self.var_type = var_type
        self.shape = shape
        size = 1
        for d in shape:
            size *= d
        if size > 2 ** 30:
            raise ValueError("Cannot allocate buffers larger than 2^30 elements!")
        self.size = size
        self.mem_size = size * var_type.item_size
        self.is_external = external is not None
        self._keepalive = None
        self.is_writable = True
        if external is None:
            self.ptr = rt.malloc(self.mem_size)
            self.owns_memory = True
        else:
            ptr, keep, writable = external
            self.ptr = int(ptr)
            self._keepalive = keep
            self.is_writable = bool(writable)
            self.owns_memory = False
        self._destroyed = False

    # -- lifetime ----------------------------------------------------------
    def destroy(self) -> None:
        if getattr(self, "_destroyed", True):
            return
        self._destroyed = True
        if self.owns_memory and self.ptr:
            try:
                rt.synchronize()
            except Exception:
                pass
            rt.free(self.ptr)
        self.ptr = 0
        self._keepalive = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __repr__(self):
        return (f"Buffer(shape={self.shape}, var_type={self.var_type.name}, mem_size={self.mem_size}, "
                f"external={self.is_external})")

    # -- host I/O ------------------------------------------------------------
    def _host_bytes(self, data) -> np.ndarray:
        if isinstance(data, (bytes, bytearray, memoryview)):
            arr = np.frombuffer(bytes(data), dtype=np.uint8)
            if arr.nbytes > self.mem_size:
                raise ValueError("Data Size must be less than buffer size")
            return arr
        arr = np.asarray(data)
        vt = self.var_type
        if vt is dtypes.complex32 and np.iscomplexobj(arr):
            packed = np.empty(arr.shape + (2,), dtype=np.float16)
            packed[..., 0] = arr.real.astype(np.float16)
            packed[..., 1] = arr.imag.astype(np.float16)
            arr = packed
        elif vt is dtypes.complex_bf16 and np.iscomplexobj(arr):
            c = arr.astype(np.complex64)
            packed = np.empty(c.shape + (2,), dtype=np.uint16)
            packed[..., 0] = dtypes.bf16_bits_from_float32(c.real)
            packed[..., 1] = dtypes.bf16_bits_from_float32(c.imag)
            arr = packed
        elif vt is dtypes.bfloat16 and arr.dtype != np.uint16:
            arr = dtypes.bf16_bits_from_float32(arr.astype(np.float32))
        arr = np.ascontiguousarray(arr)
        if arr.nbytes != self.mem_size:
            raise ValueError("Numpy buffer sizes must match!")
        return arr

    def write(self, data, index: typing.Optional[int] = None) -> None:
        """Upload host data (array of exactly ``mem_size`` bytes, or bytes no larger)."""
        if index is not None:
            if not isinstance(index, int):
                raise ValueError("Index must be an integer or None!")
            if index != 0:
                raise ValueError("Index must be a valid device index within the context!")
        if not self.is_writable:
            raise ValueError("Cannot write to a read-only buffer alias.")
        arr = self._host_bytes(data)
        rt.htod(self.ptr, arr.ctypes.data, arr.nbytes)

    def _read_raw(self, np_dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=np_dtype)
        rt.dtoh(out.ctypes.data, self.ptr, out.nbytes)
        return out

    def read(self, index: typing.Optional[int] = None) -> np.ndarray:
        """Download to a numpy array; ``read(0)`` returns the array, ``read()`` stacks
        one entry per device (a leading axis of length 1), as in the reference."""
        if index is not None:
            if not isinstance(index, int):
                raise ValueError("Index must be an integer or None!")
            if index != 0:
                raise ValueError("Index must be a valid device index within the context!")
        vt = self.var_type
        if vt is dtypes.complex32:
            raw = self._read_raw(np.float16, self.shape + (2,)).astype(np.float32)
            host = (raw[..., 0] + 1j * raw[..., 1]).astype(np.complex64)
        elif vt is dtypes.complex_bf16:
            raw = self._read_raw(np.uint16, self.shape + (2,))
            host = (dtypes.float32_from_bf16_bits(raw[..., 0]) + 1j * dtypes.float32_from_bf16_bits(raw[..., 1]))
            host = host.astype(np.complex64)
        elif vt is dtypes.bfloat16:
            host = dtypes.float32_from_bf16_bits(self._read_raw(np.uint16, self.shape))
        else:
            host = self._read_raw(vt.numpy, self.shape)
        return host if index is not None else host[None, ...]

    # -- interop ---------------------------------------------------------------
    @property
    def __cuda_array_interface__(self):
        if self.var_type in (dtypes.complex32, dtypes.complex_bf16, dtypes.bfloat16):
            raise TypeError(f"{self.var_type.name} buffers have no __cuda_array_interface__ typestr")
        return {"shape": self.shape, "typestr": np.dtype(self.var_type.numpy).str,
                "data": (self.ptr, not self.is_writable), "version": 3, "strides": None}

    def to_torch(self):
        """Zero-copy torch view of this buffer (complex32/bf16 as raw scalar pairs)."""
        import torch
        vt = self.var_type
        if vt in (dtypes.complex32,):
            t_dtype, shape = torch.float16, self.shape + (2,)
        elif vt is dtypes.complex_bf16:
            t_dtype, shape = torch.bfloat16, self.shape + (2,)
        elif vt is dtypes.bfloat16:
            t_dtype, shape = torch.bfloat16, self.shape
        else:
            t_dtype = {np.complex64: torch.complex64, np.complex128: torch.complex128, np.float32: torch.float32,
                       np.float64: torch.float64, np.float16: torch.float16, np.int32: torch.int32,
                       np.int64: torch.int64, np.int16: torch.int16}[vt.numpy]
            shape = self.shape
        return _torch_from_ptr(self.ptr, shape, t_dtype, self)


def _torch_from_ptr(ptr, shape, t_dtype, owner):
    import torch

    class _CAI:
        def __init__(self):
            typestr = {torch.float16: "<f2", torch.bfloat16: "<f2", torch.float32: "<f4",
                       torch.float64: "<f8", torch.complex64: "<c8", torch.complex128: "<c16",
                       torch.int32: "<i4", torch.int64: "<i8", torch.int16: "<i2"}[t_dtype]
            self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (ptr, False),
                                             "version": 3, "strides": None}
            self.owner = owner

    t = torch.as_tensor(_CAI(), device="cuda")
    if t_dtype is torch.bfloat16:
        t = t.view(torch.bfloat16)
    return t


def _cai_external(obj, var_type=None, require_contiguous=True, writable=None, keepalive=True):
    iface = obj.__cuda_array_interface__
    if not isinstance(iface, dict):
        raise TypeError("__cuda_array_interface__ must be a dictionary")
    for k in ("shape", "typestr", "data"):
        if k not in iface:
            raise ValueError("__cuda_array_interface__ is missing required fields (shape/typestr/data)")
    shape = tuple(int(d) for d in iface["shape"]) or (1,)
    ptr, read_only = int(iface["data"][0]), bool(iface["data"][1])
    np_dtype = np.dtype(iface["typestr"])
    inferred = dtypes.from_numpy_dtype(np_dtype)
    if var_type is None:
        var_type = inferred
    if var_type is not inferred:
        raise ValueError(f"CAI dtype ({np_dtype}) does not match requested vd dtype ({var_type.name}).")
    if require_contiguous and iface.get("strides") is not None:
        expected, stride = [], np_dtype.itemsize
        for d in reversed(shape):
            expected.insert(0, stride)
            stride *= d
        if tuple(int(x) for x in iface["strides"]) != tuple(expected):
            raise ValueError("Only contiguous C-order device arrays are supported in from_cuda_array().")
    w = (not read_only) if writable is None else bool(writable)
    if w and read_only:
        raise ValueError("Requested writable=True for a read-only device array.")
    return shape, var_type, (ptr, obj if keepalive else None, w)


def from_cuda_array(obj, var_type=None, require_contiguous=True, writable=None, keepalive=True) -> Buffer:
    """Alias a device array exposing ``__cuda_array_interface__`` (no copy)."""
    if not hasattr(obj, "__cuda_array_interface__"):
        raise TypeError("Expected an object with __cuda_array_interface__")
    rt.ensure_device()
    shape, vt, ext = _cai_external(obj, var_type, require_contiguous, writable, keepalive)
    return Buffer(shape, vt, external=ext)


def from_torch(tensor, var_type=None) -> Buffer:
    """Alias a contiguous PyTorch-ROCm device tensor (complex64/128, float32/64, ...)."""
    if not tensor.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return from_cuda_array(tensor, var_type=var_type)


def asbuffer(array) -> Buffer:
    """Cast an array-like object to a buffer (device arrays are aliased, host arrays copied)."""
    if hasattr(array, "__cuda_array_interface__"):
        return from_cuda_array(array)
    arr = np.asarray(array)
    buf = Buffer(arr.shape, dtypes.from_numpy_dtype(arr.dtype))
    buf.write(arr)
    return buf


class RFFTBuffer(Buffer):
    """Buffer of shape ``shape[:-1] + (shape[-1]//2 + 1,)`` complex values holding either a
    real signal of ``real_shape`` (rows of pitch 2*(N/2+1) reals) or its half spectrum."""

    def __init__(self, shape, fourier_type=complex64, external=None):
        if not dtypes.is_complex(fourier_type):
            raise ValueError("RFFTBuffer fourier_type must be complex32, complex64, or complex128")
        shape = tuple(int(s) for s in shape)
        super().__init__(tuple(shape[:-1]) + (shape[-1] // 2 + 1,), fourier_type, external=external)
        self.real_shape = shape
        self.fourier_shape = self.shape
        self.real_type = fourier_type.child_type

    def read_real(self, index: typing.Optional[int] = None) -> np.ndarray:
        packed_shape = tuple(self.shape[:-1]) + (self.shape[-1] * 2,)
        st = dtypes.storage_scalar_numpy(self.var_type)
        raw = self._read_raw(st, packed_shape)
        if self.var_type is dtypes.complex_bf16:
            raw = dtypes.float32_from_bf16_bits(raw)
        out = raw[..., : self.real_shape[-1]]
        return out if index is not None else out[None, ...]

    def read_fourier(self, index: typing.Optional[int] = None) -> np.ndarray:
        return self.read(index)

    def write_real(self, data, index: typing.Optional[int] = None) -> None:
        data = np.asarray(data)
        if data.shape != self.real_shape:
            raise ValueError(f"Data shape {data.shape} must match real shape {self.real_shape}!")
        if np.iscomplexobj(data):
            raise ValueError("Data dtype must be real, not complex!")
        st = dtypes.storage_scalar_numpy(self.var_type)
        true_data = np.zeros(tuple(self.shape[:-1]) + (self.shape[-1] * 2,), dtype=st)
        if self.var_type is dtypes.complex_bf16:
            true_data[..., : self.real_shape[-1]] = dtypes.bf16_bits_from_float32(data.astype(np.float32))
        else:
            true_data[..., : self.real_shape[-1]] = data
        if not self.is_writable:
            raise ValueError("Cannot write to a read-only buffer alias.")
        rt.htod(self.ptr, true_data.ctypes.data, true_data.nbytes)

    def write_fourier(self, data, index: typing.Optional[int] = None) -> None:
        data = np.asarray(data)
        if data.shape != self.fourier_shape:
            raise ValueError(f"Data shape {data.shape} must match fourier shape {self.fourier_shape}!")
        if not np.iscomplexobj(data):
            raise ValueError("Data dtype must be complex!")
        if self.var_type in (dtypes.complex32, dtypes.complex_bf16):
            self.write(data, index)
        else:
            self.write(np.ascontiguousarray(data.astype(self.var_type.numpy)), index)


def rfftbuffer_from_cuda_array(obj, real_shape, keepalive=True) -> RFFTBuffer:
    """Alias a device array (torch-ROCm tensor, any ``__cuda_array_interface__`` object)
    as an RFFTBuffer of ``real_shape`` without a copy: a complex array of the half
    spectrum shape ``real_shape[:-1] + (N//2 + 1,)``, or a real array of the packed
    layout ``real_shape[:-1] + (2 * (N//2 + 1),)`` (the reference's rows of pitch
    2(N/2+1) reals, vkdispatch/base/buffer.py:377-448)."""
    if not hasattr(obj, "__cuda_array_interface__"):
        raise TypeError("Expected an object with __cuda_array_interface__")
    rt.ensure_device()
    real_shape = tuple(int(s) for s in real_shape)
    fshape = real_shape[:-1] + (real_shape[-1] // 2 + 1,)
    shape, vt, ext = _cai_external(obj, None, True, None, keepalive)
    if dtypes.is_complex(vt):
        if shape != fshape:
            raise ValueError(f"a complex alias must have the half-spectrum shape {fshape}, got {shape}")
        ftype = vt
    else:
        if shape != real_shape[:-1] + (2 * fshape[-1],):
            raise ValueError(f"a real alias must have the packed shape {real_shape[:-1] + (2 * fshape[-1],)}, "
                             f"got {shape}")
        ftype = dtypes.complex_from_float(vt)
    return RFFTBuffer(real_shape, ftype, external=ext)


def asrfftbuffer(data, fourier_type=None) -> RFFTBuffer:
    data = np.asarray(data)
    if np.iscomplexobj(data):
        raise ValueError("Input data to asrfftbuffer must be real-valued!")
    if fourier_type is None:
        scalar = dtypes.make_floating_dtype(dtypes.from_numpy_dtype(data.dtype))
        fourier_type = dtypes.complex_from_float(scalar)
    buf = RFFTBuffer(data.shape, fourier_type=fourier_type)
    buf.write_real(data)
    return buf
