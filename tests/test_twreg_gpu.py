"""Twiddles in registers (the TWREG plan option) against the table form of the same plan.

With TWREG a kernel loads each thread's twiddles once, together with its data,
instead of reading the twiddle table after every LDS exchange.  The values and the
operations are the same; the compiler may contract the twiddle products into FMAs
differently, so the two forms agree to FP32 rounding (a few ulps), not bit for bit.
Each form must meet the FP32 bound against the FP64 oracle.  The plans are the
ones the planner picks for each shape (tuned table included), with only TWREG
switched.
"""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd
from tests._gpu_common import assert_close

pytestmark = pytest.mark.gpu

CASES = [
    ("fft", (8, 4096), 1),         # contiguous rows, three radix-16 passes
    ("fft", (4, 1024, 33), 1),     # strided column transform
    ("fft", (16, 256, 16), 1),
    ("conv", (4, 1024, 33), 1),    # fused FFT -> kernel -> IFFT along a strided axis
    ("conv", (4, 1024, 513), 1),   # the tuned differentiable-convolution column plan
    ("rfft", (8, 1024), 1),        # half-length real rows
    ("irfft", (8, 1024), 1),
]


def _kernel(kind, shape, axis, twreg):
    from vkdispatch_amd.fft import autotune as AT
    from vkdispatch_amd.fft import planner as P
    geom, tgeom = AT._tuning_geometry(kind, shape, axis)
    plan = P.make_plan(tgeom, "float", kind=kind)
    return AT._build(kind, shape, axis, geom, "float", (plan.radices, plan.TT, plan.B, twreg)), plan


def _run(kind, shape, axis, twreg, data):
    """Launch the (kind, shape, axis) kernel on a fresh copy of ``data``; returns the result."""
    k, plan = _kernel(kind, shape, axis, twreg)
    assert k.plan.TWREG is twreg and k.plan.radices == plan.radices
    if kind in ("rfft", "irfft"):
        buf = vd.RFFTBuffer(tuple(shape))
        if kind == "rfft":
            buf.write_real(data)
        else:
            buf.write_fourier(data)
        k.launch([buf])
        return buf.read_fourier(0) if kind == "rfft" else buf.read_real(0)
    buf = vd.asbuffer(data[0] if kind == "conv" else data)
    vals = [buf]
    if kind == "conv":
        vals.append(vd.asbuffer(data[1]))
    k.launch(vals)
    return buf.read(0)


@pytest.mark.parametrize("kind,shape,axis", CASES)
def test_twreg_matches_table_form(kind, shape, axis):
    rng = np.random.default_rng(sum(shape) + len(kind))
    if kind == "rfft":
        data = rng.standard_normal(shape).astype(np.float32)
    elif kind == "irfft":
        half = shape[:-1] + (shape[-1] // 2 + 1,)
        data = (rng.standard_normal(half) + 1j * rng.standard_normal(half)).astype(np.complex64)
    elif kind == "conv":
        x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        kk = (rng.standard_normal(shape[axis:]) + 1j * rng.standard_normal(shape[axis:])).astype(np.complex64)
        data = (x, kk)
    else:
        data = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    off = _run(kind, shape, axis, False, data)
    on = _run(kind, shape, axis, True, data)
    assert off.dtype == on.dtype and off.shape == on.shape
    assert O.rel_l2(on, off) <= 1e-6, f"{kind} {shape}: register and table twiddles disagree"
    if kind == "fft":
        ref = O.fft_axis(data, axis)
    elif kind == "rfft":
        ref = np.fft.rfft(data.astype(np.float64), axis=-1)
    elif kind == "irfft":
        ref = np.fft.irfft(data.astype(np.complex128), n=shape[-1], axis=-1)
    else:  # conv: default kernel map (conj(K)), K broadcast over the outer axis
        ref = O.convolve_axis(data[0], data[1], axis, kernel_inner_only=True)
    for got in (off, on):
        assert_close(got, ref)
