"""Differentiable FFT-convolution with a hand-written adjoint.

``fft_convolve2d(x, k)`` computes, per image, the vd.fft.convolve2DR result

    y = irfft2(rfft2(x) * conj(rfft2(k)))     i.e.  y[n] = sum_m x[n + m] k[m]

(circular cross-correlation, the reference's default convolution kernel map,
vkdispatch/fft/shader_factories.py:136-145) for real ``x`` of shape
(..., H, W) and a real kernel ``k`` of shape (H, W) broadcast over the batch.

Backward (adjoint of a cross-correlation):
    dx = irfft2(rfft2(g) * K)                              (convolution with k)
    dk = sum_b irfft2(rfft2(x_b) * conj(rfft2(g_b)))       (correlation of x with g)
with K = rfft2(k) and g = dL/dy.

Everything runs on the gfx950 kernels of this package, on PyTorch's current
stream, with PyTorch providing only the tensors (storage) and autograd:
  forward  : rows r2c (out of place) -> column FFT * conj(K) -> column IFFT (one
             fused kernel) -> rows c2r (out of place)
  backward : G = rfft2(g); dk: rows r2c(x) -> column FFT * conj(G_b) -> column
             IFFT -> batch sum -> rows c2r;  dx: column IFFT of G * K (a mapping
             hook multiplies by K on load) -> rows c2r.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache

from .. import codegen as vc
from .. import dtypes as dt
from .. import runtime as rt
from ..buffer import Buffer
from ..shader_map import map as vd_map
from . import functions as F
from .shader_factories import make_convolution_shader, make_fft_shader, make_real_rows_shader, register_plan_cache


def _torch():
    import torch
    return torch


def _ctype(real_dtype):
    torch = _torch()
    if real_dtype == torch.float32:
        return dt.complex64, torch.complex64
    if real_dtype == torch.float64:
        return dt.complex128, torch.complex128
    raise TypeError(f"fft_convolve2d supports float32 and float64 tensors, got {real_dtype}")


def _wrap(t, var_type, shape=None) -> Buffer:
    """Zero-copy Buffer over a contiguous CUDA tensor."""
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("fft_convolve2d needs contiguous device tensors")
    return Buffer(tuple(shape if shape is not None else t.shape), var_type, external=(t.data_ptr(), t, True))


@lru_cache(maxsize=None)
def _mult_kernel_map(img_elems: int, ctype):
    """Input map: load the spectrum element and multiply by K (not conjugated), K
    broadcast over the batch (index = io_index mod H*(W/2+1))."""
    from .hooks import read_op

    def mult_k(spec, kern):
        op = read_op()
        op.read_from_buffer(spec)
        kv = vc.to_dtype(op.register.var_type, kern[op.io_index % img_elems])
        op.register[:] = vc.mult_complex(op.register, kv)
    return vd_map(mult_k, input_types=[vc.Buffer[ctype], vc.Buffer[ctype]])


class _Plans:
    def __init__(self, B, H, W, ctype):
        self.B, self.H, self.W = B, H, W
        Wh = W // 2 + 1
        self.Wh = Wh
        self.ctype = ctype
        c = ctype
        self.rows_fwd = make_real_rows_shader((B, H, W), False, True, c, c, True, W, Wh)
        self.rows_inv = make_real_rows_shader((B, H, W), True, True, c, c, True, W, Wh)
        self.rows_fwd1 = make_real_rows_shader((1, H, W), False, True, c, c, True, W, Wh)
        self.rows_inv1 = make_real_rows_shader((1, H, W), True, True, c, c, True, W, Wh)
        self.col_fft1 = make_fft_shader((1, H, Wh), 1, input_type=c, output_type=c, compute_type=c)
        self.col_fft = make_fft_shader((B, H, Wh), 1, input_type=c, output_type=c, compute_type=c)
        self.conv_bcast = make_convolution_shader((B, H, Wh), None, 1, 1, kernel_inner_only=True, input_type=c,
                                                  output_type=c, kernel_type=c, compute_type=c)
        self.conv_full = make_convolution_shader((B, H, Wh), None, 1, 1, input_type=c, output_type=c,
                                                 kernel_type=c, compute_type=c)
        self.kmap = _mult_kernel_map(H * Wh, c)

    def kernels(self):
        return [self.rows_fwd, self.rows_inv, self.rows_fwd1, self.rows_inv1, self.col_fft1, self.col_fft,
                self.conv_bcast, self.conv_full]


@lru_cache(maxsize=64)
def plans_for(B, H, W, ctype) -> _Plans:
    if W % 2:
        raise ValueError("fft_convolve2d needs an even last dimension")
    return _Plans(B, H, W, ctype)


register_plan_cache(plans_for.cache_clear)


def kernel_spectrum(k, plans: _Plans):
    """K = rfft2(k) for a (H, W) real tensor -> (H, W/2+1) complex tensor."""
    torch = _torch()
    tct = torch.complex128 if plans.ctype is dt.complex128 else torch.complex64
    K = torch.empty((plans.H, plans.Wh), dtype=tct, device=k.device)
    Kb = _wrap(K, plans.ctype)
    plans.rows_fwd1.launch([Kb, _wrap(k, plans.ctype.child_type)])
    plans.col_fft1.launch([Kb])
    return K


def _forward(x, k, plans: _Plans):
    torch = _torch()
    tct = torch.complex128 if plans.ctype is dt.complex128 else torch.complex64
    K = kernel_spectrum(k, plans)
    ws = torch.empty((plans.B, plans.H, plans.Wh), dtype=tct, device=x.device)
    wsb = _wrap(ws, plans.ctype)
    plans.rows_fwd.launch([wsb, _wrap(x, plans.ctype.child_type)])
    plans.conv_bcast.launch([wsb, _wrap(K, plans.ctype)])
    y = torch.empty_like(x)
    plans.rows_inv.launch([_wrap(y, plans.ctype.child_type), wsb])
    return y, K


def _backward(x, K, g, plans: _Plans, need_dx: bool, need_dk: bool):
    torch = _torch()
    tct = torch.complex128 if plans.ctype is dt.complex128 else torch.complex64
    real_t = plans.ctype.child_type
    dbl = 1 if plans.ctype is dt.complex128 else 0
    Gs = torch.empty((plans.B, plans.H, plans.Wh), dtype=tct, device=g.device)
    Gb = _wrap(Gs, plans.ctype)
    plans.rows_fwd.launch([Gb, _wrap(g, real_t)])
    plans.col_fft.launch([Gb])                           # Gs = rfft2(g)
    dx = dk = None
    if need_dk:
        wsx = torch.empty_like(Gs)
        wb = _wrap(wsx, plans.ctype)
        plans.rows_fwd.launch([wb, _wrap(x, real_t)])
        plans.conv_full.launch([wb, Gb])                 # col-IFFT(rfft2(x_b) * conj(G_b))
        P = torch.empty((plans.H, plans.Wh), dtype=tct, device=g.device)
        rt.check(rt.lib().vkx_batch_sum(P.data_ptr(), wsx.data_ptr(), plans.B, 2 * plans.H * plans.Wh, dbl,
                                        rt.current_stream()), "vkx_batch_sum")
        dk = torch.empty((plans.H, plans.W), dtype=g.dtype, device=g.device)
        plans.rows_inv1.launch([_wrap(dk, real_t), _wrap(P, plans.ctype)])
    if need_dx:
        # column IFFT of G * K (K applied by the input map), in place, then rows c2r
        F.fft(Gb, Gb, _wrap(K, plans.ctype), axis=1, inverse=True, input_map=plans.kmap)
        dx = torch.empty_like(g)
        plans.rows_inv.launch([_wrap(dx, real_t), Gb])
    return dx, dk


def _run(fn, *args):
    torch = _torch()
    with rt.stream_scope(torch.cuda.current_stream()):
        return fn(*args)


def _autograd_function():
    torch = _torch()

    class FFTConvolve2DFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, k):
            shape = x.shape
            H, W = shape[-2], shape[-1]
            if tuple(k.shape) != (H, W):
                raise ValueError(f"kernel shape {tuple(k.shape)} must equal the image shape {(H, W)}")
            if x.dtype != k.dtype:
                raise TypeError("x and k must have the same dtype")
            ctype, _ = _ctype(x.dtype)
            x3 = x.contiguous().reshape(-1, H, W)
            kc = k.contiguous()
            plans = plans_for(x3.shape[0], H, W, ctype)
            y, K = _run(_forward, x3, kc, plans)
            ctx.save_for_backward(x3, K)
            ctx.plans = plans
            ctx.in_shape = shape
            return y.reshape(shape)

        @staticmethod
        def backward(ctx, gy):
            x3, K = ctx.saved_tensors
            plans = ctx.plans
            g3 = gy.contiguous().reshape(x3.shape)
            dx, dk = _run(_backward, x3, K, g3, plans, ctx.needs_input_grad[0], ctx.needs_input_grad[1])
            if dx is not None:
                dx = dx.reshape(ctx.in_shape)
            return dx, dk

    return FFTConvolve2DFunction


_FN = None


def fft_convolve2d(x, k):
    """Differentiable ``irfft2(rfft2(x) * conj(rfft2(k)))`` over the last two axes of ``x``
    (float32 or float64 CUDA tensors, even last dimension, ``k`` of shape x.shape[-2:])."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(x, k)


def FFTConvolve2D(H: int, W: int, dtype=None, init_std: float = 0.01):
    """``torch.nn.Module`` holding a learnable (H, W) kernel applied with ``fft_convolve2d``."""
    torch = _torch()

    class _M(torch.nn.Module):
        def __init__(self):
            super().__init__()
            self.kernel = torch.nn.Parameter(init_std * torch.randn(H, W, dtype=dtype or torch.float32,
                                                                    device="cuda"))

        def forward(self, x):
            return fft_convolve2d(x, self.kernel)

    return _M()
