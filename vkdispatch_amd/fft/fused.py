"""2-D real FFT-convolution (convolve2DR) execution paths and their profiling.

``try_convolve2DR`` runs the pair-packed kernels (``csrc/device/vkx_conv2d.h``)
when the call is eligible: a batch of square real images (256, 512 or 1024 on a
side) in complex64 RFFTBuffers, one kernel spectrum broadcast over the batch,
the default kernel map.  It returns False otherwise, and the caller falls back
to the reference's three-dispatch composition (rfft -> convolve(axis ndim-2) ->
irfft, vkdispatch/fft/functions.py:614-646).  Both paths compute
irfft2(rfft2(x) . conj(K)) per image (vkdispatch/fft/shader_factories.py:136-145).

The pair-packed form is ``Conv2DPhases``: three kernels in stream order through a
scratch image per pair.  VKX_FUSED_CONV=0 forces the three-dispatch composition.
"""
from __future__ import annotations

import os
import threading
from functools import lru_cache
from typing import Dict, List, Optional, Sequence, Tuple

from .. import dtypes as dt
from .. import runtime as rt
from ..buffer import Buffer
from ..graph import dispatch
from .hooks import Arg
from .kernels import CompiledKernel, check_no_spill, kernel_symbol, source_symbol
from .planner import KernelPlan
from .shader_factories import make_convolution_shader, make_fft_shader, register_plan_cache

_FUSED_RADICES = {256: (4, 4, 4, 4), 512: (8, 8, 8), 1024: (16, 16, 4)}
# two waves (128 threads) per transform: every pass keeps 8 points per lane (radices <= 8)
_FUSED_RADICES_TT128 = {1024: (8, 8, 4, 4)}


def _fused_enabled() -> bool:
    return os.environ.get("VKX_FUSED_CONV", "1") != "0"


COL_TILE = 16      # columns per column task (128-byte tile rows)
KPREP_BLOCK = 64   # kernel-prep task: 64 x 64 elements of Kt


def _conv_cfg(batch: int, H: int, W: int, normalize: bool, waves: int, rows: int, layout: int = 0,
              cols_per_wave: int = 1, pairs_per_task: int = 1, extra: Optional[Dict[str, int]] = None,
              tt: int = 64) -> Tuple[str, KernelPlan]:
    """``struct Cfg`` of the phase kernels and the matching KernelPlan
    (radices, twiddle tables).  ``tt``: threads per transform (64: one wave; 128: two
    waves, column-major layout only)."""
    if H != W or W not in _FUSED_RADICES:
        raise ValueError("the pair-packed 2-D convolution takes square 256, 512 or 1024 images")
    units = waves * 64 // tt
    if tt not in (64, 128) or (tt == 128 and (layout != 1 or W not in _FUSED_RADICES_TT128)):
        raise ValueError("two-wave transforms take 1024-point images and the column-major layout")
    if waves not in (1, 2, 4, 8, 16) or units < 1 or rows % units or H % rows:
        raise ValueError("tiling does not divide the image")
    radices = _FUSED_RADICES_TT128[W] if tt == 128 else _FUSED_RADICES[W]
    NP = W + W // 32 + 1  # odd stride between LDS column regions: conflict-free tile transposes
    pitch = 2 * (W // 2 + 1)
    npairs = (batch + 1) // 2
    scale = 1.0 / (H * W) if normalize else 1.0
    plan = KernelPlan(real="float", N=W, TT=64, B=waves, S=1, OUTER=batch, radices=tuple(radices), COLFAST=False,
                      WAVE_LOCAL=True, PAD_SHIFT=5, NP=NP, LDS_ELEMS=max(COL_TILE * NP, waves * NP), NWG=1,
                      NWG_INNER=1)
    pl = lambda rs: "vkx::PL<" + ", ".join(f"vkx::P<{r}, {ns}>" for r, ns in rs) + ">"  # noqa: E731
    cfg = "\n".join([
        "struct Cfg {",
        f"  static constexpr int H = {H}, W = {W}, NW = {waves}, RA = {rows}, NP = {NP}, LAYOUT = {layout}, "
        f"CB = {cols_per_wave}, PP = {pairs_per_task}, TP = {rows + 1};",
        f"  static constexpr vkx::u32 PITCH = {pitch}u, KW = {W // 2 + 1}u, NIMG = {batch}u, NPAIRS = {npairs}u;",
        f"  static constexpr vkx::u64 IMG = {H * pitch}ull;",
        f"  static constexpr double SCALE = {scale!r};",
        *([f"  static constexpr int TT = {tt};"] if tt != 64 else []),
        *[f"  static constexpr int {k} = {v};" for k, v in (extra or {}).items()],
        f"  using fwd = {pl(plan.fwd)};",
        f"  using inv = {pl(plan.inv)};",
        "};",
    ])
    return cfg, plan


def _check_no_spill(kernels: Sequence[CompiledKernel]) -> None:
    """Refused at plan time, before any launch (the same check every kernel gets when
    it is loaded)."""
    for k in kernels:
        check_no_spill(k)


def _conv_tu(cfg: str, symbol: str, block: int, waves_per_simd: Optional[int], body: str) -> str:
    occ = f" __attribute__((amdgpu_waves_per_eu({waves_per_simd}, {waves_per_simd})))" if waves_per_simd else ""
    return "\n".join(['#include "vkx_conv2d.h"', "", cfg, "",
                      f'extern "C" __global__ void __launch_bounds__({block}){occ} {symbol}(vkx::KParams prm) {{',
                      f"  {body}", "}", ""])


class Conv2DPhases:
    """The same computation as three ordinary kernels in stream order: row FFTs of the
    image pairs (plus the kernel prep), column FFT * Kt * IFFT, row IFFTs; each pair's
    spectrum goes through a column-major scratch image (64 MiB for 64 images of
    512 x 512) and Kt is one 512 x 512 complex image."""

    def __init__(self, batch: int, H: int, W: int, normalize: bool = True, waves: int = 8, rows: int = 16,
                 col_waves: int = 8, layout: str = "column", cols_per_wave: int = 1, pairs_per_task: int = 4,
                 waves_per_simd: Tuple[Optional[int], Optional[int], Optional[int]] = (None, None, None),
                 irow_waves: Optional[int] = None, tt=64, irow_rows: Optional[int] = None):
        self.batch, self.H, self.W = int(batch), int(H), int(W)
        self.npairs = (self.batch + 1) // 2
        self.layout = layout
        lay = {"tile": 0, "column": 1}[layout]
        self.kernels_ = []
        na = H // rows
        pp = pairs_per_task if lay == 1 else 1
        tts = tuple(tt) if isinstance(tt, (tuple, list)) else (tt, tt, tt)  # threads per transform per phase
        if len(tts) != 3:
            raise ValueError("tt is one value or one per phase")
        nb = W // COL_TILE if lay == 0 else W // (col_waves * 64 // tts[1] * cols_per_wave)
        npg = (self.npairs + pp - 1) // pp
        nk = (H // KPREP_BLOCK) * (W // KPREP_BLOCK)
        for phase, tag in enumerate(("rows", "cols", "irows")):
            nw = col_waves if phase == 1 else (irow_waves or waves) if phase == 2 else waves
            nrows = (rows, 16, irow_rows or rows)[phase]
            if H % nrows:
                raise ValueError("tiling does not divide the image")
            cfg, plan = _conv_cfg(batch, H, W, normalize, nw, nrows, lay, cols_per_wave, pp, tt=tts[phase])
            nwg = (self.npairs * na + nk, npg * nb, self.npairs * (H // nrows))[phase]
            plan = KernelPlan(**{**plan.__dict__, "NWG": nwg})
            symbol = kernel_symbol("conv2d", plan, f"pairs_{tag}" + ("_cm" if lay else "") +
                                   ("_t128" if tts[phase] == 128 else ""))
            src = _conv_tu(cfg, symbol, 64 * nw, waves_per_simd[phase],
                           f"vkx::conv2d_phase_body<Cfg, {phase}>(prm);")
            args = [Arg("output", "buffer", dt.complex64, 0), Arg("kernel", "buffer", dt.complex64, 1),
                    Arg("scratch", "buffer", dt.complex64, 2), Arg("unused", "buffer", dt.complex64, 3),
                    Arg("unused", "buffer", dt.complex64, 4), Arg("kernel_t", "buffer", dt.complex64, 5)]
            radices = list(plan.radices)
            self.kernels_.append(CompiledKernel(src, plan, args, radices, radices[::-1],
                                                name=f"conv2d_{tag}_{batch}x{H}x{W}"))
        self._dev: Optional[Tuple[Buffer, Buffer]] = None
        self._checked = False
        self._lock = threading.Lock()

    def kernels(self) -> List[CompiledKernel]:
        return list(self.kernels_)

    def resources(self) -> Tuple[Buffer, Buffer]:
        if self._dev is None:
            with self._lock:
                if self._dev is None:
                    rt.require_not_capturing("the first call of a 2-D convolution plan")
                    self._dev = (Buffer((self.npairs * self.H, self.W), dt.complex64),
                                 Buffer((self.H, self.W), dt.complex64))
        return self._dev

    def scratch(self) -> Buffer:
        return self.resources()[0]

    def values(self, buffer, kernel) -> list:
        if not self._checked:
            _check_no_spill(self.kernels())
            self._checked = True
        scr, kt = self.resources()
        return [buffer, kernel, scr, scr, scr, kt]

    def __call__(self, buffer, kernel, graph=None) -> None:
        vals = self.values(buffer, kernel)
        for k in self.kernels_:
            dispatch(k, vals, graph)


# phase-kernel tilings: for 512 the row kernels run one row per wave (16 waves, 64 VGPRs,
# no spills), so two workgroups per CU hold 8 waves per SIMD while they read HBM.  For
# 1024 the forward rows and the columns run each transform on two waves (8 points per
# lane: 88 / 123 VGPRs instead of 153 / 176; columns 144.8 -> 73.9 us, rows 71.3 ->
# 66.1 us on 32 images), the inverse rows stay on one wave (59.0 vs 68.6 us).
_PHASE_TILING = {512: dict(waves=16, irow_waves=8),
                 1024: dict(waves=16, rows=8, tt=(128, 128, 64), col_waves=4, cols_per_wave=1, pairs_per_task=4,
                            irow_waves=8, irow_rows=16)}


@lru_cache(maxsize=32)
def fused_plan(batch: int, H: int, W: int, normalize: bool = True):
    return Conv2DPhases(batch, H, W, normalize, **_PHASE_TILING.get(W, {}))


register_plan_cache(fused_plan.cache_clear)


def plan_for_shape(real_shape, kernel_map=None, transposed_kernel=False, kernel_inner_only=False, normalize=True,
                   compute_type=None):
    """The pair-packed plan for a convolve2DR over complex64 RFFTBuffers of ``real_shape``
    with a kernel spectrum of one image, or None when the call takes the reference's
    three-dispatch composition."""
    if not _fused_enabled() or kernel_map is not None or transposed_kernel:
        return None
    if compute_type is not None and compute_type is not dt.complex64:
        return None
    rs = tuple(real_shape)
    if len(rs) not in (2, 3):
        return None
    H, W = rs[-2], rs[-1]
    batch = rs[0] if len(rs) == 3 else 1
    if H != W or W not in _FUSED_RADICES:
        return None
    if batch > 1 and not kernel_inner_only:
        return None  # the reference then reads one kernel spectrum per image
    return fused_plan(batch, H, W, bool(normalize))


def fused_eligible(buffer, kernel, kernel_map=None, buffer_shape=None, transposed_kernel=False,
                   kernel_inner_only=False, normalize=True, compute_type=None, **_):
    """The pair-packed plan (Conv2DPhases) for this convolve2DR call, or None when the
    call needs the three-dispatch composition."""
    if getattr(buffer, "var_type", None) is not dt.complex64 or getattr(kernel, "var_type", None) is not dt.complex64:
        return None
    rs = tuple(getattr(buffer, "real_shape", ()))
    if len(rs) not in (2, 3) or (buffer_shape is not None and tuple(buffer_shape) != tuple(buffer.shape)):
        return None
    H, W = rs[-2], rs[-1]
    wh = W // 2 + 1
    if kernel.size < H * wh:
        return None
    if len(rs) == 3 and rs[0] > 1 and tuple(kernel.shape[-2:]) != (H, wh):
        return None
    return plan_for_shape(rs, kernel_map, transposed_kernel, kernel_inner_only, normalize, compute_type)


def try_convolve2DR(buffer, kernel, graph=None, print_shader=False, **kw) -> bool:
    plan = fused_eligible(buffer, kernel, **kw)
    if plan is None:
        return False
    if print_shader:
        for k in plan.kernels():
            print(k.source)
    plan(buffer, kernel, graph)
    return True


def composed_kernels(real_shape, kernel_inner_only=True, compute_type=dt.complex64):
    """The three kernels convolve2DR dispatches for ``real_shape`` (..., H, W)."""
    real_shape = tuple(real_shape)
    half = real_shape[:-1] + (real_shape[-1] // 2 + 1,)
    ct = dt.complex64
    r = make_fft_shader(real_shape, None, r2c=True, output_type=ct, input_type=ct, compute_type=compute_type)
    c = make_convolution_shader(half, None, 1, len(half) - 2, kernel_inner_only=kernel_inner_only, normalize=True,
                                input_type=ct, output_type=ct, kernel_type=ct, compute_type=compute_type)
    i = make_fft_shader(real_shape, None, inverse=True, normalize_inverse=True, r2c=True, output_type=ct,
                        input_type=ct, compute_type=compute_type)
    return r, c, i


_STANDARD_FUSED = ((64, 512, True), (8, 512, True), (16, 512, True), (6, 512, True), (4, 512, True),
                   (1, 512, True), (17, 512, True), (5, 256, True), (4, 256, True), (3, 256, True),
                   (3, 256, False), (40, 256, True), (2, 1024, True), (3, 1024, True), (32, 1024, True))


def standard_kernels() -> List:
    ks = list(composed_kernels((64, 512, 512))) + list(composed_kernels((16, 512, 512)))
    for batch, n, norm in _STANDARD_FUSED:
        ks += fused_plan(batch, n, n, norm).kernels()
    return ks


def profile_convolve2DR(buffer, kernel, reps: int = 20) -> List[Dict]:
    """Average duration (HIP events on the launch stream) and algorithmic bytes of each
    kernel convolve2DR runs for ``buffer`` (broadcast kernel)."""
    shape = buffer.real_shape
    batch = 1
    for s in shape[:-2]:
        batch *= s
    H, W = shape[-2], shape[-1]
    Wh = W // 2 + 1
    real_b = batch * H * W * 4
    spec_b = batch * H * Wh * 8
    kb = H * Wh * 8
    plan = fused_eligible(buffer, kernel, kernel_inner_only=True)
    if plan is not None:
        vals = plan.values(buffer, kernel)
        ks = plan.kernels()
        if len(ks) == 1:
            return [{"name": source_symbol(ks[0].source), "ms": _event_ms(ks[0], vals, reps),
                     "bytes": 2 * real_b + kb}]
        cplx_b = (batch + 1) // 2 * H * W * 8
        kt_b = H * W * 8
        return [{"name": source_symbol(ks[0].source), "ms": _event_ms(ks[0], vals, reps),
                 "bytes": real_b + cplx_b + kb + kt_b},
                {"name": source_symbol(ks[1].source), "ms": _event_ms(ks[1], vals, reps),
                 "bytes": 2 * cplx_b + kt_b},
                {"name": source_symbol(ks[2].source), "ms": _event_ms(ks[2], vals, reps),
                 "bytes": cplx_b + real_b}]
    r, c, i = composed_kernels(shape)
    out = []
    for k, vals, nbytes in ((r, [buffer], real_b + spec_b), (c, [buffer, kernel], 2 * spec_b + kb),
                            (i, [buffer], spec_b + real_b)):
        out.append({"name": source_symbol(k.source), "ms": _event_ms(k, vals, reps), "bytes": nbytes})
    return out


def _event_ms(kernel, values, reps: int) -> float:
    e0, e1 = rt.Event(), rt.Event()
    kernel.launch(values)
    e0.record()
    for _ in range(reps):
        kernel.launch(values)
    e1.record()
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps
