"""Mixed-precision FFTs: 16-bit storage, radix-16 stages on MFMA.

``fft3_mixed(out, inp)`` computes the forward (or inverse) 3-D C2C FFT of a
256^3-type volume stored as complex_bf16 or complex32 (fp16 pairs) into a
complex64 buffer.  Each axis pass is a 256-point transform done as two
radix-16 complex GEMMs on the matrix cores with hi/lo-split bf16 operands and
FP32 accumulation (``fft256_mfma_tw_body`` in csrc/device/vkx_mfma.h).  Between
passes the volume is stored as fp16 scaled by an exact power of two (1/16 per
axis, undone on the last pass), so the intermediate bytes stay at 16 bits per
real while the values keep 11 significant bits and stay inside fp16 range for
inputs of the N(0,1) scale.  Twiddles and DFT-16 constants come from a table
generated in FP64 on the GPU.

Each pass reads its transforms as contiguous 1 KB rows and writes the result
frequency-major, so three passes return the volume to natural order.
Bytes per 256^3 forward transform: 64 MiB (bf16 in) + 2 x 128 MiB (fp16
intermediates, read and written) + 64 MiB + 128 MiB (fp32 out) = 448 MiB,
against 768 MiB for the all-FP32 three-pass transform.
"""
from __future__ import annotations

import os
from functools import lru_cache
from typing import Optional, Tuple

from .. import dtypes as dt
from .. import runtime as rt
from ..buffer import Buffer
from ..graph import CommandGraph, dispatch
from .hooks import Arg
from .kernels import CompiledKernel, kernel_symbol, translation_unit
from .planner import KernelPlan
from .shader_factories import register_plan_cache

_SCALAR_CPP = {dt.complex_bf16: "vkx::bf16", dt.complex32: "_Float16", dt.complex64: "float"}


def make_tw_pass(nb: int, in_type, out_type, inverse: bool = False, out_scale: float = 1.0,
                 T: Optional[int] = None, block: int = 256) -> CompiledKernel:
    """256-point MFMA pass over ``nb`` contiguous transforms ([nb][256], slot 1, in_type)
    written frequency-major ([256][nb], slot 0, out_type)."""
    if in_type not in (dt.complex_bf16, dt.complex32) or out_type not in _SCALAR_CPP:
        raise ValueError("MFMA passes take complex_bf16 / complex32 input and complex32 / complex64 output")
    if T is None:
        T = 128 // out_type.item_size  # 128-byte output runs per frequency
    if block % 64 or T % (block // 64) or nb % T:
        raise ValueError("tile does not divide the batch / waves")
    plan = KernelPlan(real="float", N=256, TT=block, B=1, S=1, OUTER=nb, radices=(16, 16), COLFAST=False,
                      WAVE_LOCAL=True, PAD_SHIFT=5, NP=0, LDS_ELEMS=0, NWG=nb // T, NWG_INNER=1)
    cfg = "\n".join([
        "struct Cfg {",
        f"  static constexpr int T = {T}, BLOCK = {block}, DIR = {1 if inverse else -1};",
        f"  static constexpr vkx::u32 NB = {nb}u;",
        f"  static constexpr double OUT_SCALE = {out_scale!r};",
        "};",
    ])
    body = f"vkx::fft256_mfma_tw_body<Cfg, {_SCALAR_CPP[in_type]}, {_SCALAR_CPP[out_type]}>(prm);"
    tag = ("inv" if inverse else "fwd") + f"_{in_type.name}_{out_type.name}"
    src = translation_unit(cfg, [], body, block, kernel_symbol("mfma256t", plan, tag))
    src = src.replace('#include "vkx_fft.h"', '#include "vkx_mfma.h"')
    args = [Arg("output", "buffer", out_type, 0), Arg("input", "buffer", in_type, 1)]
    return CompiledKernel(src, plan, args, None, None, name=f"mfma256t_{nb}", roots_n=256)


def _mfma_tu(cfg: str, body: str, block: int, symbol: str, waves_per_simd: Optional[int]) -> str:
    occ = f" __attribute__((amdgpu_waves_per_eu({waves_per_simd}, {waves_per_simd})))" if waves_per_simd else ""
    return "\n".join(['#include "vkx_mfma.h"', "", cfg, "",
                      f'extern "C" __global__ void __launch_bounds__({block}){occ} {symbol}(vkx::KParams prm) {{',
                      f"  {body}", "}", ""])


def make_cube_pass(which: int, in_type, inverse: bool = False, out_scale: float = 1.0, prefetch: int = 4,
                   waves_per_simd: Optional[int] = 2, tiles: Optional[int] = None, run: int = 4,
                   half_tiles: bool = False, nt_input: bool = False) -> CompiledKernel:
    """One pass of the two-pass 256^3 transform (``fft3_p1_body`` / ``fft3_p2_body``):
    pass 1 reads the 16-bit volume (slot 1) and writes the fp16 intermediate (slot 0),
    pass 2 reads the intermediate (slot 1) and writes complex64 in natural order (slot 0).
    1024 workgroups of 256 threads each."""
    if which == 1:
        if in_type not in (dt.complex_bf16, dt.complex32):
            raise ValueError("pass 1 takes complex_bf16 / complex32 input")
        out_type = dt.complex32
        body = f"vkx::fft3_p1_body<Cfg, {_SCALAR_CPP[in_type]}>(prm);"
    elif which == 2:
        in_type, out_type = dt.complex32, dt.complex64
        body = "vkx::fft3_p2_body<Cfg>(prm);"
    else:
        raise ValueError("pass is 1 or 2")
    if run not in (4, 8):
        raise ValueError("run (k2 per intermediate run) is 4 or 8")
    block, nwg = (256, 1024) if which == 1 else (64 * run, 16 * (256 // run))
    if tiles is None:
        tiles = 2 if which == 1 else 1  # tiles per plane exchange (LDS: 2 workgroups per CU)
    plan = KernelPlan(real="float", N=256, TT=block, B=1, S=1, OUTER=256 * 256, radices=(16, 16, 16), COLFAST=False,
                      WAVE_LOCAL=True, PAD_SHIFT=5, NP=0, LDS_ELEMS=0, NWG=nwg, NWG_INNER=1)
    cfg = "\n".join([
        "struct Cfg {",
        f"  static constexpr int BLOCK = {block}, DIR = {1 if inverse else -1}, PF = {int(prefetch)}, NT = {int(tiles)},",
        f"                       P = {int(run)};",
        f"  static constexpr bool VH = {'true' if half_tiles else 'false'};",
        *([f"  static constexpr int NTIN = 1;"] if nt_input else []),
        f"  static constexpr double OUT_SCALE = {out_scale!r};",
        "};",
    ])
    tag = f"p{which}_" + ("inv" if inverse else "fwd") + f"_{in_type.name}" + ("" if run == 4 else f"_r{run}") + \
        ("_h" if half_tiles else "")
    src = _mfma_tu(cfg, body, block, kernel_symbol("mfma3d", plan, tag), waves_per_simd)
    args = [Arg("output", "buffer", out_type, 0), Arg("input", "buffer", in_type, 1)]
    return CompiledKernel(src, plan, args, None, None, name=f"mfma3d_p{which}", roots_n=256)


class Mixed3D:
    """Plan for the 3-D mixed-precision transform of one 256 x 256 x 256 volume, complex64
    result in natural order.  ``passes=2`` (default): the two-pass form, 16-bit input ->
    fp16 intermediate -> complex64 (three radix-16 MFMA stages per pass); ``run`` = 4 or 8
    consecutive k2 per intermediate run (8: one 8-wave pass-2 workgroup per CU).
    ``passes=3``: three transposed-write 256-point passes with 16-bit storage in between."""

    def __init__(self, shape, in_type=dt.complex_bf16, inverse: bool = False, normalize: bool = True,
                 block: int = 256, T16: Optional[int] = None, passes: int = 2, run: int = 4,
                 half_tiles_p2: bool = False, nt_input: Optional[bool] = None):
        shape = tuple(int(s) for s in shape)
        if len(shape) != 3 or any(s != 256 for s in shape):
            raise ValueError("fft3_mixed handles 256 x 256 x 256 volumes")
        if in_type not in (dt.complex_bf16, dt.complex32):
            raise ValueError("input storage must be complex_bf16 or complex32")
        if passes not in (2, 3):
            raise ValueError("passes is 2 or 3")
        self.shape = shape
        self.in_type = in_type
        self.inverse = inverse
        self.npasses = passes
        self.half_tiles_p2 = bool(half_tiles_p2)
        norm = 1.0 / float(256 ** 3) if (inverse and normalize) else 1.0
        mid = dt.complex32
        if passes == 2:
            # 1/64 after pass 1 (three radix-16 stages) keeps N(0,1)-scale data near unit size in fp16.
            # Pass 1 holds its tile results as fp16 pairs (one extra 11-bit rounding, fewer
            # registers and a two-product cross stage).  Pass 2 keeps them in FP32 by default
            # (rel-L2 2.9e-4 vs FP64 of the quantized input); half_tiles_p2=True makes them
            # fp16 too (3.6e-4): 1.2 us faster for pass 2 alone, no faster for the whole
            # transform (interleaved medians 79.2 vs 78.3 us).
            if nt_input is None:
                nt_input = os.environ.get("VKX_MIXED_NTIN", "0") == "1"
            self.passes = [make_cube_pass(1, in_type, inverse, 1.0 / 64.0, run=run, half_tiles=True,
                                          nt_input=bool(nt_input)),
                           make_cube_pass(2, mid, inverse, 64.0 * norm, run=run, half_tiles=half_tiles_p2)]
        else:
            nb = 256 * 256
            s_mid = 1.0 / 16.0
            self.passes = [
                make_tw_pass(nb, in_type, mid, inverse, s_mid, T16, block),
                make_tw_pass(nb, mid, mid, inverse, s_mid, T16, block),
                make_tw_pass(nb, mid, dt.complex64, inverse, 256.0 * norm, None, block),
            ]
        self._scratch: Optional[Tuple[Buffer, Buffer]] = None

    def kernels(self):
        return list(self.passes)

    def mfma_count(self) -> int:
        """v_mfma_f32_16x16x32 instructions (16384 flops each) per transform.  A 256-point
        tile costs 4 products for the stage reading 16-bit storage (exact operand: hi and
        lo parts of the DFT matrix, real and imaginary output) plus 6 for the stage on
        split FP32 data; a cross stage tile costs 6 (FP32 tile results) or 4 (fp16)."""
        tile = 4 + 6
        if self.npasses == 3:
            return 3 * 65536 * tile
        cubes = 4096  # 16 x 16 x 16 blocks per pass, 16 tiles + 16 cross tiles each
        return cubes * (16 * tile + 16 * 4) + cubes * (16 * tile + 16 * (4 if self.half_tiles_p2 else 6))

    def scratch(self) -> Tuple[Buffer, ...]:
        if self._scratch is None:
            n = 1 if self.npasses == 2 else 2
            self._scratch = tuple(Buffer(self.shape, dt.complex32) for _ in range(n))
        return self._scratch

    def __call__(self, out: Buffer, inp: Buffer, graph: Optional[CommandGraph] = None) -> None:
        if out.var_type is not dt.complex64 or tuple(out.shape) != self.shape:
            raise ValueError("output must be a complex64 buffer of the volume's shape")
        if inp.var_type is not self.in_type or tuple(inp.shape) != self.shape:
            raise ValueError(f"input must be a {self.in_type.name} buffer of the volume's shape")
        if self.npasses == 2:
            (m,) = self.scratch()
            dispatch(self.passes[0], [m, inp], graph)    # -> M[c1][k2/4][i0][a1][k2%4]
            dispatch(self.passes[1], [out, m], graph)    # -> [k0][k1][k2]
            return
        s1, s2 = self.scratch()
        dispatch(self.passes[0], [s1, inp], graph)   # [i0][i1][i2] -> [k2][i0][i1]
        dispatch(self.passes[1], [s2, s1], graph)    # -> [k1][k2][i0]
        dispatch(self.passes[2], [out, s2], graph)   # -> [k0][k1][k2]


@lru_cache(maxsize=8)
def _plan(shape, in_type, inverse, normalize) -> Mixed3D:
    return Mixed3D(shape, in_type, inverse, normalize)


register_plan_cache(_plan.cache_clear)


def fft3_mixed(out: Buffer, inp: Buffer, inverse: bool = False, normalize: bool = True,
               graph: Optional[CommandGraph] = None) -> None:
    """3-D C2C FFT of a 16-bit (complex_bf16 / complex32) volume into a complex64 buffer,
    radix-16 stages on MFMA with split-precision operands."""
    _plan(tuple(inp.shape), inp.var_type, bool(inverse), bool(normalize))(out, inp, graph)
