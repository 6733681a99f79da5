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

from functools import lru_cache
from typing import Optional, Tuple

from .. import dtypes as dt
from .. import runtime as rt
from ..buffer import Buffer
from ..graph import CommandGraph, dispatch
from .hooks import Arg
from .kernels import CompiledKernel, kernel_symbol, translation_unit
from .planner import KernelPlan

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


class Mixed3D:
    """Plan for the 3-D mixed-precision transform of one 256 x 256 x 256 volume: three
    transposed-write MFMA passes (each along the then-contiguous axis), 16-bit storage
    in between, complex64 result in natural order."""

    def __init__(self, shape, in_type=dt.complex_bf16, inverse: bool = False, normalize: bool = True,
                 block: int = 256, T16: Optional[int] = None):
        shape = tuple(int(s) for s in shape)
        if len(shape) != 3 or any(s != 256 for s in shape):
            raise ValueError("fft3_mixed handles 256 x 256 x 256 volumes")
        if in_type not in (dt.complex_bf16, dt.complex32):
            raise ValueError("input storage must be complex_bf16 or complex32")
        self.shape = shape
        self.in_type = in_type
        self.inverse = inverse
        nb = 256 * 256
        s_mid = 1.0 / 16.0
        s_last = 256.0
        if inverse and normalize:
            s_last = 256.0 / float(256 ** 3)
        mid = dt.complex32
        self.passes = [
            make_tw_pass(nb, in_type, mid, inverse, s_mid, T16, block),
            make_tw_pass(nb, mid, mid, inverse, s_mid, T16, block),
            make_tw_pass(nb, mid, dt.complex64, inverse, s_last, None, block),
        ]
        self._scratch: Optional[Tuple[Buffer, Buffer]] = None

    def kernels(self):
        return list(self.passes)

    def scratch(self) -> Tuple[Buffer, Buffer]:
        if self._scratch is None:
            self._scratch = (Buffer(self.shape, dt.complex32), Buffer(self.shape, dt.complex32))
        return self._scratch

    def __call__(self, out: Buffer, inp: Buffer, graph: Optional[CommandGraph] = None) -> None:
        if out.var_type is not dt.complex64 or tuple(out.shape) != self.shape:
            raise ValueError("output must be a complex64 buffer of the volume's shape")
        if inp.var_type is not self.in_type or tuple(inp.shape) != self.shape:
            raise ValueError(f"input must be a {self.in_type.name} buffer of the volume's shape")
        s1, s2 = self.scratch()
        dispatch(self.passes[0], [s1, inp], graph)   # [i0][i1][i2] -> [k2][i0][i1]
        dispatch(self.passes[1], [s2, s1], graph)    # -> [k1][k2][i0]
        dispatch(self.passes[2], [out, s2], graph)   # -> [k0][k1][k2]


@lru_cache(maxsize=8)
def _plan(shape, in_type, inverse, normalize) -> Mixed3D:
    return Mixed3D(shape, in_type, inverse, normalize)


def fft3_mixed(out: Buffer, inp: Buffer, inverse: bool = False, normalize: bool = True,
               graph: Optional[CommandGraph] = None) -> None:
    """3-D C2C FFT of a 16-bit (complex_bf16 / complex32) volume into a complex64 buffer,
    radix-16 stages on MFMA with split-precision operands."""
    _plan(tuple(inp.shape), inp.var_type, bool(inverse), bool(normalize))(out, inp, graph)
