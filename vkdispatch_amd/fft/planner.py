"""Plans for the gfx950 FFT kernels.

A plan fixes everything the kernel template needs at compile time: the radix
pass list (Stockham, each radix done in registers), the threads per transform
``TT``, the transforms per workgroup ``B``, the lane mapping (transform slot
fastest across lanes for strided axes, so every load is a contiguous run of
``B`` elements), the LDS layout and the launch grid.

Geometry follows the reference (vkdispatch/fft/config.py:217-233): a buffer of
shape ``shape`` transformed along ``axis`` is ``OUTER`` batches of ``N``
elements with stride ``S = prod(shape[axis+1:])`` (the ``S`` inner batches are
interleaved).  The plan choices themselves are this build's, made for wave64,
160 KB LDS and 256 CUs: the reference's register-budget search
(config.py:102-190) targets 32-wide subgroups and 32-64 KB of shared memory.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional, Tuple

from .prime_utils import prime_factors

LDS_BYTES_MAX = 160 * 1024     # gfx950 LDS a workgroup may allocate
LDS_BYTES_TARGET = 64 * 1024   # keep >= 2 workgroups per CU when possible
MAX_THREADS = 1024
MAX_DIRECT_PRIME = 31          # larger primes go through Bluestein
WAVE = 64

# hand-picked power-of-two plans: N -> (radices, threads per transform)
_POW2 = {
    2: ([2], 1), 4: ([4], 1), 8: ([8], 1), 16: ([16], 1),
    32: ([8, 4], 4), 64: ([8, 8], 8), 128: ([16, 8], 8), 256: ([16, 16], 16),
    512: ([8, 8, 8], 64), 1024: ([16, 16, 4], 64), 2048: ([16, 16, 8], 128),
    4096: ([16, 16, 16], 256), 8192: ([16, 16, 16, 2], 512), 16384: ([16, 16, 16, 4], 1024),
}


def _wave_local_default() -> bool:
    return os.environ.get("VKX_WAVE_LOCAL", "1") != "0"


def radix_plan(N: int) -> Tuple[List[int], int]:
    """(radices, threads-per-transform) for a length-N transform done in one workgroup."""
    if N < 1:
        raise ValueError("FFT length must be positive")
    if N == 1:
        return [1], 1
    if N in _POW2:
        r, t = _POW2[N]
        return list(r), t
    primes = prime_factors(N)
    if max(primes) > MAX_DIRECT_PRIME:
        raise ValueError(f"A prime factor of {N} is {max(primes)}, above the direct radix limit {MAX_DIRECT_PRIME}")
    if N <= 16:
        return [N], 1
    # first-fit-decreasing packing of the prime factors into radices <= 16
    radices: List[int] = []
    for p in sorted(primes, reverse=True):
        best = -1
        for i, r in enumerate(radices):
            if r * p <= 16 and (best < 0 or radices[i] > radices[best]):
                best = i
        if best >= 0:
            radices[best] *= p
        else:
            radices.append(p)
    radices.sort(reverse=True)
    tt = N // radices[0]
    return radices, max(1, tt)


def pass_list(radices: List[int]) -> List[Tuple[int, int]]:
    out, ns = [], 1
    for r in radices:
        out.append((r, ns))
        ns *= r
    return out


def slots(N: int, R: int, TT: int) -> int:
    W = N // R
    return -(-W // TT) * R


@dataclasses.dataclass(frozen=True)
class AxisGeometry:
    shape: Tuple[int, ...]
    axis: int
    N: int
    S: int       # stride between consecutive elements of one transform
    OUTER: int   # number of outer batches

    @staticmethod
    def of(shape, axis=None) -> "AxisGeometry":
        shape = tuple(int(s) for s in shape)
        if axis is None:
            axis = len(shape) - 1
        if axis < 0:
            axis += len(shape)
        if not 0 <= axis < len(shape):
            raise ValueError(f"axis {axis} out of range for shape {shape}")
        N = shape[axis]
        S = int(math.prod(shape[axis + 1:]))
        total = int(math.prod(shape))
        return AxisGeometry(shape, axis, N, S, total // (N * S))


@dataclasses.dataclass(frozen=True)
class KernelPlan:
    """Compile-time configuration of one FFT-family kernel (becomes ``struct Cfg``)."""
    real: str            # "float" | "double"
    N: int
    TT: int
    B: int
    S: int
    OUTER: int
    radices: Tuple[int, ...]
    COLFAST: bool
    WAVE_LOCAL: bool
    PAD_SHIFT: int
    NP: int
    LDS_ELEMS: int
    NWG: int
    NWG_INNER: int
    TWREG: bool = False  # twiddles in registers, loaded with the data (a tuned option)

    @property
    def BLOCK(self) -> int:
        return self.B * self.TT

    @property
    def fwd(self) -> List[Tuple[int, int]]:
        return pass_list(list(self.radices))

    @property
    def inv(self) -> List[Tuple[int, int]]:
        return pass_list(list(reversed(self.radices)))

    @property
    def VMAX(self) -> int:
        return max(slots(self.N, r, self.TT) for r in self.radices)

    @property
    def KSLOTS(self) -> int:
        return slots(self.N, self.radices[-1], self.TT)

    @property
    def lds_bytes(self) -> int:
        return self.LDS_ELEMS * (16 if self.real == "double" else 8)

    @property
    def grid(self) -> Tuple[int, int, int]:
        return (self.NWG, 1, 1)


_OVERRIDES: dict = {}
_TUNED: Optional[dict] = None


def plan_key(geom: AxisGeometry, real: str, kind: str) -> str:
    """Tuned plans depend on the transform length and stride, not on the batch count."""
    return f"{kind}|{geom.N}|{geom.S}|{real}"


def set_override(geom: AxisGeometry, real: str, kind: str, radices=None, tt=None, B=None, twreg=False) -> None:
    """Force a plan (used by the autotuner); ``radices=None`` removes the override.
    Kernel factories are cached: call ``vd.fft.cache_clear()`` after changing overrides."""
    key = plan_key(geom, real, kind)
    if radices is None:
        _OVERRIDES.pop(key, None)
    else:
        _OVERRIDES[key] = {"radices": list(radices), "tt": int(tt), "B": None if B is None else int(B),
                           "twreg": bool(twreg)}


def _tuned_table() -> dict:
    global _TUNED
    if _TUNED is None:
        import json
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_plans.json")
        try:
            with open(path) as f:
                _TUNED = json.load(f).get("plans", {})
        except (OSError, ValueError):
            _TUNED = {}
    return _TUNED


def lookup(geom: AxisGeometry, real: str, kind: str) -> Optional[dict]:
    key = plan_key(geom, real, kind)
    if key in _OVERRIDES:
        return _OVERRIDES[key]
    if os.environ.get("VKX_NO_TUNED") == "1":
        return None
    return _tuned_table().get(key)


def make_plan(geom: AxisGeometry, real: str = "float", batch_hint: Optional[int] = None,
              radices: Optional[List[int]] = None, tt: Optional[int] = None,
              need_lds: bool = False, kind: str = "fft", B_force: Optional[int] = None) -> KernelPlan:
    """Plan for ``geom``.  ``need_lds``: the kernel exchanges a full transform through
    LDS after its passes (real-FFT post-processing), so reserve B padded rows.
    ``kind`` selects the tuned-plan entry (autotuner results / overrides) if any."""
    twreg = False
    if radices is None:
        hit = lookup(geom, real, kind)
        if hit is not None:
            radices, tt = hit["radices"], hit["tt"]
            B_force = hit.get("B") or B_force
            twreg = bool(hit.get("twreg", False))
    N, S, OUTER = geom.N, geom.S, geom.OUTER
    elem = 16 if real == "double" else 8
    strided = S > 1
    if radices is None:
        if strided and N <= 32 and max(prime_factors(N) if N > 1 else [1]) <= MAX_DIRECT_PRIME:
            radices, tt = [N], 1
        else:
            radices, tt = radix_plan(N)
    assert tt is not None
    radices = list(radices)
    if math.prod(radices) != N:
        raise ValueError(f"radices {radices} do not multiply to {N}")
    if tt > MAX_THREADS:
        raise ValueError(f"FFT of length {N} needs {tt} threads per transform (max {MAX_THREADS})")
    single_pass = len(radices) == 1
    pad_shift = 4 if elem == 16 else 5
    NP = N + (N >> pad_shift) + 1

    if strided:
        if tt == 1:
            B = min(S, 256)
        else:
            B = min(S, 16, max(1, MAX_THREADS // tt))
            while B > 1 and B * N * elem > LDS_BYTES_TARGET:
                B //= 2
        colfast = B > 1
    else:
        if tt == 1:
            B = 128 if N <= 8 else 64
        else:
            B = max(1, 256 // tt)
        B = max(1, min(B, OUTER))
        while B > 1 and (need_lds or not single_pass) and B * NP * elem > LDS_BYTES_TARGET:
            B //= 2
        colfast = False
    if batch_hint is not None:
        B = max(1, min(B, batch_hint))
    if B_force is not None:
        B = max(1, min(int(B_force), S if strided else OUTER, max(1, MAX_THREADS // tt)))
        colfast = strided and B > 1

    if need_lds and not colfast:
        lds_elems = B * NP
    elif single_pass:
        lds_elems = 1
    elif colfast:
        lds_elems = N * B
    else:
        lds_elems = B * NP
    if lds_elems * elem > LDS_BYTES_MAX:
        raise ValueError(f"FFT of length {N} ({real}) needs {lds_elems * elem} bytes of LDS (max {LDS_BYTES_MAX})")

    wave_local = (not colfast) and tt <= WAVE and WAVE % tt == 0 and _wave_local_default()
    if strided:
        nwg_inner = -(-S // B)
        nwg = OUTER * nwg_inner
    else:
        nwg_inner = 1
        nwg = -(-OUTER // B)
    if nwg >= 2 ** 31:
        raise ValueError("launch grid too large")
    return KernelPlan(real=real, N=N, TT=tt, B=B, S=S, OUTER=OUTER, radices=tuple(radices), COLFAST=colfast,
                      WAVE_LOCAL=wave_local, PAD_SHIFT=pad_shift, NP=NP, LDS_ELEMS=lds_elems, NWG=nwg,
                      NWG_INNER=nwg_inner, TWREG=twreg)


def transposed_size(plan: KernelPlan) -> int:
    """Elements of the transposed-kernel layout consumed by a convolution with this plan."""
    return plan.NWG * plan.KSLOTS * plan.BLOCK
