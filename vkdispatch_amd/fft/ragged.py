"""Ragged batches: many 1-D C2C transforms of different lengths in one launch.

The transforms are packed back to back in one complex64 buffer (transform i
occupies ``offsets[i] : offsets[i] + sizes[i]``).  ``RaggedFFT`` builds one
task per transform -- a Stockham radix list for the length (radices 2..13 and
16), or, for lengths with a prime factor above 13, a Bluestein plan (FP64 chirp,
power-of-two chirp spectrum) -- and runs them all in one persistent launch of
the ragged kernel (csrc/ragged.hip), longest first.  ``tune()`` picks the
workgroup size and the largest radix by timing the actual batch.

This path also serves ``vd.fft.fft`` for lengths the single-workgroup kernels do
not cover (prime factors above 31, or more than 16384 points).
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .. import runtime as rt
from .prime_utils import prime_factors

RTASK_DTYPE = np.dtype([
    ("in", "<u8"), ("out", "<u8"), ("tw", "<u8"), ("chirp", "<u8"), ("bspec", "<u8"),
    ("n", "<i4"), ("m", "<i4"), ("npass", "<i4"), ("inverse", "<i4"),
    ("radix", "<i4", (12,)), ("scale", "<f4"), ("count", "<i4"),
])
assert RTASK_DTYPE.itemsize == 112

LDS_PINGPONG = 4096
DIRECT_PRIMES = (2, 3, 5, 7, 11, 13)


def runtime_radices(n: int, max_radix: int = 16) -> List[int]:
    """Radix list for the ragged kernel: prime factors packed first-fit-decreasing
    into radices <= max_radix (from {2..13, 16}); largest first."""
    if n == 1:
        return []
    allowed = {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16}
    rads: List[int] = []
    for p in sorted(prime_factors(n), reverse=True):
        for i, r in enumerate(rads):
            if r * p <= max_radix and (r * p) in allowed:
                rads[i] = r * p
                break
        else:
            rads.append(p)
    rads.sort(reverse=True)
    if len(rads) > 12:
        raise ValueError(f"too many passes for n={n}")
    return rads


def is_direct(n: int) -> bool:
    return n == 1 or max(prime_factors(n)) <= 13


def bluestein_length(n: int) -> int:
    m = 1
    while m < 2 * n - 1:
        m *= 2
    return m


class _Tables:
    """Device-resident FP64-generated tables shared by all ragged plans."""

    def __init__(self):
        self.tw: Dict[Tuple[int, Tuple[int, ...]], int] = {}
        self.chirp: Dict[int, int] = {}
        self.bspec: Dict[Tuple[int, int, Tuple[int, ...]], int] = {}
        self._keep = []

    def twiddles(self, n: int, radices: Sequence[int]) -> int:
        key = (n, tuple(radices))
        if key not in self.tw:
            if n <= 1:
                self.tw[key] = 0
            else:
                p = rt.malloc((n - 1) * 8)
                arr = (ctypes.c_int * len(radices))(*radices)
                rt.check(rt.lib().vkx_twiddles(p, n, arr, len(radices), 0, rt.current_stream()), "vkx_twiddles")
                self.tw[key] = p
        return self.tw[key]

    def chirp_table(self, n: int) -> int:
        if n not in self.chirp:
            p = rt.malloc(n * 8)
            rt.check(rt.lib().vkx_chirp(p, n, n, -1, 0, rt.current_stream()), "vkx_chirp")
            self.chirp[n] = p
        return self.chirp[n]

    def chirp_spectrum(self, n: int, m: int, radices: Sequence[int]) -> int:
        key = (n, m, tuple(radices))
        if key not in self.bspec:
            p = rt.malloc(m * 8)
            rt.check(rt.lib().vkx_bluestein_b(p, n, m, rt.current_stream()), "vkx_bluestein_b")
            # B = FFT_m(b) with the ragged kernel itself
            t = np.zeros(1, RTASK_DTYPE)
            t["in"], t["out"] = p, p
            t["tw"] = self.twiddles(m, radices)
            t["n"], t["m"], t["npass"], t["inverse"] = m, 0, len(radices), 0
            t["radix"][0, :len(radices)] = radices
            t["scale"], t["count"] = 1.0, 1
            _launch_tasks(t, max(1, 2 * m) if m > LDS_PINGPONG else 1, grid=1, block=256)
            self.bspec[key] = p
        return self.bspec[key]


_tables: Optional[_Tables] = None


def tables() -> _Tables:
    global _tables
    if _tables is None:
        rt.ensure_device()
        _tables = _Tables()
    return _tables


class _Arena:
    def __init__(self):
        self.ptr = 0
        self.size = 0
        self.counter = 0

    def ensure(self, nbytes: int) -> int:
        if nbytes > self.size:
            if self.ptr:
                rt.synchronize()
                rt.free(self.ptr)
            self.ptr = rt.malloc(nbytes)
            self.size = nbytes
        return self.ptr

    def counter_ptr(self) -> int:
        if not self.counter:
            self.counter = rt.malloc(256)
        return self.counter


_arena = _Arena()


def _launch_tasks(tasks: np.ndarray, scratch_elems_per_wg: int, grid: int, block: int, stream=None,
                  task_buf=None) -> None:
    s = rt.current_stream() if stream is None else int(stream)
    if task_buf is None:
        task_ptr = rt.malloc(tasks.nbytes)
        rt.htod(task_ptr, tasks.ctypes.data, tasks.nbytes, s)
    else:
        task_ptr = task_buf
    scr = _arena.ensure(max(1, scratch_elems_per_wg) * 8 * grid)
    rc = rt.lib().vkx_ragged_exec(task_ptr, len(tasks), _arena.counter_ptr(), scr, scratch_elems_per_wg, grid, block,
                                  s)
    if rc != 0:
        raise RuntimeError("vkx_ragged_exec launch failed")
    if task_buf is None:
        rt.synchronize(s)
        rt.free(task_ptr)


def _task_cost(n: int, m: int) -> float:
    L = m if m else n
    return L * max(1.0, math.log2(max(L, 2))) * (3.0 if m else 1.0)


class RaggedFFT:
    """Plan for a packed batch of transforms of lengths ``sizes`` (forward or inverse)."""

    def __init__(self, sizes: Sequence[int], inverse: bool = False, normalize: bool = True, block: int = 256,
                 max_radix: int = 16, grid: Optional[int] = None):
        self.sizes = np.asarray(sizes, dtype=np.int64)
        if np.any(self.sizes < 1):
            raise ValueError("transform lengths must be positive")
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        self.total = int(self.sizes.sum())
        self.inverse = bool(inverse)
        self.normalize = bool(normalize)
        self.block = int(block)
        self.max_radix = int(max_radix)
        dev = rt.ensure_device()
        self.grid = max(1, min(int(grid) if grid else 2 * int(dev.compute_units), len(self.sizes)))
        T = tables()
        self._template = np.zeros(len(self.sizes), RTASK_DTYPE)
        scratch = 1
        for i, n in enumerate(self.sizes.tolist()):
            t = self._template[i]
            t["n"], t["inverse"], t["count"] = n, int(self.inverse), 1
            t["scale"] = (1.0 / n) if (self.inverse and self.normalize) else 1.0
            if is_direct(n):
                rads = runtime_radices(n, self.max_radix)  # [] for n == 1: in place, nothing to do
                t["m"], t["npass"] = 0, len(rads)
                t["radix"][:len(rads)] = rads
                t["tw"] = T.twiddles(n, rads) if rads else 0
                if n > LDS_PINGPONG:
                    scratch = max(scratch, 2 * n)
            else:
                m = bluestein_length(n)
                rads = runtime_radices(m, self.max_radix)
                t["m"], t["npass"] = m, len(rads)
                t["radix"][:len(rads)] = rads
                t["tw"] = T.twiddles(m, rads)
                t["chirp"] = T.chirp_table(n)
                t["bspec"] = T.chirp_spectrum(n, m, rads)
                scratch = max(scratch, 3 * m)
        self.scratch_per_wg = scratch
        cost = np.array([_task_cost(int(t["n"]), int(t["m"])) for t in self._template])
        self.order = np.argsort(-cost, kind="stable")
        self._task_cache: Dict[Tuple[int, int], int] = {}

    def _tasks_for(self, in_ptr: int, out_ptr: int) -> int:
        key = (in_ptr, out_ptr)
        if key not in self._task_cache:
            t = self._template[self.order].copy()
            off = self.offsets[self.order] * 8
            t["in"] = in_ptr + off
            t["out"] = out_ptr + off
            p = rt.malloc(t.nbytes)
            rt.htod(p, t.ctypes.data, t.nbytes)
            self._task_cache[key] = p
        return self._task_cache[key]

    def bind(self, out_buf, in_buf=None):
        """Upload the task table for this (input, output) pair and size the scratch
        arena; returns ``launch(stream)`` that allocates nothing (safe under graph capture)."""
        if in_buf is None:
            in_buf = out_buf
        for b in (out_buf, in_buf):
            if getattr(b, "size", 0) < self.total:
                raise ValueError(f"buffer of {getattr(b, 'size', 0)} elements, ragged batch needs {self.total}")
        if any(int(n) == 1 for n in self.sizes) and in_buf is not out_buf:
            raise ValueError("out-of-place ragged batches may not contain length-1 transforms")
        tp = self._tasks_for(int(in_buf.ptr), int(out_buf.ptr))
        _arena.ensure(max(1, self.scratch_per_wg) * 8 * self.grid)
        _arena.counter_ptr()
        keep = (out_buf, in_buf)

        def launch(stream=None):
            _ = keep
            _launch_tasks(self._template, self.scratch_per_wg, self.grid, self.block, stream, task_buf=tp)
        return launch

    def __call__(self, out_buf, in_buf=None, stream=None) -> None:
        """Transform ``in_buf`` (default: ``out_buf``, in place) into ``out_buf``;
        both complex64 buffers of ``total`` elements."""
        self.bind(out_buf, in_buf)(stream)

    def flops(self) -> float:
        return float(np.sum(5.0 * self.sizes * np.log2(np.maximum(self.sizes, 2))))

    def breakdown(self) -> Dict[str, Dict[str, float]]:
        classes = {"pow2": [], "smooth": [], "bluestein": []}
        for n in self.sizes.tolist():
            if n & (n - 1) == 0:
                classes["pow2"].append(n)
            elif is_direct(n):
                classes["smooth"].append(n)
            else:
                classes["bluestein"].append(n)
        return {k: {"count": len(v), "points": int(sum(v)),
                    "flops": float(sum(5.0 * n * math.log2(max(n, 2)) for n in v))} for k, v in classes.items()}


def mixed_sizes(count: int, seed: int = 0, lo: int = 64, hi: int = 65536) -> List[int]:
    """Synthetic ragged size mix: per transform u = U(0,1); u < 0.6: 2^k with k uniform in
    [log2 lo, log2 hi]; u < 0.9: the {2,3,5,7}-smooth number nearest a log-uniform
    target in [lo, hi]; else the prime nearest such a target."""
    rng = np.random.default_rng(seed)
    span = np.arange(lo, hi + 1)
    smooth_mask = np.zeros(span.size, bool)
    r = span.copy()
    for p in (2, 3, 5, 7):
        while True:
            div = (r % p == 0)
            if not div.any():
                break
            r = np.where(div, r // p, r)
    smooth = span[r == 1]
    sieve = np.ones(hi + 1, bool)
    sieve[:2] = False
    for i in range(2, int(hi ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    primes = np.nonzero(sieve)[0]
    primes = primes[primes >= lo]
    out = []
    for _ in range(count):
        u = rng.random()
        if u < 0.6:
            out.append(int(2 ** rng.integers(int(math.log2(lo)), int(math.log2(hi)) + 1)))
            continue
        target = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        pool = smooth if u < 0.9 else primes
        i = int(np.searchsorted(pool, target))
        cands = [int(pool[j]) for j in (i - 1, i) if 0 <= j < len(pool)]
        out.append(min(cands, key=lambda c: abs(c - target)))
    return out


def packed_normal(sizes: Sequence[int], seed: int = 0):
    """Per-transform N(0,1) real then imaginary parts, in size order, packed complex64."""
    rng = np.random.default_rng(seed)
    parts = [(rng.standard_normal(n) + 1j * rng.standard_normal(n)) for n in sizes]
    return parts, np.concatenate(parts).astype(np.complex64)


def tune(sizes: Sequence[int], buf, reps: int = 5) -> Dict:
    """Pick the workgroup size and largest radix for this batch by timing."""
    best = None
    results = []
    for block in (256, 512):
        for max_radix in (8, 16):
            plan = RaggedFFT(sizes, block=block, max_radix=max_radix)
            plan(buf)
            e0, e1 = rt.Event(), rt.Event()
            e0.record()
            for _ in range(reps):
                plan(buf)
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / reps
            results.append({"block": block, "max_radix": max_radix, "ms": ms})
            if best is None or ms < best["ms"]:
                best = results[-1]
    return {"best": best, "results": results}
