"""Plan autotuner.

For a kernel kind (``fft``, ``conv``, ``rfft``, ``irfft``) and a buffer
shape/axis it enumerates candidate plans (radix order, threads per
transform, transforms per workgroup) that fit gfx950 limits, compiles them
(in parallel, hiprtc), times each on the device with HIP events and keeps the
fastest.  Results go into a table keyed ``kind|N|stride|real`` that the
planner consults (``tuned_plans.json`` next to this file holds the table
measured on MI355X; ``VKX_NO_TUNED=1`` ignores it).  Tuning is deterministic
in its candidate order; ties keep the earlier candidate.
"""
from __future__ import annotations

import itertools
import json
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

from .. import dtypes as dt
from . import planner as P
from .prime_utils import prime_factors

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_plans.json")


def _orders(factors: Sequence[int], max_radix: int = 64) -> List[Tuple[int, ...]]:
    """Distinct radix sequences: groupings of the prime factors into radices <= max_radix."""
    out = set()
    facs = sorted(factors)

    def rec(rem: Tuple[int, ...], acc: Tuple[int, ...]):
        if not rem:
            out.add(acc)
            return
        seen = set()
        # choose a sub-multiset of rem as the next radix
        for k in range(1, len(rem) + 1):
            for combo in itertools.combinations(range(len(rem)), k):
                r = math.prod(rem[i] for i in combo)
                if r > max_radix or r in seen:
                    continue
                seen.add(r)
                left = tuple(rem[i] for i in range(len(rem)) if i not in combo)
                if len(acc) < 5:
                    rec(left, acc + (r,))
    rec(tuple(facs), ())
    return sorted(out, key=lambda t: (len(t), t))


def candidates(N: int, strided: bool, elem: int = 8, limit: int = 24) -> List[Tuple[Tuple[int, ...], int, int]]:
    """(radices, threads-per-transform, transforms-per-workgroup) candidates for length N."""
    if N <= 1:
        return [((1,), 1, 64)]
    orders = [o for o in _orders(prime_factors(N)) if len(o) <= 4 and max(o) <= 32]
    # prefer few passes with radices near 16
    orders.sort(key=lambda o: (len(o), sum(abs(math.log2(r) - 4) for r in o), o))
    out = []
    for rads in orders[:8]:
        tts = sorted({N // max(rads), max(1, N // max(rads) // 2), N // min(rads)}) if len(rads) > 1 else [1]
        for tt in tts:
            if tt < 1 or tt > 1024 or max(P.slots(N, r, tt) for r in rads) > 64:
                continue
            Bs = (8, 16, 32) if strided else sorted({max(1, 256 // tt), max(1, 512 // tt)})
            for B in Bs:
                if B * tt > 1024:
                    continue
                lds = (N * B) if strided else B * (N + (N >> 5) + 1)
                if lds * elem > P.LDS_BYTES_MAX:
                    continue
                out.append((tuple(rads), tt, B))
    return out[:limit] if limit else out


def load_table(path: str = TABLE_PATH) -> Dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {"plans": {}, "measurements": {}}


def save_table(table: Dict, path: str = TABLE_PATH) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _build(kind: str, shape, axis, geom, real: str, cand):
    from . import shader_factories as SF
    rads, tt, B = cand[:3]
    twreg = bool(cand[3]) if len(cand) > 3 else False
    c64 = dt.complex128 if real == "double" else dt.complex64
    if kind in ("rfft", "irfft"):
        M = geom.N // 2
        g = P.AxisGeometry((geom.OUTER, M), 1, M, 1, geom.OUTER)
    else:
        g = geom
    P.set_override(g, real, kind, rads, tt, B, twreg)
    try:
        SF.cache_clear()
        if kind == "conv":
            return SF.make_convolution_shader(tuple(shape), None, 1, axis, kernel_inner_only=True, input_type=c64,
                                              output_type=c64, kernel_type=c64, compute_type=c64)
        if kind in ("rfft", "irfft"):
            return SF.make_fft_shader(tuple(shape), None, inverse=(kind == "irfft"), r2c=True, input_type=c64,
                                      output_type=c64, compute_type=c64)
        return SF.make_fft_shader(tuple(shape), axis, input_type=c64, output_type=c64, compute_type=c64)
    finally:
        P.set_override(g, real, kind)


def _tuning_geometry(kind, shape, axis):
    geom = P.AxisGeometry.of(shape, axis)
    if kind in ("rfft", "irfft"):
        if geom.N % 2 or geom.S != 1:
            raise ValueError("real-row tuning needs even contiguous rows")
        M = geom.N // 2
        return geom, P.AxisGeometry((geom.OUTER, M), 1, M, 1, geom.OUTER)
    return geom, geom


def tune(kind: str, shape, axis=None, real: str = "float", cands=None, reps: int = 20,
         compile_only: bool = False, table: Optional[Dict] = None, twreg=(False, True)) -> Dict:
    """Time every candidate for (kind, shape, axis) and record the fastest in ``table``
    (default: the committed table).  A candidate is (radices, threads per transform,
    transforms per workgroup[, twiddles in registers]); the default set takes every
    plan of ``candidates`` with each ``twreg`` choice.  Returns {"key", "best", "results"}."""
    from . import kernels as K
    geom, tgeom = _tuning_geometry(kind, shape, axis)
    elem = 16 if real == "double" else 8
    if cands is None:
        cands = [c + (bool(tw),) for c in candidates(tgeom.N, tgeom.S > 1, elem) for tw in twreg]
    built = []
    for c in cands:
        try:
            built.append((c, _build(kind, shape, axis, geom, real, c)))
        except ValueError:
            continue
    K.precompile([k for _, k in built])
    # a plan that spills registers to scratch is never picked (a spilling kernel
    # has produced wrong results on this part)
    built = [(c, k) for c, k in built if k.code_object_info()["private_bytes"] == 0]
    key = P.plan_key(tgeom, real, kind)
    if compile_only:
        return {"key": key, "best": None, "results": [], "compiled": len(built)}
    from .. import buffer as B_
    from .. import runtime as rt
    c64 = dt.complex128 if real == "double" else dt.complex64
    if kind in ("rfft", "irfft"):
        vals = [B_.RFFTBuffer(tuple(shape), c64)]
    else:
        buf = B_.Buffer(tuple(shape), c64)
        vals = [buf]
        if kind == "conv":
            vals.append(B_.Buffer(tuple(shape[axis:]), c64))
    results = []
    for c, k in built:
        e0, e1 = rt.Event(), rt.Event()
        k.launch(vals)
        e0.record()
        for _ in range(reps):
            k.launch(vals)
        e1.record()
        e1.synchronize()
        results.append((e0.elapsed_ms(e1) / reps, c))
    best_ms, best = min(results, key=lambda r: r[0])
    if table is None:
        table = load_table()
    table.setdefault("plans", {})[key] = {"radices": list(best[0]), "tt": best[1], "B": best[2],
                                          "twreg": bool(best[3]) if len(best) > 3 else False}
    table.setdefault("measurements", {})[key] = {"us": round(best_ms * 1e3, 3), "shape": list(shape),
                                                 "axis": axis, "candidates": len(results)}
    return {"key": key, "best": best, "best_us": best_ms * 1e3,
            "results": sorted([(round(ms * 1e3, 3), list(c[0]), c[1], c[2], bool(c[3]) if len(c) > 3 else False)
                               for ms, c in results])}


# Shapes whose plans the committed table covers.
STANDARD_JOBS = [
    ("conv", (64, 512, 257), 1),
    ("rfft", (64, 512, 512), 2),
    ("irfft", (64, 512, 512), 2),
    ("fft", (1024, 4096), 1),
    ("conv", (32, 1024, 513), 1),
    ("rfft", (32, 1024, 1024), 2),
    ("irfft", (32, 1024, 1024), 2),
    ("fft", (256, 256, 256), 0),
    ("fft", (256, 256, 256), 1),
    ("fft", (256, 256, 256), 2),
]


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="autotune vd.fft plans on the current device")
    ap.add_argument("--compile-only", action="store_true")
    ap.add_argument("--out", default=TABLE_PATH)
    args = ap.parse_args(argv)
    table = load_table(args.out)
    for kind, shape, axis in STANDARD_JOBS:
        r = tune(kind, shape, axis, compile_only=args.compile_only, table=table)
        print(json.dumps({k: v for k, v in r.items() if k != "results"}), flush=True)
    if not args.compile_only:
        save_table(table, args.out)


if __name__ == "__main__":
    main()
