"""Ragged batches: many 1-D C2C transforms of different lengths in one launch.

The transforms are packed back to back in one complex64 buffer (transform i
occupies ``offsets[i] : offsets[i] + sizes[i]``).  ``RaggedFFT`` turns the batch
into a queue of work items for one persistent launch of the ragged kernel
(csrc/ragged.hip):

* lengths up to one LDS buffer (4096 points): one item per transform, a
  Stockham radix list (radices 2..13 and 16) run in LDS;
* longer lengths that split as L = n1 * n2 with both factors <= 512: the
  four-step algorithm as column-tile items (length-n1 transforms of a run of
  columns, times W_L^(k1 j2)) and row-tile items (length-n2 transforms of a run
  of rows), so one long transform spreads over every workgroup instead of
  running on one.  The levels run as consecutive launches of the same
  persistent kernel (default), or in one launch where a row tile waits on its
  transform's done counter;
* lengths with a prime factor above 13: Bluestein (FP64 chirp, power-of-two
  chirp spectrum), as four tile levels when the power of two splits;
* anything else (longer than 512^2 without a split): one item with radix
  passes through a per-workgroup scratch slice.

With one launch the queue holds all first-level tiles, then each later level,
then the whole transforms longest first; an item only waits for items queued
before it.  With a launch per level the whole transforms head the first level's queue.

A length that many transforms share leaves the queue: power-of-two lengths of
CLASS_LENGTHS (LengthClass) and Bluestein sizes m of BLUE_CLASS_LENGTHS
(BlueClass) run as one virtual batch each through the vd.fft kernels, with
gather / scatter hooks that map transform b to its offset in the packed
buffers, on a side stream forked from and joined back into the caller's stream.
``tune()`` picks the workgroup size, the largest radix, the direct length
from which transforms are tiled and whether classes are used by timing the
actual batch; ``tuned_plan()``
keeps the pick in ``tuned_ragged.json`` (keyed by device arch and a hash of the
length list) so a batch is timed once per device.

This path also serves ``vd.fft.fft`` for lengths the single-workgroup kernels do
not cover (prime factors above 31, or more than 16384 points).
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import math
import os
from typing import Dict, List, NamedTuple, Optional, Sequence, Tuple

import numpy as np

from .. import codegen as vc
from .. import runtime as rt
from .prime_utils import prime_factors

RTASK_DTYPE = np.dtype([
    ("in", "<u8"), ("out", "<u8"), ("tw", "<u8"), ("chirp", "<u8"), ("bspec", "<u8"),
    ("n", "<i4"), ("m", "<i4"), ("npass", "<i4"), ("inverse", "<i4"),
    ("radix", "<i4", (12,)), ("scale", "<f4"), ("count", "<i4"),
    ("tile0", "<i4"), ("dep", "<i4"), ("target", "<i4"), ("done", "<i4"),
])
assert RTASK_DTYPE.itemsize == 128

# item kinds (npass bits 16-19), the conjugate flag (bit 20) and log2 of the tile width
# (bits 21-24), as in csrc/ragged.hip
K_WHOLE, K_COL, K_ROW, K_BLUE_COL, K_BLUE_ROW, K_BLUE_OUT = range(6)
CONJ = 1 << 20

LDS_PINGPONG = 4096
LDS_TILE = 4096 + 128  # complex64 elements of one LDS buffer (a tile at pitch n + 1)
FOUR_STEP_MAX = 512  # longest four-step factor: tiles of >= 8 padded columns fit one LDS buffer
TILE_MIN_DIRECT = 1 << 17  # direct lengths from here on are tiled (shorter ones run whole on one workgroup)
DIRECT_PRIMES = (2, 3, 5, 7, 11, 13)
# power-of-two length classes (LengthClass): the lengths, the two-launch splits, and how
# many transforms of a length make a class
CLASS_LENGTHS = tuple(2 ** k for k in range(6, 17))  # 64 .. 65536
CLASS_SPLIT = {32768: (128, 256), 65536: (256, 256), 131072: (256, 512)}  # (n1, n2): column, row length
CLASS_MIN_COUNT = 16  # fewer transforms of a length stay items of the persistent kernel
BLUE_CLASS_LENGTHS = tuple(2 ** k for k in range(7, 18))  # Bluestein sizes m of Bluestein classes: 128 .. 131072
CLASS_STREAMS = 2  # side streams the classes are spread over


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


def four_step_split(L: int, max_radix: int = 16) -> Optional[Tuple[List[int], List[int]]]:
    """Radix lists of L = n1 * n2 (n1: column length, n2: row length, both <= 512) for
    the four-step path of lengths above one LDS buffer; None when L fits in LDS or has
    no such split.  Prime factors go to the smaller side, largest first."""
    if L <= LDS_PINGPONG:
        return None
    if max(prime_factors(L)) > 13:
        return None
    a = b = 1
    for p in sorted(prime_factors(L), reverse=True):
        if a <= b:
            a *= p
        else:
            b *= p
    if max(a, b) > FOUR_STEP_MAX:
        return None
    r1, r2 = runtime_radices(a, max_radix), runtime_radices(b, max_radix)
    if len(r1) + len(r2) > 12:
        return None
    return r1, r2


def _encode_passes(t, radices: Sequence[int], split: Optional[Tuple[List[int], List[int]]], kind: int = K_WHOLE,
                   flags: int = 0) -> None:
    """Item pass fields: radices and npass, or for a four-step split the n1 list then the
    n2 list with npass = total | (n1 passes << 8) | (kind << 16) | flags."""
    if split is None:
        t["npass"] = len(radices) | (kind << 16) | flags
        t["radix"][:len(radices)] = radices
    else:
        r1, r2 = split
        rads = list(r1) + list(r2)
        t["npass"] = len(rads) | (len(r1) << 8) | (kind << 16) | flags
        t["radix"][:len(rads)] = rads


def auto_tile_min(sizes: Sequence[int], resident_wgs: int) -> int:
    """Default direct length from which transforms are tiled: 16384 when fewer such
    transforms than resident workgroups exist (whole, most of the GPU would idle:
    2 x 65536 points take 133 us whole and 40 us tiled), else never (a batch with
    enough long transforms keeps every workgroup busy on whole transforms, which
    then cost less than the tiles' two extra passes)."""
    n_long = sum(1 for n in sizes if n >= 16384 and is_direct(int(n)))
    return 16384 if n_long < resident_wgs else TILE_MIN_DIRECT


def tile_log2_width(n: int, npass: int = 3) -> int:
    """log2 of the tile width for length-n transforms: the largest power of two W with
    W (n + 1) <= one LDS buffer, or two buffers when the tile has at most two passes
    (its first pass reads global memory and its last writes it, so one LDS copy is
    live at a time; the LDS pitch is n + 1)."""
    w = (LDS_TILE * (2 if npass <= 2 else 1)) // (n + 1)
    if w < 1:
        raise ValueError(f"length {n} does not fit an LDS tile")
    return w.bit_length() - 1


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

    def twiddles_split(self, split: Tuple[List[int], List[int]]) -> int:
        """n1's Stockham table followed by n2's (one allocation)."""
        r1, r2 = (tuple(r) for r in split)
        key = ("split", r1, r2)
        if key not in self.tw:
            n1, n2 = int(np.prod(r1)), int(np.prod(r2))
            p = rt.malloc((n1 - 1 + n2 - 1) * 8)
            for off, n, rads in ((0, n1, r1), ((n1 - 1) * 8, n2, r2)):
                arr = (ctypes.c_int * len(rads))(*rads)
                rt.check(rt.lib().vkx_twiddles(p + off, n, arr, len(rads), 0, rt.current_stream()), "vkx_twiddles")
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
            t["scale"], t["count"], t["dep"], t["done"] = 1.0, 1, -1, -1
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
                  task_buf=None, counters: Optional[Tuple[int, int]] = None, spin_limit: int = 0,
                  scratch: int = 0) -> None:
    """Launch a task table; ``scratch`` (device pointer, scratch_elems_per_wg complex64
    per workgroup) is the caller's, else the shared arena, which only synchronous
    one-off launches use (it may be reallocated by the next such launch)."""
    s = rt.current_stream() if stream is None else int(stream)
    if task_buf is None:
        task_ptr = rt.malloc(tasks.nbytes)
        rt.htod(task_ptr, tasks.ctypes.data, tasks.nbytes, s)
    else:
        task_ptr = task_buf
    scr = scratch or _arena.ensure(max(1, scratch_elems_per_wg) * 8 * grid)
    ctr, nctr = counters if counters is not None else (_arena.counter_ptr(), 2)
    rc = rt.lib().vkx_ragged_exec(task_ptr, len(tasks), ctr, nctr, scr, scratch_elems_per_wg, grid, block,
                                  int(spin_limit), s)
    rt.check(rc, "vkx_ragged_exec")
    if task_buf is None:
        rt.synchronize(s)
        rt.free(task_ptr)


def check_queue_order(items: np.ndarray) -> None:
    """Every item that waits on a counter comes after all the items that raise that
    counter to its target (so a wait is only ever for work fetched earlier)."""
    seen: Dict[int, int] = {}
    for k, t in enumerate(items):
        dep = int(t["dep"])
        if dep >= 0 and seen.get(dep, 0) != int(t["target"]):
            raise AssertionError(f"item {k} waits for counter {dep} = {int(t['target'])}, "
                                 f"{seen.get(dep, 0)} producers queued before it")
        done = int(t["done"])
        if done >= 0:
            seen[done] = seen.get(done, 0) + 1


def _task_cost(n: int, m: int) -> float:
    L = m if m else n
    return L * max(1.0, math.log2(max(L, 2))) * (3.0 if m else 1.0)


SRC_IN, SRC_OUT, SRC_SCR = 0, 1, 2  # item pointer bases: the input, the output, the plan's scratch


class ItemQueue(NamedTuple):
    items: np.ndarray      # RTASK_DTYPE records in queue order (in / out filled at bind time)
    src: np.ndarray        # (base, element offset) of each item's input
    dst: np.ndarray        # (base, element offset) of each item's output
    ncounters: int         # queue position + abort flag + done counters
    scratch_per_wg: int    # complex64 scratch per workgroup for whole-transform items
    scratch_elems: int     # complex64 scratch of the plan's tiled transforms
    segments: List[Tuple[int, int]]  # item ranges launched one after another


def plan_items(sizes: Sequence[int], inverse: bool, normalize: bool, mr: int, four_step: bool, T,
               sync: str = "launch", tile_min: int = TILE_MIN_DIRECT, fill: int = 0,
               offsets: Optional[Sequence[int]] = None) -> ItemQueue:
    """Work-item queue of a ragged batch.  ``T`` provides the device tables
    (twiddles, twiddles_split, chirp_table, chirp_spectrum) as addresses.
    sync="launch" (default): one launch per tile level (the kernel boundary orders the
    levels, no waits), whole transforms at the head of the first;
    sync="counter": one launch, later tile levels wait on done counters (each item's
    agent-scope release / acquire writes back and invalidates L2, which costs more
    than the launches it saves on this 8-XCD device).  Direct lengths >= tile_min and
    Bluestein lengths are tiled when four_step is set and the length splits.
    offsets: element offset of each transform in the packed buffers (default: packed
    back to back in the order given)."""
    if sync not in ("counter", "launch"):
        raise ValueError("sync is 'counter' or 'launch'")
    sizes = np.asarray(sizes, dtype=np.int64)
    if offsets is None:
        offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    offsets = np.asarray(offsets, dtype=np.int64)

    def counter() -> int:
        nctr[0] += 1
        return nctr[0] - 3

    levels: List[list] = [[], [], [], []]  # (record, (src, off), (dst, off)) per tile level
    whole: list = []                       # (cost, record, offset)
    nctr = [2]
    scr_elems = 0
    scratch = 1
    cost = [_task_cost(n, 0 if is_direct(n) else bluestein_length(n)) for n in sizes.tolist()]
    I, O, S = SRC_IN, SRC_OUT, SRC_SCR
    for i in sorted(range(len(sizes)), key=lambda j: -cost[j]):
        n, off = int(sizes[i]), int(offsets[i])
        base = np.zeros((), RTASK_DTYPE)
        base["n"], base["inverse"], base["count"], base["dep"], base["done"] = n, int(inverse), 1, -1, -1
        base["scale"] = (1.0 / n) if (inverse and normalize) else 1.0
        direct = is_direct(n)
        L = n if direct else bluestein_length(n)
        rads = runtime_radices(L, mr)  # [] for n == 1: in place, nothing to do
        split = four_step_split(L, mr) if four_step and (not direct or n >= tile_min) else None
        if not direct:
            base["m"] = L
            base["chirp"] = T.chirp_table(n)
            base["bspec"] = T.chirp_spectrum(n, L, rads)
        if split is None:
            rec = base.copy()
            _encode_passes(rec, rads, None)
            rec["tw"] = T.twiddles(L, rads) if rads else 0
            if not direct:
                scratch = max(scratch, 3 * L)
            elif n > LDS_PINGPONG:
                scratch = max(scratch, 2 * n)
            whole.append((cost[i], rec, off))
            continue
        r1, r2 = split
        n1 = int(np.prod(r1))
        n2 = L // n1
        lc, lr = tile_log2_width(n1, len(r1)), tile_log2_width(n2, len(r2))
        cw, rw = 1 << lc, 1 << lr
        ncol, nrow = -(-n2 // cw), -(-n1 // rw)
        base["tw"] = T.twiddles_split(split)

        def tiles(kind, flags, src, dst, dep, target, done, scale=None, cols=True):
            span, width = (n2, cw) if cols else (n1, rw)
            for t0 in range(0, span, width):
                rec = base.copy()
                _encode_passes(rec, rads, split, kind, flags | ((lc if cols else lr) << 21))
                rec["tile0"], rec["count"] = t0, min(width, span - t0)
                rec["dep"], rec["target"], rec["done"] = dep, target, done
                if scale is not None:
                    rec["scale"] = scale
                yield rec, src, dst

        conj = CONJ if inverse else 0
        if direct:
            c1 = counter()
            y = scr_elems
            scr_elems += L
            levels[0] += tiles(K_COL, conj, (I, off), (S, y), -1, 0, c1)
            levels[1] += tiles(K_ROW, conj, (S, y), (O, off), c1, ncol, -1, cols=False)
        else:
            c1, c2, c3 = counter(), counter(), counter()
            a, b = scr_elems, scr_elems + L
            scr_elems += 2 * L
            levels[0] += tiles(K_BLUE_COL, 0, (I, off), (S, a), -1, 0, c1)
            levels[1] += tiles(K_BLUE_ROW, 0, (S, a), (S, b), c1, ncol, c2, cols=False)
            levels[2] += tiles(K_COL, 0, (S, b), (S, a), c2, nrow, c3)
            levels[3] += tiles(K_BLUE_OUT, 0, (S, a), (O, off), c3, ncol, -1,
                               scale=float(base["scale"]) / L, cols=False)
    wq = [(rec, (I, off), (O, off)) for _, rec, off in sorted(whole, key=lambda w: -w[0])]
    if sync == "counter":
        queue = [x for lv in levels for x in lv] + wq
        segments = [(0, len(queue))]
    else:
        for lv in levels:
            for rec, _, _ in lv:
                rec["dep"], rec["target"], rec["done"] = -1, 0, -1
        nctr[0] = 2
        # the whole transforms go first: the longest of them (one workgroup each) set the
        # first launch's critical path, the short first-level tiles fill in behind them.
        # fill: whole transforms of <= fill points are dealt round-robin to the tails of
        # the later (tile-only) launches instead, where workgroups otherwise idle while
        # the last tiles finish
        later = [lv for lv in levels[1:] if lv]
        first = wq
        tails: List[list] = [[] for _ in later]
        if fill and later:
            first = [x for x in wq if int(x[0]["n"]) > fill or int(x[0]["m"]) != 0]
            small = [x for x in wq if not (int(x[0]["n"]) > fill or int(x[0]["m"]) != 0)]
            for i, x in enumerate(small):
                tails[i % len(later)].append(x)
        groups = [first + levels[0]] + [lv + tl for lv, tl in zip(later, tails)]
        queue = [x for g in groups for x in g]
        bounds = np.cumsum([0] + [len(g) for g in groups]).tolist()
        segments = [(a, b) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
    items = np.zeros(len(queue), RTASK_DTYPE)
    src_ref = np.zeros((len(queue), 2), np.int64)
    dst_ref = np.zeros((len(queue), 2), np.int64)
    for k, (rec, src, dst) in enumerate(queue):
        items[k] = rec
        src_ref[k], dst_ref[k] = src, dst
    check_queue_order(items)
    return ItemQueue(items, src_ref, dst_ref, nctr[0], scratch, scr_elems, segments)


class RaggedFFT:
    """Plan for a packed batch of transforms of lengths ``sizes`` (forward or inverse).
    ``four_step`` (default): lengths above one LDS buffer that split into two factors
    <= 512 run as column-tile and row-tile items spread over all workgroups; without
    it they run whole on one workgroup through a scratch slice.  ``fill``: whole
    transforms of at most this length fill the tails of the later tile launches
    (4096-transform mix: 1191 -> 1168 us).  ``classes``: a power-of-two length of
    CLASS_LENGTHS that occurs at least this many times runs as one length class
    (LengthClass), and a Bluestein size m of BLUE_CLASS_LENGTHS shared by that many as
    one Bluestein class (BlueClass), instead of as items of the persistent kernel; 0
    turns classes off.  ``class_streams``: side streams the classes are spread over;
    ``persistent_first``: enqueue the persistent launches before the classes."""

    def __init__(self, sizes: Sequence[int], inverse: bool = False, normalize: bool = True, block: int = 256,
                 max_radix: int = 16, grid: Optional[int] = None, four_step: bool = True, spin_limit: int = 0,
                 sync: str = "launch", tile_min: Optional[int] = None, block_first: Optional[int] = None,
                 fill: int = LDS_PINGPONG, classes: int = CLASS_MIN_COUNT, class_streams: int = CLASS_STREAMS,
                 persistent_first: bool = False):
        self.sizes = np.asarray(sizes, dtype=np.int64)
        if np.any(self.sizes < 1):
            raise ValueError("transform lengths must be positive")
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        self.total = int(self.sizes.sum())
        self.inverse = bool(inverse)
        self.normalize = bool(normalize)
        self.block = int(block)
        self.block_first = int(block_first) if block_first else self.block  # first launch (holds whole items)
        self.max_radix = int(max_radix)
        self.spin_limit = int(spin_limit)
        dev = rt.ensure_device()
        if tile_min is None:
            tile_min = auto_tile_min(self.sizes, 2 * int(dev.compute_units))
        self.tile_min = int(tile_min)
        self.sync = sync
        self.classes = int(classes)
        if int(class_streams) < 1:
            raise ValueError("class_streams must be >= 1")
        self.class_streams = int(class_streams)
        self.persistent_first = bool(persistent_first)
        groups = class_groups(self.sizes, self.classes)
        bgroups = blue_class_groups(self.sizes, self.classes)
        rest = np.ones(len(self.sizes), bool)
        for idx in list(groups.values()) + list(bgroups.values()):
            rest[idx] = False
        self._classes = [BlueClass(m, self.sizes[idx], self.offsets[idx], self.inverse, self.normalize, self.max_radix)
                         for m, idx in sorted(bgroups.items(), reverse=True)]
        self._classes += [LengthClass(L, self.offsets[idx], self.inverse, self.normalize)
                          for L, idx in sorted(groups.items(), reverse=True)]
        q = plan_items(self.sizes[rest], self.inverse, self.normalize, self.max_radix, four_step, tables(), sync,
                       self.tile_min, int(fill), offsets=self.offsets[rest])
        self._items, self._src, self._dst = q.items, q.src, q.dst
        self._segments = q.segments
        self._nctr = q.ncounters
        scratch, scr_elems = q.scratch_per_wg, q.scratch_elems
        self.scratch_per_wg = scratch
        self.scratch_elems = scr_elems
        self.grid = max(1, min(int(grid) if grid else 2 * int(dev.compute_units), len(self._items)))
        self._scr = rt.malloc(max(1, scr_elems) * 8)
        self._ctr = rt.malloc(4 * self._nctr)
        self._wg_scr = 0
        self._task_cache: Dict[Tuple[int, int], int] = {}

    @property
    def n_items(self) -> int:
        return len(self._items)

    def _tasks_for(self, in_ptr: int, out_ptr: int) -> int:
        key = (in_ptr, out_ptr)
        if key not in self._task_cache:
            t = self._items.copy()
            bases = np.array([in_ptr, out_ptr, self._scr], dtype=np.uint64)
            t["in"] = bases[self._src[:, 0]] + (self._src[:, 1] * 8).astype(np.uint64)
            t["out"] = bases[self._dst[:, 0]] + (self._dst[:, 1] * 8).astype(np.uint64)
            p = rt.malloc(t.nbytes)
            rt.htod(p, t.ctypes.data, t.nbytes)
            self._task_cache[key] = p
        return self._task_cache[key]

    def bind(self, out_buf, in_buf=None):
        """Upload the item table for this (input, output) pair and allocate the plan's
        scratch; returns ``launch(stream)`` that allocates nothing (safe under graph
        capture: every pointer it launches with belongs to the plan and lives as long as
        it).  Launches of one plan share its counters and scratch: they must not overlap."""
        if in_buf is None:
            in_buf = out_buf
        for b in (out_buf, in_buf):
            if getattr(b, "size", 0) < self.total:
                raise ValueError(f"buffer of {getattr(b, 'size', 0)} elements, ragged batch needs {self.total}")
        if any(int(n) == 1 for n in self.sizes) and in_buf is not out_buf:
            raise ValueError("out-of-place ragged batches may not contain length-1 transforms")
        tp = self._tasks_for(int(in_buf.ptr), int(out_buf.ptr)) if len(self._items) else 0
        if not self._wg_scr:  # per-workgroup scratch of whole transforms, owned by the plan
            self._wg_scr = rt.malloc(max(1, self.scratch_per_wg) * 8 * self.grid)
        wg_scr = self._wg_scr
        # the length classes and the persistent kernel's items touch disjoint transforms,
        # and so do the classes among themselves: the classes run on side streams (lanes,
        # balanced by the bytes they move), forked from and joined back into the caller's
        # stream, so a class launch of few workgroups shares the GPU with the others and
        # with the persistent launches
        lanes = class_lanes(self._classes, self.class_streams)
        runs = [[self._classes[i].bind(out_buf, in_buf) for i in lane] for lane in lanes]
        side = rt.side_streams(len(runs)) if runs else []
        fork = rt.Event() if runs else None
        joins = [rt.Event() for _ in runs]
        keep = (out_buf, in_buf, fork, joins)

        def launch(stream=None):
            _ = keep
            s = rt.current_stream() if stream is None else int(stream)
            if runs:
                fork.record(s)

            def persistent():
                if self.sync == "counter":
                    rt.note_async_check(self, s)
                for i, (a, b) in enumerate(self._segments):
                    _launch_tasks(self._items[a:b], self.scratch_per_wg, self.grid,
                                  self.block_first if i == 0 else self.block, s,
                                  task_buf=tp + a * RTASK_DTYPE.itemsize, counters=(self._ctr, self._nctr),
                                  spin_limit=self.spin_limit, scratch=wg_scr)

            # a replayed graph dispatches its nodes in the order they were recorded: the
            # part enqueued second starts once the first part has been issued
            if self.persistent_first:
                persistent()
            for lane_runs, ls, join in zip(runs, side, joins):
                rt.stream_wait(ls, fork)
                for run in lane_runs:
                    run(ls)
                join.record(ls)
            if not self.persistent_first:
                persistent()
            for join in joins:
                rt.stream_wait(s, join)
        return launch

    def __call__(self, out_buf, in_buf=None, stream=None) -> None:
        """Transform ``in_buf`` (default: ``out_buf``, in place) into ``out_buf``;
        both complex64 buffers of ``total`` elements."""
        self.bind(out_buf, in_buf)(stream)

    def __del__(self):
        try:
            ptrs = [getattr(self, "_scr", 0), getattr(self, "_ctr", 0), getattr(self, "_wg_scr", 0)] + \
                list(getattr(self, "_task_cache", {}).values())
            if any(ptrs):
                rt.synchronize()
            for p in ptrs:
                if p:
                    rt.free(p)
        except Exception:
            pass

    def _abort_flag(self) -> bool:
        flag = np.zeros(1, np.int32)
        rt.dtoh(flag.ctypes.data, self._ctr + 4, 4)
        return bool(flag[0])

    def aborted(self) -> bool:
        """True when the last launch gave up a dependency wait (its result is invalid).
        Synchronises the stream; the flag is then consumed here instead of being raised by
        the next synchronize()."""
        s = rt.current_stream()
        rt.check(rt.lib().vkx_stream_synchronize(s), "vkx_stream_synchronize")
        rt.forget_async_check(self)
        return self._abort_flag()

    def async_error(self):
        """Message for synchronize() when the last counter-mode launch gave up a wait."""
        if self._abort_flag():
            return (f"ragged batch of {len(self.sizes)} transforms: a dependency wait gave up after "
                    f"{self.spin_limit or 1 << 20} polls; the result is invalid")
        return None

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


# ---- power-of-two length classes -------------------------------------------------
# A batch with many transforms of one power-of-two length runs them through the
# specialised kernels of vd.fft instead of through the persistent kernel: all of them
# as one virtual (count, L) batch whose read and write hooks map transform b to its
# offset in the packed buffers.  Lengths up to 16384 take one single-workgroup kernel
# (the whole transform in LDS, one pass over memory); 32768 and 65536 take two: column
# transforms of length n1 (strided by n2) with the twiddles W_L^(k1 j2) applied in the
# write hook, into a plan-owned scratch batch, then row transforms of length n2 written
# to X[k1 + n1 k2].  The persistent kernel's whole-transform path crosses memory once
# per radix pass (four times for these lengths); measured on uniform batches of the
# 4096-transform mix, the class kernels take 15.5 / 21.1 / 42.4 / 81.8 us where the
# persistent kernel takes 47.7 / 72.6 / 127.7 / 347.8 us (8192 x 235, 16384 x 210,
# 32768 x 195, 65536 x 236).


def class_lanes(classes: Sequence, streams: int) -> List[List[int]]:
    """Indices of ``classes`` per side stream: largest first (bytes_moved()), each to the
    least-loaded of at most ``streams`` lanes; empty lanes are dropped."""
    lanes: List[List[int]] = [[] for _ in range(max(1, int(streams)))]
    load = [0] * len(lanes)
    for i in sorted(range(len(classes)), key=lambda i: -classes[i].bytes_moved()):
        j = min(range(len(lanes)), key=lambda j: (load[j], j))
        lanes[j].append(i)
        load[j] += classes[i].bytes_moved()
    return [lane for lane in lanes if lane]


def class_groups(sizes: Sequence[int], min_count: int) -> Dict[int, np.ndarray]:
    """Indices of the transforms of each length in CLASS_LENGTHS that occurs at least
    ``min_count`` times (no classes when min_count <= 0)."""
    if min_count <= 0:
        return {}
    sizes = np.asarray(sizes)
    groups = {}
    for L in CLASS_LENGTHS:
        idx = np.nonzero(sizes == L)[0]
        if len(idx) >= min_count:
            groups[L] = idx
    return groups


def _class_hooks(L: int, n1: int, n2: int):
    """Read / write hooks of a length class: gather (transform b's element r at
    offs[b] + r), twiddle store (y[b][k1][j2] = v W^(k1 j2) from a length-L table) and
    transposed scatter (X[offs[b] + k1 + n1 k2])."""
    from ..shader_map import map as vd_map
    from .hooks import read_op, write_op

    @vd_map
    def gather(x: vc.Buffer[vc.c64], offs: vc.Buffer[vc.i32]):
        op = read_op()
        io = op.io_index
        op.read_from_buffer(x, io_index=offs[io // L] + io % L)

    @vd_map
    def scatter(out: vc.Buffer[vc.c64], offs: vc.Buffer[vc.i32]):
        op = write_op()
        io = op.io_index
        op.write_to_buffer(out, io_index=offs[io // L] + io % L)

    @vd_map
    def twiddle_store(y: vc.Buffer[vc.c64], tw: vc.Buffer[vc.c64]):
        op = write_op()
        w = vc.new_register(vc.c64)
        w[:] = tw[op.io_index % L]  # split_twiddles: W^((r / n2) (r % n2)) at r, read in order
        op.register[:] = vc.mult_complex(op.register, w)
        op.write_to_buffer(y)

    @vd_map
    def scatter_t(out: vc.Buffer[vc.c64], offs: vc.Buffer[vc.i32]):
        op = write_op()
        io = op.io_index
        r = io % L
        op.write_to_buffer(out, io_index=offs[io // L] + r // n2 + (r % n2) * n1)

    return gather, scatter, twiddle_store, scatter_t


def split_exponents(L: int) -> np.ndarray:
    """Exponent of W_L held at position r = a n2 + c of the split (n1, n2) =
    CLASS_SPLIT[L] twiddle table: (a c) mod L."""
    n1, n2 = CLASS_SPLIT[L]
    r = np.arange(L, dtype=np.int64)
    return ((r // n2) * (r % n2)) % L


def split_twiddles(L: int, inverse: bool = False) -> np.ndarray:
    """Twiddles of the split (n1, n2) = CLASS_SPLIT[L] in the order a launch over the
    (n1, n2) view meets them: entry r = a n2 + c holds exp(-+2 pi i a c / L) (FP64,
    rounded once), so the write hook reads the table at its own position, in order.
    Host statement of the table split_twiddle_buffer() builds on the device."""
    return np.exp((2j if inverse else -2j) * np.pi * split_exponents(L) / L).astype(np.complex64)


def split_twiddle_buffer(L: int, inverse: bool = False):
    """Device table of split_twiddles(L, inverse): the roots W_L^m generated in FP64 on
    the GPU (vkx_roots, rounded once to complex64), permuted to the split order (the
    inverse takes their conjugates, which is exact)."""
    from .. import dtypes as dt
    from ..buffer import Buffer
    roots = Buffer((L,), dt.complex64)
    rt.check(rt.lib().vkx_roots(roots.ptr, L, 0, rt.current_stream()), "vkx_roots")
    w = roots.read(0)[split_exponents(L)]
    out = Buffer((L,), dt.complex64)
    out.write(np.conj(w) if inverse else w)
    return out


def class_kernels(L: int, count: int, inverse: bool = False, normalize: bool = True) -> list:
    """The vd.fft kernels of a length class of ``count`` transforms of length L (no
    device needed: the code objects can be built ahead of time)."""
    from .. import dtypes as dt
    from .shader_factories import make_fft_shader
    c64 = dt.complex64
    common = dict(inverse=bool(inverse), normalize_inverse=bool(normalize), input_type=c64, output_type=c64,
                  compute_type=c64)
    if L not in CLASS_SPLIT:
        gather, scatter, _, _ = _class_hooks(L, L, 1)
        return [make_fft_shader((count, L), 1, input_map=gather, output_map=scatter, **common)]
    n1, n2 = CLASS_SPLIT[L]
    gather, _, twiddle_store, scatter_t = _class_hooks(L, n1, n2)
    shape = (count, n1, n2)
    return [make_fft_shader(shape, 1, input_map=gather, output_map=twiddle_store, **common),
            make_fft_shader(shape, 2, output_map=scatter_t, **common)]


class LengthClass:
    """All transforms of one power-of-two length L of a ragged batch (their element
    offsets ``offsets`` in the packed buffers), run through vd.fft kernels; see
    CLASS_LENGTHS.  bind() returns a launch(stream) that allocates nothing."""

    def __init__(self, L: int, offsets: np.ndarray, inverse: bool, normalize: bool):
        from .. import dtypes as dt
        from ..buffer import Buffer
        self.L, self.count = int(L), len(offsets)
        if self.L not in CLASS_LENGTHS:
            raise ValueError(f"no length class for {L}")
        if int(np.max(offsets)) + self.L > 2 ** 31 - 1:
            raise ValueError("length-class offsets must fit int32")
        self.offs = Buffer((self.count,), dt.int32)
        self.offs.write(np.asarray(offsets, np.int32))
        self.kernels = class_kernels(self.L, self.count, inverse, normalize)
        c64 = dt.complex64
        if self.L not in CLASS_SPLIT:
            self.scratch = self.tw = None
        else:
            self.scratch = Buffer((self.count * self.L,), c64)
            # W_L^(+-m) for m < L, generated in FP64 on the GPU (the inverse uses the conjugate)
            self.tw = split_twiddle_buffer(self.L, inverse)

    def bytes_moved(self) -> int:
        """Global-memory bytes of the class's launches (each reads and writes every point)."""
        return 16 * self.count * self.L * (1 if self.scratch is None else 2)

    def values(self, out_buf, in_buf) -> List[list]:
        if self.scratch is None:
            return [[out_buf, self.offs, in_buf, self.offs]]
        return [[self.scratch, self.tw, in_buf, self.offs], [out_buf, self.offs, self.scratch]]

    def bind(self, out_buf, in_buf):
        """Pointers and constants of every launch for this (output, input) pair, resolved
        now (code objects loaded, spill check done), so the returned launch(stream)
        makes only launch calls (safe under graph capture)."""
        packed = []
        for k, vals in zip(self.kernels, self.values(out_buf, in_buf)):
            k.handle()
            packed.append((k,) + tuple(k.bind(vals)))
        keep = (out_buf, in_buf, self)

        def launch(stream=None):
            _ = keep
            for k, ptrs, consts in packed:
                k.launch_raw(ptrs, consts, stream)
        return launch


# ---- Bluestein length classes -------------------------------------------------------
# Lengths with a prime factor above 13 (X_k = w_k sum_j x_j w_j conj(w_(k-j)),
# w_j = exp(-i pi j^2 / n)) of one Bluestein size m (the power of two >= 2n - 1) run as
# one class: a forward FFT_m of the chirped, zero-padded inputs times the chirp
# spectrum B (per transform, packed [count][m]), then an inverse FFT_m whose write
# hook keeps the first n outputs times the chirp.  The same semantics as the
# persistent kernel's Bluestein items (csrc/ragged.hip run_task): FP64 chirp tables,
# an inverse transform conjugates the chirp and uses conj(B[(m - k) mod m]).  m up to
# 16384 is one convolution launch (vd.fft.convolve's kernel: chirped gather, FFT_m,
# times the spectrum, inverse FFT_m, chirped output, all in LDS).  Larger m use the
# length-class split and three launches: forward columns with the twiddles, then one
# convolution along the rows (FFT over k2, times the spectrum, inverse FFT, conjugate
# twiddles), then inverse columns with the chirped output.


def blue_class_groups(sizes: Sequence[int], min_count: int) -> Dict[int, np.ndarray]:
    """Indices of the Bluestein transforms of each size m in BLUE_CLASS_LENGTHS shared by
    at least ``min_count`` transforms (no classes when min_count <= 0)."""
    if min_count <= 0:
        return {}
    sizes = np.asarray(sizes)
    ms = np.array([0 if is_direct(int(n)) else bluestein_length(int(n)) for n in sizes])
    groups = {}
    for m in BLUE_CLASS_LENGTHS:
        idx = np.nonzero(ms == m)[0]
        if len(idx) >= min_count:
            groups[m] = idx
    return groups


def _blue_hooks(m: int, inverse: bool):
    """Hooks of a Bluestein class of size m: chirped gather (j < n: x[offs + j] w_j, else
    0; indices clamped in range) and the chirped output (k < n: out[offs + k] =
    c_k w_k scale).  Both act at their own position r = io mod m: the first launch
    reads time order, and the last launch's positions are the natural time order (the
    spectrum multiply between them is the convolution's kernel map)."""
    from ..shader_map import map as vd_map
    from .hooks import read_op, write_op

    def chirp(ch, coff, b, j):
        w = vc.new_register(vc.c64)
        w[:] = ch[coff[b] + j]
        return w.conjugate() if inverse else w

    @vd_map
    def gather(x: vc.Buffer[vc.c64], offs: vc.Buffer[vc.i32], nn: vc.Buffer[vc.i32], ch: vc.Buffer[vc.c64],
               coff: vc.Buffer[vc.i32]):
        op = read_op()
        io = op.io_index
        b = (io // m).to_register()
        j = (io % m).to_register()
        n = nn[b].to_register()
        jj = vc.min(j, n - 1).to_register()
        op.read_from_buffer(x, io_index=offs[b] + jj)
        op.register[:] = vc.mult_complex(op.register, chirp(ch, coff, b, jj))
        with vc.if_block(j >= n):
            op.register[:] = vc.to_complex(0.0, 0.0)

    @vd_map
    def chirp_out(out: vc.Buffer[vc.c64], offs: vc.Buffer[vc.i32], nn: vc.Buffer[vc.i32], ch: vc.Buffer[vc.c64],
                  coff: vc.Buffer[vc.i32], scl: vc.Buffer[vc.f32]):
        op = write_op()
        io = op.io_index
        b = (io // m).to_register()
        k = (io % m).to_register()
        with vc.if_block(k < nn[b]):
            op.register[:] = vc.mult_complex(op.register, chirp(ch, coff, b, k)) * scl[b]
            op.write_to_buffer(out, io_index=offs[b] + k)

    return gather, chirp_out


def blue_spectrum_order(m: int) -> np.ndarray:
    """Frequency index held at position r of a class's forward FFT_m output: r itself
    for one launch; k1 + n1 k2 at r = k1 n2 + k2 after the columns and rows launches."""
    r = np.arange(m, dtype=np.int64)
    if m not in CLASS_SPLIT:
        return r
    n1, n2 = CLASS_SPLIT[m]
    return r // n2 + (r % n2) * n1


def blue_class_kernels(m: int, count: int, inverse: bool = False) -> list:
    """The vd.fft kernels of a Bluestein class (no device needed).  The spectrum goes in
    as the kernel of vd.fft.convolve's default map X . conj(K), so a class holds
    K = conj(spectrum)."""
    from .. import dtypes as dt
    from .shader_factories import make_convolution_shader, make_fft_shader
    c64 = dt.complex64
    fwd = dict(inverse=False, input_type=c64, output_type=c64, compute_type=c64)
    bwd = dict(inverse=True, normalize_inverse=True, input_type=c64, output_type=c64, compute_type=c64)
    conv = dict(normalize=True, input_type=c64, output_type=c64, kernel_type=c64, compute_type=c64)
    gather, chirp_out = _blue_hooks(m, inverse)
    if m not in CLASS_SPLIT:
        return [make_convolution_shader((count, m), None, 1, 1, input_map=gather, output_map=chirp_out, **conv)]
    n1, n2 = CLASS_SPLIT[m]
    _, _, tw_store, _ = _class_hooks(m, n1, n2)
    shape = (count, n1, n2)
    # forward columns over j1 (j = j1 n2 + j2) and twiddles; then the rows over j2, whose
    # output k = k1 + n1 k2 sits at r = k1 n2 + k2, times the spectrum packed in that
    # order, the inverse rows from that order and the conjugate twiddles; then inverse
    # columns over k1 -> c at r = j1' n2 + j2' = its natural index
    return [make_fft_shader(shape, 1, input_map=gather, output_map=tw_store, **fwd),
            make_convolution_shader(shape, None, 1, 2, output_map=tw_store, **conv),
            make_fft_shader(shape, 1, output_map=chirp_out, **bwd)]


class BlueClass:
    """The Bluestein transforms (lengths ``sizes``, element offsets ``offsets``) of one
    Bluestein size m of a ragged batch, run through vd.fft kernels; see
    BLUE_CLASS_LENGTHS.  bind() returns a launch(stream) that allocates nothing."""

    def __init__(self, m: int, sizes: np.ndarray, offsets: np.ndarray, inverse: bool, normalize: bool,
                 max_radix: int = 16):
        from .. import dtypes as dt
        from ..buffer import Buffer
        self.m, self.count = int(m), len(sizes)
        self.sizes = np.asarray(sizes, np.int64)
        sizes = np.asarray(sizes, np.int64)
        if self.m not in BLUE_CLASS_LENGTHS or any(bluestein_length(int(n)) != self.m for n in sizes):
            raise ValueError(f"not a Bluestein class of size {m}")
        if int(np.max(offsets + sizes)) > 2 ** 31 - 1 or int(sizes.sum()) > 2 ** 31 - 1:
            raise ValueError("Bluestein-class offsets must fit int32")
        T = tables()
        c64 = dt.complex64
        self.offs = Buffer((self.count,), dt.int32)
        self.offs.write(np.asarray(offsets, np.int32))
        self.nn = Buffer((self.count,), dt.int32)
        self.nn.write(sizes.astype(np.int32))
        coff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
        self.coff = Buffer((self.count,), dt.int32)
        self.coff.write(coff)
        self.scl = Buffer((self.count,), dt.float32)
        self.scl.write(((1.0 / sizes) if (inverse and normalize) else np.ones(self.count)).astype(np.float32))
        # packed FP64-generated chirps (sum of n); chirp spectra ([count][m]) in the order
        # the forward FFT_m leaves its output, conjugated and reflected for an inverse
        self.ch = Buffer((int(sizes.sum()),), c64)
        rads = runtime_radices(self.m, max_radix)
        s = rt.current_stream()
        order = blue_spectrum_order(self.m)
        if inverse:
            order = (self.m - order) % self.m
        spec = np.empty((self.count, self.m), np.complex64)
        for i, n in enumerate(sizes.tolist()):
            rt.dtod(self.ch.ptr + int(coff[i]) * 8, T.chirp_table(n), n * 8, s)
            B = np.empty(self.m, np.complex64)
            rt.dtoh(B.ctypes.data, T.chirp_spectrum(n, self.m, rads), self.m * 8, s)
            rt.synchronize(s)
            spec[i] = np.conj(B[order]) if inverse else B[order]
        # the kernel of the convolution launch: X . conj(K) with K = conj(spectrum)
        self.ks = Buffer((self.count * self.m,), c64)
        self.ks.write(np.conj(spec).reshape(-1))
        self.kernels = blue_class_kernels(self.m, self.count, inverse)
        self.y = self.z = self.tw = self.twi = None
        if self.m in CLASS_SPLIT:
            self.y = Buffer((self.count * self.m,), c64)
            self.z = Buffer((self.count * self.m,), c64)
            self.tw = split_twiddle_buffer(self.m)
            self.twi = split_twiddle_buffer(self.m, inverse=True)
        rt.synchronize()

    def bytes_moved(self) -> int:
        """Global-memory bytes of the class's launches: the n inputs and outputs and the
        kernel spectrum, plus two passes over the m-point intermediate per split launch
        boundary."""
        n_sum = int(self.sizes.sum())
        base = 16 * n_sum + 8 * self.count * self.m
        return base + (32 * self.count * self.m if self.m in CLASS_SPLIT else 0)

    def values(self, out_buf, in_buf) -> List[list]:
        io = [self.offs, self.nn, self.ch, self.coff]
        if self.m not in CLASS_SPLIT:  # output-map args, input-map args, kernel
            return [[out_buf] + io + [self.scl, in_buf] + io + [self.ks]]
        return [[self.y, self.tw, in_buf] + io,          # chirped columns, twiddles
                [self.z, self.twi, self.y, self.ks],     # rows * spectrum (rows order) * inverse rows, conj twiddles
                [out_buf] + io + [self.scl, self.z]]     # inverse columns, chirped first n outputs

    def bind(self, out_buf, in_buf):
        packed = []
        for k, vals in zip(self.kernels, self.values(out_buf, in_buf)):
            k.handle()
            packed.append((k,) + tuple(k.bind(vals)))
        keep = (out_buf, in_buf, self)

        def launch(stream=None):
            _ = keep
            for k, ptrs, consts in packed:
                k.launch_raw(ptrs, consts, stream)
        return launch


def packed_normal(sizes: Sequence[int], seed: int = 0):
    """Per-transform N(0,1) real then imaginary parts, in size order, packed complex64."""
    rng = np.random.default_rng(seed)
    parts = [(rng.standard_normal(n) + 1j * rng.standard_normal(n)) for n in sizes]
    return parts, np.concatenate(parts).astype(np.complex64)


TUNE_TILE_MIN = (16385, 65536, TILE_MIN_DIRECT)  # direct lengths tiled from ... / never
TUNE_KEYS = ("block", "block_first", "max_radix", "tile_min", "classes", "grid")  # RaggedFFT arguments chosen by tune()
TUNE_TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_ragged.json")


def batch_key(sizes: Sequence[int]) -> str:
    """Cache key of a ragged batch: its length and a hash of its length list (in order)."""
    a = np.ascontiguousarray(np.asarray(sizes, dtype=np.int64))
    return f"{a.size}:{hashlib.sha1(a.tobytes()).hexdigest()[:16]}"


def load_tune_table(path: str = TUNE_TABLE_PATH) -> Dict:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_tune_table(table: Dict, path: str = TUNE_TABLE_PATH) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def tune(sizes: Sequence[int], buf, reps: int = 5, cache: bool = False, save: bool = False,
         path: str = TUNE_TABLE_PATH) -> Dict:
    """Pick the workgroup sizes (first launch, which holds the whole transforms, and
    the tile launches), the largest radix, the direct length from which transforms
    are tiled, whether length classes are used and the persistent grid (half a
    workgroup per CU, leaving CUs to the class launches, one or two), for this batch by
    timing each candidate plan (deterministic candidate order; ties keep the earlier
    candidate).  The best entry's keys are RaggedFFT arguments.  cache: return the
    entry stored for (device arch, batch) in the table at ``path`` without timing,
    when there is one; save: store the result."""
    key = f"{rt.device_arch()}|{batch_key(sizes)}"
    if cache:
        hit = load_tune_table(path).get(key)
        if hit is not None:
            return {"best": dict(hit), "results": [], "cached": True}
    best = None
    results = []
    cus = int(rt.ensure_device().compute_units)
    for classes in (CLASS_MIN_COUNT, 0):
        for block, block_first in ((256, 256), (512, 512)):
            for max_radix in (8, 16):
                for tile_min in TUNE_TILE_MIN:
                    for grid in (cus // 2, cus, 2 * cus):
                        plan = RaggedFFT(sizes, block=block, max_radix=max_radix, tile_min=tile_min,
                                         block_first=block_first, classes=classes, grid=grid)
                        run = plan.bind(buf)
                        run()
                        e0, e1 = rt.Event(), rt.Event()
                        e0.record()
                        for _ in range(reps):
                            run()
                        e1.record()
                        e1.synchronize()
                        ms = e0.elapsed_ms(e1) / reps
                        results.append({"block": block, "block_first": block_first, "max_radix": max_radix,
                                        "tile_min": tile_min, "classes": classes, "grid": grid, "ms": ms})
                        if best is None or ms < best["ms"]:
                            best = results[-1]
                        del run, plan
    if save:
        table = load_tune_table(path)
        table[key] = {k: best[k] for k in TUNE_KEYS}
        save_tune_table(table, path)
    return {"best": best, "results": results, "cached": False}


def tuned_plan(sizes: Sequence[int], buf, inverse: bool = False, normalize: bool = True) -> "RaggedFFT":
    """RaggedFFT with the cached (or freshly measured and stored) tuning for this batch."""
    best = tune(sizes, buf, cache=True, save=True)["best"]
    return RaggedFFT(sizes, inverse=inverse, normalize=normalize, **tuned_args(best))


def tuned_args(best: Dict) -> Dict:
    """RaggedFFT keyword arguments of a tuning entry (keys an older entry lacks keep
    their defaults)."""
    return {k: best[k] for k in TUNE_KEYS if k in best}
