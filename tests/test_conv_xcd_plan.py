"""CPU checks of the single-launch 2-D convolution (conv2d_xcd_body in
csrc/device/vkx_conv2d.h) before it runs on a GPU:

* the task queues (fft/fused.py:xcd_task_table) under an emulation of the
  kernel's scheduling -- workgroups take tickets in queue order and hold the next
  ticket while they run a task, tasks wait on done counters -- for any number of
  workgroups per XCD and random task durations: every launch drains (no deadlock)
  and no task reads a slot block that is not in the state it expects or touches a
  block another running task writes;
* every address the kernel forms (slot scratch, images, half spectrum K, control
  block, task table) for the shipped configurations, computed with the kernel's own
  index formulas, stays inside its buffer;
* the data flow itself (pair packing, tile-major slot layout, Kt from the half
  spectrum, slot reuse by the CA tasks), executed task by task in NumPy FP64 in an
  emulated order, reproduces the oracle's convolve2DR.
"""
import random

import numpy as np
import pytest

import oracle as O
from vkdispatch_amd.fft import fused

XT_A, XT_B, XT_CA, XT_C = 0, 1, 2, 3
NUM_XCD = 8


def _queues(table):
    offs = table[:9]
    return [[int(w) for w in table[9 + offs[q]:9 + offs[q + 1]]] for q in range(NUM_XCD)]


def _decode(w):
    return w >> 30, (w >> 12) & 0x3FFFF, w & 0xFFF


class _Slot:
    """State of one slot image as (slice, tile) blocks: (local pair, 'R' rows done |
    'B' columns done) or None."""

    def __init__(self, ns, nb):
        self.v = [[None] * nb for _ in range(ns)]
        self.busy = {}  # block -> 'r' | 'w' holders count


def _task_blocks(kind, s, ns, nb):
    if kind == XT_B:
        return [(r, s) for r in range(ns)]
    return [(s, c) for c in range(nb)]


def _emulate(npairs, ns, nb, nslot, wg_per_queue, seed):
    table = fused.xcd_task_table(npairs, ns, nb, nslot)
    rnd = random.Random(seed)
    queues = _queues(table)
    for q, tasks in enumerate(queues):
        n = len(range(q, npairs, NUM_XCD))
        slots = [_Slot(ns, nb) for _ in range(nslot)]
        rows_done = [0] * n
        tiles_done = [0] * n
        ticket = 0
        # workgroup: [current task index or None, next ticket, state, remaining, exited]
        wgs = [{"cur": None, "nxt": None, "state": "idle", "left": 0} for _ in range(wg_per_queue)]
        finished = 0
        running = {}  # task idx -> wg
        steps = 0
        while finished < len(tasks):
            steps += 1
            assert steps < 10_000_000
            progress = False
            order = list(range(len(wgs)))
            rnd.shuffle(order)
            for wi in order:
                wg = wgs[wi]
                if wg["state"] == "exit":
                    continue
                if wg["state"] == "idle":
                    if wg["nxt"] is None:  # first ticket
                        wg["nxt"] = ticket
                        ticket += 1
                    cur = wg["nxt"]
                    if cur >= len(tasks):
                        wg["state"] = "exit"
                        progress = True
                        continue
                    wg["cur"] = cur
                    wg["nxt"] = ticket  # the next ticket is requested before the task runs
                    ticket += 1
                    wg["state"] = "wait"
                    progress = True
                if wg["state"] == "wait":
                    kind, i, s = _decode(tasks[wg["cur"]])
                    ok = (kind == XT_A or (kind == XT_B and rows_done[i] >= ns) or
                          (kind in (XT_CA, XT_C) and tiles_done[i] >= nb))
                    if not ok:
                        continue
                    # start: check block states and that nobody else is using the blocks
                    slot = slots[i % nslot]
                    blocks = _task_blocks(kind, s, ns, nb)
                    for b in blocks:
                        assert b not in slot.busy, f"queue {q}: task {wg['cur']} overlaps a running task"
                        st = slot.v[b[0]][b[1]]
                        if kind == XT_A:
                            assert st is None or st == (i - nslot, "B"), f"A({i}) over {st}"
                        elif kind == XT_B:
                            assert st == (i, "R"), f"B({i}) reads {st}"
                        else:
                            assert st == (i, "B"), f"C/CA({i}) reads {st}"
                        slot.busy[b] = True
                    wg["state"] = "run"
                    wg["left"] = rnd.randint(1, 6)
                    running[wg["cur"]] = wi
                    progress = True
                    continue
                if wg["state"] == "run":
                    wg["left"] -= 1
                    progress = True
                    if wg["left"] > 0:
                        continue
                    kind, i, s = _decode(tasks[wg["cur"]])
                    slot = slots[i % nslot]
                    for b in _task_blocks(kind, s, ns, nb):
                        del slot.busy[b]
                        if kind == XT_A:
                            slot.v[b[0]][b[1]] = (i, "R")
                        elif kind == XT_B:
                            slot.v[b[0]][b[1]] = (i, "B")
                        elif kind == XT_CA:
                            slot.v[b[0]][b[1]] = (i + nslot, "R")
                        else:
                            slot.v[b[0]][b[1]] = ("out", i)
                    if kind == XT_A:
                        rows_done[i] += 1
                    elif kind == XT_B:
                        tiles_done[i] += 1
                    elif kind == XT_CA:
                        rows_done[i + nslot] += 1
                    del running[wg["cur"]]
                    finished += 1
                    wg["state"] = "idle"
            assert progress, f"queue {q}: no workgroup can progress (deadlock) with {wg_per_queue} workgroups"
        # every pair went through rows, tiles and out
        assert all(r == ns for r in rows_done) and all(t == nb for t in tiles_done)


@pytest.mark.parametrize("npairs", [1, 5, 8, 9, 32, 33])
@pytest.mark.parametrize("wg", [1, 2, 3, 32])
def test_xcd_queues_drain_without_hazards(npairs, wg):
    _emulate(npairs, ns=8, nb=8, nslot=2, wg_per_queue=wg, seed=npairs * 100 + wg)


@pytest.mark.parametrize("seed", range(3))
def test_xcd_queues_full_size_emulation(seed):
    """The 64 x 512^2 workload's queues with the shipped tiling (64 row slices, 64 tiles
    per pair, 128 workgroups per XCD), as launched."""
    cfg = fused._XCD_TILING[512]
    _emulate(32, ns=512 // cfg["rows"], nb=512 // cfg["ct"], nslot=cfg["nslot"], wg_per_queue=cfg["wg_per_xcd"],
             seed=seed)


def test_xcd_task_words_fit_their_fields():
    t = fused.xcd_task_table(32, 64, 64, 2)
    q = _queues(t)
    # per pair: 64 row tasks, 64 tiles, 64 row-inverse tasks, where the row tasks of all
    # but the first two pairs of a queue ride in the CA tasks of the pair before
    assert sum(len(x) for x in q) == 32 * 64 * 3 - 16 * 64
    for tasks in q:
        for w in tasks:
            kind, i, s = _decode(w)
            assert kind in (XT_A, XT_B, XT_CA, XT_C) and i < 4 and s < 64


# ---------------------------------------------------------------------------
# index bounds of every access, with the kernel's formulas
# ---------------------------------------------------------------------------
def _tile_index(H, ky, kx, ct):
    return (kx // ct) * (ct * H) + ky * ct + (kx % ct)


@pytest.mark.parametrize("batch", [64, 15, 16, 17, 1, 40])
@pytest.mark.parametrize("H", sorted(fused._XCD_TILING))
def test_xcd_addresses_stay_in_bounds(batch, H):
    W = H
    cfg = fused._XCD_TILING[H]
    nw, rows, ct, nslot = cfg["waves"], cfg["rows"], cfg["ct"], cfg["nslot"]
    ns, nb = H // rows, W // ct
    npairs = (batch + 1) // 2
    qmax = -(-npairs // 8)
    pitch, img, kw = 2 * (W // 2 + 1), H * 2 * (W // 2 + 1), W // 2 + 1
    scratch_elems = 8 * nslot * H * W
    image_floats = batch * img
    ctl_words = 512 + 64 * 8 * qmax
    table = fused.xcd_task_table(npairs, ns, nb, nslot)
    lanes = np.arange(64)
    tids = np.arange(64 * nw)
    np_ = W + W // 32 + 1
    lds_elems = max(ct, nw) * np_
    for q, tasks in enumerate(_queues(table)):
        for w in tasks:
            kind, i, s = _decode(w)
            pair = q + 8 * i
            base = (q * nslot + i % nslot) * H * W
            assert 0 <= pair < npairs
            # done counters (A at +0, B at +32; CA signals pair i + nslot)
            for ii in ([i + nslot] if kind == XT_CA else [i]):
                assert ii < qmax
                assert 512 + 64 * (q * qmax + ii) + 32 < ctl_words
            if kind in (XT_A, XT_CA, XT_C):
                assert s < ns
                for wave in range(nw):
                    for j in range(rows // nw):
                        y = s * rows + wave + nw * j
                        assert y < H
                        kx = lanes[:, None] + 64 * np.arange(W // 64)[None, :]
                        idx = base + _tile_index(H, y, kx, ct)
                        assert idx.min() >= 0 and idx.max() < scratch_elems
                        for p in ([pair, pair + 8 * nslot] if kind == XT_CA else [pair]):
                            assert p < npairs
                            fa = 2 * p * img + y * pitch + kx
                            assert fa.max() < image_floats
                            if 2 * p + 1 < batch:
                                assert (fa + img).max() < image_floats
                    assert wave * np_ + np_ <= lds_elems  # row exchange region
            else:
                assert s < nb
                e = tids[:, None] + np.arange(ct * H // (64 * nw))[None, :] * 64 * nw
                idx = base + s * ct * H + e
                assert idx.max() < scratch_elems
                ky, c = e // ct, e % ct
                assert (c * np_ + ky + (ky >> 5)).max() < lds_elems
                # Kt columns of this tile (column-major H x W)
                kx = s * ct + np.arange(ct)
                assert (kx[:, None] * H + np.arange(H)[None, :]).max() < H * W


def test_kt_prep_covers_every_element_once():
    """conv2d_kt_body: block b (64 x 64) of the column-major Kt, one per workgroup
    (conv_kprep's index formulas), writes every element of H x W exactly once."""
    for H in sorted(fused._XCD_TILING):
        W, KB, nt = H, 64, 64 * fused._XCD_TILING[H]["waves"]
        hits = np.zeros(H * W, np.int32)
        for b in range((H // KB) * (W // KB)):
            ky0, kx0 = (b // (W // KB)) * KB, (b % (W // KB)) * KB
            for i in range(KB * KB // nt):
                e = np.arange(nt) + i * nt
                ky, kx = ky0 + e % KB, kx0 + e // KB
                np.add.at(hits, kx * H + ky, 1)
        assert (hits == 1).all()


# ---------------------------------------------------------------------------
# the data flow in NumPy FP64, task by task
# ---------------------------------------------------------------------------
def _kt(K, H, W, scale):
    Kw = W // 2 + 1
    Kf = np.zeros((H, W), np.complex128)
    ky = np.arange(H)[:, None]
    kym = (H - ky) % H
    for kx in range(W):
        if kx <= W // 2:
            v = K[ky[:, 0], kx]
        else:
            v = np.conj(K[kym[:, 0], W - kx])
        if kx in (0, W // 2):
            v = 0.5 * (K[ky[:, 0], kx] + np.conj(K[kym[:, 0], kx]))
        Kf[:, kx] = v
    assert Kf.shape == (H, W) and K.shape == (H, Kw)
    return np.conj(Kf) * scale


def test_xcd_dataflow_in_numpy_matches_oracle():
    rng = np.random.default_rng(3)
    batch, H = 21, 64          # 11 pairs: queues 0..2 hold two pairs, the rest one
    W = H
    cfg = fused._XCD_TILING[512]
    rows, ct, nslot = cfg["rows"], cfg["ct"], cfg["nslot"]
    ns, nb = H // rows, W // ct
    x = rng.uniform(-1, 1, (batch, H, W))
    K = (rng.standard_normal((H, W // 2 + 1)) + 1j * rng.standard_normal((H, W // 2 + 1)))
    Kt = _kt(K, H, W, 1.0 / (H * W))
    npairs = (batch + 1) // 2
    out = np.zeros_like(x)
    slots = np.zeros((8, nslot, W // ct, H, ct), np.complex128)  # tile-major: [tile][ky][kx % ct]

    def rows_fwd(q, i, s):
        p = q + 8 * i
        for y in range(s * rows, (s + 1) * rows):
            z = x[2 * p, y] + 1j * (x[2 * p + 1, y] if 2 * p + 1 < batch else 0)
            Z = np.fft.fft(z)
            for kx in range(W):
                slots[q, i % nslot, kx // ct, y, kx % ct] = Z[kx]

    def tile(q, i, c):
        T = slots[q, i % nslot, c]          # (H, 16)
        F = np.fft.fft(T, axis=0) * Kt[:, c * ct:(c + 1) * ct]
        slots[q, i % nslot, c] = np.fft.ifft(F, axis=0) * H

    def rows_inv(q, i, s):
        p = q + 8 * i
        for y in range(s * rows, (s + 1) * rows):
            Z = np.array([slots[q, i % nslot, kx // ct, y, kx % ct] for kx in range(W)])
            z = np.fft.ifft(Z) * W
            out[2 * p, y] = z.real
            if 2 * p + 1 < batch:
                out[2 * p + 1, y] = z.imag

    table = fused.xcd_task_table(npairs, ns, nb, nslot)
    for q, tasks in enumerate(_queues(table)):
        for w in tasks:   # queue order is a valid execution order (one workgroup)
            kind, i, s = _decode(w)
            if kind == XT_A:
                rows_fwd(q, i, s)
            elif kind == XT_B:
                tile(q, i, s)
            elif kind == XT_CA:
                rows_inv(q, i, s)
                rows_fwd(q, i + nslot, s)
            else:
                rows_inv(q, i, s)
    ref = O.convolve2DR(x, K.astype(np.complex128), kernel_inner_only=True)
    assert O.rel_l2(out, ref) < 1e-12
