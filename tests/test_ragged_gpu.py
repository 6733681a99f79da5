"""Ragged mixed-size batches, Bluestein and long transforms (GPU), against FP64."""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd
from vkdispatch_amd.fft import ragged as R

pytestmark = pytest.mark.gpu


def ragged_sizes(count, seed=0, lo=64, hi=65536):
    return R.mixed_sizes(count, seed=seed, lo=lo, hi=hi)


def _packed(sizes, seed):
    return R.packed_normal(sizes, seed)


def _check(sizes, parts, out, inverse=False, tol=1e-5):
    off = 0
    worst = 0.0
    for n, x in zip(sizes, parts):
        ref = np.fft.ifft(x.astype(np.complex64).astype(np.complex128)) if inverse else \
            np.fft.fft(x.astype(np.complex64).astype(np.complex128))
        e = O.rel_l2(out[off:off + n], ref)
        worst = max(worst, e)
        assert e <= tol, f"n={n}: rel-L2 {e:.2e}"
        off += n
    return worst


@pytest.mark.parametrize("n", [67, 97, 101, 257, 1031, 4099, 65521])
def test_bluestein_primes(n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    buf = vd.asbuffer(x)
    vd.fft.fft(buf)
    assert O.rel_l2(buf.read(0), np.fft.fft(x.astype(np.complex128), axis=1)) < 1e-5
    vd.fft.ifft(buf)
    assert O.rel_l2(buf.read(0), x) < 2e-5


@pytest.mark.parametrize("n", [32768, 65536, 59049, 2 * 3 * 5 * 7 * 11 * 13 * 2 ** 4, 1 << 20])
def test_long_transforms(n):
    rng = np.random.default_rng(n % 1000)
    x = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(np.complex64)
    buf = vd.asbuffer(x)
    vd.fft.fft(buf)
    assert O.rel_l2(buf.read(0), np.fft.fft(x.astype(np.complex128), axis=1)) < 1e-5


def test_ragged_batch_small():
    sizes = ragged_sizes(300, seed=3, hi=8192)
    parts, packed = _packed(sizes, 4)
    buf = vd.asbuffer(packed)
    plan = R.RaggedFFT(sizes)
    plan(buf)
    assert not plan.aborted()
    _check(sizes, parts, buf.read(0))
    inv = R.RaggedFFT(sizes, inverse=True)
    inv(buf)
    assert not inv.aborted()
    assert O.rel_l2(buf.read(0), packed) < 1e-5


def test_ragged_out_of_place_and_graph():
    sizes = [64, 100, 4096, 7 * 11 * 13, 97, 65536, 3 ** 8]
    parts, packed = _packed(sizes, 5)
    src = vd.asbuffer(packed)
    dst = vd.Buffer((packed.size,), vd.complex64)
    plan = R.RaggedFFT(sizes)
    plan(dst, src)
    _check(sizes, parts, dst.read(0))
    assert np.array_equal(src.read(0), packed)
    g = vd.CommandGraph()
    g.record_callable(plan.bind(dst, src))
    for _ in range(3):
        g.submit()
        vd.synchronize()
        _check(sizes, parts, dst.read(0))


@pytest.mark.parametrize("block", [256, 512])
def test_ragged_4096_mixed_sizes_every_transform(block):
    """4096 transforms, lengths 64..65536 (60% pow2, 30% 7-smooth, 10% primes via
    Bluestein), N(0,1) data from seed 0: every transform within rel-L2 1e-5 of FP64,
    with the persistent kernel's 256- and 512-thread forms (the latter keeps whole
    transforms up to 10232 points in LDS)."""
    sizes = ragged_sizes(4096, seed=0)
    parts, packed = _packed(sizes, 0)
    buf = vd.asbuffer(packed)
    plan = R.RaggedFFT(sizes, block=block)
    plan(buf)
    assert not plan.aborted()
    worst = _check(sizes, parts, buf.read(0))
    assert worst < 1e-5


@pytest.mark.parametrize("block", [256, 320, 512])
def test_ragged_whole_transforms_lds_and_global(block):
    """Whole transforms (no classes, no tiles) on either side of the LDS limits (4224 /
    10232 points) and long ones through global passes that alternate with the output,
    first passes staged through LDS: forward out of place, inverse in place."""
    sizes = [4096, 4225, 5000, 6561, 10000, 10232, 10240, 12288, 19683, 40000, 60060, 65536, 11 * 13 * 256]
    parts, packed = _packed(sizes, 31)
    src = vd.asbuffer(packed)
    dst = vd.Buffer((packed.size,), vd.complex64)
    plan = R.RaggedFFT(sizes, block=block, classes=0, four_step=False)
    plan(dst, src)
    assert not plan.aborted()
    assert _check(sizes, parts, dst.read(0)) < 1e-5
    inv = R.RaggedFFT(sizes, inverse=True, block=block, classes=0, four_step=False)
    inv(dst)
    assert O.rel_l2(dst.read(0), packed) < 1e-5


@pytest.mark.parametrize("four_step,sync", [(False, "launch"), (True, "launch"), (True, "counter")])
def test_ragged_long_and_bluestein_both_paths(four_step, sync):
    """Long smooth lengths and Bluestein primes through the tiled items (one launch per
    level, or one launch with counter waits) and through whole-transform items,
    forward and inverse."""
    sizes = [8192, 65536, 4099, 59049, 12288, 65521, 100, 5003, 131072, 60060]
    parts, packed = _packed(sizes, 9)
    src = vd.asbuffer(packed)
    dst = vd.Buffer((packed.size,), vd.complex64)
    plan = R.RaggedFFT(sizes, four_step=four_step, max_radix=8 if four_step else 16, sync=sync, tile_min=4097)
    plan(dst, src)
    assert not plan.aborted()
    _check(sizes, parts, dst.read(0))
    inv = R.RaggedFFT(sizes, inverse=True, four_step=four_step, sync=sync, tile_min=4097)
    inv(dst)
    assert not inv.aborted()
    assert O.rel_l2(dst.read(0), packed) < 1e-5


def test_ragged_length_classes_forward_inverse_and_graph():
    """Power-of-two lengths with at least CLASS_MIN_COUNT transforms run as length
    classes (vd.fft kernels with gather / scatter hooks; two launches for 32768 and
    65536), interleaved in the packed buffer with items of the persistent kernel:
    forward out of place, graph replays, inverse in place, against FP64."""
    rng = np.random.default_rng(17)
    sizes = [8192] * 16 + [16384] * 16 + [32768] * 16 + [65536] * 16 + [100, 4099, 60060, 8192, 12288] * 4
    sizes = [int(n) for n in rng.permutation(sizes)]
    parts, packed = _packed(sizes, 17)
    src = vd.asbuffer(packed)
    dst = vd.Buffer((packed.size,), vd.complex64)
    plan = R.RaggedFFT(sizes)
    print(f"length classes {[c.L for c in plan._classes]}, {plan.n_items} persistent-kernel items", flush=True)
    assert sorted(c.L for c in plan._classes) == [8192, 16384, 32768, 65536]
    assert {int(n) for n in plan._items["n"]} == {100, 4099, 60060, 12288}  # every 8192 is in its class
    plan(dst, src)
    assert not plan.aborted()
    _check(sizes, parts, dst.read(0))
    assert np.array_equal(src.read(0), packed)
    g = vd.CommandGraph()
    g.record_callable(plan.bind(dst, src))
    for _ in range(3):
        dst.write(np.zeros_like(packed))
        g.submit()
        vd.synchronize()
        _check(sizes, parts, dst.read(0))
    inv = R.RaggedFFT(sizes, inverse=True)
    inv(dst)
    assert not inv.aborted()
    assert O.rel_l2(dst.read(0), packed) < 1e-5


def _primes_between(lo, hi, count, rng):
    sieve = np.ones(hi + 1, bool)
    sieve[:2] = False
    for i in range(2, int(hi ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    ps = np.nonzero(sieve)[0]
    ps = ps[ps > lo]
    return [int(p) for p in rng.choice(ps, count, replace=False)]


def test_ragged_bluestein_classes_forward_inverse_and_graph():
    """Bluestein transforms sharing a size m (16 each for m = 8192, 65536 and 131072)
    run as Bluestein classes (chirped gather, spectrum multiply, chirped output around
    two FFT_m of the vd.fft kernels), next to items of the persistent kernel: forward
    out of place, graph replays, inverse in place, against FP64."""
    rng = np.random.default_rng(29)
    blue = _primes_between(2048, 4096, 16, rng) + _primes_between(16384, 32768, 16, rng) + \
        _primes_between(32768, 65536, 16, rng)
    sizes = [int(n) for n in rng.permutation(blue + [100, 4099, 60060, 12288, 97] * 3)]
    parts, packed = _packed(sizes, 29)
    src = vd.asbuffer(packed)
    dst = vd.Buffer((packed.size,), vd.complex64)
    plan = R.RaggedFFT(sizes)
    ms = sorted(c.m for c in plan._classes if isinstance(c, R.BlueClass))
    print(f"Bluestein classes {ms}, {plan.n_items} persistent-kernel items", flush=True)
    assert ms == [8192, 65536, 131072]
    plan(dst, src)
    assert not plan.aborted()
    _check(sizes, parts, dst.read(0))
    g = vd.CommandGraph()
    g.record_callable(plan.bind(dst, src))
    for _ in range(2):
        dst.write(np.zeros_like(packed))
        g.submit()
        vd.synchronize()
        _check(sizes, parts, dst.read(0))
    inv = R.RaggedFFT(sizes, inverse=True)
    inv(dst)
    assert not inv.aborted()
    assert O.rel_l2(dst.read(0), packed) < 1e-5


@pytest.mark.parametrize("class_streams,persistent_first", [(1, False), (2, True), (3, False)])
def test_ragged_class_schedules(class_streams, persistent_first):
    """The same batch on one, two or three side streams, with the persistent launch
    recorded before or after the classes: short power-of-two classes (64..4096),
    small-m Bluestein classes (m = 256 and 4096) and persistent-kernel items, forward
    out of place, graph replays and inverse in place, every transform against FP64."""
    rng = np.random.default_rng(41 + class_streams)
    blue = _primes_between(100, 127, 5, rng) * 4 + _primes_between(1030, 2047, 16, rng)  # m = 256, 4096
    pow2 = [64] * 20 + [512] * 17 + [4096] * 16 + [1024] * 5  # 1024: too few for a class
    smooth = [100, 4099, 60060, 12288, 97, 3 ** 9, 5 ** 5]
    sizes = [int(n) for n in rng.permutation(blue + pow2 + smooth * 2)]
    parts, packed = _packed(sizes, 41)
    src = vd.asbuffer(packed)
    dst = vd.Buffer((packed.size,), vd.complex64)
    plan = R.RaggedFFT(sizes, class_streams=class_streams, persistent_first=persistent_first)
    kinds = sorted((type(c).__name__, getattr(c, "L", getattr(c, "m", 0))) for c in plan._classes)
    print(f"classes {kinds}, {plan.n_items} persistent-kernel items", flush=True)
    assert kinds == [("BlueClass", 256), ("BlueClass", 4096), ("LengthClass", 64), ("LengthClass", 512),
                     ("LengthClass", 4096)]
    plan(dst, src)
    assert not plan.aborted()
    _check(sizes, parts, dst.read(0))
    g = vd.CommandGraph()
    g.record_callable(plan.bind(dst, src))
    for _ in range(2):
        dst.write(np.zeros_like(packed))
        g.submit()
        vd.synchronize()
        _check(sizes, parts, dst.read(0))
    inv = R.RaggedFFT(sizes, inverse=True, class_streams=class_streams, persistent_first=persistent_first)
    inv(dst)
    assert not inv.aborted()
    assert O.rel_l2(dst.read(0), packed) < 1e-5


def test_uniform_long_batch_runs_as_a_length_class():
    """vd.fft.fft on (32, 65536): the length is beyond the single-workgroup kernels, the
    batch is one length class."""
    rng = np.random.default_rng(23)
    x = (rng.standard_normal((32, 65536)) + 1j * rng.standard_normal((32, 65536))).astype(np.complex64)
    buf = vd.asbuffer(x)
    vd.fft.fft(buf)
    assert O.rel_l2(buf.read(0), np.fft.fft(x.astype(np.complex128), axis=1)) < 1e-5
    vd.fft.ifft(buf)
    assert O.rel_l2(buf.read(0), x) < 1e-5


def test_ragged_wait_gives_up_and_the_launch_ends():
    """An item whose wait target is never reached (set from the host) polls its
    counter spin_limit times, sets the abort flag and the launch still ends."""
    sizes = [8192, 64]
    _, packed = _packed(sizes, 2)
    buf = vd.asbuffer(packed)
    plan = R.RaggedFFT(sizes, spin_limit=1000, sync="counter", tile_min=4097)
    waits = np.nonzero(plan._items["dep"] >= 0)[0]
    assert waits.size > 0
    plan._items["target"][waits[0]] += 1000
    plan(buf)
    with pytest.raises(RuntimeError, match="gave up"):
        vd.synchronize()  # the deferred check reports the abandoned wait
    assert plan.aborted()
    ok = R.RaggedFFT(sizes, spin_limit=1000, sync="counter", tile_min=4097)
    buf.write(packed)
    ok(buf)
    assert not ok.aborted()


def test_tune_is_cached_per_batch(tmp_path):
    sizes = ragged_sizes(48, seed=5, hi=20000)
    _, packed = _packed(sizes, 5)
    buf = vd.asbuffer(packed)
    path = str(tmp_path / "tuned.json")
    first = R.tune(sizes, buf, reps=1, cache=True, save=True, path=path)
    # classes on / off x two block sizes x two largest radices x tile thresholds x two grids
    assert not first["cached"] and len(first["results"]) == 2 * 2 * 2 * len(R.TUNE_TILE_MIN) * 3
    again = R.tune(sizes, buf, reps=1, cache=True, path=path)
    assert again["cached"] and {k: again["best"][k] for k in R.TUNE_KEYS} == {k: first["best"][k] for k in R.TUNE_KEYS}


def test_graph_replay_after_another_plan_needs_more_scratch():
    """A captured plan launches with its own scratch: a later plan with larger
    per-workgroup scratch does not move it, and replays stay correct."""
    sizes = [65536, 100, 8192]
    parts, packed = _packed(sizes, 12)
    src = vd.asbuffer(packed)
    dst = vd.Buffer((packed.size,), vd.complex64)
    plan_a = R.RaggedFFT(sizes, four_step=False)  # whole transforms: per-workgroup scratch
    g = vd.CommandGraph()
    g.record_callable(plan_a.bind(dst, src))
    g.capture()
    g.submit()
    vd.synchronize()
    _check(sizes, parts, dst.read(0))
    big = [1 << 18]
    bparts, bpacked = _packed(big, 13)
    bbuf = vd.asbuffer(bpacked)
    plan_b = R.RaggedFFT(big, four_step=False)
    assert plan_b.scratch_per_wg > plan_a.scratch_per_wg
    plan_b(bbuf)
    vd.synchronize()
    _check(big, bparts, bbuf.read(0))
    dst.write(np.zeros_like(packed))
    for _ in range(2):
        g.submit()
        vd.synchronize()
        _check(sizes, parts, dst.read(0))
