"""CPU checks of the ragged work-item queue (vkdispatch_amd/fft/ragged.py): each item
of a planned batch is executed in NumPy FP64 with the semantics the ragged kernel
gives its kind (csrc/ragged.hip: column tile, row tile, the Bluestein levels), with
every index checked against its region, every wait checked against the counters
raised before it, and the result compared with numpy.fft.  No device is used."""
import numpy as np
import pytest

from vkdispatch_amd.fft import ragged as R


class _FakeTables:
    """Stands in for the device tables: addresses are never dereferenced here."""

    def twiddles(self, n, radices):
        return 1

    def twiddles_split(self, split):
        return 2

    def chirp_table(self, n):
        return 3

    def chirp_spectrum(self, n, m, radices):
        return 4


def _chirp(n, inverse):
    j = np.arange(n, dtype=np.int64)
    w = np.exp(-1j * np.pi * ((j * j) % (2 * n)) / n)
    return np.conj(w) if inverse else w


def _kernel_spectrum(n, m):
    b = np.zeros(m, np.complex128)
    j = np.arange(m, dtype=np.int64)
    jj = np.where(j < n, j, np.where(j > m - n, m - j, -1))
    ok = jj >= 0
    b[ok] = np.exp(1j * np.pi * ((jj[ok] * jj[ok]) % (2 * n)) / n)
    return np.fft.fft(b)


def emulate(q, sizes, x, inverse, normalize=True):
    sizes = list(sizes)
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int)
    bufs = {R.SRC_IN: x.astype(np.complex128).copy(), R.SRC_OUT: np.zeros(x.size, np.complex128),
            R.SRC_SCR: np.full(max(1, q.scratch_elems), np.nan + 0j)}
    size_at = dict(zip(offsets.tolist(), sizes))
    done = {}
    seg_of = np.zeros(len(q.items), int)
    for s_i, (a, b) in enumerate(q.segments):
        seg_of[a:b] = s_i
    wseg = np.full(max(1, q.scratch_elems), -1)  # launch segment that wrote each scratch element
    for k, t in enumerate(q.items):
        v = int(t["npass"])
        npass, np1, kind, cj, lw = v & 0xFF, (v >> 8) & 0xFF, (v >> 16) & 0xF, (v >> 20) & 1, (v >> 21) & 0xF
        n, m = int(t["n"]), int(t["m"])
        (sb, so), (db, do) = q.src[k], q.dst[k]
        src, dst = bufs[sb], bufs[db]
        if t["dep"] >= 0:
            assert done.get(int(t["dep"]), 0) == int(t["target"]), f"item {k}: wait not satisfied"
        if kind == R.K_WHOLE:
            seg = src[so:so + n]
            res = (np.fft.ifft(seg) * n if inverse else np.fft.fft(seg)) * float(t["scale"])
            dst[do:do + n] = res
            continue
        L = m if m else n
        rads = [int(r) for r in t["radix"][:npass]]
        n1 = int(np.prod(rads[:np1]))
        n2 = L // n1
        assert n1 * n2 == L and int(np.prod(rads)) == L
        # region limits: the user buffers hold n points at the transform's offset, scratch L
        src_len = n if sb != R.SRC_SCR else L
        dst_len = n if db != R.SRC_SCR else L
        if sb != R.SRC_SCR:
            assert size_at[so] == n
        cnt, t0 = int(t["count"]), int(t["tile0"])
        assert cnt >= 1
        seg = seg_of[k]

        def read_scr(ix):  # scratch written by an earlier launch, or by a counter-ordered item
            if sb == R.SRC_SCR:
                assert (wseg[so + ix] >= 0).all() and (t["dep"] >= 0 or (wseg[so + ix] < seg).all())
            return src[so + ix]

        def note_write(ix):
            if db == R.SRC_SCR:
                wseg[do + ix] = seg
        if kind in (R.K_COL, R.K_BLUE_COL):
            # tiles sit in LDS at pitch n1 + 1: one buffer, or two for a pass list of <= 2
            assert t0 + cnt <= n2 and cnt <= (1 << lw)
            assert (1 << lw) * (n1 + 1) <= R.LDS_TILE * (2 if np1 <= 2 else 1)
            j2 = t0 + np.arange(cnt)
            idx = np.arange(n1)[:, None] * n2 + j2[None, :]
            if kind == R.K_COL:
                assert idx.max() < src_len
                col = read_scr(idx)
                if cj:
                    col = np.conj(col)
            else:
                w = _chirp(n, inverse)
                col = np.zeros(idx.shape, np.complex128)
                ok = idx < n
                col[ok] = src[so + idx[ok]] * w[idx[ok]]
            f = np.fft.fft(col, axis=0)
            k1 = np.arange(n1)[:, None]
            oidx = k1 * n2 + j2[None, :]
            assert oidx.max() < dst_len
            dst[do + oidx] = f * np.exp(-2j * np.pi * ((k1 * j2[None, :]) % L) / L)
            note_write(oidx)
        else:
            np2 = npass - np1
            assert t0 + cnt <= n1 and cnt <= (1 << lw)
            assert (1 << lw) * (n2 + 1) <= R.LDS_TILE * (2 if np2 <= 2 else 1)
            r = t0 + np.arange(cnt)
            idx = r[:, None] * n2 + np.arange(n2)[None, :]
            assert idx.max() < src_len
            f = np.fft.fft(read_scr(idx), axis=1)
            kk = r[:, None] + n1 * np.arange(n2)[None, :]
            if kind == R.K_ROW:
                assert kk.max() < dst_len
                dst[do + kk] = (np.conj(f) if cj else f) * float(t["scale"])
            elif kind == R.K_BLUE_ROW:
                assert kk.max() < dst_len
                B = _kernel_spectrum(n, m)
                b = np.conj(B[(m - kk) & (m - 1)]) if inverse else B[kk]
                dst[do + kk] = np.conj(f * b)
                note_write(kk)
            else:
                assert kind == R.K_BLUE_OUT
                ok = kk < n
                w = _chirp(n, inverse)
                dst[do + kk[ok]] = np.conj(f[ok]) * w[kk[ok]] * float(t["scale"])
        if t["done"] >= 0:
            done[int(t["done"])] = done.get(int(t["done"]), 0) + 1
    return bufs[R.SRC_OUT]


SIZES = [64, 97, 100, 4096, 4099, 5003, 8192, 12288, 59049, 65536, 65521, 60060, 131072, 3 * 5 * 7 * 11 * 13 * 8]


@pytest.mark.parametrize("sync,fill", [("counter", 0), ("launch", 0), ("launch", 4096)])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("max_radix", [8, 16])
def test_item_queue_computes_the_dft(inverse, max_radix, sync, fill):
    rng = np.random.default_rng(11)
    x = rng.standard_normal(sum(SIZES)) + 1j * rng.standard_normal(sum(SIZES))
    q = R.plan_items(SIZES, inverse, True, max_radix, True, _FakeTables(), sync, tile_min=4097, fill=fill)
    if fill:  # the short whole transforms sit in the later launches
        later = set(range(q.segments[1][0], len(q.items)))
        small = [k for k, t in enumerate(q.items) if ((int(t["npass"]) >> 16) & 0xF) == R.K_WHOLE and t["n"] <= fill
                 and t["m"] == 0]
        assert small and set(small) <= later
    if sync == "launch":
        assert len(q.segments) == 4 and (q.items["dep"] < 0).all() and (q.items["done"] < 0).all()
    kinds = {(int(v) >> 16) & 0xF for v in q.items["npass"]}
    assert kinds == {R.K_WHOLE, R.K_COL, R.K_ROW, R.K_BLUE_COL, R.K_BLUE_ROW, R.K_BLUE_OUT}
    out = emulate(q, SIZES, x, inverse)
    off = 0
    for n in SIZES:
        seg = x[off:off + n]
        ref = np.fft.ifft(seg) if inverse else np.fft.fft(seg)
        err = np.linalg.norm(out[off:off + n] - ref) / np.linalg.norm(ref)
        assert err < 2e-7, f"n={n}: rel-L2 {err:.2e}"  # the records hold 1/n as float32
        off += n


def test_tile_min_keeps_shorter_direct_lengths_whole():
    q = R.plan_items(SIZES, False, True, 16, True, _FakeTables(), "launch", tile_min=60000)
    whole_n = {int(t["n"]) for t in q.items if ((int(t["npass"]) >> 16) & 0xF) == R.K_WHOLE}
    tiled_direct = {int(t["n"]) for t in q.items if ((int(t["npass"]) >> 16) & 0xF) in (R.K_COL, R.K_ROW)
                    and int(t["m"]) == 0}
    assert tiled_direct == {n for n in SIZES if n >= 60000 and R.is_direct(n) and R.four_step_split(n)}
    assert {8192, 12288, 59049} <= whole_n


def test_item_queue_without_four_step_is_whole_transforms():
    q = R.plan_items(SIZES, False, True, 16, False, _FakeTables())
    assert all(((int(v) >> 16) & 0xF) == R.K_WHOLE for v in q.items["npass"])
    assert len(q.items) == len(SIZES) and q.ncounters == 2


def test_queue_order_check_rejects_a_wait_on_later_work():
    q = R.plan_items([8192], False, True, 16, True, _FakeTables(), "counter", tile_min=4097)
    items = q.items.copy()
    R.check_queue_order(items)
    with pytest.raises(AssertionError):
        R.check_queue_order(items[::-1])


def test_large_batch_tiles_fit_and_counters_are_dense():
    sizes = R.mixed_sizes(512, seed=1)
    q = R.plan_items(sizes, False, True, 16, True, _FakeTables(), "counter", tile_min=4097)
    deps = q.items["dep"][q.items["dep"] >= 0]
    dones = q.items["done"][q.items["done"] >= 0]
    assert set(deps.tolist()) <= set(dones.tolist())
    assert q.ncounters == 2 + len(set(dones.tolist()))
    assert (q.items["count"] >= 1).all()


def test_tune_table_key_and_round_trip(tmp_path):
    a = R.mixed_sizes(64, seed=3)
    assert R.batch_key(a) == R.batch_key(list(a)) and R.batch_key(a) != R.batch_key(a[::-1])
    path = str(tmp_path / "t.json")
    assert R.load_tune_table(path) == {}
    R.save_tune_table({"gfx950|" + R.batch_key(a): {"block": 256, "block_first": 512, "max_radix": 16,
                                                   "tile_min": 4097}}, path)
    assert R.load_tune_table(path)["gfx950|" + R.batch_key(a)]["tile_min"] == 4097
    shipped = R.load_tune_table()
    assert all(set(v) == set(R.TUNE_KEYS) for v in shipped.values())


def test_auto_tile_min():
    assert R.auto_tile_min([65536] * 2, 512) == 16384
    assert R.auto_tile_min([8192] * 32, 512) > 8192  # short lengths stay whole
    assert R.auto_tile_min(R.mixed_sizes(4096, seed=0), 512) == R.TILE_MIN_DIRECT
    assert R.auto_tile_min([65536] * 600, 512) == R.TILE_MIN_DIRECT  # enough long transforms to fill the GPU


def test_class_groups_pick_frequent_power_of_two_lengths():
    sizes = [8192] * 16 + [16384] * 15 + [65536] * 20 + [4096] * 40 + [12288] * 30 + [32768] * 16 + [32] * 20
    g = R.class_groups(sizes, 16)
    assert sorted(g) == [4096, 8192, 32768, 65536]  # 16384: too few; 12288: not a power of two; 32: too short
    assert list(g[8192]) == list(range(16)) and len(g[65536]) == 20
    assert R.class_groups(sizes, 0) == {}


@pytest.mark.parametrize("L", [32768, 65536])
@pytest.mark.parametrize("inverse", [False, True])
def test_length_class_split_computes_the_dft(L, inverse):
    """The two-launch length class as its hooks index it (csrc-independent restatement
    of _class_hooks): transform b is gathered from offs[b], column transforms of length
    n1 over j1 (stride n2), times tw[k1 j2] (exp(-+2 pi i m / L)), row transforms of
    length n2, scattered to offs[b] + k1 + n1 k2; each stage normalised by its own
    length on the inverse.  Checked against numpy.fft on a packed batch with gaps."""
    n1, n2 = R.CLASS_SPLIT[L]
    rng = np.random.default_rng(L % 97 + inverse)
    count = 3
    offs = np.array([5, 5 + L + 7, 5 + 2 * L + 20])
    total = offs[-1] + L + 3
    buf = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    tw = R.split_twiddles(L, inverse).astype(complex)
    io = np.arange(count * L)
    b, r = io // L, io % L
    x = buf[offs[b] + r].reshape(count, n1, n2)                      # gather
    y = (np.fft.ifft(x, axis=1) if inverse else np.fft.fft(x, axis=1)).reshape(-1)
    y = y * tw[r]                                                      # twiddle store at its own position
    z = (np.fft.ifft(y.reshape(count, n1, n2), axis=2) if inverse
         else np.fft.fft(y.reshape(count, n1, n2), axis=2)).reshape(-1)
    out = np.zeros(total, complex)
    out[offs[b] + r // n2 + (r % n2) * n1] = z                          # transposed scatter
    for k in range(count):
        seg = buf[offs[k]:offs[k] + L]
        ref = np.fft.ifft(seg) if inverse else np.fft.fft(seg)
        got = out[offs[k]:offs[k] + L]
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-7  # complex64 twiddle table


@pytest.mark.parametrize("L", [8192, 16384, 32768, 65536])
def test_length_class_kernels_compile(L):
    """Each class's kernels (gather / scatter hooks, twiddle store) build for gfx950
    without spills."""
    from vkdispatch_amd.fft import kernels as K
    for k in R.class_kernels(L, 17) + R.class_kernels(L, 17, inverse=True):
        assert K.compile_source(k.source)[:4] == b"\x7fELF"
        assert k.code_object_info()["private_bytes"] == 0, k.name


def test_blue_class_groups_pick_frequent_bluestein_sizes():
    sizes = [4093] * 16 + [4099] * 9 + [8191] * 11 + [65521] * 3 + [30011] * 16 + [4096] * 50
    g = R.blue_class_groups(sizes, 16)
    assert sorted(g) == [8192, 16384, 65536]  # m >= 2n - 1: 4093 -> 8192; 4099, 8191 -> 16384; 30011 -> 65536
    assert len(g[16384]) == 20 and R.blue_class_groups(sizes, 0) == {}


@pytest.mark.parametrize("m,ns", [(8192, (4093, 4001, 2053)), (131072, (65521, 40009, 32771))])
@pytest.mark.parametrize("inverse", [False, True])
def test_bluestein_class_pipeline_computes_the_dft(m, ns, inverse):
    """A Bluestein class as its hooks compute it (restated with numpy): chirped, clamped
    and masked gather; forward FFT_m (for a split size: columns over j1, twiddles read
    at their own position, rows over j2, output left in the rows order r = k1 n2 + k2);
    times the spectrum packed in that order (conjugated and reflected on the inverse;
    the class stores its conjugate as the convolution kernel K of X . conj(K)); inverse
    FFT_m from that order (rows, conjugate twiddles, columns), normalised by 1/m, which
    lands in natural order; the first n outputs times the (conjugate) chirp and the
    transform's scale.  One convolution launch for m <= 16384; for a split size the
    rows FFT, the product and the inverse rows are one convolution launch between the
    two column launches.  Against numpy.fft for primes with gaps between the
    transforms in the packed buffer."""
    rng = np.random.default_rng(m % 89 + inverse)
    sizes = np.array(ns)
    assert all(R.bluestein_length(int(n)) == m for n in sizes)
    offs = np.concatenate([[3], 3 + np.cumsum(sizes + 5)[:-1]])
    total = int(offs[-1] + sizes[-1] + 2)
    x = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    count = len(sizes)
    io = np.arange(count * m)
    b, j = io // m, io % m
    jj = np.minimum(j, sizes[b] - 1)
    w = [_chirp(int(n), inverse) for n in sizes]
    a = np.array([x[offs[bb] + q] * w[bb][q] for bb, q in zip(b, jj)])
    a[j >= sizes[b]] = 0
    order = R.blue_spectrum_order(m)
    if inverse:
        order = (m - order) % m
    B = np.stack([_kernel_spectrum(int(n), m) for n in sizes])
    bs = np.conj(B[:, order]) if inverse else B[:, order]
    if m not in R.CLASS_SPLIT:
        C = np.fft.fft(a.reshape(count, m), axis=1) * bs
        c = np.fft.ifft(C, axis=1)
    else:
        n1, n2 = R.CLASS_SPLIT[m]
        tw, twi = R.split_twiddles(m).astype(complex), R.split_twiddles(m, inverse=True).astype(complex)
        y = np.fft.fft(a.reshape(count, n1, n2), axis=1).reshape(count, m) * tw
        C = np.fft.fft(y.reshape(count, n1, n2), axis=2).reshape(count, m) * bs
        y = np.fft.ifft(C.reshape(count, n1, n2), axis=2).reshape(count, m) * twi
        c = np.fft.ifft(y.reshape(count, n1, n2), axis=1).reshape(count, m)
    out = np.zeros(total, complex)
    for bb, n in enumerate(sizes):
        scale = 1.0 / n if inverse else 1.0
        out[offs[bb]:offs[bb] + n] = c[bb, :n] * w[bb][:n] * scale
    for bb, n in enumerate(sizes):
        seg = x[offs[bb]:offs[bb] + n]
        ref = np.fft.ifft(seg) if inverse else np.fft.fft(seg)
        got = out[offs[bb]:offs[bb] + n]
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-7  # complex64 tables


@pytest.mark.parametrize("m", [8192, 32768, 131072])
def test_bluestein_class_kernels_compile(m):
    from vkdispatch_amd.fft import kernels as K
    for k in R.blue_class_kernels(m, 17) + R.blue_class_kernels(m, 17, inverse=True):
        assert K.compile_source(k.source)[:4] == b"\x7fELF"
        assert k.code_object_info()["private_bytes"] == 0, k.name


def test_class_lanes_balance_by_bytes():
    """Classes go largest first to the least-loaded lane (ties: the lower lane); every
    class lands in exactly one lane, at most ``streams`` lanes, none empty."""
    class Fake:
        def __init__(self, b):
            self.b = b

        def bytes_moved(self):
            return self.b

    cs = [Fake(b) for b in (10, 70, 30, 30, 5, 40)]
    lanes = R.class_lanes(cs, 2)
    assert sorted(i for lane in lanes for i in lane) == list(range(len(cs)))
    # 70 -> lane 0, 40 -> lane 1, 30 -> lane 1 (70 = 70), 30 -> lane 0 (tie), 10 and 5 -> lane 1
    assert lanes == [[1, 3], [5, 2, 0, 4]]
    assert [sum(cs[i].b for i in lane) for lane in lanes] == [100, 85]
    assert R.class_lanes(cs[:1], 3) == [[0]]  # empty lanes are dropped
    assert R.class_lanes([], 2) == []
    assert len(R.class_lanes(cs, 8)) == len(cs)
