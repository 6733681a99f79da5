"""CPU-only tests: the C-ABI library, the planner, precision rules, hook tracing and
kernel compilation (hiprtc cross-compiles gfx950 without a GPU).  No test here
launches device work."""
import os
import re
from types import SimpleNamespace

import numpy as np
import pytest

import vkdispatch_amd as vd
import vkdispatch_amd.codegen as vc
import vkdispatch_amd.fft.functions as fft_functions
from vkdispatch_amd import runtime as rt
from vkdispatch_amd.fft import kernels as K
from vkdispatch_amd.fft import planner as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vkdfft.h")


def _header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(vkx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    names = _header_functions()
    assert len(names) >= 30
    lib = rt.lib()
    for n in names:
        assert hasattr(lib, n), f"libvkdfft.so does not export {n}"
    assert set(names) == set(rt.exported_symbols())


def test_library_basic_calls_without_device():
    lib = rt.lib()
    assert lib.vkx_version() == 100
    lib.vkx_clear_error()
    assert lib.vkx_last_error() == b""


# ---------------------------------------------------------------------------
# planner
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [n for n in range(1, 65) if n == 1 or max(P.prime_factors(n)) <= 31] +
                         [81, 100, 128, 162, 243, 256, 500, 512, 720, 1000, 1024, 1859, 2048, 4096, 8192, 10000,
                          16384])
def test_radix_plan_invariants(n):
    radices, tt = P.radix_plan(n)
    assert int(np.prod(radices)) == n
    assert 1 <= tt <= 1024
    assert all(r <= 31 or n == r for r in radices)
    vmax = max(P.slots(n, r, tt) for r in radices)
    assert vmax <= 36
    # every butterfly of every pass is covered by some thread
    for r in radices:
        assert -(-(n // r) // tt) * tt >= n // r


def test_plan_geometry_matches_reference_axis_rules():
    g = P.AxisGeometry.of((64, 512, 257), 1)
    assert (g.N, g.S, g.OUTER) == (512, 257, 64)
    g = P.AxisGeometry.of((7, 5, 3), 0)
    assert (g.N, g.S, g.OUTER) == (7, 15, 1)
    g = P.AxisGeometry.of((1024, 4096))
    assert (g.N, g.S, g.OUTER) == (4096, 1, 1024)


def test_plan_lds_and_threads_within_gfx950_limits():
    for shape, axis in [((1024, 4096), 1), ((64, 512, 257), 1), ((256, 256, 256), 0), ((2, 16384), 1),
                        ((3, 4096, 7), 1), ((10, 10000), 1)]:
        plan = P.make_plan(P.AxisGeometry.of(shape, axis))
        assert plan.BLOCK <= 1024
        assert plan.lds_bytes <= 160 * 1024
        assert plan.NWG >= 1


def test_xcd_tile_order_only_for_rows_off_the_cache_line():
    """Kernels whose rows or strides start off a 128-byte line, or that carry hooks, take
    their tiles in XCD-contiguous order (grid a multiple of 8); aligned ones keep the
    plain order.  The logical order is a permutation of the grid."""
    f = vd.fft
    on = lambda k: "XCDMAP = true" in k.source  # noqa: E731
    assert not on(f.make_fft_shader((1024, 4096), None))            # 32 KB rows
    assert not on(f.make_fft_shader((256, 256, 256), 1))            # 2 KB stride
    assert on(f.make_fft_shader((32, 1024, 513), 1))                # 4104-byte stride
    assert on(f.make_fft_shader((2000, 100), None))                 # 800-byte rows, 80 workgroups
    assert not on(f.make_fft_shader((64, 100), None))               # 3 workgroups: not a multiple of 8
    assert on(f.make_convolution_shader((32, 1024, 513), None, 1, 1, kernel_inner_only=True))
    assert not on(f.make_convolution_shader((64, 512, 256), None, 1, 1, kernel_inner_only=True))
    from vkdispatch_amd.fft import autograd
    p = autograd.plans_for(32, 1024, 1024, vd.complex64)
    assert not on(p.rows_fwd) and not on(p.rows_inv)                # real rows: plain order (measured)
    for k in (f.make_fft_shader((32, 1024, 513), 1), f.make_fft_shader((2000, 100), None)):
        G = k.plan.NWG
        assert G % 8 == 0 and f"NWG = {G}u" in k.source
        b = np.arange(G)
        assert sorted(((b & 7) * (G // 8) + (b >> 3)).tolist()) == list(range(G))


def test_pad_dim_and_primes_like_reference():
    from vkdispatch_amd.fft.prime_utils import pad_dim, prime_factors, group_primes
    assert prime_factors(360) == [2, 2, 2, 3, 3, 5]
    assert group_primes([2, 2, 2, 3, 3, 5], 16) == [[2, 2, 2], [3, 3], [5]]
    assert pad_dim(17) == 18 and pad_dim(1021) == 1024 and pad_dim(512) == 512
    with pytest.raises(ValueError):
        pad_dim(0)


def test_large_prime_needs_bluestein():
    with pytest.raises(ValueError):
        P.radix_plan(2 * 37)


@pytest.mark.parametrize("L", [4096, 4200, 8192, 59049, 60060, 65536, 131072, 120120, 2 * 3 * 5 * 7 * 11 * 13 * 16,
                               1 << 18, 1 << 20, 4099 * 2, 19 * 4096])
def test_four_step_split(L):
    """Lengths above one LDS buffer split into n1 * n2 with both factors <= 512 (so a
    tile holds >= 8 padded columns or rows) and at most 12 radices in all."""
    from vkdispatch_amd.fft import ragged as R
    s = R.four_step_split(L)
    if L <= R.LDS_PINGPONG:
        assert s is None
        return
    if L > R.FOUR_STEP_MAX ** 2 or L in (4099 * 2, 19 * 4096):  # too long / a prime above 13: no split
        assert s is None
        return
    assert s is not None
    r1, r2 = s
    n1, n2 = int(np.prod(r1)), int(np.prod(r2))
    assert n1 * n2 == L and max(n1, n2) <= R.FOUR_STEP_MAX and len(r1) + len(r2) <= 12
    assert all(r in (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16) for r in r1 + r2)
    t = np.zeros(1, R.RTASK_DTYPE)[0]
    R._encode_passes(t, r1 + r2, s, R.K_ROW, R.CONJ)
    v = int(t["npass"])
    assert v & 0xff == len(r1) + len(r2) and (v >> 8) & 0xff == len(r1) and (v >> 16) & 0xf == R.K_ROW
    assert (v >> 20) & 1 == 1
    assert list(t["radix"][:len(r1) + len(r2)]) == r1 + r2


# ---------------------------------------------------------------------------
# precision rules (ports of the reference's fake-buffer tests,
# tests/test_fft_mixed_precision.py:270-320)
# ---------------------------------------------------------------------------
class _FakeBuffer:
    def __init__(self, var_type):
        self.var_type = var_type


def test_resolve_input_precision_output_map_infers_input_from_post_map_argument(monkeypatch):
    monkeypatch.setattr(fft_functions.prec, "ensure_supported_complex_precision", lambda dtype, role: None)
    output_map = SimpleNamespace(buffer_types=[vc.Buffer[vd.complex64], vc.Buffer[vd.float32]])
    resolved = fft_functions._resolve_input_precision(
        (_FakeBuffer(vd.complex64), _FakeBuffer(vd.float32), _FakeBuffer(vd.complex128)),
        input_map=None, output_map=output_map, input_type=None, output_precision=None)
    assert resolved is vd.complex128


def test_resolve_input_precision_output_map_requires_input_buffer_after_map_args(monkeypatch):
    monkeypatch.setattr(fft_functions.prec, "ensure_supported_complex_precision", lambda dtype, role: None)
    output_map = SimpleNamespace(buffer_types=[vc.Buffer[vd.complex64]])
    with pytest.raises(ValueError, match="input buffer argument must be provided"):
        fft_functions._resolve_input_precision((_FakeBuffer(vd.complex64),), input_map=None, output_map=output_map,
                                               input_type=None, output_precision=None)


def test_compute_precision_defaults():
    from vkdispatch_amd.fft import precision as prec
    assert prec.resolve_compute_precision([vd.complex32], None) is vd.complex64
    assert prec.resolve_compute_precision([vd.complex32, vd.complex128], None) is vd.complex128
    assert prec.resolve_compute_precision([], None) is vd.complex64
    assert prec.resolve_compute_precision([vd.complex64], vd.complex128) is vd.complex128
    with pytest.raises(ValueError):
        prec.resolve_compute_precision([vd.float32], None)


def test_in_place_type_mismatch_rejected():
    with pytest.raises(ValueError, match="in-place"):
        fft_functions._resolve_input_precision((_FakeBuffer(vd.complex64),), None, None, vd.complex128, vd.complex64)


# ---------------------------------------------------------------------------
# hook tracing and kernel compilation
# ---------------------------------------------------------------------------
def test_trace_default_and_hooked_sources():
    src = vd.fft.make_fft_shader((8, 256), None).source
    assert "vkx::fft_body<Cfg, HookIn, HookOut>" in src and "DefaultIn<0, float>" in src

    @vd.map
    def scale(s: vc.Var[vc.f32]):
        op = vd.fft.read_op()
        op.register[:] = op.register * s

    src = vd.fft.make_convolution_shader((8, 256), scale, 1, None).source
    assert "vkx::cval<float>(prm, 0)" in src and "conv_body" in src


def test_hook_argument_order_matches_reference():
    """output-map args, then input-map args, then kernel-map args (io_manager.py:86-102)."""
    @vd.map
    def om(a: vc.Buffer[vd.complex64], s: vc.Var[vd.float32]):
        vd.fft.write_op().write_to_buffer(a)

    @vd.map
    def im(b: vc.Buffer[vd.float32]):
        op = vd.fft.read_op()
        op.register.real = b[op.io_index]

    @vd.map
    def km(c: vc.Buffer[vd.complex64], t: vc.Const[vd.int32]):
        pass

    k = vd.fft.make_convolution_shader((4, 64), km, 1, None, input_map=im, output_map=om)
    kinds = [(a.role, a.kind, a.dtype.name, a.slot) for a in k.args]
    assert kinds == [("output", "buffer", "complex64", 0), ("output", "scalar", "float32", 0),
                     ("input", "buffer", "float32", 1), ("kernel", "buffer", "complex64", 2),
                     ("kernel", "scalar", "int32", 4)]


def test_tracer_expressions():
    with vc.trace_scope() as b:
        r = vc.new_register(vd.complex64)
        u = vc.new_uint_register(5)
        r[:] = vc.mult_complex(r, r.conjugate())
        r.real = r.imag * 2
        idx = vc.ravel_index(u, (3, 4, 5))
        flat = vc.unravel_index(idx, (3, 4, 5))
        v = flat.to_register()
    text = "\n".join(b.lines)
    assert "vkx::cmul" in text and "vkx::conj" in text and ".x =" in text
    assert v.var_type is vd.uint32


def test_bind_packs_buffers_and_scalars():
    @vd.map
    def km(s: vc.Var[vd.float32], c: vc.Const[vd.complex64]):
        pass

    k = vd.fft.make_convolution_shader((4, 64), km, 1, None)
    buf = SimpleNamespace(ptr=0x1000, size=256, var_type=vd.complex64)
    ptrs, consts = k.bind([buf, 2.5, 1 - 2j])
    assert ptrs == [0x1000]
    assert np.frombuffer(consts[:4], np.float32)[0] == 2.5
    assert np.allclose(np.frombuffer(consts[8:16], np.float32), [1, -2])
    with pytest.raises(ValueError):
        k.bind([buf])


@pytest.mark.parametrize("make", [
    lambda: vd.fft.make_fft_shader((1024, 4096), None),
    lambda: vd.fft.make_fft_shader((64, 512, 257), 1, inverse=True),
    lambda: vd.fft.make_fft_shader((6, 162), None, r2c=True, inverse=True),
    lambda: vd.fft.make_fft_shader((16, 720), None, compute_type=vd.complex128, output_type=vd.complex128,
                                   input_type=vd.complex128),
    lambda: vd.fft.make_convolution_shader((64, 512, 257), None, 1, 1, kernel_inner_only=True),
    lambda: vd.fft.make_transpose_shader((3, 64, 33), 1),
])
def test_kernels_compile_for_gfx950(make, tmp_path, monkeypatch):
    monkeypatch.setenv("VKX_KERNEL_CACHE", str(tmp_path))
    code = K.compile_source(make().source)
    assert code[:4] == b"\x7fELF"


def test_pair_conv_kernels_do_not_spill():
    """Every standard pair-packed convolution kernel keeps its values in registers
    (a spilling 16-wave variant gave wrong rows on the MI355X; the plans refuse
    spilling kernels at launch, so the defaults must not spill)."""
    from vkdispatch_amd.fft import fused
    for batch, n, norm in fused._STANDARD_FUSED:
        for k in fused.fused_plan(batch, n, n, norm).kernels():
            info = k.code_object_info()
            assert info["private_bytes"] == 0, (k.name, info)


def test_map_resolves_postponed_annotations():
    """Hook parameters annotated under ``from __future__ import annotations`` (strings)
    resolve to the same buffer types."""
    src = ("from __future__ import annotations\n"
           "import vkdispatch_amd as vd\n"
           "import vkdispatch_amd.codegen as vc\n"
           "def hook(buf: vc.Buffer[vd.complex64], s: vc.Var[vd.float32]):\n"
           "    pass\n")
    ns = {}
    exec(compile(src, "<postponed>", "exec"), ns)
    m = vd.map(ns["hook"])
    import vkdispatch_amd.codegen as vc
    assert m.buffer_types == [vc.Buffer[vd.complex64], vc.Var[vd.float32]]


def test_tuned_plans_do_not_spill():
    """Every plan of the committed tuned table compiles without register spills (the
    tuner refuses spilling candidates; a spilling kernel has given wrong results on
    the MI355X), and every table entry belongs to a shape listed here."""
    from vkdispatch_amd.fft import autotune as AT
    from vkdispatch_amd.fft import kernels as K
    from vkdispatch_amd.fft import planner as P
    plans = AT.load_table().get("plans", {})
    jobs = list(AT.STANDARD_JOBS) + [("fft", (32, 1024, 513), 1)]
    seen, built = set(), []
    for kind, shape, axis in jobs:
        geom, tgeom = AT._tuning_geometry(kind, shape, axis)
        key = P.plan_key(tgeom, "float", kind)
        if key not in plans or key in seen:
            continue
        seen.add(key)
        e = plans[key]
        cand = (tuple(e["radices"]), e["tt"], e.get("B"), bool(e.get("twreg", False)))
        built.append((key, AT._build(kind, shape, axis, geom, "float", cand)))
    assert seen == set(plans), f"tuned entries without a shape here: {sorted(set(plans) - seen)}"
    K.precompile([k for _, k in built])
    for key, k in built:
        assert k.code_object_info()["private_bytes"] == 0, (key, k.code_object_info())


def test_vkfft_zero_blocks_map_to_read_ranges_and_compile(tmp_path, monkeypatch):
    """vd.vkfft padding (VkFFT's zero block [left, right) per axis): blocks at an end of the
    axis become a kept range, interior blocks a wrapped kept range; both compile."""
    from vkdispatch_amd import vkfft
    from vkdispatch_amd.fft.shader_factories import ZeroBlock, _resolve_signal_range
    assert vkfft._signal_range(None, 32) is None and vkfft._signal_range((5, 5), 32) is None
    assert vkfft._signal_range((24, 32), 32) == (0, 24)
    assert vkfft._signal_range((0, 8), 32) == (8, 32)
    zb = vkfft._signal_range((8, 16), 32)
    assert isinstance(zb, ZeroBlock) and _resolve_signal_range(zb, 32) == (16, 8)
    with pytest.raises(ValueError):
        vkfft._signal_range((8, 40), 32)
    with pytest.raises(ValueError):
        _resolve_signal_range((16, 8), 32)  # the public argument stays an ordered range
    monkeypatch.setenv("VKX_KERNEL_CACHE", str(tmp_path))
    k = vd.fft.make_convolution_shader((3, 32, 32), None, 1, 1, input_signal_range=zb)
    assert "SIG_LO = 16, SIG_HI = 8" in k.source
    assert K.compile_source(k.source)[:4] == b"\x7fELF"
