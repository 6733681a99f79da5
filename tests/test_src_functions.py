"""CPU tests of the ``*_src`` functions (vkdispatch/fft/src_functions.py:7-351): each returns
the translation units of exactly the kernels the matching call launches (captured here
by recording the dispatches of the call on fake device buffers, nothing launched), and
every returned source compiles for gfx950 with hiprtc."""
import pytest

import vkdispatch_amd as vd
from vkdispatch_amd import dtypes as dt
from vkdispatch_amd.fft import fused
from vkdispatch_amd.fft import functions as F
from vkdispatch_amd.fft import kernels as K


def _fake(shape, vt=dt.complex64):
    return vd.Buffer(shape, vt, external=(0x10000000, None, True))


def _fake_rfft(real_shape):
    return vd.RFFTBuffer(real_shape, dt.complex64, external=(0x20000000, None, True))


@pytest.fixture
def launched(monkeypatch):
    """Sources of the kernels a vd.fft call dispatches (recorded, not launched)."""
    got = []

    def record(kernel, values, graph=None):
        got.append(kernel.source)
    monkeypatch.setattr(F, "dispatch", record)
    monkeypatch.setattr(fused, "dispatch", record)
    return got


def _as_tuple(x):
    return tuple(x) if isinstance(x, tuple) else (x,)


@pytest.mark.parametrize("name,shape", [("fft", (8, 512)), ("ifft", (8, 512)), ("fft2", (4, 64, 96)),
                                        ("ifft2", (4, 64, 96)), ("fft3", (16, 24, 32)), ("ifft3", (16, 24, 32))])
def test_complex_src_is_what_the_call_launches(name, shape, launched):
    getattr(vd.fft, name)(_fake(shape))
    assert _as_tuple(getattr(vd.fft, name + "_src")(shape)) == tuple(launched)


@pytest.mark.parametrize("name,shape", [("rfft", (8, 512)), ("irfft", (8, 512)), ("rfft2", (4, 64, 96)),
                                        ("irfft2", (4, 64, 96)), ("rfft3", (8, 16, 32)), ("irfft3", (8, 16, 32))])
def test_real_src_is_what_the_call_launches(name, shape, launched):
    getattr(vd.fft, name)(_fake_rfft(shape))
    assert _as_tuple(getattr(vd.fft, name + "_src")(shape)) == tuple(launched)


@pytest.mark.parametrize("shape,path", [((64, 512, 512), "pairs"), ((3, 256, 256), "pairs"),
                                        ((4, 64, 96), "composition"), ((2, 1024, 1024), "pairs")])
def test_convolve2DR_src_is_what_the_call_launches(shape, path, launched, monkeypatch):
    # the pair plans would allocate their scratch at the first call: record instead
    monkeypatch.setattr(fused.Conv2DPhases, "values", lambda self, b, k: [b, k])
    buf = _fake_rfft(shape)
    ker = _fake_rfft(shape[-2:])
    vd.fft.convolve2DR(buf, ker, kernel_inner_only=True)
    srcs = vd.fft.convolve2DR_src(shape, kernel_inner_only=True)
    assert srcs == tuple(launched)
    assert (path == "pairs") == ("conv2d" in srcs[0])


def test_convolve_and_transpose_src_match_their_calls(launched):
    buf, ker = _fake((4, 64, 33)), _fake((4, 64, 33))
    vd.fft.convolve(buf, ker, axis=1)
    assert vd.fft.convolve_src((4, 64, 33), axis=1) == launched[-1]
    src = vd.fft.transpose_src((4, 64, 33), axis=1)
    assert "transpose_body" in src


def test_every_src_compiles_for_gfx950(tmp_path, monkeypatch):
    monkeypatch.setenv("VKX_KERNEL_CACHE", str(tmp_path))
    srcs = []
    srcs += vd.fft.fft2_src((4, 64, 96)) + vd.fft.irfft3_src((8, 16, 32)) + vd.fft.convolve2D_src((4, 64, 64))
    srcs += vd.fft.convolve2DR_src((64, 512, 512), kernel_inner_only=True)
    srcs += (vd.fft.transpose_src((4, 64, 33), axis=1), vd.fft.fft_src((3, 1000), line_numbers=False))
    K.precompile(srcs)
    for s in srcs:
        assert K.compile_source(s)[:4] == b"\x7fELF"
        meta = K.parse_launch_comment(s)
        assert meta["symbol"] == K.source_symbol(s) and meta["grid"][0] >= 1


def test_print_src_prints_every_stage(capsys):
    vd.fft.convolve2DR_print_src((64, 512, 512), kernel_inner_only=True)
    out = capsys.readouterr().out
    assert out.count("// vkx-launch") == len(vd.fft.convolve2DR_src((64, 512, 512), kernel_inner_only=True))
    vd.fft.fft3_print_src((16, 24, 32))
    assert capsys.readouterr().out.count("FFT Stage (axis") == 3
    numbered = vd.fft.fft_src((8, 512), line_numbers=True)
    assert numbered.split("\n")[0].startswith("   1: // vkx-launch")


def test_src_dimension_checks_like_the_reference():
    with pytest.raises(ValueError):
        vd.fft.fft3_src((8, 512))
    with pytest.raises(ValueError):
        vd.fft.convolve2DR_src((512,))
