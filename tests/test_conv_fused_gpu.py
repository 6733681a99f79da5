"""Pair-packed 2-D FFT-convolution (convolve2DR with one broadcast kernel spectrum,
three phase kernels) against the FP64 oracle and against the reference's
three-dispatch composition."""
import contextlib
import os

import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd
from vkdispatch_amd.fft import fused

from _gpu_common import assert_close

pytestmark = pytest.mark.gpu


def _gauss_kernel(rng, n, sigma=4.0):
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    dy, dx = np.minimum(yy, n - yy), np.minimum(xx, n - xx)
    g = np.exp(-(dy ** 2 + dx ** 2) / (2 * sigma ** 2))
    return (g + 0.01 * rng.standard_normal((n, n))).astype(np.float32)


@contextlib.contextmanager
def _env(name, value):
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        yield
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


def _run(x, K, kernel_inner_only=True, normalize=True):
    buf = vd.asrfftbuffer(x)
    kb = vd.RFFTBuffer(x.shape[-2:])
    kb.write_fourier(K)
    plan = fused.fused_eligible(buf, kb, kernel_inner_only=kernel_inner_only, normalize=normalize)
    assert plan is not None, "expected the pair-packed path"
    assert isinstance(plan, fused.Conv2DPhases)
    vd.fft.convolve2DR(buf, kb, kernel_inner_only=kernel_inner_only, normalize=normalize)
    vd.synchronize()
    return buf.read_real(0)


@pytest.mark.parametrize("batch,n", [(8, 512), (5, 256), (2, 1024), (3, 1024), (1, 512), (17, 512), (40, 256)])
def test_fused_matches_oracle(batch, n):
    print(f"pair-packed convolve2DR batch={batch} n={n}", flush=True)
    rng = np.random.default_rng(batch * 1000 + n)
    x = rng.uniform(-1, 1, (batch, n, n)).astype(np.float32)
    k = _gauss_kernel(rng, n)
    K = np.fft.rfft2(k.astype(np.float64)).astype(np.complex64)
    if batch == 1:
        x = x[0]
    got = _run(x, K)
    ref = O.convolve2DR(x, K, kernel_inner_only=True)
    e = assert_close(got, ref, rel=1e-5)
    print(f"  rel-L2 {e:.3e}", flush=True)


def test_fused_non_hermitian_half_spectrum():
    """A half spectrum that is not Hermitian on the DC and Nyquist columns: the
    pair-packed path must reproduce what rfft -> convolve -> irfft does with it."""
    print("pair-packed convolve2DR non-Hermitian K", flush=True)
    rng = np.random.default_rng(7)
    n = 256
    x = rng.uniform(-1, 1, (4, n, n)).astype(np.float32)
    K = (rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))).astype(np.complex64)
    got = _run(x, K)
    ref = O.convolve2DR(x, K, kernel_inner_only=True)
    assert_close(got, ref, rel=1e-5)


def test_fused_unnormalized():
    print("pair-packed convolve2DR normalize=False", flush=True)
    rng = np.random.default_rng(8)
    n = 256
    x = rng.uniform(-1, 1, (3, n, n)).astype(np.float32)
    K = np.fft.rfft2(_gauss_kernel(rng, n).astype(np.float64)).astype(np.complex64)
    got = _run(x, K, normalize=False)
    ref = O.convolve2DR(x, K, normalize=False, kernel_inner_only=True)
    assert_close(got, ref, rel=1e-5)


def test_fused_agrees_with_three_dispatch_path():
    print("pair-packed vs reference composition 16 x 512^2", flush=True)
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (16, 512, 512)).astype(np.float32)
    K = np.fft.rfft2(_gauss_kernel(rng, 512).astype(np.float64)).astype(np.complex64)
    got = _run(x, K)
    with _env("VKX_FUSED_CONV", "0"):
        buf = vd.asrfftbuffer(x)
        kb = vd.RFFTBuffer((512, 512))
        kb.write_fourier(K)
        assert fused.fused_eligible(buf, kb, kernel_inner_only=True) is None
        vd.fft.convolve2DR(buf, kb, kernel_inner_only=True)
        three = buf.read_real(0)
    assert O.rel_l2(got, three) <= 2e-6


def test_fused_repeated_launches_and_graph_replays():
    """Back-to-back launches and graph replays give the same result as the first
    launch (the plan's scratch carries nothing from one call to the next)."""
    print("pair-packed convolve2DR repeated launches / graph replays", flush=True)
    rng = np.random.default_rng(10)
    n = 512
    x = rng.uniform(-1, 1, (6, n, n)).astype(np.float32)
    K = np.fft.rfft2(_gauss_kernel(rng, n).astype(np.float64)).astype(np.complex64)
    ref = O.convolve2DR(x, K, kernel_inner_only=True)
    first = _run(x, K)
    assert_close(first, ref, rel=1e-5)
    for _ in range(3):
        assert np.array_equal(_run(x, K), first)
    # graph: the recorded dispatches replayed on fresh copies of the input
    buf = vd.asrfftbuffer(x)
    kb = vd.RFFTBuffer((n, n))
    kb.write_fourier(K)
    g = vd.CommandGraph()
    vd.fft.convolve2DR(buf, kb, kernel_inner_only=True, graph=g)
    for _ in range(3):
        buf.write_real(x)
        g.submit()
        assert np.array_equal(buf.read_real(0), first)


def test_fused_not_used_for_per_image_kernels():
    """Without kernel_inner_only a batched buffer reads one kernel spectrum per image
    (the reference's io_index rule): that call keeps the three-dispatch path."""
    buf = vd.RFFTBuffer((4, 256, 256))
    kb = vd.RFFTBuffer((4, 256, 256))
    assert fused.fused_eligible(buf, kb, kernel_inner_only=False) is None


def test_headline_workload_every_image():
    """The benchmark's workload exactly: 64 x 512^2 U(-1, 1) images from seed 0, one
    512^2 kernel (Gaussian sigma 4 + 1% noise) as a pre-transformed half spectrum,
    default convolve2DR with the broadcast kernel; all 64 images against the FP64
    oracle (reference composition: vkdispatch/fft/functions.py:614-646,
    tests/test_conv.py:132-157)."""
    print("headline 64 x 512^2 convolve2DR, all images", flush=True)
    import bench
    x, k = bench.conv_inputs(64, 512, 0)
    K = np.fft.rfft2(k.astype(np.float64)).astype(np.complex64)
    got = _run(x, K)
    ref = O.convolve2DR(x, K, kernel_inner_only=True)
    worst = max(O.rel_l2(got[b], ref[b]) for b in range(64))
    print(f"  worst per-image rel-L2 {worst:.3e}, whole batch {O.rel_l2(got, ref):.3e}", flush=True)
    assert worst <= 1e-5


def test_timed_launches_match_plain_launches():
    """vkx_plan_exec_timed (each dispatch records its own start / stop events, as
    bench.py times the kernels) computes exactly what the plain launches do and
    reports a positive duration for every kernel."""
    print("pair-packed convolve2DR through timed launches", flush=True)
    rng = np.random.default_rng(21)
    n = 512
    x = rng.uniform(-1, 1, (4, n, n)).astype(np.float32)
    K = np.fft.rfft2(_gauss_kernel(rng, n).astype(np.float64)).astype(np.complex64)
    plain = _run(x, K)
    buf = vd.asrfftbuffer(x)
    kb = vd.RFFTBuffer((n, n))
    kb.write_fourier(K)
    plan = fused.fused_eligible(buf, kb, kernel_inner_only=True)
    vals = plan.values(buf, kb)
    evs = [(vd.Event(), vd.Event()) for _ in plan.kernels()]
    for kk, (a, b) in zip(plan.kernels(), evs):
        kk.launch_timed(vals, a, b)
    vd.synchronize()
    assert np.array_equal(buf.read_real(0), plain)
    ms = [a.elapsed_ms(b) for a, b in evs]
    print(f"  kernel ms {ms}", flush=True)
    assert all(0.0 < t < 100.0 for t in ms)
