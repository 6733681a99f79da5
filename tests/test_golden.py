"""Committed fixtures (tests/golden/fft_golden.npz, numpy.fft in FP64, written by
tests/golden/make_golden.py): the FP64 oracle on the CPU and the HIP path on the GPU
must both reproduce them."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")
FIXTURE_REL = 2e-7   # the expected outputs are stored rounded to complex64 / float32


@pytest.fixture(scope="module")
def g():
    with np.load(GOLD, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


# ---------------------------------------------------------------- oracle (CPU)
@pytest.mark.parametrize("n", [64, 60, 17, 1024])
def test_oracle_c2c(g, n):
    x = g[f"c2c_n{n}_x"]
    assert O.rel_l2(O.dft(x), g[f"c2c_n{n}_fwd"]) <= FIXTURE_REL
    assert O.rel_l2(O.idft(x), g[f"c2c_n{n}_inv"]) <= FIXTURE_REL


def test_oracle_axes_and_multidim(g):
    assert O.rel_l2(O.fft_axis(g["c2c_axis1_x"], axis=1), g["c2c_axis1_fwd"]) <= FIXTURE_REL
    assert O.rel_l2(O.fftn_axes(g["fft2_x"], (1, 2)), g["fft2_fwd"]) <= FIXTURE_REL
    assert O.rel_l2(O.fftn_axes(g["fft3_x"], (0, 1, 2)), g["fft3_fwd"]) <= FIXTURE_REL


def test_oracle_real(g):
    assert O.rel_l2(O.rfft_rows(g["rfft_x"]), g["rfft_fwd"]) <= FIXTURE_REL
    r2 = O.fft_axis(O.rfft_rows(g["rfft2_x"]), axis=1)
    assert O.rel_l2(r2, g["rfft2_fwd"]) <= FIXTURE_REL


def test_oracle_signal_range_and_convolutions(g):
    lo, hi = (int(v) for v in g["range_lo_hi"])
    assert O.rel_l2(O.fft_axis(g["range_x"], signal_range=(lo, hi)), g["range_fwd"]) <= FIXTURE_REL
    assert O.rel_l2(O.convolve_axis(g["conv_x"], g["conv_k"]), g["conv_out"]) <= FIXTURE_REL
    for n in (16, 64, 128):
        got = O.convolve2DR(g[f"conv2dr_n{n}_x"], g[f"conv2dr_n{n}_k"], kernel_inner_only=True)
        assert O.rel_l2(got, g[f"conv2dr_n{n}_out"]) <= FIXTURE_REL


@pytest.mark.parametrize("n", [97, 4099])
def test_oracle_primes(g, n):
    x = g[f"prime_n{n}_x"]
    assert O.rel_l2(O.dft(x), g[f"prime_n{n}_fwd"]) <= FIXTURE_REL
    assert O.rel_l2(O.idft(x), g[f"prime_n{n}_inv"]) <= FIXTURE_REL
    assert O.rel_l2(O.bluestein_dft(x.astype(np.complex128)), g[f"prime_n{n}_fwd"]) <= FIXTURE_REL


def test_oracle_ragged_batch(g):
    sizes, x = g["ragged_sizes"], g["ragged_x"]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    out = np.concatenate([O.dft(x[o:o + n]) for o, n in zip(offs, sizes)])
    assert O.rel_l2(out, g["ragged_fwd"]) <= FIXTURE_REL


# ---------------------------------------------------------------- HIP path (GPU)
GPU_REL = 5e-6


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 60, 17, 1024])
def test_gpu_c2c(g, n):
    import vkdispatch_amd as vd
    x = g[f"c2c_n{n}_x"]
    buf = vd.asbuffer(x)
    vd.fft.fft(buf)
    assert O.rel_l2(buf.read(0), g[f"c2c_n{n}_fwd"]) <= GPU_REL
    buf = vd.asbuffer(x)
    vd.fft.ifft(buf)
    assert O.rel_l2(buf.read(0), g[f"c2c_n{n}_inv"]) <= GPU_REL


@pytest.mark.gpu
def test_gpu_axes_multidim_real(g):
    import vkdispatch_amd as vd
    buf = vd.asbuffer(g["c2c_axis1_x"])
    vd.fft.fft(buf, axis=1)
    assert O.rel_l2(buf.read(0), g["c2c_axis1_fwd"]) <= GPU_REL
    buf = vd.asbuffer(g["fft2_x"])
    vd.fft.fft2(buf)
    assert O.rel_l2(buf.read(0), g["fft2_fwd"]) <= GPU_REL
    buf = vd.asbuffer(g["fft3_x"])
    vd.fft.fft3(buf)
    assert O.rel_l2(buf.read(0), g["fft3_fwd"]) <= GPU_REL
    rb = vd.asrfftbuffer(g["rfft_x"])
    vd.fft.rfft(rb)
    assert O.rel_l2(rb.read_fourier(0), g["rfft_fwd"]) <= GPU_REL
    rb = vd.asrfftbuffer(g["rfft2_x"])
    vd.fft.rfft2(rb)
    assert O.rel_l2(rb.read_fourier(0), g["rfft2_fwd"]) <= GPU_REL


@pytest.mark.gpu
def test_gpu_signal_range_and_convolutions(g):
    import vkdispatch_amd as vd
    lo, hi = (int(v) for v in g["range_lo_hi"])
    buf = vd.asbuffer(g["range_x"])
    vd.fft.fft(buf, input_signal_range=(lo, hi))
    assert O.rel_l2(buf.read(0), g["range_fwd"]) <= GPU_REL
    buf = vd.asbuffer(g["conv_x"])
    vd.fft.convolve(buf, vd.asbuffer(g["conv_k"]))
    assert O.rel_l2(buf.read(0), g["conv_out"]) <= 1e-5
    for n in (16, 64, 128):
        rb = vd.asrfftbuffer(g[f"conv2dr_n{n}_x"])
        kb = vd.RFFTBuffer((n, n))
        kb.write_fourier(g[f"conv2dr_n{n}_k"])
        vd.fft.convolve2DR(rb, kb, kernel_inner_only=True)
        assert O.rel_l2(rb.read_real(0), g[f"conv2dr_n{n}_out"]) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("n", [97, 4099])
def test_gpu_primes(g, n):
    """Bluestein through vd.fft (m = 256 whole; m = 16384 as tiles)."""
    import vkdispatch_amd as vd
    x = g[f"prime_n{n}_x"]
    buf = vd.asbuffer(x)
    vd.fft.fft(buf)
    assert O.rel_l2(buf.read(0), g[f"prime_n{n}_fwd"]) <= 1e-5
    buf = vd.asbuffer(x)
    vd.fft.ifft(buf)
    assert O.rel_l2(buf.read(0), g[f"prime_n{n}_inv"]) <= 1e-5


@pytest.mark.gpu
def test_gpu_ragged_batch(g):
    import vkdispatch_amd as vd
    from vkdispatch_amd.fft.ragged import RaggedFFT
    sizes, x = [int(n) for n in g["ragged_sizes"]], g["ragged_x"]
    src, dst = vd.asbuffer(x), vd.Buffer((x.size,), vd.complex64)
    RaggedFFT(sizes)(dst, src)
    out, ref = dst.read(0), g["ragged_fwd"]
    off = 0
    for n in sizes:
        assert O.rel_l2(out[off:off + n], ref[off:off + n]) <= 1e-5, n
        off += n
