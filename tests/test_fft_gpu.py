"""GPU parity of vd.fft.fft / ifft / fft2 / fft3 / rfft / irfft against the FP64 oracle.

Mirrors the reference's tests/test_fft.py (growing random shapes over primes
2..13, every axis, 1-D/2-D/3-D) with seeded inputs; each result must be within
rel-L2 5e-6 of the FP64 oracle and within the reference's own atol.
"""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd
from tests._gpu_common import assert_close, crand, precompile, shape_sequence

pytestmark = pytest.mark.gpu


def _warm(shapes, axes_fn, **kw):
    ks = []
    for s in shapes:
        for a in axes_fn(s):
            ks.append(vd.fft.make_fft_shader(tuple(s), a, compute_type=vd.complex64, output_type=vd.complex64,
                                             input_type=vd.complex64, **kw))
    precompile(ks)


def test_fft_1d_every_axis():
    rng = np.random.default_rng(1)
    shapes = shape_sequence(11, 1)
    _warm(shapes, lambda s: range(len(s)))
    for shape in shapes:
        data = crand(rng, shape)
        buf = vd.Buffer(shape, vd.complex64)
        for axis in range(len(shape)):
            buf.write(data)
            vd.fft.fft(buf, axis=axis)
            assert_close(buf.read(0), O.fft_axis(data, axis), atol=1e-3)


def test_ifft_1d_every_axis():
    rng = np.random.default_rng(2)
    shapes = shape_sequence(12, 1)
    _warm(shapes, lambda s: range(len(s)), inverse=True)
    for shape in shapes:
        data = crand(rng, shape)
        buf = vd.Buffer(shape, vd.complex64)
        for axis in range(len(shape)):
            buf.write(data)
            vd.fft.ifft(buf, axis=axis)
            assert_close(buf.read(0), O.fft_axis(data, axis, inverse=True), atol=1e-3)


@pytest.mark.parametrize("inverse", [False, True])
def test_fft2(inverse):
    rng = np.random.default_rng(3 + inverse)
    shapes = shape_sequence(13 + inverse, 2)
    _warm(shapes, lambda s: (len(s) - 2, len(s) - 1), inverse=inverse)
    for shape in shapes:
        data = crand(rng, shape)
        buf = vd.asbuffer(data)
        (vd.fft.ifft2 if inverse else vd.fft.fft2)(buf)
        n = len(shape)
        ref = O.fftn_axes(data, (n - 2, n - 1), inverse=inverse)
        assert_close(buf.read(0), ref, atol=1e-2)


@pytest.mark.parametrize("inverse", [False, True])
def test_fft3(inverse):
    rng = np.random.default_rng(5 + inverse)
    shapes = shape_sequence(15 + inverse, 3)
    _warm(shapes, lambda s: (0, 1, 2), inverse=inverse)
    for shape in shapes:
        data = crand(rng, shape)
        buf = vd.asbuffer(data)
        (vd.fft.ifft3 if inverse else vd.fft.fft3)(buf)
        assert_close(buf.read(0), O.fftn_axes(data, (0, 1, 2), inverse=inverse), atol=5e-2)


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_rfft(dims):
    rng = np.random.default_rng(20 + dims)
    shapes = [s for s in shape_sequence(30 + dims, dims, max_dims=dims)]
    for shape in shapes:
        data = rng.random(shape).astype(np.float32)
        buf = vd.RFFTBuffer(shape)
        buf.write_real(data)
        {1: vd.fft.rfft, 2: vd.fft.rfft2, 3: vd.fft.rfft3}[dims](buf)
        X = O.rfft_rows(data)
        for a in range(dims - 1):
            X = O.fft_axis(X, a)
        assert_close(buf.read_fourier(0), X, atol={1: 1e-3, 2: 1e-2, 3: 5e-2}[dims])


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_irfft_roundtrip(dims):
    rng = np.random.default_rng(40 + dims)
    for shape in shape_sequence(50 + dims, dims, max_dims=dims):
        data = rng.random(shape).astype(np.float32)
        buf = vd.asrfftbuffer(data)
        {1: vd.fft.rfft, 2: vd.fft.rfft2, 3: vd.fft.rfft3}[dims](buf)
        {1: vd.fft.irfft, 2: vd.fft.irfft2, 3: vd.fft.irfft3}[dims](buf)
        assert_close(buf.read_real(0), data, atol=1e-3)


def test_irfft_matches_oracle_on_given_spectrum():
    """c2r of an arbitrary (non-Hermitian DC/Nyquist) spectrum: imaginary DC/Nyquist ignored."""
    rng = np.random.default_rng(7)
    for n in (8, 9, 162, 256, 1000):
        X = (rng.standard_normal((3, n // 2 + 1)) + 1j * rng.standard_normal((3, n // 2 + 1))).astype(np.complex64)
        buf = vd.RFFTBuffer((3, n))
        buf.write_fourier(X)
        vd.fft.irfft(buf)
        assert_close(buf.read_real(0), O.irfft_rows(X, n))


def test_c2c_4096x1024_forward_inverse():
    """C2C N=4096, batch 1024, complex64, N(0,1) re/im from seed 0: forward then inverse."""
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((1024, 4096)) + 1j * rng.standard_normal((1024, 4096))).astype(np.complex64)
    buf = vd.asbuffer(x)
    vd.fft.fft(buf)
    X = buf.read(0)
    ref = O.dft(x.astype(np.complex128))
    assert_close(X, ref, rel=5e-6)
    vd.fft.ifft(buf)
    assert_close(buf.read(0), x, rel=1e-5)


@pytest.mark.parametrize("n", [1, 2, 3, 16, 17, 31, 64, 29 * 2, 2048, 8192, 16384, 10000, 13 * 13 * 11])
def test_assorted_lengths(n):
    rng = np.random.default_rng(n)
    batch = max(1, min(64, (1 << 18) // n))
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    buf = vd.asbuffer(x)
    vd.fft.fft(buf)
    assert_close(buf.read(0), O.dft(x))
    vd.fft.ifft(buf)
    assert_close(buf.read(0), x, rel=1e-5)


@pytest.mark.parametrize("shape,axis", [((256, 256, 256), 0), ((64, 512, 257), 1), ((3, 1024, 5), 1),
                                        ((1000, 33), 0), ((7, 4096, 2), 1)])
def test_strided_axes(shape, axis):
    rng = np.random.default_rng(sum(shape))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    buf = vd.asbuffer(x)
    vd.fft.fft(buf, axis=axis)
    assert_close(buf.read(0), O.fft_axis(x, axis))


def test_complex128_compute_and_storage():
    rng = np.random.default_rng(9)
    x = rng.standard_normal((16, 720)) + 1j * rng.standard_normal((16, 720))
    buf = vd.asbuffer(x)
    vd.fft.fft(buf)
    assert_close(buf.read(0), O.dft(x), rel=1e-13)
    vd.fft.ifft(buf, axis=0)
    assert_close(buf.read(0), O.idft(O.dft(x), 0), rel=1e-13)


def test_unnormalized_inverse():
    rng = np.random.default_rng(10)
    x = crand(rng, (4, 360))
    buf = vd.asbuffer(x)
    vd.fft.ifft(buf, normalize=False)
    assert_close(buf.read(0), O.idft(x, -1, normalize=False))
