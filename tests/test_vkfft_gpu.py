"""vd.vkfft (the reference's VkFFT-style API, vkdispatch/vkfft/vkfft_dispatcher.py) on the
HIP kernels: the cases of the reference's tests/test_vkfft.py and tests/test_vkfft_conv.py
(mixed-radix shapes along every axis, 2-D/3-D, r2c round trips, batched 2-D convolution
with transposed kernels), on seeded inputs, against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd

pytestmark = pytest.mark.gpu

REL = 5e-6
RNG = np.random.default_rng(77)


def crand(*shape):
    return (RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)).astype(np.complex64)


SHAPES = [(35,), (6, 22), (4, 9, 10), (13, 12, 7), (2, 3, 120)]


@pytest.mark.parametrize("shape", SHAPES)
def test_fft_ifft_each_axis(shape):
    x = crand(*shape)
    buf = vd.Buffer(x.shape, vd.complex64)
    for axis in range(len(shape)):
        buf.write(x)
        vd.vkfft.fft(buf, axis=axis)
        assert O.rel_l2(buf.read(0), O.fft_axis(x, axis)) <= REL
        buf.write(x)
        vd.vkfft.ifft(buf, axis=axis)
        assert O.rel_l2(buf.read(0), O.fft_axis(x, axis, inverse=True)) <= REL
    vd.vkfft.clear_plan_cache()


def test_fft2_fft3_and_inverses():
    x = crand(6, 20, 12)
    buf = vd.asbuffer(x)
    vd.vkfft.fft2(buf)
    assert O.rel_l2(buf.read(0), O.fftn_axes(x, (1, 2))) <= REL
    vd.vkfft.ifft2(buf)
    assert O.rel_l2(buf.read(0), x) <= REL
    buf.write(x)
    vd.vkfft.fft3(buf)
    assert O.rel_l2(buf.read(0), O.fftn_axes(x, (0, 1, 2))) <= REL
    vd.vkfft.ifft3(buf)
    assert O.rel_l2(buf.read(0), x) <= REL
    # an explicit axis list and no normalisation on the inverse
    buf.write(x)
    vd.vkfft.fft(buf, axis=[0, 2])
    assert O.rel_l2(buf.read(0), O.fftn_axes(x, (0, 2))) <= REL
    vd.vkfft.ifft(buf, axis=[0, 2], normalize=False)
    assert O.rel_l2(buf.read(0), x * (6 * 12)) <= REL


@pytest.mark.parametrize("shape", [(30,), (7, 24), (5, 6, 16)])
def test_rfft_irfft(shape):
    x = RNG.standard_normal(shape).astype(np.float32)
    rb = vd.asrfftbuffer(x)
    nd = len(shape)
    [vd.vkfft.rfft, vd.vkfft.rfft2, vd.vkfft.rfft3][nd - 1](rb)
    ref = O.rfft_rows(x)
    if nd > 1:
        ref = O.fftn_axes(ref, tuple(range(nd - 1)) if nd == 3 else (0,))
    assert O.rel_l2(rb.read_fourier(0), ref) <= REL
    [vd.vkfft.irfft, vd.vkfft.irfft2, vd.vkfft.irfft3][nd - 1](rb)
    assert O.rel_l2(rb.read_real(0), x) <= REL


def test_out_of_place_and_graph():
    x = crand(8, 48)
    src = vd.asbuffer(x)
    dst = vd.Buffer(x.shape, vd.complex64)
    g = vd.CommandGraph()
    vd.vkfft.fft(dst, input_buffer=src, graph=g)
    for _ in range(2):
        g.submit()
        assert O.rel_l2(dst.read(0), O.fft_axis(x, 1)) <= REL
    assert np.array_equal(src.read(0), x)


@pytest.mark.parametrize("shape", [(64, 16, 16), (16, 32, 32), (4, 64, 64), (1, 96, 80)])
def test_convolve2D_transposed_kernels(shape):
    """tests/test_vkfft_conv.py: one kernel per image, normalize=True."""
    x, k = crand(*shape), crand(*shape)
    buf, kb = vd.asbuffer(x), vd.asbuffer(k)
    vd.vkfft.transpose_kernel2D(kb)
    assert O.rel_l2(kb.read(0), O.fftn_axes(k, (1, 2))) <= REL
    vd.vkfft.convolve2D(buf, kb, normalize=True)
    ref = O.vkfft_convolve2D(x, O.fftn_axes(k, (1, 2)), normalize=True)
    assert O.rel_l2(buf.read(0), ref) <= 1e-5


def test_convolve2D_broadcast_conjugate_unnormalised():
    x = crand(5, 24, 40)
    k = crand(24, 40)
    K = O.fftn_axes(k, (0, 1))
    kb = vd.asbuffer(k)
    vd.vkfft.transpose_kernel2D(kb)
    for conj in (False, True):
        buf = vd.asbuffer(x)
        vd.vkfft.convolve2D(buf, kb, conjugate_kernel=conj)
        ref = O.vkfft_convolve2D(x, K, normalize=False, conjugate_kernel=conj)
        assert O.rel_l2(buf.read(0), ref) <= 1e-5
    # a 2-D buffer is one image
    buf = vd.asbuffer(x[0])
    vd.vkfft.convolve2D(buf, kb, normalize=True)
    assert O.rel_l2(buf.read(0), O.vkfft_convolve2D(x[0], K, normalize=True)) <= 1e-5


def test_convolve2D_zero_padding_ranges():
    x = crand(3, 32, 32)
    k = crand(32, 32)
    K = O.fftn_axes(k, (0, 1))
    kb = vd.asbuffer(k)
    vd.vkfft.transpose_kernel2D(kb)
    buf = vd.asbuffer(x)
    vd.vkfft.convolve2D(buf, kb, normalize=True, padding=((24, 32), (0, 8)))
    xz = x.astype(np.complex128).copy()
    xz[:, 24:, :] = 0
    xz[:, :, :8] = 0
    assert O.rel_l2(buf.read(0), O.vkfft_convolve2D(xz, K, normalize=True)) <= 1e-5
    # interior blocks on either axis and on both (VkFFT's fft_zeropad_left/right with
    # 0 < left < right < N): the block is read as zeros
    for pad in (((8, 16), (0, 0)), ((0, 0), (5, 27)), ((3, 29), (10, 11)), ((1, 31), (16, 32))):
        buf = vd.asbuffer(x)
        vd.vkfft.convolve2D(buf, kb, normalize=True, padding=pad)
        xz = x.astype(np.complex128).copy()
        xz[:, pad[0][0]:pad[0][1], :] = 0
        xz[:, :, pad[1][0]:pad[1][1]] = 0
        assert O.rel_l2(buf.read(0), O.vkfft_convolve2D(xz, K, normalize=True)) <= 1e-5, pad
    with pytest.raises(ValueError):
        vd.vkfft.convolve2D(vd.asbuffer(x), kb, padding=((8, 40), (0, 0)))
    with pytest.raises(ValueError):
        vd.vkfft.convolve2D(vd.asbuffer(x), vd.asbuffer(crand(7, 7)))
