"""GPU parity of vd.fft.convolve / convolve2D / convolve2DR / transpose.

Mirrors the reference's tests/test_conv.py: the kernel is pre-transformed with
vd.fft itself and the reference result is ifft(fft(x) * conj(fft(k)))
(tests/test_conv.py:9-21); here the FP64 oracle supplies it.
"""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd
from tests._gpu_common import assert_close, crand, shape_sequence

pytestmark = pytest.mark.gpu


def test_convolution_1d():
    rng = np.random.default_rng(100)
    for shape in shape_sequence(101, 2):
        data = crand(rng, shape)
        kern = crand(rng, shape)
        buf = vd.asbuffer(data)
        kbuf = vd.asbuffer(kern)
        vd.fft.fft(kbuf)
        vd.fft.convolve(buf, kbuf)
        ref = O.convolve_axis(data, O.dft(kern, -1), axis=len(shape) - 1)
        assert_close(buf.read(0), ref, atol=1e-3)


def test_convolution_2d():
    rng = np.random.default_rng(102)
    for shape in shape_sequence(103, 2):
        data = crand(rng, shape)
        kern = crand(rng, shape)
        buf = vd.asbuffer(data)
        kbuf = vd.asbuffer(kern)
        vd.fft.fft2(kbuf)
        vd.fft.convolve2D(buf, kbuf)
        n = len(shape)
        K = O.fftn_axes(kern, (n - 2, n - 1))
        assert_close(buf.read(0), O.convolve2D(data, K), atol=1e-3)


def test_convolution_2d_transposed_kernel():
    rng = np.random.default_rng(104)
    for shape in shape_sequence(105, 2):
        data = crand(rng, shape)
        kern = crand(rng, shape)
        buf = vd.asbuffer(data)
        kbuf = vd.asbuffer(kern)
        size = vd.fft.get_transposed_size(tuple(shape), axis=len(shape) - 2)
        kt = vd.Buffer((size,), var_type=vd.complex64)
        vd.fft.fft2(kbuf)
        vd.fft.transpose(kbuf, out_buffer=kt, axis=len(shape) - 2)
        vd.fft.convolve2D(buf, kt, transposed_kernel=True)
        n = len(shape)
        K = O.fftn_axes(kern, (n - 2, n - 1))
        assert_close(buf.read(0), O.convolve2D(data, K), atol=1e-3)


def test_convolution_2d_real():
    rng = np.random.default_rng(106)
    for shape in shape_sequence(107, 2):
        data = rng.random(shape).astype(np.float32)
        kern = rng.random(shape).astype(np.float32)
        buf = vd.asrfftbuffer(data)
        kbuf = vd.asrfftbuffer(kern)
        vd.fft.rfft2(kbuf)
        vd.fft.convolve2DR(buf, kbuf)
        K = np.fft.rfft2(kern.astype(np.float64))
        assert_close(buf.read_real(0), O.convolve2DR(data, K), atol=1e-3)


def test_convolution_2d_real_register_shuffle_edge_case():
    """Shape (162, 13) from the reference (tests/test_conv.py:159-181)."""
    rng = np.random.default_rng(108)
    shape = (162, 13)
    data = rng.random(shape).astype(np.float32)
    kern = rng.random(shape).astype(np.float32)
    buf = vd.asrfftbuffer(data)
    kbuf = vd.asrfftbuffer(kern)
    vd.fft.rfft2(kbuf)
    vd.fft.convolve2DR(buf, kbuf)
    assert_close(buf.read_real(0), O.convolve2DR(data, np.fft.rfft2(kern.astype(np.float64))), atol=1e-3)


@pytest.mark.parametrize("transposed", [False, True])
def test_convolution_2d_inner_only(transposed):
    """One kernel broadcast over the outer batch (the reference's commented-out
    kernel_inner_only tests, tests/test_conv.py:183-263)."""
    rng = np.random.default_rng(109 + transposed)
    for shape in [(3, 12, 10), (4, 64, 33), (2, 210, 7)]:
        data = crand(rng, shape)
        kern = crand(rng, shape[1:])
        buf = vd.asbuffer(data)
        kbuf = vd.asbuffer(kern)
        vd.fft.fft2(kbuf)
        K = O.fftn_axes(kern, (0, 1))
        if transposed:
            size = vd.fft.get_transposed_size(tuple(shape), axis=1)
            kt = vd.Buffer((size,), var_type=vd.complex64)
            vd.fft.transpose(kbuf, conv_shape=shape, out_buffer=kt, axis=1, kernel_inner_only=True)
            vd.fft.convolve2D(buf, kt, transposed_kernel=True, kernel_inner_only=True)
        else:
            vd.fft.convolve2D(buf, kbuf, kernel_inner_only=True)
        assert_close(buf.read(0), O.convolve2D(data, K, kernel_inner_only=True), atol=1e-3)


def test_convolve_rejects_short_kernel_buffer():
    buf = vd.Buffer((4, 32), vd.complex64)
    k = vd.Buffer((32,), vd.complex64)
    with pytest.raises(ValueError):
        vd.fft.convolve(buf, k, axis=0)


def test_convolve2DR_broadcast_kernel_512():
    """The FP32 2-D FFT-convolution workload at reduced batch: 4 images of 512x512,
    U(-1,1), one Gaussian(sigma=4)+1% noise kernel broadcast (kernel_inner_only)."""
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (4, 512, 512)).astype(np.float32)
    yy, xx = np.meshgrid(np.arange(512), np.arange(512), indexing="ij")
    dy, dx = np.minimum(yy, 512 - yy), np.minimum(xx, 512 - xx)
    g = np.exp(-(dy ** 2 + dx ** 2) / (2 * 4.0 ** 2))
    k = (g + 0.01 * rng.standard_normal((512, 512))).astype(np.float32)
    K = np.fft.rfft2(k.astype(np.float64)).astype(np.complex64)
    buf = vd.asrfftbuffer(x)
    kbuf = vd.RFFTBuffer((512, 512))
    kbuf.write_fourier(K)
    vd.fft.convolve2DR(buf, kbuf, kernel_inner_only=True)
    ref = O.convolve2DR(x, K, kernel_inner_only=True)
    assert_close(buf.read_real(0), ref, rel=1e-5)


def test_kernel_num_two_with_output_map():
    """kernel_num > 1: each kernel's result goes where the output map's mapped_kernel_index() says."""
    import vkdispatch_amd.codegen as vc
    rng = np.random.default_rng(111)
    n, b = 64, 3
    x = crand(rng, (b, n))
    k0, k1 = crand(rng, (b, n)), crand(rng, (b, n))
    buf = vd.asbuffer(x)
    kb = vd.asbuffer(np.concatenate([np.fft.fft(k0, axis=1), np.fft.fft(k1, axis=1)]).astype(np.complex64))
    out = vd.Buffer((2 * b, n), vd.complex64)

    @vd.map
    def kmap(kbuf: vc.Buffer[vd.complex64]):
        op = vd.fft.read_op()
        kv = vc.new_register(vd.complex64)
        op.read_from_buffer(kbuf, register=kv, io_index=op.io_index + vd.fft.mapped_kernel_index() * b * n)
        op.register[:] = vc.mult_complex(op.register, kv.conjugate())

    @vd.map
    def omap(obuf: vc.Buffer[vd.complex64]):
        op = vd.fft.write_op()
        op.write_to_buffer(obuf, io_index=op.io_index + vd.fft.mapped_kernel_index() * b * n)

    vd.fft.convolve(out, buf, kb, kernel_map=kmap, kernel_num=2, output_map=omap)
    res = out.read(0)
    assert_close(res[:b], O.convolve_axis(x, np.fft.fft(k0, axis=1), axis=1))
    assert_close(res[b:], O.convolve_axis(x, np.fft.fft(k1, axis=1), axis=1))
