"""GPU tests of mapping hooks, signal ranges and mixed precision.

Ports of the reference's tests/test_fft_padded.py and
tests/test_fft_mixed_precision.py (same seeds and tolerances), checked
against the FP64 oracle, plus the docs' hook examples
(docs/tutorials/reductions_and_fft.rst:147-205).
"""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd
import vkdispatch_amd.codegen as vc
from tests._gpu_common import assert_close, crand, shape_sequence

pytestmark = pytest.mark.gpu


def _zero_outside(data, axis, s, e):
    z = data.copy()
    sl = [slice(None)] * data.ndim
    sl[axis] = slice(0, s)
    z[tuple(sl)] = 0
    sl[axis] = slice(e, data.shape[axis])
    z[tuple(sl)] = 0
    return z


def test_fft_signal_range_every_axis():
    rng = np.random.default_rng(200)
    for shape in shape_sequence(201, 1):
        data = crand(rng, shape)
        buf = vd.Buffer(shape, vd.complex64)
        for axis in range(len(shape)):
            buf.write(data)
            s = int(rng.integers(0, shape[axis] - 1)) if shape[axis] > 1 else 0
            e = int(rng.integers(s + 1, shape[axis] + 1))
            vd.fft.fft(buf, axis=axis, input_signal_range=(s, e))
            ref = O.fft_axis(_zero_outside(data, axis, s, e), axis)
            assert np.allclose(buf.read(0), ref, atol=1e-3)
            assert_close(buf.read(0), ref, rel=1e-5)


def test_rfft_signal_range():
    rng = np.random.default_rng(202)
    for shape in shape_sequence(203, 1):
        data = rng.random(shape).astype(np.float32)
        buf = vd.RFFTBuffer(shape)
        buf.write_real(data)
        s = int(rng.integers(0, shape[-1] - 1)) if shape[-1] > 1 else 0
        e = int(rng.integers(s + 1, shape[-1] + 1))
        vd.fft.fft(buf, buffer_shape=buf.real_shape, r2c=True, input_signal_range=(s, e))
        ref = O.rfft_rows(_zero_outside(data, data.ndim - 1, s, e))
        assert np.allclose(buf.read_fourier(0), ref, atol=1e-3)


def _write_complex32(buffer, values):
    packed = np.empty(values.shape + (2,), dtype=np.float16)
    packed[..., 0] = values.real.astype(np.float16)
    packed[..., 1] = values.imag.astype(np.float16)
    buffer.write(np.ascontiguousarray(packed))


def _quantize_c32(values):
    return (values.real.astype(np.float16).astype(np.float32)
            + 1j * values.imag.astype(np.float16).astype(np.float32)).astype(np.complex64)


def test_fft_complex32_io_with_complex64_compute():
    rng = np.random.default_rng(7)
    data = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(np.complex64)
    q = _quantize_c32(data)
    buf = vd.Buffer(data.shape, vd.complex32)
    _write_complex32(buf, data)
    vd.fft.fft(buf, compute_type=vd.complex64)
    res = buf.read(0).astype(np.complex64)
    ref = O.dft(q)
    assert np.allclose(res, ref, atol=3e-1, rtol=2e-2)
    # stored result is the fp16 rounding of an fp32-accurate spectrum
    assert O.rel_l2(res, ref) < 2e-3


def test_fft_map_complex32_input_to_complex128_output_auto_compute():
    rng = np.random.default_rng(11)
    data = (rng.standard_normal(32) + 1j * rng.standard_normal(32)).astype(np.complex64)
    q = _quantize_c32(data)
    inb = vd.Buffer(data.shape, vd.complex32)
    _write_complex32(inb, data)
    outb = vd.Buffer(data.shape, vd.complex128)

    def input_map(buffer: vc.Buffer[vd.complex32]):
        vd.fft.read_op().read_from_buffer(buffer)

    def output_map(buffer: vc.Buffer[vd.complex128]):
        vd.fft.write_op().write_to_buffer(buffer)

    vd.fft.fft(outb, inb, input_map=vd.map(input_map), output_map=vd.map(output_map))
    res = outb.read(0)
    assert np.allclose(res, O.dft(q), atol=3e-1, rtol=2e-2)
    assert O.rel_l2(res, O.dft(q)) < 1e-6


def test_fft_input_output_maps_allow_float32_buffers():
    rng = np.random.default_rng(23)
    data = rng.standard_normal(64).astype(np.float32)
    inb = vd.asbuffer(data)
    outb = vd.Buffer(data.shape, vd.float32)

    def input_map(buffer: vc.Buffer[vd.float32]):
        op = vd.fft.read_op()
        value = vc.to_dtype(op.register.var_type.child_type, buffer[op.io_index])
        op.register.real = value
        op.register.imag = vc.to_dtype(op.register.var_type.child_type, 0)

    def output_map(buffer: vc.Buffer[vd.float32]):
        op = vd.fft.write_op()
        buffer[op.io_index] = vc.to_dtype(buffer.var_type, op.register.real)

    vd.fft.fft(outb, inb, input_map=vd.map(input_map), output_map=vd.map(output_map))
    res = outb.read(0).astype(np.float32)
    ref = O.dft(data.astype(np.complex128)).real
    assert np.allclose(res, ref, atol=2e-3, rtol=1e-3)


def test_convolve_kernel_map_allows_float32_buffer():
    rng = np.random.default_rng(31)
    data = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(np.complex64)
    scale = np.float32(0.5)
    sig = vd.asbuffer(data.copy())
    sbuf = vd.asbuffer(np.full(data.shape, scale, dtype=np.float32))

    def kernel_map(scale_values: vc.Buffer[vd.float32]):
        op = vd.fft.read_op()
        sv = vc.to_dtype(op.register.var_type, vc.to_complex(scale_values[op.io_index]))
        op.register[:] = vc.mult_complex(op.register, sv)

    vd.fft.convolve(sig, sbuf, kernel_map=vd.map(kernel_map))
    res = sig.read(0).astype(np.complex64)
    assert np.allclose(res, data * scale, atol=2e-3, rtol=1e-3)


def test_fft_output_map_without_input_map_uses_explicit_input_buffer():
    rng = np.random.default_rng(37)
    data = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(np.complex64)
    inb = vd.asbuffer(data.copy())
    outb = vd.Buffer(data.shape, vd.complex64)

    @vd.map
    def output_map(buffer: vc.Buffer[vd.complex64]):
        vd.fft.write_op().write_to_buffer(buffer)

    vd.fft.fft(outb, inb, output_map=output_map)
    assert np.allclose(outb.read(0), O.dft(data), atol=2e-3, rtol=1e-3)
    assert np.array_equal(inb.read(0), data)  # input left untouched


def test_convolve_identity_kernel_map_with_output_map():
    """The reference disables this case (tests/test_fft_mixed_precision.py:217-249); it works here."""
    rng = np.random.default_rng(41)
    data = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(np.complex64)
    inb = vd.asbuffer(data.copy())
    outb = vd.Buffer(data.shape, vd.complex64)

    @vd.map
    def kernel_map():
        return

    @vd.map
    def output_map(buffer: vc.Buffer[vd.complex64]):
        vd.fft.write_op().write_to_buffer(buffer)

    vd.fft.convolve(outb, inb, kernel_map=kernel_map, output_map=output_map)
    assert np.allclose(outb.read(0), data, atol=2e-3, rtol=1e-3)


def test_fft_complex64_io_with_complex128_compute():
    rng = np.random.default_rng(29)
    data = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(np.complex64)
    buf = vd.asbuffer(data)
    vd.fft.fft(buf, compute_type=vd.complex128)
    res = buf.read(0).astype(np.complex64)
    ref = O.dft(data)
    assert np.allclose(res, ref, atol=2e-3, rtol=1e-3)
    assert O.rel_l2(res, ref) < 1e-7  # only the final fp32 rounding remains


def test_var_scaling_hook_from_docs():
    """docs/tutorials/reductions_and_fft.rst:147-153: convolve(buf, 0.5, kernel_map=scale)."""
    rng = np.random.default_rng(50)
    x = (rng.random(256) + 1j * rng.random(256)).astype(np.complex64)
    buf = vd.asbuffer(x)

    @vd.map
    def scale_spectrum(scale_factor: vc.Var[vc.f32]):
        op = vd.fft.read_op()
        op.register[:] = op.register * scale_factor

    vd.fft.convolve(buf, np.float32(0.5), kernel_map=scale_spectrum)
    assert_close(buf.read(0), 0.5 * x.astype(np.complex128), rel=1e-6)


def test_padded_axis_remap_hook_from_docs():
    """docs/tutorials/reductions_and_fft.rst:177-205: padded axis FFT via ravel/unravel remaps."""
    rng = np.random.default_rng(51)
    B, R, C, signal_cols = 2, 16, 8, 6
    x = (rng.standard_normal((B, R, C)) + 1j * rng.standard_normal((B, R, C))).astype(np.complex64)
    buf = vd.asbuffer(x)
    trimmed = (B, signal_cols, C)

    def remap(io_index):
        return vc.unravel_index(vc.ravel_index(io_index, trimmed).to_register(), buf.shape)

    @vd.map
    def input_map(input_buffer: vc.Buffer[vc.c64]):
        op = vd.fft.read_op()
        op.read_from_buffer(input_buffer, io_index=remap(op.io_index))

    @vd.map
    def output_map(output_buffer: vc.Buffer[vc.c64]):
        op = vd.fft.write_op()
        op.write_to_buffer(output_buffer, io_index=remap(op.io_index))

    vd.fft.fft(buf, buf, buffer_shape=trimmed, axis=1, input_map=input_map, output_map=output_map,
               input_signal_range=(0, signal_cols))
    ref = x.astype(np.complex128).copy()
    ref[:, :signal_cols, :] = np.fft.fft(x[:, :signal_cols, :], axis=1)
    assert_close(buf.read(0), ref)


def test_fft_context_procedural():
    """docs/tutorials/reductions_and_fft.rst:225-247."""
    rng = np.random.default_rng(52)
    x = (rng.standard_normal(1024) + 1j * rng.standard_normal(1024)).astype(np.complex64)
    buf = vd.asbuffer(x)
    with vd.fft.fft_context(buffer_shape=(1024,), axis=0) as ctx:
        args = ctx.declare_shader_args([vc.Buffer[vc.c64]])
        for read_op in vd.fft.global_reads_iterator(ctx.registers):
            read_op.read_from_buffer(args[0])
        ctx.execute(inverse=False)
        for write_op in vd.fft.global_writes_iterator(ctx.registers):
            write_op.write_to_buffer(args[0])
    kern = ctx.get_callable()
    kern(buf)
    assert_close(buf.read(0), O.dft(x))


def test_complex_bf16_storage():
    rng = np.random.default_rng(53)
    x = (rng.standard_normal((8, 256)) + 1j * rng.standard_normal((8, 256))).astype(np.complex64)
    buf = vd.Buffer(x.shape, vd.complex_bf16)
    buf.write(x)
    xq = buf.read(0)
    vd.fft.fft(buf)
    ref = O.dft(xq)
    # bf16 storage of the result: unit roundoff 2^-8
    assert O.rel_l2(buf.read(0), ref) < 4e-3


def test_command_graph_replay():
    rng = np.random.default_rng(54)
    x = crand(rng, (32, 512))
    buf = vd.asbuffer(x)
    g = vd.CommandGraph()
    vd.fft.fft(buf, graph=g)
    vd.fft.ifft(buf, graph=g)
    for i in range(3):
        g.submit()
        vd.synchronize()
        assert_close(buf.read(0), x, rel=1e-5)
    g2 = vd.CommandGraph()
    vd.fft.fft(buf, graph=g2)
    g2.submit()
    assert_close(buf.read(0), O.dft(x))
    g2.submit()
    assert_close(buf.read(0), O.dft(O.dft(x)), rel=1e-5)


def test_const_hook_arguments_run_on_the_gpu():
    """Const[...] hook arguments (vkdispatch/codegen/arguments.py:10-20) carry their value
    into the kernel like Var: a complex constant multiplies the spectrum and an integer
    constant shifts the kernel-buffer index (X[k] * w * K[(k + s) mod N])."""
    rng = np.random.default_rng(55)
    N, B = 512, 3
    x = (rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))).astype(np.complex64)
    kv = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    buf = vd.asbuffer(x)
    kbuf = vd.asbuffer(kv)

    @vd.map
    def km(kern: vc.Buffer[vc.c64], shift: vc.Const[vc.u32], w: vc.Const[vc.c64]):
        op = vd.fft.read_op()
        kk = vc.new_register(vc.c64)
        op.read_from_buffer(kern, register=kk, io_index=(op.fft_index + shift) % N)
        op.register[:] = vc.mult_complex(vc.mult_complex(op.register, w), kk)

    w = 0.5 - 0.25j
    vd.fft.convolve(buf, kbuf, np.uint32(7), w, kernel_map=km)
    X = np.fft.fft(x.astype(np.complex128), axis=-1)
    ref = np.fft.ifft(X * w * np.roll(kv.astype(np.complex128), -7)[None, :], axis=-1)
    assert_close(buf.read(0), ref)
