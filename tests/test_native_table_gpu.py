"""The reference-side backend module (vkdispatch_amd/native_table.py) driven the way
vkdispatch drives a backend (context -> buffers -> stage_compute_plan_create ->
descriptor set -> stage_compute_record -> command_list_submit -> buffer_read,
vkdispatch/backends/opencl_backend.py:1268-1898), with the planner's HIP translation
units as shader sources; results against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd
import vkdispatch_amd.codegen as vc
from vkdispatch_amd import native_table as nt
from vkdispatch_amd.fft import kernels as K

from _gpu_common import assert_close, crand

pytestmark = pytest.mark.gpu


def _upload(ctx, arr):
    b = nt.buffer_create(ctx, arr.nbytes, 1)
    assert b, nt.get_error_string()
    nt.buffer_write_staging(b, 0, arr.tobytes(), arr.nbytes)
    nt.buffer_write(b, 0, arr.nbytes, -1)
    return b


def _download(ctx, b, like):
    nt.buffer_read(b, 0, like.nbytes, 0)
    assert nt.context_synchronize(ctx)
    return np.frombuffer(nt.buffer_read_staging(b, 0, like.nbytes), dtype=like.dtype).reshape(like.shape)


def _run(ctx, kernel, buffers, consts=b"", instances=1):
    plan = nt.stage_compute_plan_create(ctx, kernel.source, list(range(len(buffers))), len(consts),
                                        K.source_symbol(kernel.source))
    assert plan, nt.get_error_string()
    ds = nt.descriptor_set_create(plan)
    for i, b in enumerate(buffers):
        nt.descriptor_set_write_buffer(ds, i, b, 0, 0, 0, 1, 1)
    cl = nt.command_list_create(ctx)
    grid = K.parse_launch_comment(kernel.source)["grid"]
    nt.stage_compute_record(cl, plan, ds, *grid)
    assert nt.get_error_string() == 0
    assert nt.command_list_submit(cl, consts * instances, instances, 0) is True
    assert nt.get_error_string() == 0, nt.get_error_string()
    return plan


def test_fft_and_ifft_through_the_native_table():
    print("native table: fft / ifft 8 x 1024 complex64", flush=True)
    nt.clear_error()
    ctx = nt.context_create([0], [[0]])
    assert ctx, nt.get_error_string()
    rng = np.random.default_rng(0)
    x = crand(rng, (8, 1024))
    b = _upload(ctx, x)
    _run(ctx, vd.fft.make_fft_shader((8, 1024), None), [b])
    assert_close(_download(ctx, b, x), O.dft(x))
    _run(ctx, vd.fft.make_fft_shader((8, 1024), None, inverse=True), [b])
    assert_close(_download(ctx, b, x), x)


def test_hook_constants_travel_as_push_constants():
    """A Var[f32] input-map hook: its value reaches the kernel through the
    command list's push-constant bytes, per instance."""
    print("native table: Var hook through push constants", flush=True)
    nt.clear_error()
    ctx = nt.context_create([0], [[0]])

    @vd.map
    def scaled(src: vc.Buffer[vd.complex64], s: vc.Var[vd.float32]):
        op = vd.fft.read_op()
        op.read_from_buffer(src)
        op.register[:] = op.register * s

    rng = np.random.default_rng(1)
    x = crand(rng, (4, 512))
    out = np.zeros_like(x)
    bo, bi = _upload(ctx, out), _upload(ctx, x)
    k = vd.fft.make_fft_shader((4, 512), None, input_map=scaled)
    _ptrs, consts = k.bind([type("B", (), {"ptr": 1})(), type("B", (), {"ptr": 2})(), 0.5])
    _run(ctx, k, [bo, bi], consts)
    assert_close(_download(ctx, bo, out), O.dft(0.5 * x.astype(np.complex128)))
