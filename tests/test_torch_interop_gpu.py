"""torch-ROCm interop: zero-copy aliasing both ways (__cuda_array_interface__,
vkdispatch/base/buffer.py:309-375), and vd launches captured into a torch CUDA graph
through the vd.cuda_graph_capture scope (vkdispatch/execution_pipeline/
cuda_graph_capture.py:22-52; the reference's examples/pytorch_cuda_graph_cuda_python.py),
replayed several times with eager work between the replays; every replay is checked
against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd

from _gpu_common import assert_close

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_from_cuda_array_and_to_torch_alias_the_same_memory():
    print("interop: from_cuda_array / to_torch aliasing", flush=True)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((16, 256)) + 1j * rng.standard_normal((16, 256))).astype(np.complex64)
    t = torch.from_numpy(x).to("cuda")
    buf = vd.from_cuda_array(t)
    assert buf.ptr == t.data_ptr() and buf.shape == (16, 256)
    torch.cuda.synchronize()
    vd.fft.fft(buf)
    vd.synchronize()
    assert_close(t.cpu().numpy(), O.dft(x))
    back = buf.to_torch()
    assert back.data_ptr() == t.data_ptr()
    back.mul_(2)
    torch.cuda.synchronize()
    assert_close(buf.read(0), 2 * O.dft(x))


def test_rfftbuffer_alias_of_a_torch_tensor():
    print("interop: RFFTBuffer alias of a packed real tensor", flush=True)
    rng = np.random.default_rng(1)
    n = 64
    x = rng.uniform(-1, 1, (3, n, n)).astype(np.float32)
    packed = torch.zeros((3, n, 2 * (n // 2 + 1)), dtype=torch.float32, device="cuda")
    packed[..., :n] = torch.from_numpy(x).to("cuda")
    buf = vd.rfftbuffer_from_cuda_array(packed, (3, n, n))
    assert buf.real_shape == (3, n, n) and buf.ptr == packed.data_ptr()
    torch.cuda.synchronize()
    vd.fft.rfft2(buf)
    vd.synchronize()
    spec = torch.view_as_complex(packed.view(3, n, n // 2 + 1, 2)).cpu().numpy()
    assert_close(spec, np.fft.rfft2(x.astype(np.float64)))


def test_convolve2DR_inside_torch_cuda_graph():
    """The captured graph replays the convolution on whatever the input tensor holds at
    replay time; eager vd work runs between replays."""
    print("interop: convolve2DR captured in torch.cuda.graph", flush=True)
    rng = np.random.default_rng(2)
    B, n = 16, 512
    xs = [rng.uniform(-1, 1, (B, n, n)).astype(np.float32) for _ in range(3)]
    k = rng.standard_normal((n, n)).astype(np.float32)
    K = np.fft.rfft2(k.astype(np.float64)).astype(np.complex64)
    x_t = torch.zeros((B, n, 2 * (n // 2 + 1)), dtype=torch.float32, device="cuda")
    k_t = torch.from_numpy(K).to("cuda")
    buf = vd.rfftbuffer_from_cuda_array(x_t, (B, n, n))
    kb = vd.rfftbuffer_from_cuda_array(k_t, (n, n))
    # warm-up outside the capture: the plan allocates its scratch on first use
    x_t[..., :n] = torch.from_numpy(xs[0]).to("cuda")
    torch.cuda.synchronize()
    vd.fft.convolve2DR(buf, kb, kernel_inner_only=True)
    vd.synchronize()
    assert_close(x_t[..., :n].cpu().numpy(), O.convolve2DR(xs[0], K, kernel_inner_only=True), rel=1e-5)
    cmd = vd.CommandGraph()
    vd.fft.convolve2DR(buf, kb, kernel_inner_only=True, graph=cmd)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        with vd.cuda_graph_capture(torch.cuda.current_stream()):
            cmd.submit()
    side = vd.asbuffer((rng.standard_normal((4, 256)) + 0j).astype(np.complex64))
    for rep in range(3):
        for x in xs:
            x_t.zero_()
            x_t[..., :n] = torch.from_numpy(x).to("cuda")
            g.replay()
            torch.cuda.synchronize()
            got = x_t[..., :n].cpu().numpy()
            assert_close(got, O.convolve2DR(x, K, kernel_inner_only=True), rel=1e-5)
            vd.fft.fft(side)  # eager vd work between replays
            vd.synchronize()


def test_plan_creation_inside_a_capture_is_refused_clearly():
    """A plan's first use allocates device memory: inside a capture that is an error with
    a message that says what to do, raised before any device call is recorded."""
    import ctypes
    print("interop: first use inside capture -> clear error", flush=True)
    vd.fft.cache_clear()
    x_t = torch.zeros((2, 128, 130), dtype=torch.float32, device="cuda")
    k_t = torch.zeros((128, 65), dtype=torch.complex64, device="cuda")
    buf = vd.rfftbuffer_from_cuda_array(x_t, (2, 128, 128))
    kb = vd.rfftbuffer_from_cuda_array(k_t, (128, 128))
    torch.cuda.synchronize()
    rt = vd.runtime
    s = rt.current_stream()
    rt.check(rt.lib().vkx_graph_begin(s), "vkx_graph_begin")
    try:
        assert rt.is_capturing(s)
        with pytest.raises(RuntimeError, match="before the capture"):
            vd.fft.convolve2DR(buf, kb, kernel_inner_only=True)
    finally:
        h = ctypes.c_uint64(0)
        rt.lib().vkx_graph_end(s, ctypes.byref(h))
        rt.lib().vkx_clear_error()
        if h.value:
            rt.lib().vkx_graph_destroy(h.value)
    assert not rt.is_capturing(s)
