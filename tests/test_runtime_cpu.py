"""CPU-only tests of the runtime's host logic: C-ABI error plumbing through ctypes
(every nonzero return carries a message), the hiprtc compile step's failure path
(bad source -> nonzero return, compiler log, error string; nothing launched), and
the CommandGraph's ownership of the buffers it recorded."""
import ctypes
import gc
import os
import weakref

import pytest

from vkdispatch_amd import runtime as rt
from vkdispatch_amd.graph import CommandGraph


def _err():
    msg = rt.lib().vkx_last_error().decode()
    rt.lib().vkx_clear_error()
    return msg


@pytest.mark.parametrize("args,needle", [
    ((0, 4, 0, 2, 0, 0, 1, 63, 0, 0), "block"),     # block not a multiple of 64
    ((0, 4, 0, 2, 0, 0, 1, 1024, 0, 0), "block"),   # block above 512
    ((0, 4, 0, 2, 0, 0, 0, 256, 0, 0), "grid"),     # no workgroups
    ((0, 4, 0, 1, 0, 0, 1, 256, 0, 0), "counters"),  # no abort-flag slot
    ((0, 4, 0, 2, 0, 0, 1, 256, 0, 0), "null"),     # null task array
])
def test_ragged_exec_rejects_bad_arguments_with_a_message(args, needle):
    rt.lib().vkx_clear_error()
    rc = rt.lib().vkx_ragged_exec(*args)
    assert rc == 1
    msg = _err()
    assert msg.startswith("vkx_ragged_exec:") and needle in msg, msg


def test_ragged_exec_with_no_tasks_is_a_no_op():
    assert rt.lib().vkx_ragged_exec(0, 0, 0, 0, 0, 0, 0, 0, 0, 0) == 0
    assert _err() == ""


@pytest.mark.parametrize("n,m,needle", [(0, 8, "n must"), (5, 8, "m must"), (4, 7, "null")])
def test_bluestein_b_rejects_bad_arguments_with_a_message(n, m, needle):
    rt.lib().vkx_clear_error()
    rc = rt.lib().vkx_bluestein_b(0, n, m, 0)
    assert rc == 1
    msg = _err()
    assert msg.startswith("vkx_bluestein_b:") and needle in msg, msg


def test_plan_create_and_exec_reject_null_handles():
    d = rt.PlanDesc()
    h = ctypes.c_uint64(123)
    assert rt.lib().vkx_plan_create(ctypes.byref(h), ctypes.byref(d)) == 1
    assert h.value == 0  # 0 = no handle, the reference's convention
    assert "null function" in _err()
    assert rt.lib().vkx_plan_exec(0, None, 0, None, 0, 0) == 1
    assert "null plan" in _err()
    assert rt.lib().vkx_plan_exec_timed(0, None, 0, None, 0, 0, 0, 0) == 1
    assert "vkx_plan_exec_timed: null event" in _err()
    assert rt.lib().vkx_plan_exec_timed(0, None, 0, None, 0, 0, 1, 2) == 1
    assert "vkx_plan_exec_timed: null plan" in _err()


def _compile(source: str, opts=("--offload-arch=gfx950",)):
    arr = (ctypes.c_char_p * len(opts))(*[o.encode() for o in opts])
    code, size, log = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_void_p()
    rc = rt.lib().vkx_compile(source.encode(), b"t.hip", arr, len(opts), ctypes.byref(code), ctypes.byref(size),
                              ctypes.byref(log))
    log_text = ""
    if log.value:
        log_text = ctypes.string_at(log.value).decode(errors="replace")
        rt.lib().vkx_free_host(log)
    if code.value:
        rt.lib().vkx_free_host(code)
    return rc, size.value, log_text, code.value


def test_compile_error_returns_nonzero_with_log_and_message():
    """The reference's stage_compute_plan_create reports a failed build with the source and
    the compiler log (vkdispatch/shader/shader_function.py:147-160); here vkx_compile
    returns 1, fills the log and sets the error string, and produces no code object."""
    bad = 'extern "C" __global__ void k(float* p) { p[0] = undefined_name; }\n'
    rc, size, log_text, code = _compile(bad)
    assert rc == 1
    assert size == 0 and not code
    assert "undefined_name" in log_text
    assert _err().startswith("hiprtcCompileProgram:")


def test_compile_good_source_gives_a_code_object():
    good = 'extern "C" __global__ void k(float* p) { p[threadIdx.x] = 1.0f; }\n'
    rc, size, _log, _code = _compile(good)
    assert rc == 0 and size > 1000
    assert _err() == ""


# ---------------------------------------------------------------------------
# CommandGraph keeps what it recorded alive
# ---------------------------------------------------------------------------
class _FakeBuffer:
    def __init__(self, ptr):
        self.ptr = ptr


class _FakeKernel:
    name = "fake"

    def bind(self, values):
        return [v.ptr for v in values], b""

    def handle(self):
        return 1


def test_graph_record_keeps_plan_scratch_alive():
    """A plan-owned scratch buffer must survive the plan's eviction from its cache while a
    graph that recorded it can still replay (the replay would otherwise write freed memory)."""
    g = CommandGraph()
    user, scratch = _FakeBuffer(0x1000), _FakeBuffer(0x2000)
    g.record_kernel(_FakeKernel(), [user, scratch])
    ref = weakref.ref(scratch)
    del scratch
    gc.collect()
    assert ref() is not None, "the graph dropped a buffer it will replay into"
    g.reset()
    gc.collect()
    assert ref() is None


def test_cuda_graph_capture_scope_routes_the_stream_and_restores_it():
    """The capture scope (vkdispatch/execution_pipeline/cuda_graph_capture.py:22-52) routes
    this thread's launches to the capturing stream and restores the previous routing."""
    import vkdispatch_amd as vd
    assert vd.get_cuda_capture() is None
    before = rt._state.stream_override
    with vd.cuda_graph_capture(0x1234) as cap:
        assert vd.get_cuda_capture() is cap and cap.stream_handle == 0x1234
        assert rt._state.stream_override == 0x1234
        with vd.suspend_cuda_capture():
            assert vd.get_cuda_capture() is None and rt._state.stream_override is None
        assert vd.get_cuda_capture() is cap and rt._state.stream_override == 0x1234
    assert vd.get_cuda_capture() is None and rt._state.stream_override == before


def test_one_hip_runtime_per_process():
    """Loading libvkdfft.so first must not leave torch with a second HIP runtime: torch's
    wheel carries its own libamdhip64 under the same SONAME, and a copy loaded after the
    system one finds no GPU.  rt.lib() imports torch first, so one copy serves both."""
    import subprocess
    import sys
    code = ("import vkdispatch_amd.runtime as rt; rt.lib(); import torch\n"
            "libs = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(libs))\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                         cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "1"
