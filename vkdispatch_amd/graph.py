"""``vd.CommandGraph`` on HIP graphs.

The reference records dispatches into a CommandGraph and submits them through
its native queues (vkdispatch/execution_pipeline/command_graph.py:38-319); the
default thread-local graph submits on record (:321-348).  Here a graph records
bound launches (kernel plan, device pointers, packed constants); ``submit``
captures them once into a hipGraph on the current stream (plans and twiddle
tables are created at record time, so capture never allocates) and replays the
instantiated graph on every later submit.
"""
from __future__ import annotations

import contextlib
import ctypes
import threading
from typing import List, Optional, Tuple

from . import runtime as rt


class CommandGraph:
    def __init__(self, reset_on_submit: bool = False, submit_on_record: bool = False):
        self.reset_on_submit = reset_on_submit
        self.submit_on_record = submit_on_record
        # (kernel or launcher, device pointers, packed constants, objects kept alive)
        self._records: List[Tuple[object, Optional[List[int]], Optional[bytes], tuple]] = []
        self._exec: Optional[int] = None
        self._exec_stream: Optional[int] = None

    def record_kernel(self, kernel, values) -> None:
        ptrs, consts = kernel.bind(values)
        kernel.handle()
        if self.submit_on_record:
            kernel.launch_raw(ptrs, consts)
            return
        # The record holds the Buffer objects themselves, not only their addresses: a
        # buffer owned by a cached plan (pair-convolution or mixed-precision scratch)
        # must outlive the plan's eviction from its cache while this graph can replay.
        keep = tuple(v for v in values if hasattr(v, "ptr"))
        self._records.append((kernel, ptrs, consts, keep))
        self._drop_exec()

    def record_callable(self, fn) -> None:
        """Record a host-side launcher ``fn(stream)`` (used by fused multi-kernel paths)."""
        if self.submit_on_record:
            fn(rt.current_stream())
            return
        self._records.append((fn, None, None, ()))
        self._drop_exec()

    def __len__(self):
        return len(self._records)

    def _drop_exec(self):
        if self._exec is not None:
            rt.lib().vkx_graph_destroy(self._exec)
            self._exec = None

    def _replay(self, stream: int) -> None:
        for k, ptrs, consts, _keep in self._records:
            if ptrs is None:
                k(stream)
            else:
                k.launch_raw(ptrs, consts, stream)

    def capture(self, stream: Optional[int] = None) -> int:
        s = rt.current_stream() if stream is None else int(stream)
        if self._exec is None or self._exec_stream != s:
            self._drop_exec()
            rt.check(rt.lib().vkx_graph_begin(s), "vkx_graph_begin")
            try:
                self._replay(s)
            finally:
                h = ctypes.c_uint64(0)
                rt.check(rt.lib().vkx_graph_end(s, ctypes.byref(h)), "vkx_graph_end")
            self._exec, self._exec_stream = h.value, s
        return self._exec

    def submit(self, instance_count: Optional[int] = None, queue_index: int = -2, stream=None,
               use_graph: bool = True) -> None:
        if not self._records:
            return
        cap = get_cuda_capture()
        if cap is not None and stream is None:
            stream = cap.stream_handle
        s = rt.current_stream() if stream is None else int(stream)
        reps = 1 if instance_count is None else int(instance_count)
        if cap is not None or rt.is_capturing(s):
            # the caller is capturing this stream (torch.cuda.graph): launch the records
            # directly so that they become nodes of the caller's graph
            for _ in range(reps):
                self._replay(s)
        elif use_graph:
            ex = self.capture(s)
            for _ in range(reps):
                rt.check(rt.lib().vkx_graph_launch(ex, s), "vkx_graph_launch")
        else:
            for _ in range(reps):
                self._replay(s)
        if self.reset_on_submit:
            self.reset()

    def reset(self) -> None:
        self._drop_exec()
        self._records = []

    def destroy(self) -> None:
        self.reset()

    def __del__(self):
        try:
            self._drop_exec()
        except Exception:
            pass


_tls = threading.local()


# ---------------------------------------------------------------------------
# capture scope for a caller's graph (torch.cuda.graph / hipGraph), the
# reference's vkdispatch/execution_pipeline/cuda_graph_capture.py:9-52
# ---------------------------------------------------------------------------
class CUDAGraphCapture:
    """State of an open capture scope: the stream vd launches go to, and the uniform
    buffers the reference would keep alive (unused here: Var / Const values travel in
    the kernel's by-value argument, which the captured launch copies)."""

    def __init__(self, cuda_stream):
        self.cuda_stream = cuda_stream
        self.uniform_buffers: List[object] = []
        if cuda_stream is None:
            import torch
            cuda_stream = torch.cuda.current_stream()
        self.stream_handle = int(cuda_stream.cuda_stream) if hasattr(cuda_stream, "cuda_stream") else int(cuda_stream)

    def add_uniform_buffer(self, buffer) -> None:
        self.uniform_buffers.append(buffer)


def get_cuda_capture() -> Optional[CUDAGraphCapture]:
    return getattr(_tls, "capture", None)


@contextlib.contextmanager
def cuda_graph_capture(cuda_stream=None):
    """Inside ``with torch.cuda.graph(g): with vd.cuda_graph_capture(torch.cuda.current_stream()):``
    every vd launch (direct calls and CommandGraph.submit) goes to ``cuda_stream`` (default:
    torch's current stream) and is recorded into the caller's graph.  Plans must have been
    used once before the capture (their first use allocates device memory)."""
    cap = CUDAGraphCapture(cuda_stream)
    prev_cap = get_cuda_capture()
    prev_stream = rt._state.stream_override
    _tls.capture = cap
    rt._state.stream_override = cap.stream_handle
    try:
        yield cap
    finally:
        _tls.capture = prev_cap
        rt._state.stream_override = prev_stream


@contextlib.contextmanager
def suspend_cuda_capture():
    """Temporarily leave the capture scope (for work that must not be captured)."""
    cap = get_cuda_capture()
    if cap is None:
        yield
        return
    prev_stream = rt._state.stream_override
    _tls.capture = None
    rt._state.stream_override = None
    try:
        yield
    finally:
        _tls.capture = cap
        rt._state.stream_override = prev_stream


def global_graph() -> CommandGraph:
    g = getattr(_tls, "graph", None)
    if g is None:
        g = _tls.graph = CommandGraph(reset_on_submit=True, submit_on_record=True)
    return g


def dispatch(kernel, values, graph: Optional[CommandGraph] = None) -> None:
    if graph is not None and not isinstance(graph, CommandGraph):
        raise ValueError("Expected a CommandGraph object for 'graph'!")
    (graph if graph is not None else global_graph()).record_kernel(kernel, values)
