"""ctypes binding of libvkdfft.so (include/vkdfft.h) and the process runtime.

The library is required: importing the package succeeds without it (so the
planner, hook tracer and oracle tests run on a CPU-only host), but every call
that touches the device raises ``RuntimeError`` when the library or a GPU is
missing.  There is no CPU fallback anywhere on the product path.
"""
from __future__ import annotations

import ctypes
import os
import threading

from . import _build

_lib = None
_lib_lock = threading.Lock()

VKX_MAX_PASSES = 24


class DeviceProps(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int),
        ("wave_size", ctypes.c_int),
        ("max_threads_per_block", ctypes.c_int),
        ("max_shared_per_block", ctypes.c_int),
        ("l2_bytes", ctypes.c_int),
        ("clock_khz", ctypes.c_int),
        ("mem_clock_khz", ctypes.c_int),
        ("mem_bus_bits", ctypes.c_int),
        ("total_mem", ctypes.c_size_t),
        ("supports_fp64", ctypes.c_int),
        ("supports_fp16", ctypes.c_int),
    ]


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("function", ctypes.c_uint64),
        ("grid", ctypes.c_uint32 * 3),
        ("block", ctypes.c_uint32 * 3),
        ("shared_bytes", ctypes.c_uint32),
        ("n", ctypes.c_int),
        ("double_precision", ctypes.c_int),
        ("num_passes_fwd", ctypes.c_int),
        ("radices_fwd", ctypes.c_int * VKX_MAX_PASSES),
        ("num_passes_inv", ctypes.c_int),
        ("radices_inv", ctypes.c_int * VKX_MAX_PASSES),
        ("roots_n", ctypes.c_int),
    ]


# (name, argtypes) of every exported entry point; tests check the export list
# against include/vkdfft.h.
_U64 = ctypes.c_uint64
_PU64 = ctypes.POINTER(ctypes.c_uint64)
_SIGS = {
    "vkx_last_error": ([], ctypes.c_char_p),
    "vkx_clear_error": ([], None),
    "vkx_version": ([], ctypes.c_int),
    "vkx_init": ([ctypes.c_int], ctypes.c_int),
    "vkx_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "vkx_device_props_get": ([ctypes.c_int, ctypes.POINTER(DeviceProps)], ctypes.c_int),
    "vkx_malloc": ([_PU64, ctypes.c_size_t], ctypes.c_int),
    "vkx_free": ([_U64], ctypes.c_int),
    "vkx_memcpy_htod": ([_U64, ctypes.c_void_p, ctypes.c_size_t, _U64], ctypes.c_int),
    "vkx_memcpy_dtoh": ([ctypes.c_void_p, _U64, ctypes.c_size_t, _U64], ctypes.c_int),
    "vkx_memcpy_dtod": ([_U64, _U64, ctypes.c_size_t, _U64], ctypes.c_int),
    "vkx_memset": ([_U64, ctypes.c_int, ctypes.c_size_t, _U64], ctypes.c_int),
    "vkx_stream_create": ([_PU64], ctypes.c_int),
    "vkx_stream_destroy": ([_U64], ctypes.c_int),
    "vkx_stream_synchronize": ([_U64], ctypes.c_int),
    "vkx_device_synchronize": ([], ctypes.c_int),
    "vkx_event_create": ([_PU64], ctypes.c_int),
    "vkx_event_destroy": ([_U64], ctypes.c_int),
    "vkx_event_record": ([_U64, _U64], ctypes.c_int),
    "vkx_stream_wait_event": ([_U64, _U64], ctypes.c_int),
    "vkx_event_synchronize": ([_U64], ctypes.c_int),
    "vkx_event_elapsed_ms": ([ctypes.POINTER(ctypes.c_float), _U64, _U64], ctypes.c_int),
    "vkx_compile": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                     ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
                     ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "vkx_compiler_version": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "vkx_free_host": ([ctypes.c_void_p], None),
    "vkx_module_load": ([_PU64, ctypes.c_void_p, ctypes.c_size_t], ctypes.c_int),
    "vkx_module_unload": ([_U64], ctypes.c_int),
    "vkx_module_function": ([_PU64, _U64, ctypes.c_char_p], ctypes.c_int),
    "vkx_function_attrs": ([_U64] + [ctypes.POINTER(ctypes.c_int)] * 4, ctypes.c_int),
    "vkx_launch": ([_U64, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                    _U64, ctypes.c_void_p, ctypes.c_size_t], ctypes.c_int),
    "vkx_plan_create": ([_PU64, ctypes.POINTER(PlanDesc)], ctypes.c_int),
    "vkx_plan_destroy": ([_U64], ctypes.c_int),
    "vkx_plan_exec": ([_U64, _PU64, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, _U64], ctypes.c_int),
    "vkx_plan_exec_timed": ([_U64, _PU64, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, _U64, _U64, _U64],
                            ctypes.c_int),
    "vkx_twiddles": ([_U64, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, _U64],
                     ctypes.c_int),
    "vkx_roots": ([_U64, ctypes.c_int, ctypes.c_int, _U64], ctypes.c_int),
    "vkx_batch_sum": ([_U64, _U64, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, _U64], ctypes.c_int),
    "vkx_chirp": ([_U64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _U64], ctypes.c_int),
    "vkx_ragged_exec": ([_U64, ctypes.c_int, _U64, ctypes.c_int, _U64, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                         ctypes.c_int, _U64],
                        ctypes.c_int),
    "vkx_bluestein_b": ([_U64, ctypes.c_int, ctypes.c_int, _U64], ctypes.c_int),
    "vkx_graph_begin": ([_U64], ctypes.c_int),
    "vkx_stream_is_capturing": ([_U64, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "vkx_graph_end": ([_U64, _PU64], ctypes.c_int),
    "vkx_graph_launch": ([_U64, _U64], ctypes.c_int),
    "vkx_graph_destroy": ([_U64], ctypes.c_int),
}


def lib_path() -> str:
    return os.environ.get("VKX_LIB", _build.LIB_PATH)


def share_torch_hip_runtime() -> None:
    """Import torch, when it is installed, before libvkdfft.so is loaded.

    torch's ROCm wheel carries its own libamdhip64 / libhsa-runtime64 / libhiprtc under
    the SONAMEs libvkdfft.so links against.  Already loaded, they are the copies
    libvkdfft.so binds to, so the process has one HIP runtime and torch tensors, streams
    and graph captures are valid handles here.  Loaded after libvkdfft.so (which then
    brings the system ROCm copies), they form a second runtime in the same process, and
    that one finds no GPU: torch raises "No HIP GPUs are available"."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load libvkdfft.so (once).  Raises RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is None:
            path = lib_path()
            if not os.path.exists(path):
                raise RuntimeError(
                    f"vkdispatch_amd native runtime not built ({path} missing); "
                    "run `python -m vkdispatch_amd._build` or __graft_entry__.build()")
            share_torch_hip_runtime()
            L = ctypes.CDLL(path)
            for name, (argtypes, restype) in _SIGS.items():
                fn = getattr(L, name)
                fn.argtypes = argtypes
                fn.restype = restype
            _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().vkx_last_error().decode(errors="replace")
        lib().vkx_clear_error()
        raise RuntimeError(f"{what}: {msg}" if what else msg)


# ---------------------------------------------------------------------------
# device state
# ---------------------------------------------------------------------------
class _State(threading.local):
    stream_override = None


_state = _State()
_device_lock = threading.Lock()
_device = None  # (index, props)
_default_stream = None


def device_count() -> int:
    c = ctypes.c_int(0)
    check(lib().vkx_device_count(ctypes.byref(c)), "vkx_device_count")
    return c.value


def ensure_device(index: int = 0):
    """Initialise the HIP device (once) and return its properties."""
    global _device, _default_stream
    if _device is not None:
        return _device[1]
    with _device_lock:
        if _device is None:
            if device_count() == 0:
                raise RuntimeError("vkdispatch_amd: no HIP device available (the FFT path runs only on the GPU)")
            check(lib().vkx_init(index), "vkx_init")
            props = DeviceProps()
            check(lib().vkx_device_props_get(index, ctypes.byref(props)), "vkx_device_props_get")
            s = ctypes.c_uint64(0)
            check(lib().vkx_stream_create(ctypes.byref(s)), "vkx_stream_create")
            _default_stream = s.value
            _device = (index, props)
    return _device[1]


def device_initialised() -> bool:
    """True once this process has initialised the HIP device (ensure_device)."""
    return _device is not None


def device_arch() -> str:
    return ensure_device().arch.decode().split(":")[0]


def current_stream() -> int:
    """Stream handle vd.fft launches go to (thread-local override, else the library stream)."""
    if _state.stream_override is not None:
        return _state.stream_override
    ensure_device()
    return _default_stream


_side_streams: list = []


def side_streams(count: int) -> list:
    """``count`` further library streams (default priority), created once per process.
    A plan whose work splits into independent parts forks the parts onto them and
    joins back with events (stream_wait), inside a graph capture too."""
    if len(_side_streams) < count:
        ensure_device()
        with _device_lock:
            while len(_side_streams) < count:
                s = ctypes.c_uint64(0)
                check(lib().vkx_stream_create(ctypes.byref(s)), "vkx_stream_create")
                _side_streams.append(s.value)
    return list(_side_streams[:count])


def side_stream() -> int:
    """The first side stream (see side_streams)."""
    return side_streams(1)[0]


def stream_wait(stream: int, event: "Event") -> None:
    """Work queued on ``stream`` from now on waits for ``event``."""
    check(lib().vkx_stream_wait_event(int(stream), event.handle), "vkx_stream_wait_event")


def set_stream(stream) -> None:
    """Route launches of this thread to ``stream``: None (library stream), an int handle,
    or any object with a ``cuda_stream`` attribute (torch.cuda.Stream)."""
    if stream is None:
        _state.stream_override = None
    elif hasattr(stream, "cuda_stream"):
        _state.stream_override = int(stream.cuda_stream)
    else:
        _state.stream_override = int(stream)


class stream_scope:
    """``with vd.stream_scope(torch.cuda.current_stream()): ...``"""

    def __init__(self, stream):
        self.stream = stream
        self.prev = None

    def __enter__(self):
        self.prev = _state.stream_override
        set_stream(self.stream)
        return self

    def __exit__(self, *exc):
        _state.stream_override = self.prev
        return False


# Launches whose kernels wait on other workgroups with a bounded spin (the ragged
# batch's counter mode, the single-launch 2-D convolution) report a wait that gave
# up through a device flag instead of hanging.  Such a launch registers its plan
# here; synchronize() reads the flags of the plans launched on the synchronised
# stream and raises RuntimeError if one is set.  Graph replays are checked the
# same way when the plan registered itself at record time.
_async_lock = threading.Lock()
_async_pending = {}  # stream -> {id(plan): plan}


def note_async_check(plan, stream) -> None:
    """Check ``plan.async_error()`` (a message or None; called after the stream is
    synchronised, may read device memory) at the next synchronize() of ``stream``."""
    with _async_lock:
        _async_pending.setdefault(int(stream), {})[id(plan)] = plan


def forget_async_check(plan) -> None:
    """Drop ``plan`` from every pending check (its caller has read the flag itself)."""
    with _async_lock:
        for pending in _async_pending.values():
            pending.pop(id(plan), None)


def _raise_async_errors(stream: int) -> None:
    with _async_lock:
        plans = list(_async_pending.pop(int(stream), {}).values())
    errors = [m for m in (p.async_error() for p in plans) if m]
    if errors:
        raise RuntimeError("; ".join(errors))


def is_capturing(stream=None) -> bool:
    """True while ``stream`` (default: the current stream) is being captured into a graph."""
    c = ctypes.c_int(0)
    check(lib().vkx_stream_is_capturing(current_stream() if stream is None else int(stream), ctypes.byref(c)),
          "vkx_stream_is_capturing")
    return c.value != 0


def require_not_capturing(what: str) -> None:
    """Raise if the current stream is capturing: ``what`` allocates or initialises device
    memory, which a graph capture must not record."""
    if is_capturing():
        raise RuntimeError(f"{what} allocates and initialises device memory, which cannot happen while the stream "
                           "is being captured into a graph: run the same call once before the capture "
                           "(as a warm-up), then capture")


def synchronize(stream=None) -> None:
    s = current_stream() if stream is None else int(stream)
    check(lib().vkx_stream_synchronize(s), "vkx_stream_synchronize")
    if _async_pending:
        _raise_async_errors(s)


# ---------------------------------------------------------------------------
# thin helpers
# ---------------------------------------------------------------------------
def malloc(nbytes: int) -> int:
    ensure_device()
    require_not_capturing("a device allocation")
    p = ctypes.c_uint64(0)
    check(lib().vkx_malloc(ctypes.byref(p), max(1, int(nbytes))), "vkx_malloc")
    return p.value


def free(ptr: int) -> None:
    if ptr and _lib is not None:
        lib().vkx_free(ptr)


def htod(dst: int, host_ptr: int, nbytes: int, stream=None) -> None:
    check(lib().vkx_memcpy_htod(dst, host_ptr, nbytes, current_stream() if stream is None else stream), "htod")


def dtoh(host_ptr: int, src: int, nbytes: int, stream=None) -> None:
    check(lib().vkx_memcpy_dtoh(host_ptr, src, nbytes, current_stream() if stream is None else stream), "dtoh")


def dtod(dst: int, src: int, nbytes: int, stream=None) -> None:
    check(lib().vkx_memcpy_dtod(dst, src, nbytes, current_stream() if stream is None else stream), "dtod")


def memset(dst: int, value: int, nbytes: int, stream=None) -> None:
    check(lib().vkx_memset(dst, value, nbytes, current_stream() if stream is None else stream), "memset")


class Event:
    def __init__(self):
        ensure_device()
        e = ctypes.c_uint64(0)
        check(lib().vkx_event_create(ctypes.byref(e)), "vkx_event_create")
        self.handle = e.value

    def record(self, stream=None):
        check(lib().vkx_event_record(self.handle, current_stream() if stream is None else int(stream)), "event_record")

    def synchronize(self):
        check(lib().vkx_event_synchronize(self.handle), "event_synchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        check(lib().vkx_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle), "event_elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle and _lib is not None:
                _lib.vkx_event_destroy(self.handle)
        except Exception:
            pass
