"""The reference's backend "native table" for the vd.fft path, over libvkdfft.so.

vkdispatch picks one native-table module per process
(vkdispatch/backends/backend_selection.py:9-14,82-91) and calls it for contexts,
buffers, command lists, descriptor sets and compute stages
(vkdispatch/backends/opencl_backend.py:1118-1898 is the OpenCL one).  This module
implements the subset the FFT path binds, with the same function names, argument
meaning and error behaviour (handles are ints, 0 = failure; failures leave a message
for get_error_string(), vkdispatch/base/errors.py:15-44), on the C-ABI declared in
include/vkdfft.h.  A maintainer installs it as vkdispatch/backends/hip_backend.py; it
imports nothing from vkdispatch_amd (only ctypes, and the library path from VKX_LIB).

What differs from the OpenCL backend: shader_source is the HIP translation unit that
vkdispatch_amd's planner emits for the dispatch (make_fft_shader(...).source, or the
vd.fft.*_src functions).  Its first line describes the launch (grid, block, radix
lists for the FP64 twiddle tables, real-FFT roots), so stage_compute_plan_create binds
exactly the geometry and tables the planner chose; stage_compute_record checks the
caller's block counts against that grid.
"""
from __future__ import annotations

import ctypes
import json
import os
import threading
from typing import Dict, List, Optional, Tuple

_LAUNCH_TAG = "// vkx-launch "
_MAX_PASSES = 24
_U64 = ctypes.c_uint64


class _PlanDesc(ctypes.Structure):
    _fields_ = [("function", _U64), ("grid", ctypes.c_uint32 * 3), ("block", ctypes.c_uint32 * 3),
                ("shared_bytes", ctypes.c_uint32), ("n", ctypes.c_int), ("double_precision", ctypes.c_int),
                ("num_passes_fwd", ctypes.c_int), ("radices_fwd", ctypes.c_int * _MAX_PASSES),
                ("num_passes_inv", ctypes.c_int), ("radices_inv", ctypes.c_int * _MAX_PASSES),
                ("roots_n", ctypes.c_int)]


_lib = None
_lock = threading.Lock()


def _L():
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                path = os.environ.get("VKX_LIB") or os.path.join(
                    os.path.dirname(os.path.abspath(__file__)), "_lib", "libvkdfft.so")
                # torch's wheel carries its own HIP runtime under the same SONAMEs: loaded
                # first, it is the one libvkdfft.so binds to (one runtime per process; a
                # second copy loaded later finds no GPU)
                try:
                    import torch  # noqa: F401
                except ImportError:
                    pass
                L = ctypes.CDLL(path)
                L.vkx_last_error.restype = ctypes.c_char_p
                L.vkx_compile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_void_p)]
                L.vkx_free_host.argtypes = [ctypes.c_void_p]
                L.vkx_module_load.argtypes = [ctypes.POINTER(_U64), ctypes.c_void_p, ctypes.c_size_t]
                L.vkx_module_function.argtypes = [ctypes.POINTER(_U64), _U64, ctypes.c_char_p]
                L.vkx_module_unload.argtypes = [_U64]
                L.vkx_plan_create.argtypes = [ctypes.POINTER(_U64), ctypes.POINTER(_PlanDesc)]
                L.vkx_plan_destroy.argtypes = [_U64]
                L.vkx_plan_exec.argtypes = [_U64, ctypes.POINTER(_U64), ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_size_t, _U64]
                L.vkx_malloc.argtypes = [ctypes.POINTER(_U64), ctypes.c_size_t]
                L.vkx_free.argtypes = [_U64]
                L.vkx_memcpy_htod.argtypes = [_U64, ctypes.c_void_p, ctypes.c_size_t, _U64]
                L.vkx_memcpy_dtoh.argtypes = [ctypes.c_void_p, _U64, ctypes.c_size_t, _U64]
                L.vkx_stream_create.argtypes = [ctypes.POINTER(_U64)]
                L.vkx_stream_synchronize.argtypes = [_U64]
                L.vkx_event_create.argtypes = [ctypes.POINTER(_U64)]
                L.vkx_event_record.argtypes = [_U64, _U64]
                L.vkx_event_synchronize.argtypes = [_U64]
                L.vkx_event_destroy.argtypes = [_U64]
                _lib = L
    return _lib


# ---- errors (vkdispatch/base/errors.py:15-44: None when there is none) ----
_error: Optional[str] = None


def _set_error(msg: str) -> None:
    global _error
    _error = msg


def _ok(rc: int, where: str) -> bool:
    if rc != 0:
        L = _L()
        _set_error(f"{where}: {L.vkx_last_error().decode(errors='replace')}")
        L.vkx_clear_error()
        return False
    return True


def get_error_string():
    return 0 if _error is None else _error


def clear_error() -> None:
    global _error
    _error = None


# ---- handles ----
_objects: Dict[int, object] = {}
_next = [1]


def _new(obj) -> int:
    with _lock:
        h = _next[0]
        _next[0] += 1
        _objects[h] = obj
    return h


def _get(h, kind):
    obj = _objects.get(int(h))
    return obj if isinstance(obj, kind) else None


class _Context:
    def __init__(self, device: int, stream: int):
        self.device, self.stream = device, stream


class _Buffer:
    def __init__(self, ctx: int, size: int, ptr: int, owned: bool):
        self.ctx, self.size, self.ptr, self.owned = ctx, size, ptr, owned
        self.staging = bytearray(size)


class _Plan:
    def __init__(self, ctx, module, function, plan, meta, pc_size, bindings):
        self.ctx, self.module, self.function, self.plan = ctx, module, function, plan
        self.meta, self.pc_size, self.bindings = meta, pc_size, bindings


class _DescriptorSet:
    def __init__(self, plan: int):
        self.plan = plan
        self.buffers: Dict[int, Tuple[int, int]] = {}  # binding -> (buffer handle, offset)


class _CommandList:
    def __init__(self, ctx: int):
        self.ctx = ctx
        self.commands: List[Tuple[int, int, Tuple[int, int, int]]] = []


class _Signal:
    def __init__(self, event: int):
        self.event = event


# ---- init / context (opencl_backend.py:1118,1268) ----
def init(debug, log_level):
    clear_error()


def context_create(device_indicies, queue_families):
    devs = [int(d) for d in device_indicies]
    if len(devs) != 1:
        _set_error("the HIP backend drives exactly one device per context")
        return 0
    if not _ok(_L().vkx_init(devs[0]), "context_create"):
        return 0
    s = _U64(0)
    if not _ok(_L().vkx_stream_create(ctypes.byref(s)), "context_create"):
        return 0
    return _new(_Context(devs[0], s.value))


def context_destroy(context):
    _objects.pop(int(context), None)


# ---- buffers (opencl_backend.py:1429-1620) ----
def buffer_create(context, size, per_device):
    if _get(context, _Context) is None:
        _set_error("Invalid context handle for buffer_create")
        return 0
    size = int(size)
    if size <= 0:
        _set_error("Buffer size must be greater than zero")
        return 0
    p = _U64(0)
    if not _ok(_L().vkx_malloc(ctypes.byref(p), size), "buffer_create"):
        return 0
    return _new(_Buffer(int(context), size, p.value, True))


def buffer_create_external(context, size, device_ptr):
    """A device pointer the caller owns (a torch tensor's storage): used as is."""
    if _get(context, _Context) is None:
        _set_error("Invalid context handle for buffer_create_external")
        return 0
    return _new(_Buffer(int(context), int(size), int(device_ptr), False))


def buffer_destroy(buffer):
    b = _objects.pop(int(buffer), None)
    if isinstance(b, _Buffer) and b.owned:
        _ok(_L().vkx_free(b.ptr), "buffer_destroy")


def buffer_write_staging(buffer, queue_index, data, size):
    b = _get(buffer, _Buffer)
    if b is None:
        return
    payload = bytes(data)
    n = min(int(size), len(payload), b.size)
    b.staging[:n] = payload[:n]


def buffer_read_staging(buffer, queue_index, size):
    b = _get(buffer, _Buffer)
    if b is None:
        return bytes(int(size))
    return bytes(b.staging[:int(size)]).ljust(int(size), b"\0")


def buffer_write(buffer, offset, size, index):
    b = _get(buffer, _Buffer)
    ctx = _get(b.ctx, _Context) if b is not None else None
    if ctx is None:
        _set_error(f"Invalid buffer handle {buffer} for buffer_write")
        return
    n = min(int(offset) + int(size), b.size) - int(offset)
    if n <= 0:
        return
    src = (ctypes.c_char * n).from_buffer(b.staging)
    _ok(_L().vkx_memcpy_htod(b.ptr + int(offset), ctypes.addressof(src), n, ctx.stream), "buffer_write")


def buffer_read(buffer, offset, size, index):
    b = _get(buffer, _Buffer)
    ctx = _get(b.ctx, _Context) if b is not None else None
    if ctx is None:
        _set_error(f"Invalid buffer handle {buffer} for buffer_read")
        return
    n = min(int(offset) + int(size), b.size) - int(offset)
    if n <= 0:
        return
    dst = (ctypes.c_char * n).from_buffer(b.staging)
    _ok(_L().vkx_memcpy_dtoh(ctypes.addressof(dst), b.ptr + int(offset), n, ctx.stream), "buffer_read")


# ---- signals (opencl_backend.py:1372-1425) ----
def signal_insert(context, queue_index):
    ctx = _get(context, _Context)
    if ctx is None:
        _set_error("Invalid context handle for signal_insert")
        return 0
    e = _U64(0)
    if not (_ok(_L().vkx_event_create(ctypes.byref(e)), "signal_insert") and
            _ok(_L().vkx_event_record(e.value, ctx.stream), "signal_insert")):
        return 0
    return _new(_Signal(e.value))


def signal_wait(signal_ptr, wait_for_timestamp, queue_index):
    s = _get(signal_ptr, _Signal)
    if s is None:
        return True
    return _ok(_L().vkx_event_synchronize(s.event), "signal_wait")


def signal_destroy(signal_ptr):
    s = _objects.pop(int(signal_ptr), None)
    if isinstance(s, _Signal):
        _L().vkx_event_destroy(s.event)


# ---- compute stages (opencl_backend.py:1823-1898) ----
def stage_compute_plan_create(context, shader_source, bindings, pc_size, shader_name):
    """Compile the HIP translation unit for gfx950, load it, and bind the launch the
    planner described in its first line (grid, block, FP64 twiddle tables)."""
    ctx = _get(context, _Context)
    if ctx is None:
        _set_error("Invalid context handle for stage_compute_plan_create")
        return 0
    src = shader_source.decode() if isinstance(shader_source, (bytes, bytearray)) else str(shader_source)
    name = shader_name.decode() if isinstance(shader_name, (bytes, bytearray)) else str(shader_name)
    first = src.split("\n", 1)[0]
    if not first.startswith(_LAUNCH_TAG):
        _set_error(f"Failed to create HIP kernel '{name}': the source has no launch description line")
        return 0
    try:
        meta = json.loads(first[len(_LAUNCH_TAG):])
    except ValueError as exc:
        _set_error(f"Failed to parse the launch description of '{name}': {exc}")
        return 0
    L = _L()
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-Xclang", b"-target-feature", b"-Xclang",
            b"-packed-fp32-ops", ("-I" + os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc",
                                                     "device")).encode()]
    arr = (ctypes.c_char_p * len(opts))(*opts)
    code, size, log = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_void_p()
    rc = L.vkx_compile(src.encode(), (name + ".hip").encode(), arr, len(opts), ctypes.byref(code),
                       ctypes.byref(size), ctypes.byref(log))
    log_text = ctypes.string_at(log.value).decode(errors="replace") if log.value else ""
    if log.value:
        L.vkx_free_host(log)
    if rc != 0:
        msg = L.vkx_last_error().decode(errors="replace")
        L.vkx_clear_error()
        _set_error(f"Failed to compile HIP kernel '{name}': {msg}\n{log_text}")
        return 0
    mod, fn = _U64(0), _U64(0)
    loaded = _ok(L.vkx_module_load(ctypes.byref(mod), code, size.value), "stage_compute_plan_create")
    L.vkx_free_host(code)
    if not loaded:
        return 0
    if not _ok(L.vkx_module_function(ctypes.byref(fn), mod.value, meta["symbol"].encode()),
               "stage_compute_plan_create"):
        L.vkx_module_unload(mod.value)
        return 0
    d = _PlanDesc()
    d.function = fn.value
    for i in range(3):
        d.grid[i], d.block[i] = int(meta["grid"][i]), int(meta["block"][i])
    d.n, d.double_precision, d.roots_n = int(meta["n"]), int(meta["double"]), int(meta["roots_n"])
    for field, key in (("fwd", "radices_fwd"), ("inv", "radices_inv")):
        rs = [int(r) for r in meta[key]]
        setattr(d, f"num_passes_{field}", len(rs))
        arrf = getattr(d, f"radices_{field}")
        for i, r in enumerate(rs):
            arrf[i] = r
    h = _U64(0)
    if not _ok(L.vkx_plan_create(ctypes.byref(h), ctypes.byref(d)), "stage_compute_plan_create"):
        L.vkx_module_unload(mod.value)
        return 0
    return _new(_Plan(int(context), mod.value, fn.value, h.value, meta, int(pc_size), [int(b) for b in bindings]))


def stage_compute_plan_destroy(plan):
    p = _objects.pop(int(plan), None)
    if isinstance(p, _Plan):
        _L().vkx_plan_destroy(p.plan)
        _L().vkx_module_unload(p.module)


def descriptor_set_create(plan):
    if _get(plan, _Plan) is None:
        _set_error("Invalid compute plan handle for descriptor_set_create")
        return 0
    return _new(_DescriptorSet(int(plan)))


def descriptor_set_destroy(descriptor_set):
    _objects.pop(int(descriptor_set), None)


def descriptor_set_write_buffer(descriptor_set, binding, object, offset, range, uniform, read_access,
                                write_access):
    ds = _get(descriptor_set, _DescriptorSet)
    if ds is None:
        _set_error("Invalid descriptor set handle for descriptor_set_write_buffer")
        return
    ds.buffers[int(binding)] = (int(object), int(offset))


def command_list_create(context):
    if _get(context, _Context) is None:
        _set_error("Invalid context handle for command_list_create")
        return 0
    return _new(_CommandList(int(context)))


def command_list_destroy(command_list):
    _objects.pop(int(command_list), None)


def command_list_reset(command_list):
    cl = _get(command_list, _CommandList)
    if cl is not None:
        cl.commands = []


def command_list_get_instance_size(command_list):
    cl = _get(command_list, _CommandList)
    if cl is None:
        return 0
    return sum(_objects[p].pc_size for p, _, _ in cl.commands if isinstance(_objects.get(p), _Plan))


def stage_compute_record(command_list, plan, descriptor_set, blocks_x, blocks_y, blocks_z):
    cl, p = _get(command_list, _CommandList), _get(plan, _Plan)
    if cl is None or p is None:
        _set_error("Invalid command list or compute plan handle for stage_compute_record")
        return
    blocks = (int(blocks_x), int(blocks_y), int(blocks_z))
    if blocks != tuple(int(v) for v in p.meta["grid"]):
        _set_error(f"stage_compute_record: {blocks} workgroups requested, the plan's kernel was built for "
                   f"{tuple(p.meta['grid'])}")
        return
    cl.commands.append((int(plan), int(descriptor_set), blocks))


def command_list_submit(command_list, data, instance_count, index):
    """Launch every recorded stage instance_count times on the context's stream, each
    with its push-constant bytes from ``data`` (instance-major, then command order).
    Returns True (False would ask the caller to retry, which never happens here)."""
    cl = _get(command_list, _CommandList)
    ctx = _get(cl.ctx, _Context) if cl is not None else None
    if ctx is None:
        _set_error(f"Invalid command list {command_list} for command_list_submit")
        return True
    count = int(instance_count)
    if count <= 0:
        return True
    payload = bytes(data) if data is not None else b""
    per_instance = command_list_get_instance_size(command_list)
    if len(payload) != per_instance * count:
        _set_error(f"Push-constant data size mismatch. Expected {per_instance * count} bytes, got {len(payload)}.")
        return True
    L = _L()
    for inst in range(count):
        off = inst * per_instance
        for plan_h, ds_h, _blocks in cl.commands:
            p = _objects[plan_h]
            ds = _get(ds_h, _DescriptorSet) if ds_h else None
            ptrs = []
            for binding in sorted(ds.buffers) if ds is not None else []:
                bh, boff = ds.buffers[binding]
                b = _get(bh, _Buffer)
                if b is None:
                    _set_error(f"Invalid buffer handle {bh} in descriptor set {ds_h}")
                    return True
                ptrs.append(b.ptr + boff)
            consts = payload[off:off + p.pc_size]
            off += p.pc_size
            arr = (_U64 * max(1, len(ptrs)))(*ptrs)
            cbuf = ctypes.create_string_buffer(consts, max(1, len(consts)))
            if not _ok(L.vkx_plan_exec(p.plan, arr, len(ptrs), cbuf, len(consts), ctx.stream),
                       "command_list_submit"):
                return True
    return True


def context_synchronize(context) -> bool:
    """Wait for everything submitted on the context's stream (the reference waits on
    per-queue signals for this)."""
    ctx = _get(context, _Context)
    return ctx is not None and _ok(_L().vkx_stream_synchronize(ctx.stream), "context_synchronize")
