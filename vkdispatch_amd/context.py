"""Device context: the limits model the FFT planner reads.

The reference planner reads ``max_shared_memory``, ``max_workgroup_size``,
``max_workgroup_invocations`` and ``subgroup_size`` from its context
(vkdispatch/base/context.py:199-254, vkdispatch/fft/config.py:93-100), so its
plans are device-dependent.  Here the limits are the gfx950 (MI355X) values;
they are constants rather than queries so that plans, kernel sources and the
transposed-kernel layout are identical on a CPU-only host (where kernels are
compiled ahead of time) and on the GPU.
"""
from __future__ import annotations

import dataclasses
import os
import threading
from typing import List, Tuple


@dataclasses.dataclass
class DeviceInfo:
    name: str = "AMD Instinct MI355X (gfx950)"
    float_16_support: int = 1
    float_64_support: int = 1
    storage_buffer_16_bit_access: int = 1
    uniform_and_storage_buffer_16_bit_access: int = 1
    sub_group_size: int = 64
    max_compute_shared_memory_size: int = 160 * 1024


@dataclasses.dataclass
class Context:
    arch: str = "gfx950"
    compute_units: int = 256
    xcds: int = 8
    subgroup_size: int = 64
    max_shared_memory: int = 160 * 1024
    max_workgroup_size: Tuple[int, int, int] = (1024, 1024, 1024)
    max_workgroup_invocations: int = 1024
    max_workgroup_count: Tuple[int, int, int] = (2 ** 31 - 1, 65535, 65535)
    queue_count: int = 1
    device_infos: List[DeviceInfo] = dataclasses.field(default_factory=lambda: [DeviceInfo()])
    hbm_peak_bytes_per_s: float = 8.0e12


_ctx = None
_lock = threading.Lock()


def get_context() -> Context:
    global _ctx
    if _ctx is None:
        with _lock:
            if _ctx is None:
                _ctx = Context(arch=os.environ.get("VKX_OFFLOAD_ARCH", "gfx950"))
    return _ctx


def make_context(device_ids=None, device_count=None, queue_counts=None, queue_families=None,
                 use_cpu=False, multi_device=False, multi_queue=False) -> Context:
    """Single-device, single-stream context (the reference's multi-queue replication is not
    part of the FFT path)."""
    if multi_device or (device_count is not None and device_count > 1):
        raise ValueError("vkdispatch_amd runs one process per GPU; use one context per process")
    if use_cpu:
        raise ValueError("vkdispatch_amd has no CPU execution path")
    return get_context()


def initialize(backend=None, debug_mode=False, log_level=None, loader_debug_logs=False) -> None:
    """Accepts the reference's signature; the only backend is HIP."""
    if backend not in (None, "hip", "rocm"):
        raise ValueError(f"unsupported backend {backend!r}: vkdispatch_amd has a single HIP backend")
    get_context()


def is_cuda() -> bool:
    return False


def is_vulkan() -> bool:
    return False


def is_dummy() -> bool:
    return False
