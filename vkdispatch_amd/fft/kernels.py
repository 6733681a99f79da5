"""Instantiation, compilation and launch of FFT-family kernels.

For each plan the host writes a short HIP translation unit: a ``struct Cfg``
of compile-time constants, the I/O functors (defaults from ``vkx_fft.h`` or
traced mapping hooks) and an ``extern "C"`` entry that calls one of the
hand-written bodies (``fft_body``, ``conv_body``, ``transpose_body``).  It is
compiled for the device arch with hiprtc through ``vkx_compile`` (with packed
FP32 disabled) and cached on disk by content hash, then loaded with
``vkx_module_load`` and bound to its twiddle tables with ``vkx_plan_create``.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import re
import threading
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .. import dtypes as dt
from .. import runtime as rt
from .planner import KernelPlan

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEVICE_INCLUDE = os.path.join(PKG_DIR, "csrc", "device")
HEADER = os.path.join(DEVICE_INCLUDE, "vkx_fft.h")

JIT_FLAGS = ["-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def arch() -> str:
    return os.environ.get("VKX_OFFLOAD_ARCH", "gfx950")


def cache_dir() -> str:
    d = os.environ.get("VKX_KERNEL_CACHE")
    if d:
        return d
    return os.path.join(PKG_DIR, "_kcache")


_header_digest = None
_compiler_tag = None


def compiler_tag() -> str:
    """Release of the HIP installation whose hiprtc compiles (torch's bundled copy when
    torch is installed, else the system ROCm)."""
    global _compiler_tag
    if _compiler_tag is None:
        v = ctypes.c_int()
        rt.check(rt.lib().vkx_compiler_version(ctypes.byref(v)), "vkx_compiler_version")
        _compiler_tag = f"hip-{v.value}"
    return _compiler_tag


def _header_hash() -> str:
    global _header_digest
    if _header_digest is None:
        h = hashlib.sha256()
        for name in sorted(os.listdir(DEVICE_INCLUDE)):
            if name.endswith(".h"):
                with open(os.path.join(DEVICE_INCLUDE, name), "rb") as f:
                    h.update(name.encode())
                    h.update(f.read())
        _header_digest = h.hexdigest()
    return _header_digest


def source_key(source: str) -> str:
    h = hashlib.sha256()
    h.update(source.encode())
    h.update(_header_hash().encode())
    h.update(arch().encode())
    h.update(" ".join(JIT_FLAGS).encode())
    h.update(compiler_tag().encode())
    return h.hexdigest()[:40]


def compile_source(source: str, name: str = "vkx_kernel.hip") -> bytes:
    """Compile (or fetch from the on-disk cache) the code object of ``source``."""
    key = source_key(source)
    d = cache_dir()
    path = os.path.join(d, key + ".hsaco")
    if os.path.exists(path):
        with open(path, "rb") as f:
            return f.read()
    opts = [f"--offload-arch={arch()}", *JIT_FLAGS, f"-I{DEVICE_INCLUDE}"]
    arr = (ctypes.c_char_p * len(opts))(*[o.encode() for o in opts])
    code = ctypes.c_void_p()
    size = ctypes.c_size_t()
    log = ctypes.c_void_p()
    L = rt.lib()
    rc = L.vkx_compile(source.encode(), name.encode(), arr, len(opts), ctypes.byref(code), ctypes.byref(size),
                       ctypes.byref(log))
    log_text = ""
    if log.value:
        log_text = ctypes.string_at(log.value).decode(errors="replace")
        L.vkx_free_host(log)
    if rc != 0:
        err = L.vkx_last_error().decode(errors="replace")
        L.vkx_clear_error()
        raise RuntimeError(f"kernel compilation failed: {err}\n{log_text}\n--- source ---\n{source}")
    data = ctypes.string_at(code.value, size.value)
    L.vkx_free_host(code)
    try:
        os.makedirs(d, exist_ok=True)
        tmp = f"{path}.{os.getpid()}.{threading.get_ident()}.tmp"
        with open(tmp, "wb") as f:
            f.write(data)
        os.replace(tmp, path)
    except OSError:
        pass
    return data


# ---------------------------------------------------------------------------
# source assembly
# ---------------------------------------------------------------------------
def _pl(passes: Sequence[Tuple[int, int]]) -> str:
    return "vkx::PL<" + ", ".join(f"vkx::P<{r}, {ns}>" for r, ns in passes) + ">"


def cfg_struct(plan: KernelPlan, *, in_os: int, out_os: int, k_os: int, sig: Tuple[int, int], r2c: bool,
               inverse: bool, normalize: bool, scale: float, knum: int = 1, k_inner_only: bool = False,
               k_transposed: bool = False, xcd_map: bool = False) -> str:
    """``struct Cfg`` of a generic FFT-family kernel.  xcd_map: deal the workgroups of one
    XCD consecutive tiles (applied when the grid is a multiple of 8 workgroups)."""
    b = lambda v: "true" if v else "false"  # noqa: E731
    lines = [
        "struct Cfg {",
        f"  using real = {plan.real};",
        f"  static constexpr int N = {plan.N}, TT = {plan.TT}, B = {plan.B}, BLOCK = {plan.BLOCK}, VMAX = {plan.VMAX};",
        f"  static constexpr bool COLFAST = {b(plan.COLFAST)}, WAVE_LOCAL = {b(plan.WAVE_LOCAL)}, "
        f"TWREG = {b(plan.TWREG)};",
        f"  static constexpr vkx::u32 S = {plan.S}u, OUTER = {plan.OUTER}u, NWG_INNER = {plan.NWG_INNER}u;",
        f"  static constexpr vkx::u32 IN_OS = {in_os}u, OUT_OS = {out_os}u, K_OS = {k_os}u;",
        f"  static constexpr int PAD_SHIFT = {plan.PAD_SHIFT}, NP = {plan.NP}, LDS_ELEMS = {plan.LDS_ELEMS};",
        f"  static constexpr int SIG_LO = {sig[0]}, SIG_HI = {sig[1]};",
        f"  static constexpr bool R2C = {b(r2c)}, INV = {b(inverse)}, NORMALIZE = {b(normalize)};",
        f"  static constexpr double SCALE = {scale!r};",
        f"  static constexpr int KNUM = {knum}, KSLOTS = {plan.KSLOTS};",
        f"  static constexpr bool KERNEL_INNER_ONLY = {b(k_inner_only)}, KERNEL_TRANSPOSED = {b(k_transposed)};",
        f"  using fwd = {_pl(plan.fwd)};",
        f"  using inv = {_pl(plan.inv)};",
    ]
    if xcd_map and plan.NWG % 8 == 0 and plan.NWG >= 16:
        lines.append(f"  static constexpr bool XCDMAP = true; static constexpr vkx::u32 NWG = {plan.NWG}u;")
    lines.append("};")
    return "\n".join(lines)


def kernel_symbol(kind: str, plan: KernelPlan, tag: str = "") -> str:
    """Descriptive entry-point name (what rocprof reports), e.g. vkx_fft_n4096_s1_o1024_fwd."""
    sym = f"vkx_{kind}_n{plan.N}_s{plan.S}_o{plan.OUTER}" + (f"_{tag}" if tag else "")
    return re.sub(r"[^A-Za-z0-9_]", "_", sym)


_SYMBOL_RE = re.compile(r'extern "C" __global__ void __launch_bounds__\(\d+\)(?: __attribute__\(\(\w+\([\d, ]+\)\)\))? (\w+)\(')


def source_symbol(source: str) -> str:
    m = _SYMBOL_RE.search(source)
    if not m:
        raise ValueError("kernel source has no extern \"C\" entry point")
    return m.group(1)


def translation_unit(cfg: str, functors: Sequence[str], body: str, block: int, symbol: str = "vkx_main") -> str:
    parts = ['#include "vkx_fft.h"', "", cfg, ""]
    parts += list(functors)
    parts += ["", f'extern "C" __global__ void __launch_bounds__({block}) {symbol}(vkx::KParams prm) {{',
              f"  {body}", "}", ""]
    return "\n".join(parts)


# ---------------------------------------------------------------------------
# loaded kernels
# ---------------------------------------------------------------------------
_modules: Dict[str, Tuple[int, int]] = {}
_mod_lock = threading.Lock()


LAUNCH_TAG = "// vkx-launch "


def launch_comment(meta: Dict) -> str:
    import json
    return LAUNCH_TAG + json.dumps(meta, separators=(",", ":")) + "\n"


def parse_launch_comment(source: str) -> Dict:
    """The launch description a kernel source carries in its first line."""
    import json
    first = source.split("\n", 1)[0]
    if not first.startswith(LAUNCH_TAG):
        raise ValueError("kernel source has no launch description line")
    return json.loads(first[len(LAUNCH_TAG):])


def load_function(source: str) -> int:
    key = source_key(source)
    with _mod_lock:
        if key in _modules:
            return _modules[key][1]
    code = compile_source(source)
    rt.ensure_device()
    L = rt.lib()
    mod = ctypes.c_uint64(0)
    buf = ctypes.create_string_buffer(code, len(code))
    rt.check(L.vkx_module_load(ctypes.byref(mod), buf, len(code)), "vkx_module_load")
    fn = ctypes.c_uint64(0)
    rt.check(L.vkx_module_function(ctypes.byref(fn), mod.value, source_symbol(source).encode()),
             "vkx_module_function")
    with _mod_lock:
        _modules[key] = (mod.value, fn.value)
    return fn.value


class CompiledKernel:
    """One FFT / convolution / transpose dispatch: source + plan + argument layout."""

    def __init__(self, source: str, plan: KernelPlan, args, twiddle_fwd: Optional[List[int]],
                 twiddle_inv: Optional[List[int]], name: str = "vkx_fft", roots_n: int = 0):
        self.plan = plan
        self.args = list(args)
        self.name = name
        self.tw_fwd = twiddle_fwd
        self.tw_inv = twiddle_inv
        self.roots_n = int(roots_n)
        # The translation unit describes its own launch in its first line, so a backend
        # that receives only the source (the reference's stage_compute_plan_create)
        # can bind the same geometry and twiddle tables (launch_metadata()).
        self.source = launch_comment(self.launch_metadata(source)) + source
        self._plan_handle = None
        self._lock = threading.Lock()

    def launch_metadata(self, source: Optional[str] = None) -> Dict:
        return {"symbol": source_symbol(source if source is not None else self.source),
                "grid": [int(v) for v in self.plan.grid], "block": [int(self.plan.BLOCK), 1, 1],
                "n": int(self.plan.N), "double": 1 if self.plan.real == "double" else 0,
                "radices_fwd": [int(r) for r in (self.tw_fwd or [])],
                "radices_inv": [int(r) for r in (self.tw_inv or [])],
                "roots_n": self.roots_n,
                "buffers": [a.role for a in self.args if a.kind == "buffer"]}

    def __str__(self):
        return self.source

    @property
    def num_buffers(self) -> int:
        return sum(1 for a in self.args if a.kind == "buffer")

    def handle(self) -> int:
        if self._plan_handle is None:
            with self._lock:
                if self._plan_handle is None:
                    check_no_spill(self)
                    rt.require_not_capturing(f"the first launch of {self.name}")
                    fn = load_function(self.source)
                    d = rt.PlanDesc()
                    d.function = fn
                    for i, v in enumerate(self.plan.grid):
                        d.grid[i] = v
                    d.block[0], d.block[1], d.block[2] = self.plan.BLOCK, 1, 1
                    d.shared_bytes = 0
                    d.n = self.plan.N
                    d.double_precision = 1 if self.plan.real == "double" else 0
                    if self.tw_fwd:
                        d.num_passes_fwd = len(self.tw_fwd)
                        for i, r in enumerate(self.tw_fwd):
                            d.radices_fwd[i] = r
                    if self.tw_inv:
                        d.num_passes_inv = len(self.tw_inv)
                        for i, r in enumerate(self.tw_inv):
                            d.radices_inv[i] = r
                    d.roots_n = self.roots_n
                    h = ctypes.c_uint64(0)
                    rt.check(rt.lib().vkx_plan_create(ctypes.byref(h), ctypes.byref(d)), "vkx_plan_create")
                    self._plan_handle = h.value
        return self._plan_handle

    def code_object_info(self) -> Dict[str, int]:
        """Resource figures from the compiled code object's metadata note (no device
        needed): VGPRs, SGPRs, static LDS bytes and private (spill) bytes per lane."""
        return code_object_info(compile_source(self.source))

    def attributes(self) -> Dict[str, int]:
        """Register count, static LDS, maximum block size and private (spill) bytes per
        lane of the loaded code object."""
        fn = load_function(self.source)
        vals = [ctypes.c_int(0) for _ in range(4)]
        rt.check(rt.lib().vkx_function_attrs(fn, *[ctypes.byref(v) for v in vals]), "vkx_function_attrs")
        return dict(zip(("num_regs", "shared_bytes", "max_threads", "private_bytes"), (v.value for v in vals)))

    def bind(self, values: Sequence) -> Tuple[List[int], bytes]:
        """Device pointers and packed constants for positional ``values``."""
        if len(values) != len(self.args):
            raise ValueError(f"{self.name} expects {len(self.args)} arguments "
                             f"({', '.join(a.kind + ':' + a.dtype.name for a in self.args)}), got {len(values)}")
        ptrs: List[int] = []
        consts = bytearray(256)
        used = 0
        for a, v in zip(self.args, values):
            if a.kind == "buffer":
                ptr = getattr(v, "ptr", None)
                if ptr is None:
                    raise ValueError(f"Expected a buffer for {a.role} argument, got {v!r}")
                ptrs.append(int(ptr))
            else:
                if hasattr(v, "ptr"):
                    raise ValueError(f"Expected a scalar value for {a.role} argument, got a buffer")
                if a.dtype.is_complex:
                    c = complex(v)
                    raw = np.array([c.real, c.imag], dtype=dt.storage_scalar_numpy(a.dtype)).tobytes()
                elif a.dtype is dt.bfloat16:
                    raw = dt.bf16_bits_from_float32(np.array([v], np.float32)).tobytes()
                else:
                    raw = np.array([v], dtype=a.dtype.numpy).tobytes()
                consts[a.slot:a.slot + len(raw)] = raw
                used = max(used, a.slot + len(raw))
        return ptrs, bytes(consts[:max(used, 0)])

    def launch(self, values: Sequence, stream: Optional[int] = None) -> None:
        ptrs, consts = self.bind(values)
        self.launch_raw(ptrs, consts, stream)

    def launch_raw(self, ptrs: Sequence[int], consts: bytes = b"", stream: Optional[int] = None) -> None:
        h = self.handle()
        arr = (ctypes.c_uint64 * max(1, len(ptrs)))(*ptrs)
        cbuf = ctypes.create_string_buffer(consts, max(1, len(consts)))
        s = rt.current_stream() if stream is None else int(stream)
        rt.check(rt.lib().vkx_plan_exec(h, arr, len(ptrs), cbuf, len(consts), s), f"{self.name} launch")

    def launch_timed(self, values: Sequence, start, stop, stream: Optional[int] = None) -> None:
        """One launch whose dispatch records its own start / stop times in two
        runtime Events (no marker between this launch and its neighbours)."""
        ptrs, consts = self.bind(values)
        h = self.handle()
        arr = (ctypes.c_uint64 * max(1, len(ptrs)))(*ptrs)
        cbuf = ctypes.create_string_buffer(consts, max(1, len(consts)))
        s = rt.current_stream() if stream is None else int(stream)
        rt.check(rt.lib().vkx_plan_exec_timed(h, arr, len(ptrs), cbuf, len(consts), s, start.handle, stop.handle),
                 f"{self.name} timed launch")


class SpillError(RuntimeError):
    """A hiprtc kernel needs private (scratch) memory for register spills."""


def check_no_spill(kernel: "CompiledKernel") -> None:
    """Refuse a kernel whose code object spills registers to private memory.  Every
    kernel the planners emit is built to stay in registers (the tuned plans, the
    pair-packed convolutions and every FFT shape of the test suite compile without
    spills); a spilling variant once produced wrong rows on the MI355X with no cause
    found, so a plan that spills is an error here, never a silent slow path."""
    info = kernel.code_object_info()
    if info["private_bytes"]:
        raise SpillError(f"{kernel.name}: {info['private_bytes']} bytes/lane of register spills "
                         f"({info['vgprs']} VGPRs); this plan is refused")


def code_object_info(code: bytes) -> Dict[str, int]:
    """Kernel resource figures from the NT_AMDGPU_METADATA note (msgpack) of an
    AMDGPU code object built here."""
    import struct

    import msgpack
    shoff = struct.unpack_from("<Q", code, 0x28)[0]
    shentsize, shnum = struct.unpack_from("<HH", code, 0x3A)
    for i in range(shnum):
        off = shoff + i * shentsize
        if struct.unpack_from("<I", code, off + 4)[0] != 7:  # SHT_NOTE
            continue
        sh_offset, sh_size = struct.unpack_from("<QQ", code, off + 0x18)
        pos = sh_offset
        while pos + 12 <= sh_offset + sh_size:
            namesz, descsz, ntype = struct.unpack_from("<III", code, pos)
            dpos = pos + 12 + ((namesz + 3) & ~3)
            if ntype == 32:
                md = msgpack.unpackb(code[dpos:dpos + descsz], raw=False)
                k = md["amdhsa.kernels"][0]
                return {"vgprs": int(k.get(".vgpr_count", 0)), "sgprs": int(k.get(".sgpr_count", 0)),
                        "lds_bytes": int(k.get(".group_segment_fixed_size", 0)),
                        "private_bytes": int(k.get(".private_segment_fixed_size", 0))}
            pos = dpos + ((descsz + 3) & ~3)
    raise ValueError("code object has no AMDGPU metadata note")


def precompile(kernels, workers: int = 0) -> int:
    """Compile the code objects of ``kernels`` (CompiledKernel or source strings) in
    parallel into the on-disk cache; returns how many were compiled."""
    from concurrent.futures import ThreadPoolExecutor
    srcs = []
    for k in kernels:
        s = k if isinstance(k, str) else k.source
        if not os.path.exists(os.path.join(cache_dir(), source_key(s) + ".hsaco")):
            srcs.append(s)
    srcs = list(dict.fromkeys(srcs))
    if not srcs:
        return 0
    if workers <= 0:
        workers = min(16, os.cpu_count() or 4)
    if rt.device_initialised():
        # never start worker processes from a process that holds a GPU context
        workers = 1
    if workers == 1 or len(srcs) == 1:
        for s in srcs:
            compile_source(s)
        return len(srcs)
    # hiprtc serialises compilations inside one process: use worker processes
    # (started with spawn, so they never inherit a GPU context)
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=min(workers, len(srcs)), mp_context=mp.get_context("spawn")) as ex:
        list(ex.map(_compile_worker, srcs, chunksize=1))
    return len(srcs)


def _compile_worker(source: str) -> int:
    return len(compile_source(source))
