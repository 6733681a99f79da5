"""vkdispatch_amd: an MI355X-native (gfx950) implementation of vkdispatch's FFT and
FFT-convolution path.

Use it as the reference's ``vd`` namespace for that path::

    import vkdispatch_amd as vd
    import vkdispatch_amd.codegen as vc

    buf = vd.asbuffer(x)            # numpy -> device
    vd.fft.fft(buf)                 # in place, exp(-2 pi i k n / N)
    vd.fft.ifft(buf)                # normalised inverse
    rbuf = vd.asrfftbuffer(img)     # packed real layout
    vd.fft.convolve2DR(rbuf, kbuf)  # irfft2(rfft2(x) * conj(K))

Host side: Python -> ctypes -> libvkdfft.so (include/vkdfft.h) -> HIP.
Kernels: hand-written templates in csrc/device/vkx_fft.h, instantiated per
plan and compiled for gfx950 with hiprtc.
"""
from . import codegen
from . import dtypes
from .buffer import (Buffer, RFFTBuffer, asbuffer, asrfftbuffer, from_cuda_array, from_torch,
                     rfftbuffer_from_cuda_array)
from .context import get_context, initialize, is_cuda, is_dummy, is_vulkan, make_context
from .dtypes import (bfloat16, complex32, complex64, complex128, complex_bf16, dtype, float16, float32, float64,
                    int16, int32, int64, uint16, uint32, uint64)
from .graph import CommandGraph, cuda_graph_capture, get_cuda_capture, global_graph, suspend_cuda_capture
from .runtime import Event, set_stream, stream_scope, synchronize
from .shader_map import MappingFunction, map
from . import fft
from . import vkfft

__version__ = "0.1.0"
