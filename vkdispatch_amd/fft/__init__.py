"""vd.fft: FFT, real FFT and FFT-convolution on gfx950 (the reference's
vkdispatch/fft/__init__.py:1-36 surface)."""
from .context import FFTContext, fft_context, global_reads_iterator, global_writes_iterator
from .functions import (convolve, convolve2D, convolve2DR, fft, fft2, fft3, ifft, ifft2, ifft3, irfft, irfft2,
                        irfft3, rfft, rfft2, rfft3, transpose)
from .hooks import read_op, write_op
from .planner import AxisGeometry, KernelPlan, make_plan, radix_plan
from .prime_utils import pad_dim, prime_factors
from .shader_factories import (cache_clear, get_cache_info, get_transposed_size, make_convolution_shader,
                               make_fft_shader, make_transpose_shader, mapped_kernel_index, print_cache_info)
from .src_functions import (convolve2D_print_src, convolve2D_src, convolve2DR_print_src, convolve2DR_src,
                            convolve_print_src, convolve_src, fft2_print_src, fft2_src, fft3_print_src, fft3_src,
                            fft_print_src, fft_src, ifft2_print_src, ifft2_src, ifft3_print_src, ifft3_src,
                            ifft_print_src, ifft_src, irfft2_print_src, irfft2_src, irfft3_print_src, irfft3_src,
                            irfft_print_src, irfft_src, rfft2_print_src, rfft2_src, rfft3_print_src, rfft3_src,
                            rfft_print_src, rfft_src, transpose_print_src, transpose_src)
