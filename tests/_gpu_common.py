"""Shared helpers of the GPU parity tests (seeded inputs, shape generators, checks)."""
import random

import numpy as np

import oracle as O

PRIMES = [2, 3, 5, 7, 11, 13]
MAX_DIM = 1024          # min(max_shared_memory / 8, max_workgroup_size[0]) on gfx950
MAX_BYTES = 2 ** 20     # reference size cap: prod(shape) * 8 < 2^20 (tests/test_fft.py:32-34)

# FP32 parity bound: each side within ~5e-6 of the FP64 truth (reference parity proxy)
REL_L2_FP32 = 5e-6


def shape_sequence(seed: int, min_dims: int, count: int = 2, max_dims: int = 3):
    """Seeded version of the reference's growing-shape generator (tests/test_fft.py:16-55)."""
    rnd = random.Random(seed)
    out = []
    for _ in range(count):
        dims = rnd.choice(list(range(min_dims, max_dims + 1)))
        shape = [rnd.choice(PRIMES) for _ in range(dims)]
        while all(d <= MAX_DIM for d in shape) and int(np.prod(shape)) * 8 < MAX_BYTES:
            out.append(tuple(shape))
            shape[rnd.randrange(dims)] *= rnd.choice(PRIMES)
    return out


def crand(rng, shape, dtype=np.complex64):
    return (rng.random(shape) + 1j * rng.random(shape)).astype(dtype)


def assert_close(got, ref, rel=REL_L2_FP32, atol=None):
    e = O.rel_l2(got, ref)
    assert e <= rel, f"rel-L2 {e:.3e} > {rel:.1e}"
    if atol is not None:
        assert np.allclose(got, ref, atol=atol), "allclose failed"
    return e


def precompile(kernels):
    from vkdispatch_amd.fft import kernels as K
    K.precompile(kernels)
