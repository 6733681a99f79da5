"""Write tests/golden/fft_golden.npz: seeded inputs and expected outputs of the vd.fft
operations, computed in FP64 with numpy.fft (pocketfft), independently of the oracle,
and stored rounded to complex64 / float32 (relative rounding ~6e-8, far below the
1e-5 parity bars) to keep the fixture small.

Semantics (the reference's, vkdispatch/fft/functions.py): forward sign exp(-2 pi i kn/N);
inverse normalised by 1/N per axis; convolve(x, K) along an axis = ifft(fft(x) * conj(K))
with K the already-transformed kernel (vkdispatch/fft/shader_factories.py:136-145);
convolve2DR(x, K) = irfft2(rfft2(x) * conj(K)) per image, K broadcast with
kernel_inner_only; input_signal_range zeroes samples outside [s, e) before the FFT
(vkdispatch/fft/global_memory_iterators.py:167-195).

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def main():
    rng = np.random.default_rng(20240)
    d = {}
    # 1-D C2C along the last axis: power of two, mixed radix, prime
    for n in (64, 60, 17, 1024):
        x = crandn(rng, (4, n)).astype(np.complex64)
        d[f"c2c_n{n}_x"] = x
        d[f"c2c_n{n}_fwd"] = np.fft.fft(x.astype(np.complex128), axis=-1)
        d[f"c2c_n{n}_inv"] = np.fft.ifft(x.astype(np.complex128), axis=-1)
    # strided axis of a 3-D buffer
    x = crandn(rng, (6, 40, 12)).astype(np.complex64)
    d["c2c_axis1_x"] = x
    d["c2c_axis1_fwd"] = np.fft.fft(x.astype(np.complex128), axis=1)
    # 2-D / 3-D
    x = crandn(rng, (3, 32, 48)).astype(np.complex64)
    d["fft2_x"] = x
    d["fft2_fwd"] = np.fft.fftn(x.astype(np.complex128), axes=(1, 2))
    x = crandn(rng, (16, 12, 20)).astype(np.complex64)
    d["fft3_x"] = x
    d["fft3_fwd"] = np.fft.fftn(x.astype(np.complex128))
    # real transforms
    x = rng.standard_normal((5, 64)).astype(np.float32)
    d["rfft_x"] = x
    d["rfft_fwd"] = np.fft.rfft(x.astype(np.float64), axis=-1)
    x = rng.standard_normal((4, 24, 30)).astype(np.float32)
    d["rfft2_x"] = x
    d["rfft2_fwd"] = np.fft.rfft2(x.astype(np.float64))
    # signal range
    x = crandn(rng, (3, 128)).astype(np.complex64)
    xs = x.astype(np.complex128).copy()
    xs[:, :20] = 0
    xs[:, 90:] = 0
    d["range_x"] = x
    d["range_lo_hi"] = np.array([20, 90])
    d["range_fwd"] = np.fft.fft(xs, axis=-1)
    # 1-D convolution along the last axis with a transformed kernel
    x = crandn(rng, (4, 96)).astype(np.complex64)
    K = np.fft.fft(crandn(rng, (4, 96)), axis=-1).astype(np.complex64)
    d["conv_x"] = x
    d["conv_k"] = K
    d["conv_out"] = np.fft.ifft(np.fft.fft(x.astype(np.complex128), axis=-1) * np.conj(K.astype(np.complex128)),
                                axis=-1)
    # 2-D real convolution, one kernel spectrum broadcast over the batch
    for n in (16, 64, 128):
        x = rng.uniform(-1, 1, (3, n, n)).astype(np.float32)
        k = rng.standard_normal((n, n))
        K = np.fft.rfft2(k).astype(np.complex64)
        d[f"conv2dr_n{n}_x"] = x
        d[f"conv2dr_n{n}_k"] = K
        d[f"conv2dr_n{n}_out"] = np.fft.irfft2(np.fft.rfft2(x.astype(np.float64)) * np.conj(K.astype(np.complex128)),
                                               s=(n, n))
    # prime lengths (Bluestein: m = 256 and m = 16384, the latter as tiles)
    for n, rows in ((97, 3), (4099, 2)):
        x = crandn(rng, (rows, n)).astype(np.complex64)
        d[f"prime_n{n}_x"] = x
        d[f"prime_n{n}_fwd"] = np.fft.fft(x.astype(np.complex128), axis=-1)
        d[f"prime_n{n}_inv"] = np.fft.ifft(x.astype(np.complex128), axis=-1)
    # a ragged batch: transforms of mixed lengths packed back to back
    sizes = np.array([64, 97, 1000, 4099, 8192])
    x = crandn(rng, (int(sizes.sum()),)).astype(np.complex64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    d["ragged_sizes"] = sizes
    d["ragged_x"] = x
    d["ragged_fwd"] = np.concatenate([np.fft.fft(x[o:o + n].astype(np.complex128)) for o, n in zip(offs, sizes)])
    d = {k: (v.astype(np.complex64) if np.iscomplexobj(v) else v.astype(np.float32) if v.dtype == np.float64 else v)
         for k, v in d.items()}
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT}: {len(d)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
