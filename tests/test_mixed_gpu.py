"""Mixed precision: 3-D C2C FFT of 256^3 complex stored as bf16 (and fp16) pairs,
radix-16 stages as split-precision MFMA GEMMs, FP32 output; compared with FP64
numpy.fft.fftn of the quantized input (rel-L2 <= 1e-3 for bf16, much lower for fp16)."""
import numpy as np
import pytest

import oracle as O
import vkdispatch_amd as vd
from vkdispatch_amd import dtypes
from vkdispatch_amd.fft.mixed import fft3_mixed

pytestmark = pytest.mark.gpu

SHAPE = (256, 256, 256)


def _volume(seed=0):
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(SHAPE)
    im = rng.standard_normal(SHAPE)
    return (re + 1j * im).astype(np.complex64)


@pytest.fixture(scope="module")
def volume():
    return _volume(0)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(1).standard_normal(10000).astype(np.float32)
    ours = dtypes.quantize_bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours, ref)


@pytest.mark.parametrize("store", ["bf16", "fp16"])
def test_fft3_mixed_256cubed(volume, store):
    vt = vd.complex_bf16 if store == "bf16" else vd.complex32
    inp = vd.Buffer(SHAPE, vt)
    inp.write(volume)
    xq = inp.read(0)                       # the quantized input actually stored
    out = vd.Buffer(SHAPE, vd.complex64)
    fft3_mixed(out, inp)
    res = out.read(0)
    ref = np.fft.fftn(xq.astype(np.complex128))
    err = O.rel_l2(res, ref)
    err_unq = O.rel_l2(res, np.fft.fftn(volume.astype(np.complex128)))
    print(f"{store}: rel-L2 vs FP64(quantized) {err:.3e}, vs FP64(unquantized) {err_unq:.3e}")
    assert err <= (1e-3 if store == "bf16" else 5e-4)  # fp16 intermediates: 2 roundings at 11 bits


def test_fft3_mixed_inverse_small_values(volume):
    inp = vd.Buffer(SHAPE, vd.complex32)
    inp.write(volume)
    xq = inp.read(0)
    out = vd.Buffer(SHAPE, vd.complex64)
    fft3_mixed(out, inp, inverse=True)
    ref = np.fft.ifftn(xq.astype(np.complex128))
    assert O.rel_l2(out.read(0), ref) <= 5e-4
