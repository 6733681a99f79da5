"""Differentiable FFT-convolution: the hand-written adjoint against FP64 central
finite differences (eps 1e-4, 64 seeded entries of each gradient), against
torch.fft autograd (FP64 and FP32), and against the FP64 oracle forward."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _fft_conv():
    from vkdispatch_amd.fft.autograd import fft_convolve2d
    return fft_convolve2d


def _torch_ref(x, k):
    H, W = x.shape[-2:]
    return torch.fft.irfft2(torch.fft.rfft2(x) * torch.conj(torch.fft.rfft2(k)), s=(H, W))


def _loss_np(x, k, t):
    y = O.convolve2DR(x, np.fft.rfft2(k), kernel_inner_only=True)
    return 0.5 * np.sum((y - t) ** 2)


def test_forward_matches_oracle_and_convolve2DR_semantics():
    f = _fft_conv()
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 64, 48))
    k = 0.1 * rng.standard_normal((64, 48))
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 2e-6)):
        xt = torch.tensor(x, dtype=dtype, device="cuda")
        kt = torch.tensor(k, dtype=dtype, device="cuda")
        y = f(xt, kt).cpu().numpy()
        ref = O.convolve2DR(x, np.fft.rfft2(k), kernel_inner_only=True)
        assert O.rel_l2(y, ref) < tol
        # y[n] = sum_m x[n+m] k[m] (circular cross-correlation)
    small_x, small_k = x[:1, :6, :4], k[:6, :4]
    ys = f(torch.tensor(small_x, device="cuda"), torch.tensor(small_k, device="cuda")).cpu().numpy()
    assert np.allclose(ys, O.correlate_bruteforce(small_x, small_k).real)


def test_gradients_vs_fp64_central_differences():
    f = _fft_conv()
    rng = np.random.default_rng(1)
    B, H, W = 2, 32, 24
    x = rng.standard_normal((B, H, W))
    k = 0.01 * rng.standard_normal((H, W))
    t = rng.standard_normal((B, H, W))
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    kt = torch.tensor(k, device="cuda", requires_grad=True)
    y = f(xt, kt)
    loss = 0.5 * ((y - torch.tensor(t, device="cuda")) ** 2).sum()
    loss.backward()
    dx = xt.grad.cpu().numpy()
    dk = kt.grad.cpu().numpy()
    eps = 1e-4
    for arr, grad, which in ((x, dx, "x"), (k, dk, "k")):
        flat = rng.choice(arr.size, size=64, replace=False)
        for idx in flat:
            pos = np.unravel_index(idx, arr.shape)
            ap, am = arr.copy(), arr.copy()
            ap[pos] += eps
            am[pos] -= eps
            if which == "x":
                fd = (_loss_np(ap, k, t) - _loss_np(am, k, t)) / (2 * eps)
            else:
                fd = (_loss_np(x, ap, t) - _loss_np(x, am, t)) / (2 * eps)
            assert abs(fd - grad[pos]) <= 1e-6 * max(1.0, abs(fd)), (which, pos, fd, grad[pos])


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-5)])
def test_gradients_vs_torch_fft_autograd(dtype, tol):
    f = _fft_conv()
    g = torch.Generator(device="cpu").manual_seed(2)
    x = torch.randn(4, 128, 96, generator=g, dtype=dtype).cuda()
    k = (0.01 * torch.randn(128, 96, generator=g, dtype=dtype)).cuda()
    t = torch.randn(4, 128, 96, generator=g, dtype=dtype).cuda()
    xa, ka = x.clone().requires_grad_(), k.clone().requires_grad_()
    xb, kb = x.clone().requires_grad_(), k.clone().requires_grad_()
    (0.5 * ((f(xa, ka) - t) ** 2).sum()).backward()
    x64, k64 = x.double().requires_grad_(), k.double().requires_grad_()
    (0.5 * ((_torch_ref(x64, k64) - t.double()) ** 2).sum()).backward()
    assert O.rel_l2(xa.grad.double().cpu().numpy(), x64.grad.cpu().numpy()) < tol
    assert O.rel_l2(ka.grad.double().cpu().numpy(), k64.grad.cpu().numpy()) < tol


def test_differentiable_convolution_32x1024x1024():
    """Batch 32 of float32 1024x1024 N(0,1) images, learnable 1024x1024 kernel N(0, 0.01),
    L = 0.5 ||y - t||^2 with t ~ N(0,1): gradients vs torch.fft autograd in FP64."""
    f = _fft_conv()
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.randn(32, 1024, 1024, generator=g).cuda()
    k = (0.01 * torch.randn(1024, 1024, generator=g)).cuda()
    t = torch.randn(32, 1024, 1024, generator=g).cuda()
    xa, ka = x.clone().requires_grad_(), k.clone().requires_grad_()
    y = f(xa, ka)
    (0.5 * ((y - t) ** 2).sum()).backward()
    x64, k64 = x.double().requires_grad_(), k.double().requires_grad_()
    y64 = _torch_ref(x64, k64)
    (0.5 * ((y64 - t.double()) ** 2).sum()).backward()
    assert O.rel_l2(y.detach().double().cpu().numpy(), y64.detach().cpu().numpy()) < 1e-5
    assert O.rel_l2(xa.grad.double().cpu().numpy(), x64.grad.cpu().numpy()) < 1e-5
    assert O.rel_l2(ka.grad.double().cpu().numpy(), k64.grad.cpu().numpy()) < 1e-5
