"""Autograd behaviour on the GPU, mirroring the reference's own tests
(tests/test_grad.py:41-171, tests/test_vmap.py:12-85 of DiffAPF/torchlpc):
gradcheck with forward AD, gradgradcheck, fp32 vs fp64, jacfwd vs backward."""
from __future__ import annotations

import pytest
import torch
import torch.nn.functional as F
from torch.autograd.gradcheck import gradcheck, gradgradcheck
from torch.func import jacfwd

from torchlpc_amd import LPC, Recurrence, sample_wise_lpc

pytestmark = pytest.mark.gpu


def random_biquads(gen, cmplx=False):
    if cmplx:
        mag = torch.rand(2, dtype=torch.double, generator=gen)
        phase = torch.rand(2, dtype=torch.double, generator=gen) * 2 * torch.pi
        roots = mag * torch.exp(1j * phase)
        return torch.stack([-roots[0] - roots[1], roots[0] * roots[1]])
    mag = torch.rand(1, dtype=torch.double, generator=gen)
    phase = torch.rand(1, dtype=torch.double, generator=gen) * torch.pi
    return torch.cat([-mag * torch.cos(phase) * 2, mag**2])


def biquad_inputs(B, T, cmplx=False, seed=0):
    """Per-sample linear interpolation between two random biquads (order 2)."""
    gen = torch.Generator().manual_seed(seed)
    c0, c1 = random_biquads(gen, cmplx), random_biquads(gen, cmplx)
    dtype = torch.complex128 if cmplx else torch.double
    w = torch.linspace(0, 1, T, dtype=torch.double)[:, None]
    A = ((1 - w) * c0[None, :] + w * c1[None, :]).to(dtype).unsqueeze(0).repeat(B, 1, 1)
    x = torch.randn(B, T, dtype=dtype, generator=gen)
    zi = torch.randn(B, 2, dtype=dtype, generator=gen)
    return x.cuda(), A.cuda(), zi.cuda()


@pytest.mark.parametrize("a_grad", [True, False])
@pytest.mark.parametrize("zi_grad", [True, False])
@pytest.mark.parametrize("cmplx", [False, True])
def test_low_order_gradcheck(a_grad, zi_grad, cmplx):
    x, A, zi = biquad_inputs(4, 32, cmplx, seed=int(a_grad) * 2 + int(zi_grad))
    x.requires_grad_(True)
    A.requires_grad_(a_grad)
    zi.requires_grad_(zi_grad)
    assert gradcheck(LPC.apply, (x, A, zi), check_forward_ad=True)
    assert gradgradcheck(LPC.apply, (x, A, zi))


@pytest.mark.parametrize("order", [3, 16])
def test_higher_order_gradcheck(order):
    gen = torch.Generator().manual_seed(order)
    B, T = 2, 40
    A = (torch.rand(B, T, order, dtype=torch.double, generator=gen) - 0.5) * (0.6 / order)
    x = torch.randn(B, T, dtype=torch.double, generator=gen)
    zi = torch.randn(B, order, dtype=torch.double, generator=gen)
    args = tuple(t.cuda().requires_grad_() for t in (x, A, zi))
    assert gradcheck(LPC.apply, args, check_forward_ad=True)


def test_time_invariant_gradcheck():
    gen = torch.Generator().manual_seed(4)
    B, T, N = 3, 30, 3
    a = ((torch.rand(B, N, dtype=torch.double, generator=gen) - 0.5) * 0.5).cuda().requires_grad_()
    x = torch.randn(B, T, dtype=torch.double, generator=gen).cuda().requires_grad_()
    zi = torch.randn(B, N, dtype=torch.double, generator=gen).cuda().requires_grad_()
    assert gradcheck(LPC.apply, (x, a, zi), check_forward_ad=True)
    assert gradgradcheck(LPC.apply, (x, a, zi))


def test_float64_vs_32():
    x, A, zi = biquad_inputs(4, 32, seed=7)
    y64 = LPC.apply(x, A, zi)
    y32 = LPC.apply(x.float(), A.float(), zi.float())
    assert torch.allclose(y64, y32.double(), atol=1e-6), (y64 - y32.double()).abs().max()


@pytest.mark.parametrize("a_grad", [True, False])
@pytest.mark.parametrize("zi_grad", [True, False])
@pytest.mark.parametrize("cmplx", [False, True])
def test_recurrence_gradcheck(a_grad, zi_grad, cmplx):
    gen = torch.Generator().manual_seed(11 + int(cmplx))
    B, T = 2, 123
    dtype = torch.complex128 if cmplx else torch.double
    x = torch.randn(B, T, dtype=dtype, generator=gen)
    if cmplx:
        A = torch.rand(B, T, dtype=torch.double, generator=gen).sqrt() * torch.exp(
            1j * torch.rand(B, T, dtype=torch.double, generator=gen) * 2 * torch.pi
        )
    else:
        A = torch.rand(B, T, dtype=dtype, generator=gen) * 2 - 1
    zi = torch.randn(B, dtype=dtype, generator=gen)
    x, A, zi = x.cuda().requires_grad_(), A.cuda().requires_grad_(a_grad), zi.cuda().requires_grad_(zi_grad)
    assert gradcheck(Recurrence.apply, (A, x, zi), check_forward_ad=True)
    assert gradgradcheck(Recurrence.apply, (A, x, zi))


def test_vmap_jacfwd_matches_backward():
    B, T = 4, 40
    x, A, zi = biquad_inputs(B, T, seed=3)
    target = torch.randn_like(x)
    a = A[:, 0, :].clone()
    args = (x.requires_grad_(), a.requires_grad_(), zi.requires_grad_())

    def func(x, a, zi):
        return F.mse_loss(LPC.apply(x, a[:, None, :].expand(-1, T, -1), zi), target)

    jacs = jacfwd(func, argnums=(0, 1, 2))(*args)
    func(*args).backward()
    for jac, arg in zip(jacs, args):
        assert torch.allclose(jac, arg.grad)


def test_recurrence_vmap_jacfwd_matches_backward():
    gen = torch.Generator().manual_seed(5)
    B, T = 3, 255
    x = torch.randn(B, T, dtype=torch.double, generator=gen).cuda()
    A = (torch.rand(B, T, dtype=torch.double, generator=gen) * 2 - 1).cuda()
    zi = torch.randn(B, dtype=torch.double, generator=gen).cuda()
    y = torch.randn(B, T, dtype=torch.double, generator=gen).cuda()
    args = (x.requires_grad_(), A.requires_grad_(), zi.requires_grad_())

    def func(x, A, zi):
        return F.mse_loss(Recurrence.apply(A, x, zi), y)

    jacs = jacfwd(func, argnums=(0, 1, 2))(*args)
    func(*args).backward()
    for jac, arg in zip(jacs, args):
        assert torch.allclose(jac, arg.grad)


def test_sample_wise_lpc_order_one_matches_lpc_math():
    gen = torch.Generator().manual_seed(9)
    B, T = 3, 500
    a = ((torch.rand(B, T, 1, generator=gen, dtype=torch.double) - 0.5) * 1.8).cuda()
    x = torch.randn(B, T, generator=gen, dtype=torch.double).cuda()
    zi = torch.randn(B, 1, generator=gen, dtype=torch.double).cuda()
    y = sample_wise_lpc(x, a, zi)
    # y[t] = x[t] - a[t] y[t-1]
    ref = torch.empty_like(x)
    prev = zi[:, 0]
    for t in range(T):
        prev = x[:, t] - a[:, t, 0] * prev
        ref[:, t] = prev
    assert torch.allclose(y, ref, rtol=1e-12, atol=1e-12)
