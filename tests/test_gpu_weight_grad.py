"""mlff_weight_grad (the Linear-layer weight and bias gradients of the
training pass) against float64 products on the same inputs, including ragged
row counts, widths that are not multiples of the 64-wide tile, two fused
terms, row weights, and bit-identical repeats."""

import pytest
import torch

from conftest import pkg_mod

train = pkg_mod("train")


def _ref(a0, b0, a1=None, b1=None, w=None):
    d = lambda t: t.double().cpu()
    wa = (lambda t: t) if w is None else (lambda t: t * d(w)[:, None])
    out = wa(d(a0)).T @ d(b0)
    if a1 is not None:
        out = out + wa(d(a1)).T @ d(b1)
    return out, wa(d(a0)).sum(0)


@pytest.mark.gpu
@pytest.mark.parametrize("K,M,N", [(0, 8, 8), (1, 4, 4), (37, 20, 20), (2736, 384, 128),
                                   (2736, 128, 256), (2736, 384, 384), (50132, 384, 128),
                                   (50132, 128, 20), (50133, 64, 36), (300000, 128, 20)])
def test_weight_grad_matches_float64(gpu, K, M, N):
    print(f"[case] weight_grad K={K} M={M} N={N}", flush=True)
    g = torch.Generator().manual_seed(K + 7 * M + N)
    mk = lambda c: torch.randn(K, c, generator=g).to(gpu)
    a0, b0, a1, b1 = mk(M), mk(N), mk(M), mk(N)
    w = torch.rand(K, generator=g).to(gpu)
    for two in (False, True):
        for rw in (None, w):
            args = (a0, b0, a1, b1) if two else (a0, b0)
            out, bias = train._wgrad(*args, w=rw, bias=True)
            again, bias2 = train._wgrad(*args, w=rw, bias=True)
            torch.cuda.synchronize()
            assert torch.equal(out, again) and torch.equal(bias, bias2)
            ref, rbias = _ref(*args, w=rw)
            # float32 products summed over K rows: error grows like sqrt(K)
            scale = max(K, 1) ** 0.5
            tol = 2e-6 * scale * (2 if two else 1)
            assert out.shape == (M, N) and bias.shape == (M,)
            assert (out.double().cpu() - ref).abs().max().item() <= tol * 4 + 1e-30
            assert (bias.double().cpu() - rbias).abs().max().item() <= tol * 2 + 1e-30


@pytest.mark.gpu
def test_weight_grad_on_strided_views(gpu):
    """Column slices of wider rows (row stride above the width) give the
    same bits as their contiguous copies."""
    print("[case] weight_grad strided views", flush=True)
    g = torch.Generator().manual_seed(3)
    full = torch.randn(5000, 3 * 128, generator=g).to(gpu)
    b = torch.randn(5000, 64, generator=g).to(gpu)
    view = full[:, 128:256]
    assert view.stride(0) == 384
    o1, _ = train._wgrad(view, b)
    o2, _ = train._wgrad(view.contiguous(), b)
    torch.cuda.synchronize()
    assert torch.equal(o1, o2)


@pytest.mark.gpu
def test_silu_chain_matches_float64(gpu):
    """mlff_silu_chain: silu, its tangent / adjoint product and the tangent
    of the adjoint against float64 autograd of x sigmoid(x)."""
    print("[case] silu chain", flush=True)
    g = torch.Generator().manual_seed(11)
    n = 4 * 12345
    a, t, gg, dg = (torch.randn(n, generator=g) * 4 for _ in range(4))
    A = a.double().requires_grad_(True)
    h = A * torch.sigmoid(A)
    (d1,) = torch.autograd.grad(h.sum(), A, create_graph=True)
    (d2,) = torch.autograd.grad(d1.sum(), A)
    d1 = d1.detach()
    dev = lambda x: x.to(gpu)
    got_h = train._silu(dev(a)).double().cpu()
    got_d = train._silu_d(dev(a), dev(t)).double().cpu()
    got_dd = train._silu_dd(dev(a), dev(gg), dev(t), dev(dg)).double().cpu()
    got_dd0 = train._silu_dd(dev(a), dev(gg), dev(t)).double().cpu()
    rel = lambda x, y: ((x - y).abs() / (y.abs() + 1)).max().item()
    assert rel(got_h, h.detach()) < 2e-6
    assert rel(got_d, d1 * t.double()) < 2e-6
    assert rel(got_dd, dg.double() * d1 + gg.double() * d2 * t.double()) < 4e-6
    assert rel(got_dd0, gg.double() * d2 * t.double()) < 4e-6


