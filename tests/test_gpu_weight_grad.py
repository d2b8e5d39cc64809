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
def test_grouped_partial_sums_equal_single_launches(gpu):
    """Weight gradients whose partial sums are added by one grouped launch
    (mlff_weight_grad_partials + mlff_weight_grad_reduce_group) have the
    bits of mlff_weight_grad, for several shapes in one group, an empty row
    range, and grids smaller than the work (the blocks then loop)."""
    print("[case] weight_grad grouped reduce", flush=True)
    g = torch.Generator().manual_seed(21)
    cases = [(50132, 384, 128, True, True), (2736, 128, 256, False, True),
             (37, 20, 20, True, False), (0, 8, 8, False, False), (50133, 64, 36, False, True)]
    ops = []
    for K, M, N, two, rw in cases:
        mk = lambda c: torch.randn(K, c, generator=g).to(gpu)
        a0, b0 = mk(M), mk(N)
        extra = (mk(M), mk(N)) if two else ()
        w = torch.rand(K, generator=g).to(gpu) if rw else None
        ops.append(((a0, b0) + extra, w))
    want = [train._wgrad(*args, w=w, bias=True) for args, w in ops]
    lib = train.C.lib()
    for n_blocks in (None, 1, 7):
        grp = train._WgGroup()
        got = [train._wgrad(*args, w=w, bias=True, group=grp) for args, w in ops]
        if n_blocks is None:
            grp.flush()
        else:   # the same table on a fixed small grid
            table = torch.tensor(grp.rows, dtype=torch.int64).to(gpu)
            train.C.check("reduce_group", lib.mlff_weight_grad_reduce_group(
                train.C.stream_handle(), len(grp.rows), table.data_ptr(), n_blocks))
        torch.cuda.synchronize()
        for (o1, b1), (o2, b2) in zip(want, got):
            assert torch.equal(o1, o2) and torch.equal(b1, b2)


@pytest.mark.gpu
@pytest.mark.parametrize("n,H", [(2736, 128), (37, 128), (0, 128), (100000, 32), (5000, 64)])
def test_mixing_matrix_gradient_matches_float64(gpu, n, H):
    """The 3x3 mixing-matrix gradient (sum over atoms and channels of the xyz
    outer products, one or two terms, with and without row weights) against
    float64; bit-identical on a repeat and through a deferred group."""
    print(f"[case] xyz_outer n={n} H={H}", flush=True)
    g = torch.Generator().manual_seed(n + H)
    mk = lambda: torch.randn(n, 3, H, generator=g).to(gpu)
    a, b, a1, b1 = mk(), mk(), mk(), mk()
    w = torch.rand(n, generator=g).to(gpu)
    d = lambda t: t.double().cpu()
    for two in (False, True):
        for rw in (None, w):
            args = (a, b, a1, b1) if two else (a, b)
            got = train._xyz_outer(*args, w=rw)
            again = train._xyz_outer(*args, w=rw)
            grp = train._WgGroup()
            late = train._xyz_outer(*args, w=rw, group=grp)
            grp.flush()
            torch.cuda.synchronize()
            assert got.shape == (3, 3)
            assert torch.equal(got, again) and torch.equal(got, late)
            wd = torch.ones(n, dtype=torch.float64) if rw is None else d(rw)
            ref = torch.einsum("n,nxh,nzh->xz", wd, d(a), d(b))
            if two:
                ref = ref + torch.einsum("n,nxh,nzh->xz", wd, d(a1), d(b1))
            scale = max(n * H, 1) ** 0.5
            assert (d(got) - ref).abs().max().item() <= 4e-6 * scale * (2 if two else 1) + 1e-30


@pytest.mark.gpu
@pytest.mark.parametrize("periodic", [False, True])
def test_pair_geometry_matches_float64(gpu, periodic):
    """The fused pair geometry and basis (vec, d, den, u, e, de/dd) and its
    tangent along a position direction (dd, du, de_t) against float64 of the
    same formulas, with pairs inside and beyond the cutoff and, for the
    periodic case, an image shift per pair."""
    print(f"[case] pair geometry periodic={periodic}", flush=True)
    g = torch.Generator().manual_seed(17 + int(periodic))
    n, P, R, rc = 400, 5000, 20, 5.0
    x = (torch.rand(n, 3, generator=g) * 8.0).to(gpu)
    lam = torch.randn(n, 3, generator=g).to(gpu)
    lo = torch.randint(0, n - 1, (P,), generator=g)
    hi = torch.minimum(lo + 1 + torch.randint(0, 30, (P,), generator=g), torch.tensor(n - 1))
    shift = (torch.randn(P, 3, generator=g) * 2.0).to(gpu) if periodic else None
    lo, hi = lo.to(gpu), hi.to(gpu)
    pg = train.PairGraph(None, None, None, lo, hi, lo.to(torch.int32), hi.to(torch.int32),
                         shift, None, None, 1)
    centers = torch.linspace(0.0, rc, R).to(gpu)
    widths = (0.5 + torch.rand(R, generator=g)).to(gpu)
    vec, d, den, u, e, de = train._geometry(pg, x, centers, widths, rc)
    dd, du, de_t = train._geometry(pg, x, centers, widths, rc, lam)
    torch.cuda.synchronize()
    D = lambda t: t.double().cpu()
    rvec = D(x)[D(hi).long()] - D(x)[D(lo).long()]
    if periodic:
        rvec = rvec + D(shift)
    rd = rvec.norm(dim=1)
    rden = rd + train.EPS
    ru = rvec / rden[:, None]
    re, rde = train._edge_features(rd, D(centers), D(widths), rc)
    rdvec = D(lam)[D(hi).long()] - D(lam)[D(lo).long()]
    rdd = (rvec * rdvec).sum(1) / rd
    rdu = rdvec / rden[:, None] - rvec * (rdd / rden ** 2)[:, None]
    rde_t = rde * rdd[:, None]
    assert bool((rd > rc).any()) and bool((rd < rc).any())
    for got, ref, name in ((vec, rvec, "vec"), (d, rd, "d"), (den, rden, "den"), (u, ru, "u"),
                           (e, re, "e"), (de, rde, "de"), (dd, rdd, "dd"), (du, rdu, "du"),
                           (de_t, rde_t, "de_t")):
        err = (D(got) - ref).abs().max().item()
        assert err <= 4e-6 * max(ref.abs().max().item(), 1.0), (name, err)


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




@pytest.mark.gpu
@pytest.mark.parametrize("H,R,P", [(128, 20, 50132), (128, 20, 37), (64, 16, 1000),
                                   (32, 16, 4099), (128, 32, 300)])
def test_train_filter_matches_float64(gpu, H, R, P):
    """The fused filter MLP of the training pass (forward: a1, silu(a1) and
    W; tangent: da1, silu'(a1) da1 and dW) against float64 products of the
    same float32 inputs, on full and ragged tiles."""
    print(f"[case] train_filter H={H} R={R} P={P}", flush=True)
    g = torch.Generator().manual_seed(H + R + P)
    pre = "f"
    p = {pre + ".0.weight": torch.randn(H, R, generator=g) / R ** 0.5,
         pre + ".0.bias": 0.1 * torch.randn(H, generator=g),
         pre + ".2.weight": torch.randn(3 * H, H, generator=g) / H ** 0.5,
         pre + ".2.bias": 0.1 * torch.randn(3 * H, generator=g)}
    p = {k: v.to(gpu) for k, v in p.items()}
    e = torch.rand(P, R, generator=g).to(gpu)
    de = torch.randn(P, R, generator=g).to(gpu)
    a1, h, W = train._filter(p, pre, e)
    da1, dh, dW = train._filter(p, pre, de, a1)
    again = train._filter(p, pre, de, a1)
    torch.cuda.synchronize()
    assert all(torch.equal(x, y) for x, y in zip((da1, dh, dW), again))
    d = {k: v.double().cpu() for k, v in p.items()}
    ra1 = e.double().cpu() @ d[pre + ".0.weight"].T + d[pre + ".0.bias"]
    rh = torch.nn.functional.silu(ra1)
    rW = rh @ d[pre + ".2.weight"].T + d[pre + ".2.bias"]
    sig = torch.sigmoid(ra1)
    rda1 = de.double().cpu() @ d[pre + ".0.weight"].T
    rdh = sig * (1 + ra1 * (1 - sig)) * rda1
    rdW = rdh @ d[pre + ".2.weight"].T
    for got, ref, name in ((a1, ra1, "a1"), (h, rh, "h"), (W, rW, "W"), (da1, rda1, "da1"),
                           (dh, rdh, "dh"), (dW, rdW, "dW")):
        err = (got.double().cpu() - ref).abs().max().item()
        scale = ref.abs().max().item()
        # float32 inputs and float32 accumulation: a few ulps of the largest term
        assert err <= 2e-6 * max(scale, 1.0), (name, err, scale)
    # a weight update is seen: the prepared weights follow the version counter
    with torch.no_grad():
        p[pre + ".2.weight"].mul_(2.0)
    _, _, W2x = train._filter(p, pre, e)
    torch.cuda.synchronize()
    err = (W2x.double().cpu() - (2.0 * (rW - d[pre + ".2.bias"]) + d[pre + ".2.bias"])).abs().max()
    assert err.item() <= 4e-6 * max(rW.abs().max().item(), 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("H,P", [(128, 50132), (128, 37), (64, 1000), (32, 4099), (128, 300)])
def test_train_filter_adjoint_matches_float64(gpu, H, P):
    """The fused reverse of the filter's first layer (ga1 = silu'(a1) gW W2
    and its tangent dga1 = silu'(a1) dgW W2 + silu''(a1) da1 gW W2) against
    float64 of the same float32 inputs, on full and ragged tiles, bit-stable
    on a repeat; the forward-only form gives the same ga1."""
    print(f"[case] train_filter_adjoint H={H} P={P}", flush=True)
    g = torch.Generator().manual_seed(5 * H + P)
    pre = "f"
    p = {pre + ".0.weight": torch.randn(H, 20, generator=g) / 20 ** 0.5,
         pre + ".0.bias": 0.1 * torch.randn(H, generator=g),
         pre + ".2.weight": torch.randn(3 * H, H, generator=g) / H ** 0.5,
         pre + ".2.bias": 0.1 * torch.randn(3 * H, generator=g)}
    p = {k: v.to(gpu) for k, v in p.items()}
    gW = torch.randn(P, 3 * H, generator=g).to(gpu)
    dgW = torch.randn(P, 3 * H, generator=g).to(gpu)
    a1 = 3 * torch.randn(P, H, generator=g).to(gpu)
    da1 = torch.randn(P, H, generator=g).to(gpu)
    ga, dga = train._filter_adjoint(p, pre, gW, a1, dgW, da1)
    ga2, dga2 = train._filter_adjoint(p, pre, gW, a1, dgW, da1)
    ga0, none = train._filter_adjoint(p, pre, gW, a1)
    torch.cuda.synchronize()
    assert none is None and torch.equal(ga, ga2) and torch.equal(dga, dga2)
    assert torch.equal(ga, ga0)
    d = lambda t: t.double().cpu()
    W2 = d(p[pre + ".2.weight"])
    x = d(a1)
    s = torch.sigmoid(x)
    d1 = s * (1 + x * (1 - s))
    d2 = s * (1 - s) * (2 + x * (1 - 2 * s))
    gh, dgh = d(gW) @ W2, d(dgW) @ W2
    rga = d1 * gh
    rdga = d1 * dgh + d2 * d(da1) * gh
    for got, ref, name in ((ga, rga, "ga"), (dga, rdga, "dga")):
        err = (d(got) - ref).abs().max().item()
        scale = ref.abs().max().item()
        assert err <= 2e-6 * max(scale, 1.0), (name, err, scale)


@pytest.mark.gpu
def test_train_filter_weights_follow_the_tensor_objects(gpu):
    """The prepared filter weights belong to the W1/W2 tensor objects: other
    tensors at the same addresses (here views of one buffer with another
    hidden size, as a later model can be allocated where a freed one was)
    get their own preparation, and the results match float64 products."""
    print("[case] train_filter weights at reused addresses", flush=True)
    g = torch.Generator().manual_seed(11)
    buf1 = torch.randn(1 << 16, generator=g).to(gpu) * 0.1
    buf2 = torch.randn(1 << 17, generator=g).to(gpu) * 0.05
    e = torch.rand(300, 16, generator=g).to(gpu)
    for H in (32, 128):
        p = {"f.0.weight": buf1[:H * 16].view(H, 16), "f.0.bias": torch.zeros(H, device=gpu),
             "f.2.weight": buf2[:3 * H * H].view(3 * H, H),
             "f.2.bias": torch.zeros(3 * H, device=gpu)}
        a1, h, W = train._filter(p, "f", e)
        torch.cuda.synchronize()
        d = {k: v.double().cpu() for k, v in p.items()}
        ra1 = e.double().cpu() @ d["f.0.weight"].T
        rW = torch.nn.functional.silu(ra1) @ d["f.2.weight"].T
        assert W.shape == (300, 3 * H)
        err = (W.double().cpu() - rW).abs().max().item()
        assert err <= 2e-6 * max(rW.abs().max().item(), 1.0), (H, err)
