"""The hand-derived training gradients (oracle/train_ref.py, the checker of
the HIP training pass) against the oracle's autograd double backward and
central finite differences, in float64 on the CPU."""

import numpy as np
import pytest
import torch

from conftest import pkg_mod
from oracle import painn_ref as R
from oracle import train_ref


def _setup(H=16, L=2, Rb=6, rc=3.5, n_mol=3, seed=0, bias=0.2):
    model_mod = pkg_mod("model")
    torch.manual_seed(seed)
    m = model_mod.StudentForceField(hidden_dim=H, num_interactions=L, num_rbf=Rb, cutoff=rc,
                                    max_z=16)
    gen = torch.Generator().manual_seed(seed + 7)
    sd = {}
    for k, v in m.state_dict().items():
        v = v.double().clone()
        if k.endswith(".bias"):
            v = torch.randn(v.shape, generator=gen, dtype=torch.float64) * bias
        sd[k] = v
    rng = np.random.default_rng(seed)
    Zs, Ps = [], []
    for b in range(n_mol):
        n = int(rng.integers(4, 9))
        Zs.append(rng.choice([1, 6, 7, 8, 16], size=n))
        Ps.append(rng.normal(size=(n, 3)) * 1.3 + np.array([30.0 * b, 0, 0]))
    Z = torch.tensor(np.concatenate(Zs))
    x = torch.tensor(np.concatenate(Ps))
    batch = torch.tensor(np.repeat(np.arange(n_mol), [len(z) for z in Zs]))
    return sd, Z, x, batch, rc


def _graph(x, batch, rc, n_mol):
    train = train_ref
    ei = R.radius_graph_dense(x, rc, batch)
    keep = ei[1] > ei[0]                       # src > dst: one entry per unordered pair
    lo, hi = ei[1][keep], ei[0][keep]
    return train.PairGraph(lo, hi, None, batch, n_mol)


def test_forward_and_forces_match_oracle():
    train = train_ref
    sd, Z, x, batch, rc = _setup()
    g = _graph(x, batch, rc, 3)
    E, cache = train.forward(sd, Z, x, g, rc)
    _, F, _ = train.reverse(sd, cache, g, torch.ones_like(E))
    Eref, Fref = R.energy_and_forces(sd, Z.numpy(), x.numpy(), cutoff=rc, batch=batch.numpy())
    assert torch.allclose(E, Eref, atol=1e-12, rtol=0)
    assert torch.allclose(F, Fref, atol=1e-11, rtol=0)


def _autograd_loss_grads(sd, Z, x, batch, rc, El, Fl, wE, wF, rbf=False):
    p = {k: v.clone().requires_grad_(True) for k, v in sd.items()}
    pos = x.clone().requires_grad_(True)
    E = R.energy(p, Z, pos, rc, batch=batch, n_structures=int(batch.max()) + 1)
    F = -torch.autograd.grad(E.sum(), pos, create_graph=True)[0]
    loss = wE * ((E - El) ** 2).mean() + wF * ((F - Fl) ** 2).mean()
    names = [k for k in p if p[k].requires_grad
             and (rbf or k not in ("rbf.centers", "rbf.widths"))]
    gr = torch.autograd.grad(loss, [p[k] for k in names], allow_unused=True)
    return loss.detach(), dict(zip(names, gr))


def test_loss_gradient_matches_autograd_double_backward():
    train = train_ref
    sd, Z, x, batch, rc = _setup()
    g = _graph(x, batch, rc, 3)
    gen = torch.Generator().manual_seed(5)
    El = torch.randn(3, generator=gen, dtype=torch.float64)
    Fl = torch.randn(x.shape, generator=gen, dtype=torch.float64)
    loss, E, F, grads = train.force_matching_grads(sd, Z, x, g, rc, El, Fl, 1.0, 100.0)
    loss_ref, ref = _autograd_loss_grads(sd, Z, x, batch, rc, El, Fl, 1.0, 100.0)
    assert abs(loss.item() - loss_ref.item()) < 1e-10 * max(1, abs(loss_ref.item()))
    for k, r in ref.items():
        if r is None:
            continue
        assert k in grads, k
        err = (grads[k] - r).abs().max().item()
        scale = max(1.0, r.abs().max().item())
        assert err < 1e-9 * scale, (k, err, scale)


def test_loss_gradient_central_differences():
    """Pins the derivation independently of autograd."""
    train = train_ref
    sd, Z, x, batch, rc = _setup(H=8, L=2, Rb=4, n_mol=2, seed=3)
    g = _graph(x, batch, rc, 2)
    El = torch.tensor([0.3, -0.2], dtype=torch.float64)
    Fl = torch.zeros(x.shape, dtype=torch.float64)
    _, _, _, grads = train.force_matching_grads(sd, Z, x, g, rc, El, Fl, 1.0, 100.0)
    picks = [("interactions.0.message.rbf_to_scalar.2.weight", (3, 2)),
             ("interactions.1.update.update_mlp.0.weight", (1, 5)),
             ("interactions.0.update.mixing_matrix", (2, 0)),
             ("energy_head.2.bias", (1,)),
             ("embedding.weight", (int(Z[0]), 4))]
    h = 1e-6
    for name, idx in picks:
        vals = []
        for sgn in (1, -1):
            q = {k: v.clone() for k, v in sd.items()}
            q[name][idx] += sgn * h
            loss, _, _, _ = train.force_matching_grads(q, Z, x, g, rc, El, Fl, 1.0, 100.0)
            vals.append(loss.item())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - grads[name][idx].item()) < 1e-6 * max(1.0, abs(fd)), (name, fd,
                                                                            grads[name][idx].item())


def test_weight_gradient_operands():
    """Operands of mlff_weight_grad: widths padded with zero columns to a
    multiple of 4, views with a usable row stride passed as they are, other
    layouts copied (host logic; the kernel is tested on the GPU)."""
    train = pkg_mod("train")
    a = torch.randn(10, 20)
    assert train._wg_operand(a).data_ptr() == a.data_ptr()
    p = train._wg_operand(torch.randn(10, 5))
    assert p.shape == (10, 8) and torch.equal(p[:, 5:], torch.zeros(10, 3))
    wide = torch.randn(10, 3 * 16)
    view = wide[:, 16:32]
    assert train._wg_operand(view).data_ptr() == view.data_ptr()
    t = torch.randn(20, 10).T                      # column stride 10
    q = train._wg_operand(t)
    assert q.stride(1) == 1 and torch.equal(q[:, :10], t.contiguous()[:, :10])
    r = train._wg_operand(torch.randn(6, 3, 8))    # [n, 3, H] rows flattened
    assert r.shape == (6, 24)
    assert train._wg_operand(torch.randn(0, 3, 5)).shape == (0, 16)


def test_weighted_gradients_equal_weighted_adjoint():
    """Structures do not interact, so the parameter gradients of sum_b w_b E_b
    from the unit adjoint with rows weighted by w (gw) equal those of the
    adjoint seeded with w; want_grads=False returns no gradients."""
    train = train_ref
    sd, Z, x, batch, rc = _setup(seed=2)
    g = _graph(x, batch, rc, 3)
    E, cache = train.forward(sd, Z, x, g, rc)
    w = torch.tensor([0.7, -1.3, 2.1], dtype=torch.float64)
    ref, _, _ = train.reverse(sd, cache, g, w, want_forces=False)
    got, _, _ = train.reverse(sd, cache, g, torch.ones_like(E), want_forces=False, gw=w)
    assert set(ref) == set(got)
    for k in ref:
        assert torch.allclose(got[k], ref[k], rtol=1e-11, atol=1e-12), k
    none, F, _ = train.reverse(sd, cache, g, torch.ones_like(E), want_grads=False)
    assert none == {} and F is not None


def test_learnable_rbf_gradients_match_autograd_double_backward():
    """Basis centres and widths as parameters (learnable_rbf,
    student_model.py:226-236): their loss gradients, first and second order,
    equal FP64 autograd's; the centres and widths are moved off their
    defaults so every term of the derivative is exercised."""
    train = train_ref
    sd, Z, x, batch, rc = _setup(seed=4)
    gen = torch.Generator().manual_seed(11)
    sd["rbf.centers"] = sd["rbf.centers"] + 0.1 * torch.randn(sd["rbf.centers"].shape,
                                                              generator=gen, dtype=torch.float64)
    sd["rbf.widths"] = sd["rbf.widths"] * (1 + 0.2 * torch.rand(sd["rbf.widths"].shape,
                                                                generator=gen,
                                                                dtype=torch.float64))
    g = _graph(x, batch, rc, 3)
    El = torch.randn(3, generator=gen, dtype=torch.float64)
    Fl = torch.randn(x.shape, generator=gen, dtype=torch.float64)
    _, _, _, grads = train.force_matching_grads(sd, Z, x, g, rc, El, Fl, 1.0, 100.0,
                                                rbf_grads=True)
    _, ref = _autograd_loss_grads(sd, Z, x, batch, rc, El, Fl, 1.0, 100.0, rbf=True)
    for k in ("rbf.centers", "rbf.widths"):
        err = (grads[k] - ref[k]).abs().max().item()
        scale = max(1.0, ref[k].abs().max().item())
        assert ref[k].abs().max().item() > 1e-3, k
        assert err < 1e-9 * scale, (k, err, scale)
    _, _, _, plain = train.force_matching_grads(sd, Z, x, g, rc, El, Fl, 1.0, 100.0)
    assert "rbf.centers" not in plain and "rbf.widths" not in plain


def test_trainer_parameter_views_follow_the_model():
    """The trainer's cached parameter views see in-place optimizer updates
    and are rebuilt when the model's tensors are replaced."""
    model_mod = pkg_mod("model")
    train = pkg_mod("train")
    torch.manual_seed(0)
    m = model_mod.StudentForceField(hidden_dim=32, num_interactions=1, num_rbf=8)
    tr = train.ForceMatchingTrainer(m, lr=1e-3)
    p1 = tr._params()
    w = m.energy_head[0].weight
    with torch.no_grad():
        w.add_(1.0)
    assert torch.equal(tr._params()["energy_head.0.weight"], w)
    assert tr._params() is p1
    m.load_state_dict({k: v.clone() for k, v in m.state_dict().items()}, assign=True)
    p2 = tr._params()
    assert p2 is not p1
    assert p2["energy_head.0.weight"].data_ptr() == m.energy_head[0].weight.data_ptr()


def test_pair_graph_without_device_reads_equals_counting_form():
    """pair_graph_from_csr given the pair count and the structure offsets
    (the molecules' form, no host read of device values) builds the same
    pairs, edge-to-pair map and offsets as the form that counts them, on
    random symmetric neighbour lists of two structures."""
    train = pkg_mod("train")
    g = torch.Generator().manual_seed(0)
    for trial in range(30):
        na, nb = (int(v) for v in torch.randint(1, 25, (2,), generator=g))
        n = na + nb
        A = torch.zeros(n, n, dtype=torch.bool)
        for lo_, hi_ in ((0, na), (na, n)):
            blk = torch.rand(hi_ - lo_, hi_ - lo_, generator=g) < 0.4
            A[lo_:hi_, lo_:hi_] = blk | blk.T
        A.fill_diagonal_(False)
        rows, cols = torch.nonzero(A, as_tuple=True)
        row_ptr = torch.zeros(n + 1, dtype=torch.int32)
        row_ptr[1:] = torch.cumsum(torch.bincount(rows, minlength=n), 0)
        col = cols.to(torch.int32)
        batch = torch.cat([torch.zeros(na), torch.ones(nb)]).long()
        a = train.pair_graph_from_csr(row_ptr, col, batch, 2)
        b = train.pair_graph_from_csr(row_ptr, col, batch, 2, n_pairs=col.numel() // 2,
                                      mol_ptr=torch.tensor([0, na, n], dtype=torch.int32))
        for f in ("row_ptr", "col", "edge_pair", "lo", "hi", "lo32", "hi32", "mol_ptr"):
            assert torch.equal(getattr(a, f), getattr(b, f)), (trial, f)
        # every upper edge (row < col) is its own pair's slot; every pair is (lo < hi)
        assert bool((b.lo < b.hi).all())
