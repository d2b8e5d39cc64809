"""The interaction submodules' own forwards (PaiNNMessage, PaiNNUpdate,
PaiNNInteraction on a caller-given directed edge list) on the HIP kernels,
against the FP64 oracle (needs an MI355X)."""

import numpy as np
import pytest
import torch

from conftest import pkg
from oracle import painn_ref as R

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)


def test_painn_interaction_reference_case(gpu):
    """tests/unit/test_student_model.py:569-599 verbatim in substance: H64,
    10 RBF, 10 atoms, 30 random directed edges (duplicates and self loops
    included), unit edge vectors: output shapes and no NaNs."""
    note("reference test_painn_interaction")
    from importlib import import_module
    model = import_module("mlff-distiller_amd.model")
    interaction = model.PaiNNInteraction(hidden_dim=64, num_rbf=10).to(gpu)
    num_atoms, num_edges = 10, 30
    s = torch.randn(num_atoms, 64, device=gpu)
    v = torch.randn(num_atoms, 3, 64, device=gpu)
    edge_index = torch.randint(0, num_atoms, (2, num_edges), device=gpu)
    edge_rbf = torch.randn(num_edges, 10, device=gpu)
    ev = torch.randn(num_edges, 3, device=gpu)
    ev = ev / torch.norm(ev, dim=1, keepdim=True)
    so, vo = interaction(s, v, edge_index, edge_rbf, ev)
    assert so.shape == s.shape and vo.shape == v.shape
    assert not torch.isnan(so).any() and not torch.isnan(vo).any()


def _case(H, R_, n, E, seed, gpu):
    g = torch.Generator().manual_seed(seed)
    s = torch.randn(n, H, generator=g)
    v = torch.randn(n, 3, H, generator=g)
    ei = torch.randint(0, n, (2, E), generator=g)
    ei[:, :3] = torch.tensor([[1, 1, 2], [4, 4, 2]])      # a duplicate edge and a self loop
    rbf = torch.rand(E, R_, generator=g)
    ev = torch.randn(E, 3, generator=g)
    ev = ev / ev.norm(dim=1, keepdim=True)
    return s, v, ei, rbf, ev


@pytest.mark.parametrize("H", [32, 128])
def test_interaction_forward_and_gradients_match_oracle(gpu, H):
    """Outputs and the gradients of a random linear functional of them, with
    respect to the features, the edge RBF inputs, the edge vectors and every
    parameter, against FP64 autograd of the oracle's message/update
    (student_model.py:346-470)."""
    note(f"PaiNNInteraction H{H} forward + backward vs fp64 oracle")
    from importlib import import_module
    model = import_module("mlff-distiller_amd.model")
    torch.manual_seed(3)
    R_ = 12
    blk = model.PaiNNInteraction(H, R_)
    with torch.no_grad():
        for name, p in blk.named_parameters():
            if name.endswith(".bias"):
                p.copy_(0.1 * torch.randn(p.shape))
    s, v, ei, rbf, ev = _case(H, R_, 40, 600, 7, gpu)
    gen = torch.Generator().manual_seed(11)
    ws = torch.randn(40, H, generator=gen)
    wv = torch.randn(40, 3, H, generator=gen)
    # device path
    dblk = blk.to(gpu)
    ins = [t.to(gpu).requires_grad_(True) for t in (s, v, rbf, ev)]
    so, vo = dblk(ins[0], ins[1], ei.to(gpu), ins[2], ins[3])
    obj = (so * ws.to(gpu)).sum() + (vo * wv.to(gpu)).sum()
    params = [p for _, p in dblk.named_parameters()]
    grads = torch.autograd.grad(obj, ins + params)
    # oracle in float64 (the reference's own message/update restated)
    sd = {f"interactions.0.{k}": p.detach().cpu().double().requires_grad_(True)
          for k, p in dblk.named_parameters()}
    rins = [t.double().requires_grad_(True) for t in (s, v, rbf, ev)]
    s1, v1 = R.message(sd, 0, H, rins[0], rins[1], ei[0], ei[1], rins[2], rins[3])
    s2, v2 = R.update(sd, 0, H, s1, v1)
    robj = (s2 * ws.double()).sum() + (v2 * wv.double()).sum()
    rgrads = torch.autograd.grad(robj, rins + list(sd.values()))
    scale = max(1.0, s2.abs().max().item(), v2.abs().max().item())
    assert (so.detach().cpu().double() - s2.detach()).abs().max().item() < 1e-5 * scale
    assert (vo.detach().cpu().double() - v2.detach()).abs().max().item() < 1e-5 * scale
    names = ["s", "v", "edge_rbf", "edge_vector"] + [k for k, _ in dblk.named_parameters()]
    for name, a, b in zip(names, grads, rgrads):
        rel = ((a.detach().cpu().double() - b).norm() / (b.norm() + 1e-30)).item()
        assert rel < 1e-4, (name, rel)
    # message block alone, and an empty edge list (the residual only)
    msg_out = dblk.message(ins[0].detach(), ins[1].detach(), ei.to(gpu), ins[2].detach(),
                           ins[3].detach())
    assert (msg_out[0].cpu().double() - s1.detach()).abs().max().item() < 1e-5 * scale
    empty = torch.zeros(2, 0, dtype=torch.long, device=gpu)
    e_out = dblk.message(ins[0].detach(), ins[1].detach(), empty,
                         torch.zeros(0, R_, device=gpu), torch.zeros(0, 3, device=gpu))
    assert torch.equal(e_out[0], ins[0].detach()) and torch.equal(e_out[1], ins[1].detach())


def test_interaction_rejects_host_tensors():
    """No CPU fallback: host tensors are refused with a clear error."""
    from importlib import import_module
    model = import_module("mlff-distiller_amd.model")
    blk = model.PaiNNUpdate(16)
    with pytest.raises(TypeError, match="ROCm device"):
        blk(torch.zeros(4, 16), torch.zeros(4, 3, 16))
