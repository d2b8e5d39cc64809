"""Training on the GPU: hand-derived gradients through the HIP neighbour list
against the float64 autograd of the oracle (needs an MI355X)."""

import numpy as np
import pytest
import torch

from conftest import pkg, pkg_mod
from oracle import painn_ref as R

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)


def _labels(sd_teacher, mols):
    synth = pkg_mod("synth")
    Z, P, mp = synth.flatten(mols)
    sizes = np.diff(mp)
    b = np.repeat(np.arange(len(mols)), sizes)
    E, F = R.energy_and_forces(sd_teacher, Z, P.astype(np.float32).astype(np.float64), batch=b)
    return Z, P.astype(np.float32), sizes, b, E, F


def _ref_grads(sd, Z, P, b, El, Fl, wE, wF):
    p = {k: v.double().clone().requires_grad_(True) for k, v in sd.items()}
    pos = torch.tensor(P, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b)
    E = R.energy(p, torch.tensor(Z), pos, 5.0, batch=bt, n_structures=int(b.max()) + 1)
    F = -torch.autograd.grad(E.sum(), pos, create_graph=True)[0]
    loss = wE * ((E - El) ** 2).mean() + wF * ((F - Fl) ** 2).mean()
    names = [k for k in p if k not in ("rbf.centers", "rbf.widths")]
    gr = torch.autograd.grad(loss, [p[k] for k in names], allow_unused=True)
    return loss.item(), {k: g for k, g in zip(names, gr) if g is not None}


def test_force_matching_gradients_vs_fp64_autograd(gpu):
    note("force-matching grads H128 vs fp64 autograd")
    synth = pkg_mod("synth")
    torch.manual_seed(1)
    teacher = pkg().StudentForceField()
    torch.manual_seed(0)
    student = pkg().StudentForceField().to(gpu)
    mols = synth.molecule_set(6, 20, 60, seed=4)
    Z, P, sizes, b, El, Fl = _labels({k: v.double() for k, v in teacher.state_dict().items()}, mols)
    tr = pkg_mod("train").ForceMatchingTrainer(student)
    loss, E, F, grads = tr.loss_and_grads(torch.tensor(Z, device=gpu), torch.tensor(P, device=gpu),
                                          sizes, El.float().to(gpu), Fl.float().to(gpu))
    sd = {k: v.detach().cpu() for k, v in student.state_dict().items()}
    loss_ref, ref = _ref_grads(sd, Z, P, b, El, Fl, 1.0, 100.0)
    note(f"loss gpu {float(loss):.6f} ref {loss_ref:.6f}")
    assert abs(float(loss) - loss_ref) < 1e-4 * abs(loss_ref)
    worst = 0.0
    for k, r in ref.items():
        gk = grads[k].double().cpu()
        rel = ((gk - r).norm() / (r.norm() + 1e-30)).item()
        worst = max(worst, rel)
        assert rel < 2e-3, (k, rel)
    note(f"worst relative gradient error {worst:.2e}")


def test_autograd_create_graph_api_matches_trainer(gpu):
    """model(...) + autograd.grad(create_graph=True) + loss.backward() gives the
    same parameter gradients as the direct trainer."""
    note("create_graph API")
    synth = pkg_mod("synth")
    torch.manual_seed(1)
    teacher = pkg().StudentForceField()
    torch.manual_seed(0)
    student = pkg().StudentForceField(hidden_dim=64, num_interactions=2).to(gpu)
    mols = synth.molecule_set(4, 20, 40, seed=8)
    Z, P, sizes, b, El, Fl = _labels({k: v.double() for k, v in teacher.state_dict().items()}, mols)
    Zt = torch.tensor(Z, device=gpu)
    pos = torch.tensor(P, device=gpu, requires_grad=True)
    bt = torch.tensor(b, device=gpu)
    E = student(Zt, pos, batch=bt)
    F = -torch.autograd.grad(E.sum(), pos, create_graph=True)[0]
    loss = ((E - El.float().to(gpu)) ** 2).mean() + 100.0 * ((F - Fl.float().to(gpu)) ** 2).mean()
    student.zero_grad()
    loss.backward()
    got = {n: p.grad.clone() for n, p in student.named_parameters()}
    tr = pkg_mod("train").ForceMatchingTrainer(student)
    _, _, _, grads = tr.loss_and_grads(Zt, pos.detach(), sizes, El.float().to(gpu),
                                       Fl.float().to(gpu))
    for n, g in got.items():
        ref = grads[n].reshape(g.shape)
        rel = ((g - ref).norm() / (ref.norm() + 1e-30)).item()
        assert rel < 1e-4, (n, rel)


def test_training_reduces_loss(gpu):
    note("20 AdamW steps")
    synth = pkg_mod("synth")
    torch.manual_seed(1)
    teacher = pkg().StudentForceField(hidden_dim=64, num_interactions=2)
    torch.manual_seed(0)
    student = pkg().StudentForceField(hidden_dim=32, num_interactions=2, num_rbf=16).to(gpu)
    mols = synth.molecule_set(16, 10, 40, seed=6)
    Z, P, sizes, b, El, Fl = _labels({k: v.double() for k, v in teacher.state_dict().items()}, mols)
    tr = pkg_mod("train").ForceMatchingTrainer(student, lr=3e-3)
    args = (torch.tensor(Z, device=gpu), torch.tensor(P, device=gpu), sizes, El.float().to(gpu),
            Fl.float().to(gpu))
    losses = [tr.step(*args) for _ in range(20)]
    note(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.7 * losses[0]


def test_distillation_wrapper_on_collated_dataset(gpu, tmp_path):
    """Frames written in the reference's dataset layout, read back, collated
    and run through the DistillationWrapper: eval mode returns the fused
    evaluation's energies and forces; training mode keeps the graph, and a
    force-matching loss gives the trainer's parameter gradients."""
    note("dataset -> collate -> DistillationWrapper")
    ds = pkg_mod("dataset")
    synth = pkg_mod("synth")
    torch.manual_seed(1)
    teacher = pkg().StudentForceField()
    torch.manual_seed(0)
    student = pkg().StudentForceField(hidden_dim=64, num_interactions=2).to(gpu)
    mols = synth.molecule_set(4, 20, 40, seed=9)
    Z, P, sizes, b, El, Fl = _labels({k: v.double() for k, v in teacher.state_dict().items()}, mols)
    ptr = np.concatenate([[0], np.cumsum(sizes)])
    path = str(tmp_path / "frames.npz")
    with ds.FrameDatasetWriter(path) as w:
        for m in range(len(mols)):
            w.add_structure(energy=float(El[m]), forces=Fl.numpy()[ptr[m]:ptr[m + 1]],
                            numbers=Z[ptr[m]:ptr[m + 1]], positions=P[ptr[m]:ptr[m + 1]])
    r = ds.FrameDataset(path)
    batch = ds.distillation_collate([r[i] for i in range(len(r))], device=gpu)
    wrap = ds.DistillationWrapper(student)
    wrap.eval()
    with torch.no_grad():
        for p in student.parameters():
            p.requires_grad_(False)
        out = wrap(batch)
        e64, f, _ = student.evaluate(batch["atomic_numbers"], batch["positions"], sizes)
        for p in student.parameters():
            p.requires_grad_(True)
    assert torch.allclose(out["energy"].double(), e64, rtol=1e-6, atol=1e-5)
    assert (out["forces"] - f).abs().max().item() < 1e-5
    wrap.train()
    out = wrap(batch)
    loss = (((out["energy"] - batch["energy"].float()) ** 2).mean()
            + 100.0 * ((out["forces"] - batch["forces"]) ** 2).mean())
    student.zero_grad()
    loss.backward()
    tr = pkg_mod("train").ForceMatchingTrainer(student)
    _, _, _, grads = tr.loss_and_grads(batch["atomic_numbers"], batch["positions"], sizes,
                                       batch["energy"].float(), batch["forces"])
    for n, p in student.named_parameters():
        ref = grads[n].reshape(p.shape)
        rel = ((p.grad - ref).norm() / (ref.norm() + 1e-30)).item()
        assert rel < 1e-4, (n, rel)


def test_training_gradients_are_bit_reproducible(gpu):
    """Fixed-order sums everywhere (row kernels, pair kernel, segment sums,
    chunked weight-gradient products): two evaluations of the loss gradient
    on the same batch are bit-identical."""
    note("training gradients: two runs bit-identical")
    synth = pkg_mod("synth")
    torch.manual_seed(1)
    teacher = pkg().StudentForceField()
    torch.manual_seed(0)
    student = pkg().StudentForceField().to(gpu)
    mols = synth.molecule_set(64, 20, 60, seed=4)
    Z, P, sizes, b, El, Fl = _labels({k: v.double() for k, v in teacher.state_dict().items()},
                                     mols[:12])
    tr = pkg_mod("train").ForceMatchingTrainer(student)
    args = (torch.tensor(Z, device=gpu), torch.tensor(P, device=gpu), sizes, El.float().to(gpu),
            Fl.float().to(gpu))
    l1, _, _, g1 = tr.loss_and_grads(*args)
    l2, _, _, g2 = tr.loss_and_grads(*args)
    assert float(l1) == float(l2)
    assert set(g1) == set(g2)
    for k in g1:
        assert torch.equal(g1[k], g2[k]), k


def test_learnable_rbf_gradients_vs_fp64_autograd(gpu):
    """learnable_rbf=True: the basis centres and widths train too; their
    gradients (and every other) match FP64 autograd double backward."""
    note("learnable RBF gradients vs fp64 autograd")
    synth = pkg_mod("synth")
    torch.manual_seed(1)
    teacher = pkg().StudentForceField(hidden_dim=64, num_interactions=2)
    torch.manual_seed(0)
    student = pkg().StudentForceField(hidden_dim=64, num_interactions=2,
                                      learnable_rbf=True).to(gpu)
    with torch.no_grad():
        g = torch.Generator().manual_seed(3)
        student.rbf.centers.add_((0.05 * torch.randn(20, generator=g)).to(gpu))
    mols = synth.molecule_set(5, 20, 40, seed=12)
    Z, P, sizes, b, El, Fl = _labels({k: v.double() for k, v in teacher.state_dict().items()}, mols)
    tr = pkg_mod("train").ForceMatchingTrainer(student)
    _, _, _, grads = tr.loss_and_grads(torch.tensor(Z, device=gpu), torch.tensor(P, device=gpu),
                                       sizes, El.float().to(gpu), Fl.float().to(gpu))
    sd = {k: v.detach().cpu() for k, v in student.state_dict().items()}
    p = {k: v.double().clone().requires_grad_(True) for k, v in sd.items()}
    pos = torch.tensor(P, dtype=torch.float64, requires_grad=True)
    E = R.energy(p, torch.tensor(Z), pos, 5.0, batch=torch.tensor(b), n_structures=len(mols))
    F = -torch.autograd.grad(E.sum(), pos, create_graph=True)[0]
    loss = ((E - El) ** 2).mean() + 100.0 * ((F - Fl) ** 2).mean()
    names = list(p)
    ref = dict(zip(names, torch.autograd.grad(loss, [p[k] for k in names], allow_unused=True)))
    for k in ("rbf.centers", "rbf.widths", "interactions.0.message.rbf_to_scalar.0.weight"):
        r = ref[k]
        rel = ((grads[k].double().cpu() - r).norm() / (r.norm() + 1e-30)).item()
        note(f"{k}: rel {rel:.2e}")
        assert rel < 2e-3, (k, rel)


def test_periodic_training_gradients_vs_fp64_autograd(gpu):
    """A periodic water box (minimum-image edges with lattice shifts): the
    training gradients match FP64 autograd double backward on the oracle's
    own edge set."""
    note("periodic training gradients vs fp64 autograd")
    from oracle import nl_ref
    synth = pkg_mod("synth")
    box = synth.water_box(4, seed=3)
    torch.manual_seed(1)
    teacher = pkg().StudentForceField(hidden_dim=64, num_interactions=2)
    torch.manual_seed(0)
    student = pkg().StudentForceField(hidden_dim=64, num_interactions=2).to(gpu)
    P32 = box.positions.astype(np.float32)
    n = len(P32)
    rp, cl, sh = nl_ref.neighbor_csr(P32, np.array([0, n]), 5.0, box.cell[None], box.pbc[None])
    ei = np.stack([cl.astype(np.int64), np.repeat(np.arange(n), np.diff(rp))])
    cell = box.cell.astype(np.float32).astype(np.float64)
    sdt = {k: v.double() for k, v in teacher.state_dict().items()}
    El, Fl = R.energy_and_forces(sdt, box.numbers, P32.astype(np.float64), edge_index=ei,
                                 shifts=sh, cell=cell)
    El = El.reshape(1)
    tr = pkg_mod("train").ForceMatchingTrainer(student)
    _, _, _, grads = tr.loss_and_grads(torch.tensor(box.numbers, device=gpu),
                                       torch.tensor(P32, device=gpu), np.array([n]),
                                       El.float().to(gpu), Fl.float().to(gpu),
                                       cells=box.cell[None], pbc=box.pbc[None])
    sd = {k: v.detach().cpu() for k, v in student.state_dict().items()}
    p = {k: v.double().clone().requires_grad_(True) for k, v in sd.items()}
    pos = torch.tensor(P32, dtype=torch.float64, requires_grad=True)
    sv = torch.tensor(sh, dtype=torch.float64) @ torch.tensor(cell)
    E = R.energy(p, torch.tensor(box.numbers), pos, 5.0, edge_index=torch.tensor(ei),
                 shift_vec=sv).reshape(1)
    F = -torch.autograd.grad(E.sum(), pos, create_graph=True)[0]
    loss = ((E - El) ** 2).mean() + 100.0 * ((F - Fl) ** 2).mean()
    names = [k for k in p if not k.startswith("rbf.")]
    ref = dict(zip(names, torch.autograd.grad(loss, [p[k] for k in names], allow_unused=True)))
    worst = 0.0
    for k, r in ref.items():
        if r is None:
            continue
        rel = ((grads[k].double().cpu() - r).norm() / (r.norm() + 1e-30)).item()
        worst = max(worst, rel)
        assert rel < 2e-3, (k, rel)
    note(f"worst relative gradient error {worst:.2e}")


def test_async_evaluation_regrows_on_overflow(gpu):
    """evaluate_async on another stream gives evaluate()'s results, and when
    the neighbour list outgrows the plan's capacity (the same batch
    compressed to half its size) the completion re-evaluates with a regrown
    plan."""
    note("evaluate_async: side stream, capacity overflow")
    synth = pkg_mod("synth")
    torch.manual_seed(0)
    model = pkg().StudentForceField().to(gpu)
    Z, P, mp = synth.flatten(synth.molecule_set(8, 30, 60, seed=9))
    sizes = np.diff(mp)
    Zt = torch.tensor(Z, device=gpu)
    for scale in (1.0, 0.5):
        Pt = torch.tensor(P * scale, dtype=torch.float32, device=gpu)
        cap0 = [int(e["plan"].ecap) for e in model._plans.values()]
        side = torch.cuda.Stream(gpu)
        side.wait_stream(torch.cuda.current_stream(gpu))
        with torch.cuda.stream(side):
            finish = model.evaluate_async(Zt, Pt, sizes)
        Ea, Fa, _ = finish()
        torch.cuda.current_stream(gpu).wait_stream(side)
        if scale < 1.0:
            assert cap0 and int(list(model._plans.values())[-1]["plan"].ecap) > cap0[-1]
        Eb, Fb, _ = model.evaluate(Zt, Pt, sizes)
        assert torch.equal(Ea, Eb) and torch.equal(Fa, Fb)


def test_trainer_recomputes_after_capacity_overflow(gpu):
    """The trainer enqueues its gradient pass behind the force evaluation
    before reading the evaluation's capacity check.  When the neighbour list
    overflows (the same batch compressed to half its size after a call that
    sized the plan), the pass runs again on the regrown evaluation: loss and
    gradients equal those of a trainer that meets the compressed batch
    first."""
    note("trainer: capacity overflow behind the enqueued gradient pass")
    synth = pkg_mod("synth")
    train = pkg_mod("train")
    Z, P, mp = synth.flatten(synth.molecule_set(8, 30, 60, seed=9))
    sizes = np.diff(mp)
    Zt = torch.tensor(Z, device=gpu)
    El = torch.zeros(len(sizes), device=gpu)
    Fl = torch.zeros(len(Z), 3, device=gpu)
    Ph = torch.tensor(P * 0.5, dtype=torch.float32, device=gpu)
    runs = []
    for warm in (True, False):
        torch.manual_seed(0)
        model = pkg().StudentForceField().to(gpu)
        tr = train.ForceMatchingTrainer(model)
        if warm:
            tr.loss_and_grads(Zt, torch.tensor(P, dtype=torch.float32, device=gpu), sizes, El, Fl)
            cap0 = int(list(model._plans.values())[-1]["plan"].ecap)
        loss, E, F, grads = tr.loss_and_grads(Zt, Ph, sizes, El, Fl)
        if warm:
            assert int(list(model._plans.values())[-1]["plan"].ecap) > cap0
        runs.append((loss, E, F, grads))
    (l1, E1, F1, g1), (l2, E2, F2, g2) = runs
    assert torch.equal(E1, E2) and torch.equal(F1, F2) and float(l1) == float(l2)
    assert g1.keys() == g2.keys()
    for k in g1:
        assert torch.equal(g1[k], g2[k]), k


def test_element_segments_follow_the_tensor(gpu):
    """The element segments of the embedding-gradient sum are reused for the
    same element tensor while it is unmodified, and remade for a written or
    a different tensor (with the same values in the same order)."""
    note("element segments cache")
    train = pkg_mod("train")
    Z = torch.tensor([6, 1, 8, 1, 6, 7, 1], device=gpu)
    p1, o1 = train._z_segments(Z, 10)
    p2, o2 = train._z_segments(Z, 10)
    assert p1 is p2 and o1 is o2
    Z.add_(0)   # a write moves the version counter
    p3, o3 = train._z_segments(Z, 10)
    assert p3 is not p1 and torch.equal(p3, p1) and torch.equal(o3, o1)
    Z2 = Z.clone()
    Z2[0] = 8
    p4, o4 = train._z_segments(Z2, 10)
    zs = Z2.cpu().tolist()
    want = [sum(1 for z in zs if z < k) for k in range(11)]
    assert p4.cpu().tolist() == want
    assert [zs[i] for i in o4.cpu().tolist()] == sorted(zs)


def test_pair_graph_from_the_neighbour_kernels_equals_the_csr_form(gpu):
    """For isolated structures the training pair graph comes from the
    neighbour-list kernels (pairs numbered by the fill pass, edges linked by
    the link pass); it equals the graph built from the plain receiver CSR by
    pair_graph_from_csr, field by field."""
    note("pair graph: neighbour kernels vs CSR form")
    synth = pkg_mod("synth")
    train = pkg_mod("train")
    engine = pkg_mod("engine")
    Z, P, mp = synth.flatten(synth.molecule_set(24, 5, 70, seed=13))
    sizes = np.diff(mp)
    x = torch.tensor(P, dtype=torch.float32, device=gpu)
    g = train.device_graph(torch.tensor(Z, device=gpu), x, sizes, 5.0)
    layout = engine.make_layout(sizes, gpu)
    row_ptr, col, _ = engine.neighbor_list(x, layout, 5.0)
    ref = train.pair_graph_from_csr(row_ptr, col, layout.atom_mol.long(), len(sizes))
    for name in ("row_ptr", "col", "edge_pair", "lo", "hi", "lo32", "hi32", "batch", "mol_ptr"):
        a, b = getattr(g, name), getattr(ref, name)
        assert a.dtype == b.dtype and torch.equal(a, b), name
    assert g.shift is None and g.n_structures == ref.n_structures
