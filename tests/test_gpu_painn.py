"""GPU parity of the HIP path against the oracle (needs an MI355X).

Tolerances (stated by the task): energy within 1e-5 eV/atom and forces
within 1e-4 eV/A max-abs of the float64 oracle; neighbour lists bit-exact.
"""

import sys

import numpy as np
import pytest
import torch

from conftest import golden, pkg, pkg_mod, tiny_state
from oracle import nl_ref
from oracle import painn_ref as R

pytestmark = pytest.mark.gpu

E_TOL = 1e-5   # eV/atom
F_TOL = 1e-4   # eV/A


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _model(H=128, L=3, R_=20, seed=0, bias_scale=0.0, device="cuda:0"):
    torch.manual_seed(seed)
    m = pkg().StudentForceField(hidden_dim=H, num_interactions=L, num_rbf=R_, cutoff=5.0)
    if bias_scale:
        g = torch.Generator().manual_seed(seed + 1)
        with torch.no_grad():
            for n, p in m.named_parameters():
                if n.endswith(".bias"):
                    p.copy_(torch.randn(p.shape, generator=g) * bias_scale)
    return m.to(device)


def _mols(n, lo, hi, seed):
    synth = pkg_mod("synth")
    return synth.molecule_set(n, lo, hi, seed)


def _csr_to_edges(row_ptr, col):
    row_ptr = row_ptr.cpu().numpy().astype(np.int64)
    col = col.cpu().numpy().astype(np.int64)
    src = np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr))
    return np.stack([src, col])


def test_neighbor_list_bit_exact_at_ulp_boundary(gpu):
    note("NL golden nl_fp32 (ulp-boundary pairs)")
    g = golden("nl_fp32.npz")
    pos = torch.tensor(g["positions"], device=gpu)
    ei = pkg().radius_graph(pos, 5.0).cpu().numpy()
    np.testing.assert_array_equal(ei, g["edge_index"])
    eib = pkg().radius_graph(pos, 5.0, torch.tensor(g["batch"], device=gpu)).cpu().numpy()
    np.testing.assert_array_equal(eib, g["edge_index_batched"])


def test_neighbor_list_batched_molecules(gpu):
    note("NL 64 molecules 10..100")
    mols = _mols(64, 10, 100, seed=11)
    synth = pkg_mod("synth")
    Z, P, mol_ptr = synth.flatten(mols)
    pos32 = P.astype(np.float32)
    row_ptr, col, _ = pkg().radius_graph_csr(
        torch.tensor(pos32, device=gpu), 5.0,
        torch.tensor(np.repeat(np.arange(len(mols)), np.diff(mol_ptr)), device=gpu))
    rp, cl, _ = nl_ref.neighbor_csr(pos32, mol_ptr, 5.0)
    np.testing.assert_array_equal(row_ptr.cpu().numpy(), rp)
    np.testing.assert_array_equal(col.cpu().numpy(), cl)


@pytest.mark.parametrize("n", [1023, 1024, 1025, 2049, 5 * 1024])
def test_neighbor_list_scan_slice_boundaries(gpu, n):
    """The row offsets are scanned per 1024-atom slice: atom counts at and
    one either side of a slice boundary give the oracle's CSR exactly."""
    note(f"NL scan slices, {n} atoms in two structures")
    rng = np.random.default_rng(n)
    side = (n / 0.03) ** (1 / 3)            # about 0.03 atoms per A^3: ~15 neighbours
    pos32 = rng.uniform(0, side, size=(n, 3)).astype(np.float32)
    cut = n // 3
    mol_ptr = np.array([0, cut, n])
    batch = np.repeat([0, 1], [cut, n - cut])
    row_ptr, col, _ = pkg().radius_graph_csr(torch.tensor(pos32, device=gpu), 5.0,
                                             torch.tensor(batch, device=gpu))
    rp, cl, _ = nl_ref.neighbor_csr(pos32, mol_ptr, 5.0)
    np.testing.assert_array_equal(row_ptr.cpu().numpy(), rp)
    np.testing.assert_array_equal(col.cpu().numpy(), cl)


def test_neighbor_list_many_slices(gpu):
    """400 molecules (about 22 scan slices): every slice's offset is the sum
    of all slice totals before it."""
    note("NL 400 molecules 10..100 (many scan slices)")
    mols = _mols(400, 10, 100, seed=13)
    synth = pkg_mod("synth")
    Z, P, mol_ptr = synth.flatten(mols)
    assert mol_ptr[-1] > 16 * 1024
    pos32 = P.astype(np.float32)
    row_ptr, col, _ = pkg().radius_graph_csr(
        torch.tensor(pos32, device=gpu), 5.0,
        torch.tensor(np.repeat(np.arange(len(mols)), np.diff(mol_ptr)), device=gpu))
    rp, cl, _ = nl_ref.neighbor_csr(pos32, mol_ptr, 5.0)
    np.testing.assert_array_equal(row_ptr.cpu().numpy(), rp)
    np.testing.assert_array_equal(col.cpu().numpy(), cl)


def test_neighbor_list_count_only(gpu):
    """The count-only call (capacity 0, no edge buffers) that sizes new plans
    reports the oracle's edge count, for molecules and a periodic box."""
    note("NL count-only mode")
    engine = pkg_mod("engine")
    synth = pkg_mod("synth")
    mols = _mols(40, 10, 100, seed=17)
    Z, P, mol_ptr = synth.flatten(mols)
    pos32 = P.astype(np.float32)
    lay = engine.make_layout(np.diff(mol_ptr), torch.device(gpu))
    n = engine.count_edges(torch.tensor(pos32, device=gpu), lay, 5.0)
    rp, _, _ = nl_ref.neighbor_csr(pos32, mol_ptr, 5.0)
    assert n == rp[-1]
    box = synth.water_box(6, seed=3)
    bpos = box.positions.astype(np.float32)
    blay = engine.make_layout([len(bpos)], torch.device(gpu), box.cell[None], box.pbc[None])
    nb = engine.count_edges(torch.tensor(bpos, device=gpu), blay, 5.0)
    rpb, _, _ = nl_ref.neighbor_csr(bpos, np.array([0, len(bpos)]), 5.0, box.cell[None],
                                    box.pbc[None])
    assert nb == rpb[-1]


def test_neighbor_list_periodic_water(gpu):
    note("NL periodic water box 1000 molecules")
    synth = pkg_mod("synth")
    box = synth.water_box(10, seed=3)
    pos32 = box.positions.astype(np.float32)
    row_ptr, col, shift = pkg().radius_graph_csr(torch.tensor(pos32, device=gpu), 5.0, None,
                                                 box.cell, box.pbc)
    rp, cl, sh = nl_ref.neighbor_csr(pos32, np.array([0, len(pos32)]), 5.0, box.cell[None],
                                     box.pbc[None])
    np.testing.assert_array_equal(row_ptr.cpu().numpy(), rp)
    np.testing.assert_array_equal(col.cpu().numpy(), cl)
    np.testing.assert_array_equal(shift.cpu().numpy(), sh)
    assert 40 < len(cl) / len(pos32) < 70


@pytest.mark.parametrize("skew", [0.0, 2.5])
def test_cell_list_equals_brute_force(gpu, skew):
    """The cell-list scan (periodic boxes >= 256 atoms) reproduces the
    brute-force edge set exactly: energies, forces and virials are
    bit-identical."""
    note(f"cell list vs brute force, skew {skew}")
    synth = pkg_mod("synth")
    engine = pkg_mod("engine")
    box = synth.water_box(10, seed=3)
    cell = box.cell.copy()
    cell[1, 0] += skew
    cell[2, 1] -= 0.5 * skew
    Z = torch.tensor(box.numbers, device=gpu)
    P = torch.tensor(box.positions.astype(np.float32), device=gpu)
    m1 = _model(64, 2, 20, seed=0)
    out_cl = m1.energy_forces_virial(Z, P, torch.tensor(cell), torch.tensor(box.pbc))
    plan = next(iter(m1._plans.values()))["plan"]
    assert plan.cell_grid[0, 0] >= 3          # the cell list was used
    old = engine.CELL_LIST_MIN_ATOMS
    engine.CELL_LIST_MIN_ATOMS = 10 ** 9
    try:
        m2 = _model(64, 2, 20, seed=0)
        out_bf = m2.energy_forces_virial(Z, P, torch.tensor(cell), torch.tensor(box.pbc))
        assert next(iter(m2._plans.values()))["plan"].cell_grid[0, 0] == 0
    finally:
        engine.CELL_LIST_MIN_ATOMS = old
    for a, b in zip(out_cl, out_bf):
        assert torch.equal(a, b)


@pytest.mark.parametrize("skew", [0.0, 2.5])
def test_cell_list_long_rows_equal_brute_force(gpu, skew):
    """A periodic box with a dense slab: rows of more than 256 neighbours
    (walked by one lane) beside shorter rows (walked by the whole wave and
    sorted in LDS); energies, forces and virials are bit-identical to the
    brute-force scan."""
    note(f"cell list, dense slab (rows above and below 256 neighbours), skew {skew}")
    engine = pkg_mod("engine")
    L = 16.5
    cell = np.diag([L, L, L]).astype(np.float64)
    cell[1, 0] += skew
    cell[2, 1] -= 0.5 * skew
    rng = np.random.default_rng(23)
    frac = np.concatenate([rng.uniform([0, 0, 0], [1 / 3, 1, 1], size=(1400, 3)),
                           rng.uniform([1 / 3, 0, 0], [1, 1, 1], size=(1100, 3))])
    pos32 = (frac @ cell).astype(np.float32)
    cell32 = cell.astype(np.float32)
    pbc = np.array([True, True, True])
    numbers = rng.choice([1, 6, 7, 8], size=len(pos32))
    rp, _, _ = nl_ref.neighbor_csr(pos32, np.array([0, len(pos32)]), 5.0, cell32[None], pbc[None])
    deg = np.diff(rp)
    assert deg.max() > 256 and deg.min() < 256, (deg.min(), deg.max())
    Z = torch.tensor(numbers, device=gpu)
    P = torch.tensor(pos32, device=gpu)
    m1 = _model(64, 2, 20, seed=0)
    out_cl = m1.energy_forces_virial(Z, P, torch.tensor(cell32), torch.tensor(pbc))
    assert next(iter(m1._plans.values()))["plan"].cell_grid[0, 0] >= 3
    old = engine.CELL_LIST_MIN_ATOMS
    engine.CELL_LIST_MIN_ATOMS = 10 ** 9
    try:
        m2 = _model(64, 2, 20, seed=0)
        out_bf = m2.energy_forces_virial(Z, P, torch.tensor(cell32), torch.tensor(pbc))
        assert next(iter(m2._plans.values()))["plan"].cell_grid[0, 0] == 0
    finally:
        engine.CELL_LIST_MIN_ATOMS = old
    for a, b in zip(out_cl, out_bf):
        assert torch.equal(a, b)


def _check_ef(model, Z, P32, sizes, E, F, label, absolute=True):
    """Compare with the float64 oracle.  ``absolute``: the task's bars
    (1e-5 eV/atom, 1e-4 eV/A).  Otherwise (weights scaled up so forces reach
    hundreds of eV/A, beyond what float32 resolves to 1e-4) the bar is three
    times the error of the reference algorithm itself run in float32."""
    sd = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    batch = np.repeat(np.arange(len(sizes)), sizes)
    Eref, Fref = R.energy_and_forces(sd, Z, P32.astype(np.float64), cutoff=model.cutoff,
                                     batch=batch, n_structures=len(sizes))
    e_err = np.max(np.abs(E - Eref.numpy()) / np.asarray(sizes))
    f_err = np.max(np.abs(F - Fref.numpy()))
    e_tol, f_tol = E_TOL, F_TOL
    if not absolute:
        E32, F32 = R.energy_and_forces(sd, Z, P32.astype(np.float64), cutoff=model.cutoff,
                                       batch=batch, n_structures=len(sizes), dtype=torch.float32)
        e_tol = max(E_TOL, 3 * np.max(np.abs(E32.numpy() - Eref.numpy()) / np.asarray(sizes)))
        f_tol = max(F_TOL, 3 * np.max(np.abs(F32.numpy() - Fref.numpy())))
    note(f"{label}: max|dE|/N={e_err:.3e} (bar {e_tol:.1e}) max|dF|={f_err:.3e} (bar {f_tol:.1e})"
         f" |F|max={np.abs(Fref.numpy()).max():.2f}")
    assert e_err < e_tol, label
    assert f_err < f_tol, label


@pytest.mark.parametrize("H,L,R_", [(128, 3, 20), (64, 2, 20), (32, 2, 16), (128, 1, 32),
                                    (64, 8, 1)])
def test_energy_forces_vs_fp64_oracle(gpu, H, L, R_):
    """Seed-0 random initialisation (zero biases), as the benchmarks use; the
    supported range's ends too (one block and 32 basis functions; eight
    blocks and one basis function).  Eight random blocks give forces of order
    1e4 eV/A, beyond what float32 resolves to 1e-4: that case takes the
    relative bar (three times the reference algorithm's own float32 error)."""
    note(f"E+F H{H} L{L} R{R_}")
    model = _model(H, L, R_, seed=0)
    mols = _mols(12, 10, 100, seed=21)
    synth = pkg_mod("synth")
    Z, P, mol_ptr = synth.flatten(mols)
    sizes = np.diff(mol_ptr)
    P32 = P.astype(np.float32)
    e64, f, _ = model.evaluate(torch.tensor(Z, device=gpu), torch.tensor(P32, device=gpu), sizes)
    _check_ef(model, Z, P32, sizes, e64.cpu().numpy(), f.cpu().numpy(), f"H{H} L{L} R{R_}",
              absolute=L <= 3)


@pytest.mark.parametrize("H,L,R_", [(128, 3, 20), (32, 2, 16)])
def test_energy_forces_with_biases(gpu, H, L, R_):
    """Non-zero biases exercise every bias term (large forces: relative bar)."""
    note(f"E+F biased H{H} L{L} R{R_}")
    model = _model(H, L, R_, seed=0, bias_scale=0.1)
    mols = _mols(12, 10, 100, seed=21)
    synth = pkg_mod("synth")
    Z, P, mol_ptr = synth.flatten(mols)
    sizes = np.diff(mol_ptr)
    P32 = P.astype(np.float32)
    e64, f, _ = model.evaluate(torch.tensor(Z, device=gpu), torch.tensor(P32, device=gpu), sizes)
    _check_ef(model, Z, P32, sizes, e64.cpu().numpy(), f.cpu().numpy(), f"biased H{H}",
              absolute=False)


def test_golden_full_seed0_molecule(gpu):
    note("golden full seed-0 molecule")
    fx = golden("full_seed0.npz")
    model = _model(128, 3, 20, seed=0)
    Z = torch.tensor(fx["full_eval/numbers"], device=gpu)
    pos = torch.tensor(fx["full_eval/positions"], dtype=torch.float32, device=gpu)
    E, F = model.predict_energy_and_forces(Z, pos)
    n = len(fx["full_eval/numbers"])
    assert abs(model.last_energy_f64.item() - float(fx["full_eval/energy"])) / n < E_TOL
    assert np.abs(F.detach().cpu().numpy() - fx["full_eval/forces"]).max() < F_TOL
    model.eval()
    E2, F2 = model.predict_energy_and_forces(Z, pos)
    assert not F2.requires_grad
    assert np.abs(F2.cpu().numpy() - fx["full_eval/forces"]).max() < F_TOL


def test_autograd_forces_and_batched_api(gpu):
    note("forward + autograd.grad, batched")
    model = _model(128, 3, 20, seed=0)
    mols = _mols(5, 10, 40, seed=31)
    synth = pkg_mod("synth")
    Z, P, mol_ptr = synth.flatten(mols)
    sizes = np.diff(mol_ptr)
    batch = torch.tensor(np.repeat(np.arange(5), sizes), device=gpu)
    pos = torch.tensor(P.astype(np.float32), device=gpu, requires_grad=True)
    Zt = torch.tensor(Z, device=gpu)
    E = model(Zt, pos, batch=batch)
    assert E.shape == (5,)
    F = -torch.autograd.grad(E, pos, grad_outputs=torch.ones_like(E))[0]
    _check_ef(model, Z, P.astype(np.float32), sizes, model.last_energy_f64.cpu().numpy(),
              F.cpu().numpy(), "autograd batched")
    # each molecule alone gives the same numbers
    for b in range(5):
        sl = slice(mol_ptr[b], mol_ptr[b + 1])
        e1, f1 = model.predict_energy_and_forces(Zt[sl], pos.detach()[sl])
        assert abs(e1.item() - E[b].item()) < 1e-4
        assert (f1 - F[sl]).abs().max().item() < 1e-5


def test_structures_without_edges_and_locality(gpu):
    """A batch with a single atom, two atoms beyond the cutoff, an atom left
    alone inside a molecule and two ordinary molecules: energies and forces
    match the oracle, atoms without neighbours feel exactly zero force, and
    every structure evaluated alone gives bit-identical forces (nothing leaks
    between the rows of a shared tile)."""
    note("sparse rows and structure locality")
    model = _model(128, 3, 20, seed=0, bias_scale=0.05)
    a, b = _mols(2, 20, 40, seed=51)
    Pb = b.positions.copy()
    Pb[-1] += np.array([25.0, 0.0, 0.0])          # an atom far from the rest of its molecule
    structs = [(np.array([8]), np.zeros((1, 3))),
               (np.array([6, 1]), np.array([[0.0, 0.0, 0.0], [6.0, 0.0, 0.0]])),
               (a.numbers, a.positions),
               (b.numbers, Pb)]
    Z = np.concatenate([z for z, _ in structs])
    P32 = np.concatenate([p for _, p in structs]).astype(np.float32)
    sizes = np.array([len(z) for z, _ in structs])
    ptr = np.concatenate([[0], np.cumsum(sizes)])
    Zt, Pt = torch.tensor(Z, device=gpu), torch.tensor(P32, device=gpu)
    e64, f, _ = model.evaluate(Zt, Pt, sizes)
    _check_ef(model, Z, P32, sizes, e64.cpu().numpy(), f.cpu().numpy(), "sparse rows",
              absolute=False)
    lonely = [0, 1, 2, ptr[4] - 1]
    assert torch.count_nonzero(f[lonely]).item() == 0
    for s in range(len(structs)):
        sl = slice(ptr[s], ptr[s + 1])
        e1, f1, _ = model.evaluate(Zt[sl], Pt[sl], sizes[s:s + 1])
        assert torch.equal(f1, f[sl]), s
        assert abs(e1.item() - e64[s].item()) <= 1e-12 * max(1.0, abs(e1.item())), s


def test_runs_are_bit_reproducible(gpu):
    note("determinism")
    model = _model(128, 3, 20, seed=0)
    mols = _mols(8, 10, 100, seed=41)
    synth = pkg_mod("synth")
    Z, P, mol_ptr = synth.flatten(mols)
    Zt = torch.tensor(Z, device=gpu)
    Pt = torch.tensor(P.astype(np.float32), device=gpu)
    a = model.evaluate(Zt, Pt, np.diff(mol_ptr))
    b = model.evaluate(Zt, Pt, np.diff(mol_ptr))
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_periodic_water_energy_forces_virial(gpu):
    note("periodic water 300 atoms E/F/virial")
    synth = pkg_mod("synth")
    box = synth.water_box(5, seed=3)        # 125 waters, 15.5 A box
    model = _model(64, 2, 20, seed=0)
    P32 = box.positions.astype(np.float32)
    e64, f, vir = model.energy_forces_virial(torch.tensor(box.numbers, device=gpu),
                                             torch.tensor(P32, device=gpu),
                                             torch.tensor(box.cell), torch.tensor(box.pbc))
    rp, cl, sh = nl_ref.neighbor_csr(P32, np.array([0, len(P32)]), 5.0, box.cell[None],
                                     box.pbc[None])
    src = cl.astype(np.int64)
    dst = np.repeat(np.arange(len(P32)), np.diff(rp))
    sd = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    # oracle edge (src=j, dst=i) has vector p_j - p_i + shift @ cell
    E, F, W = R.energy_and_forces(sd, box.numbers, P32.astype(np.float64), cutoff=5.0,
                                  edge_index=np.stack([src, dst]), shifts=sh,
                                  cell=box.cell.astype(np.float32).astype(np.float64),
                                  virial=True)
    n = len(P32)
    assert abs(e64.item() - E.item()) / n < E_TOL
    assert np.abs(f.cpu().numpy() - F.numpy()).max() < F_TOL
    assert np.abs(vir.cpu().numpy()[0] - W.numpy()[0]).max() < 1e-3 * max(1.0, np.abs(W.numpy()).max())


def test_mixed_periodic_batch(gpu):
    """One batch with per-structure cells and pbc: a periodic water box (cell
    list), a molecule, a slab periodic in x and y only, and a skewed box
    (27-image minimum image): energies, forces and virials of every
    structure match the oracle run on that structure alone."""
    note("mixed batch: periodic box, molecule, slab, skewed box")
    synth = pkg_mod("synth")
    model = _model(64, 2, 20, seed=0)
    box = synth.water_box(5, seed=3)
    slab = synth.water_box(4, seed=5)
    skew = synth.water_box(4, seed=7)
    skew_cell = skew.cell.copy()
    skew_cell[1, 0] += 2.0
    skew_cell[2, 1] -= 1.0
    mol = _mols(1, 30, 30, seed=61)[0]
    structs = [(box.numbers, box.positions, box.cell, np.array([True, True, True])),
               (mol.numbers, mol.positions, np.zeros((3, 3)), np.array([False, False, False])),
               (slab.numbers, slab.positions, slab.cell, np.array([True, True, False])),
               (skew.numbers, skew.positions, skew_cell, np.array([True, True, True]))]
    Z = np.concatenate([s[0] for s in structs])
    P32 = np.concatenate([s[1] for s in structs]).astype(np.float32)
    cells = np.stack([s[2] for s in structs]).astype(np.float32)
    pbc = np.stack([s[3] for s in structs])
    sizes = np.array([len(s[0]) for s in structs])
    ptr = np.concatenate([[0], np.cumsum(sizes)])
    e64, f, vir = model.evaluate(torch.tensor(Z, device=gpu), torch.tensor(P32, device=gpu),
                                 sizes, cells, pbc, need_virial=True)
    sd = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    for s, (z, _, cell, per) in enumerate(structs):
        p = P32[ptr[s]:ptr[s + 1]]
        periodic = bool(per.any())
        rp, cl, sh = nl_ref.neighbor_csr(p, np.array([0, len(p)]), 5.0,
                                         cells[s][None] if periodic else None,
                                         per[None] if periodic else None)
        ei = np.stack([cl.astype(np.int64), np.repeat(np.arange(len(p)), np.diff(rp))])
        E, F, W = R.energy_and_forces(sd, z, p.astype(np.float64), cutoff=5.0, edge_index=ei,
                                      shifts=sh if periodic else None,
                                      cell=cells[s].astype(np.float64) if periodic else None,
                                      virial=True)
        e_err = abs(e64[s].item() - E.item()) / len(p)
        f_err = np.abs(f[ptr[s]:ptr[s + 1]].cpu().numpy() - F.numpy()).max()
        w_err = np.abs(vir[s].cpu().numpy() - W.numpy()[0]).max()
        note(f"structure {s}: |dE|/N={e_err:.2e} |dF|={f_err:.2e} |dW|={w_err:.2e}")
        assert e_err < E_TOL and f_err < F_TOL, s
        assert w_err < 1e-3 * max(1.0, np.abs(W.numpy()).max()), s


def test_filter_stream_overlap_is_bit_identical(gpu):
    """The filter kernels on a second stream (fork/join by events) give the
    same bits as every kernel on one stream: same kernels, same inputs."""
    note("two streams vs one stream")
    engine = pkg_mod("engine")
    synth = pkg_mod("synth")
    mols = _mols(8, 10, 100, seed=43)
    Z, P, mol_ptr = synth.flatten(mols)
    Zt = torch.tensor(Z, device=gpu)
    Pt = torch.tensor(P.astype(np.float32), device=gpu)
    box = synth.water_box(5, seed=3)
    bz = torch.tensor(box.numbers, device=gpu)
    bp = torch.tensor(box.positions.astype(np.float32), device=gpu)
    outs = {}
    old = engine.DEFAULT_STREAMS
    try:
        for n in (2, 1):
            engine.DEFAULT_STREAMS = n
            m = _model(128, 3, 20, seed=0)
            mol = m.evaluate(Zt, Pt, np.diff(mol_ptr))
            water = m.energy_forces_virial(bz, bp, torch.tensor(box.cell), torch.tensor(box.pbc))
            assert all(v["plan"].streams == n for v in m._plans.values())
            outs[n] = [mol[0], mol[1]] + list(water)
    finally:
        engine.DEFAULT_STREAMS = old
    for a, b in zip(outs[2], outs[1]):
        assert torch.equal(a, b)


def test_persistent_grid_size_is_bit_identical(gpu):
    """The autotuner picks the persistent grid of the pair kernels; every
    pair's arithmetic is independent of which workgroup runs its tile, so
    any grid (including sizes that are not a multiple of 8, which take the
    plain strided tile order) gives the same bits."""
    note("persistent grid 8 / 100 / 1536 / 4096")
    engine = pkg_mod("engine")
    synth = pkg_mod("synth")
    mols = _mols(24, 10, 100, seed=53)
    Z, P, mol_ptr = synth.flatten(mols)
    sizes = np.diff(mol_ptr)
    dev = torch.device(gpu)
    lay = engine.make_layout(sizes, dev)
    Zt = torch.tensor(Z, dtype=torch.int64, device=dev)
    Pt = torch.tensor(P.astype(np.float32), device=dev)
    model = _model(64, 2, 20, seed=0)
    w = model.packed_weights(dev)
    cap = engine.edge_capacity_bound(sizes, False)
    outs = []
    C = pkg_mod("_capi")
    for blocks in (8, 100, 1536, 4096):
        plan = engine.Plan(64, 2, 20, model.max_z, 5.0, lay, Zt, w, cap)
        C.check("set_blocks", C.lib().mlff_plan_set_blocks(plan.handle, blocks))
        e = torch.empty(len(sizes), dtype=torch.float64, device=dev)
        f = torch.empty(len(Z), 3, dtype=torch.float32, device=dev)
        plan.energy_forces(Pt, e, f)
        torch.cuda.synchronize()
        outs.append((e.clone(), f.clone()))
    for e, f in outs[1:]:
        assert torch.equal(e, outs[0][0]) and torch.equal(f, outs[0][1])


def test_weight_update_refreshes_split_planes(gpu):
    """The filter's second-layer weights also live as bf16 planes split at
    bind; changing the weights in place must rebind the plan and refresh
    them (a stale copy would give the old model's energies and forces)."""
    note("in-place weight update -> new results")
    model = _model(128, 3, 20, seed=0)
    mols = _mols(6, 10, 60, seed=47)
    synth = pkg_mod("synth")
    Z, P, mol_ptr = synth.flatten(mols)
    sizes = np.diff(mol_ptr)
    P32 = P.astype(np.float32)
    Zt, Pt = torch.tensor(Z, device=gpu), torch.tensor(P32, device=gpu)
    e0, f0, _ = model.evaluate(Zt, Pt, sizes)
    gen = torch.Generator().manual_seed(5)
    with torch.no_grad():
        for l in range(3):
            w = model.interactions[l].message.rbf_to_scalar[2].weight
            w.mul_(1.25).add_(0.02 * torch.randn(w.shape, generator=gen).to(w.device))
    e1, f1, _ = model.evaluate(Zt, Pt, sizes)
    assert not torch.equal(f0, f1)
    _check_ef(model, Z, P32, sizes, e1.cpu().numpy(), f1.cpu().numpy(), "after weight update")


def test_invariances_and_central_differences_on_the_kernels(gpu):
    """The reference's model checks (tests/unit/test_student_model.py:205-221,
    294-346, 377-411) on the HIP path itself: translation, permutation,
    extensivity over far-apart copies, and forces against central differences
    of the kernels' own energy.  Positions sit on a 2^-10 A grid, so shifts by
    whole multiples of 16 A are exact in float32 and leave every edge vector
    unchanged: those cases must agree bit for bit."""
    note("invariances and finite differences on the kernels")
    model = _model(128, 3, 20, seed=0, bias_scale=0.05)
    mol = _mols(1, 24, 24, seed=71)[0]
    Z = mol.numbers
    P = (np.round(mol.positions * 1024.0) / 1024.0).astype(np.float32)
    n = len(Z)

    def ef(z, p):
        e64, f, _ = model.evaluate(torch.tensor(z, device=gpu),
                                   torch.tensor(np.ascontiguousarray(p, dtype=np.float32),
                                                device=gpu), np.array([len(z)]))
        return e64.item(), f.cpu().numpy()

    e0, f0 = ef(Z, P)
    exact = np.array([64.0, -32.0, 16.0], dtype=np.float32)
    assert np.array_equal((P + exact) - exact, P)
    ex, fx = ef(Z, P + exact)
    assert ex == e0 and np.array_equal(fx, f0)
    et, ft = ef(Z, P + np.array([3.1, -2.2, 7.7], dtype=np.float32))
    assert abs(et - e0) / n < 1e-6 and np.abs(ft - f0).max() < 1e-4
    perm = np.random.default_rng(5).permutation(n)
    ep, fp = ef(Z[perm], P[perm])
    assert abs(ep - e0) / n < 1e-6 and np.abs(fp - f0[perm]).max() < 1e-4
    e2, f2 = ef(np.concatenate([Z, Z]), np.concatenate([P, P + np.float32(64.0)]))
    assert abs(e2 - 2.0 * e0) <= 1e-12 * abs(e0)
    assert np.array_equal(f2[:n], f0) and np.array_equal(f2[n:], f0)
    h = np.float32(2e-3)
    for atom, k in [(0, 0), (5, 1), (11, 2), (23, 0)]:
        pp, pm = P.copy(), P.copy()
        pp[atom, k] += h
        pm[atom, k] -= h
        step = float(pp[atom, k]) - float(pm[atom, k])
        fd = -(ef(Z, pp)[0] - ef(Z, pm)[0]) / step
        assert abs(fd - f0[atom, k]) < 5e-3 * max(1.0, abs(f0[atom, k])), (atom, k, fd, f0[atom, k])


def test_reference_memory_and_speed_budgets(gpu):
    """The reference's budgets (tests/unit/test_student_model.py:448-505):
    a forward pass over 1000 random atoms within 500 MB of peak device
    memory (measured here above what is already allocated), and under 50 ms
    per inference of 100 atoms."""
    import gc
    import time
    note("1000-atom memory budget, 100-atom inference time")
    gc.collect()
    torch.cuda.synchronize()
    model = _model(128, 3, 20, seed=0)
    g = torch.Generator().manual_seed(0)
    Z = torch.randint(1, 10, (1000,), generator=g).to(gpu)
    P = (torch.randn(1000, 3, generator=g) * 10.0).to(gpu)
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    with torch.no_grad():
        E = model(Z, P)
    torch.cuda.synchronize()
    assert torch.isfinite(E).all()
    peak_mb = (torch.cuda.max_memory_allocated() - base) / 2 ** 20
    note(f"peak above baseline: {peak_mb:.1f} MB")
    assert peak_mb < 500, peak_mb
    Z = torch.randint(1, 10, (100,), generator=g).to(gpu)
    P = (torch.randn(100, 3, generator=g) * 5.0).to(gpu)
    with torch.no_grad():
        for _ in range(10):
            model(Z, P)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(100):
            model(Z, P)
        torch.cuda.synchronize()
    ms = (time.perf_counter() - t) / 100 * 1e3
    note(f"100 atoms: {ms:.2f} ms per inference")
    assert ms < 50, ms


def test_coincident_atoms_follow_the_reference(gpu):
    """Two atoms at the same position: the pair is inside the cutoff (d = 0)
    and the reference's direction u = vec / (d + 1e-8) makes its forces huge
    (1e6 eV/A) but finite.  The kernels reproduce them to float32 relative
    precision, and the other atoms' forces to the usual absolute bar."""
    note("coincident atoms (d = 0 pair)")
    model = _model(128, 3, 20, seed=0)
    Z = np.array([6, 1, 8, 1])
    P = np.array([[0., 0., 0.], [0., 0., 0.], [1.2, 0.1, -0.3], [2.0, 1.0, 0.5]])
    e64, f, _ = model.evaluate(torch.tensor(Z, device=gpu),
                               torch.tensor(P.astype(np.float32), device=gpu), np.array([4]))
    sd = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    E, F = R.energy_and_forces(sd, Z, P.astype(np.float32).astype(np.float64))
    F = F.numpy()
    f = f.cpu().numpy()
    assert np.isfinite(f).all()
    assert abs(e64.item() - E.item()) / 4 < E_TOL
    big = np.abs(F[:2]).max()
    assert big > 1e4
    assert np.abs(f[:2] - F[:2]).max() < 1e-5 * big
    assert np.abs(f[2:] - F[2:]).max() < F_TOL


def test_empty_structure_gives_zero_energy(gpu):
    """No atoms: the reference's forward sums over nothing and returns 0
    (student_model.py:734-757); forces are an empty [0, 3] array, also
    through autograd in training mode."""
    note("empty structure")
    model = _model(32, 2, 16, seed=0).eval()
    Z = torch.zeros(0, dtype=torch.int64, device=gpu)
    P = torch.zeros(0, 3, dtype=torch.float32, device=gpu)
    with torch.no_grad():
        assert model(Z, P).item() == 0.0
    E, F = model.predict_energy_and_forces(Z, P)
    assert E.item() == 0.0 and tuple(F.shape) == (0, 3)
    model.train()
    P2 = P.clone().requires_grad_(True)
    E2 = model(Z, P2)
    g = torch.autograd.grad(E2, P2)[0]
    assert E2.item() == 0.0 and tuple(g.shape) == (0, 3)


def test_batch_with_an_empty_structure(gpu):
    """A batch index that skips a structure (batch = [0, 0, 0, 2, 2, 2, 2]):
    the reference's per-structure loop (student_model.py:734-757) gives the
    skipped structure energy 0; the others equal their own evaluations."""
    note("batch with an empty structure")
    model = _model(64, 2, 20, seed=0).eval()
    Z = torch.tensor([6, 1, 1, 8, 1, 1, 6], device=gpu)
    P = torch.tensor([[0.0, 0.0, 0.0], [1.1, 0.0, 0.0], [-0.4, 1.0, 0.0],
                      [10.0, 0.0, 0.0], [10.9, 0.3, 0.0], [9.6, 0.8, 0.2], [11.5, -1.2, 0.4]],
                     dtype=torch.float32, device=gpu)
    batch = torch.tensor([0, 0, 0, 2, 2, 2, 2], device=gpu)
    with torch.no_grad():
        E = model(Z, P, batch=batch)
        E0 = model(Z[:3], P[:3])
        E2 = model(Z[3:], P[3:])
    assert tuple(E.shape) == (3,)
    assert E[1].item() == 0.0
    assert abs(E[0].item() - E0.item()) < 1e-6 and abs(E[2].item() - E2.item()) < 1e-6


def test_periodic_water_full_model_energy_forces_virial(gpu):
    """The periodic bulk workload's model at its own size class: H128/L3/R20
    (seed-0 weights) on a 7^3-water box (1029 atoms, 21.8 A, cell-list
    neighbour search): energy, forces and virial against the FP64 oracle on
    the oracle's own minimum-image edges."""
    note("periodic water 7^3 H128/L3/R20 E/F/virial")
    synth = pkg_mod("synth")
    box = synth.water_box(7, seed=3)
    model = _model(128, 3, 20, seed=0)
    P32 = box.positions.astype(np.float32)
    e64, f, vir = model.energy_forces_virial(torch.tensor(box.numbers, device=gpu),
                                             torch.tensor(P32, device=gpu),
                                             torch.tensor(box.cell), torch.tensor(box.pbc))
    rp, cl, sh = nl_ref.neighbor_csr(P32, np.array([0, len(P32)]), 5.0, box.cell[None],
                                     box.pbc[None])
    ei = np.stack([cl.astype(np.int64), np.repeat(np.arange(len(P32)), np.diff(rp))])
    sd = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    E, F, W = R.energy_and_forces(sd, box.numbers, P32.astype(np.float64), cutoff=5.0,
                                  edge_index=ei, shifts=sh,
                                  cell=box.cell.astype(np.float32).astype(np.float64),
                                  virial=True)
    n = len(P32)
    e_err = abs(e64.item() - E.item()) / n
    f_err = np.abs(f.cpu().numpy() - F.numpy()).max()
    w_err = np.abs(vir.cpu().numpy()[0] - W.numpy()[0]).max()
    note(f"N={n} E={ei.shape[1]} |dE|/N={e_err:.2e} |dF|={f_err:.2e} |dW|={w_err:.2e}")
    assert n == 1029 and ei.shape[1] > 40 * n
    assert e_err < E_TOL and f_err < F_TOL
    assert w_err < 1e-3 * max(1.0, np.abs(W.numpy()).max())


def test_forward_with_analytical_forces_matches_oracle(gpu):
    """forward_with_analytical_forces (student_model.py:797-932): energy and
    forces of one molecule and of a shuffled three-molecule batch against the
    FP64 oracle, and equal to predict_energy_and_forces bit for bit."""
    note("forward_with_analytical_forces single + batched")
    model = _model(128, 3, 20, seed=0).eval()
    sd = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    mols = _mols(3, 15, 60, seed=23)
    m0 = mols[0]
    z = torch.tensor(m0.numbers, device=gpu)
    x32 = m0.positions.astype(np.float32)
    x = torch.tensor(x32, device=gpu)
    E, F = model.forward_with_analytical_forces(z, x)
    Eref, Fref = R.energy_and_forces(sd, m0.numbers, x32.astype(np.float64))
    assert abs(E.item() - Eref.item()) / len(z) < E_TOL
    assert np.abs(F.cpu().numpy() - Fref.numpy()).max() < F_TOL
    E2, F2 = model.predict_energy_and_forces(z, x)
    assert torch.equal(F, F2) and E.item() == E2.item()
    synth = pkg_mod("synth")
    Z, P, mol_ptr = synth.flatten(mols)
    b = np.repeat(np.arange(3), np.diff(mol_ptr))
    perm = np.random.default_rng(4).permutation(len(Z))       # atoms interleaved
    Zp, Pp, bp = Z[perm], P[perm].astype(np.float32), b[perm]
    Eb, Fb = model.forward_with_analytical_forces(torch.tensor(Zp, device=gpu),
                                                  torch.tensor(Pp, device=gpu),
                                                  batch=torch.tensor(bp, device=gpu))
    Eref, Fref = R.energy_and_forces(sd, Zp, Pp.astype(np.float64), batch=bp)
    sizes = np.bincount(bp)
    assert Eb.shape == (3,)
    assert np.max(np.abs(Eb.detach().cpu().numpy() - Eref.numpy()) / sizes) < E_TOL
    assert np.abs(Fb.cpu().numpy() - Fref.numpy()).max() < F_TOL


@pytest.mark.parametrize("H,L,R_", [(128, 3, 20), (32, 2, 16)])
def test_every_launch_shape_is_bit_identical(gpu, H, L, R_):
    """The launch-shape tuner picks pairs per filter tile (forward and
    adjoint), edges in flight per message row and the filter adjoint's
    variant among pre-instantiated kernels; each output element sums
    in the same order in all of them, so
    every combination gives the bits of the default (energies, forces and
    virials); a shape that was not instantiated is refused."""
    note(f"launch shapes H{H}: every combination vs default")
    engine = pkg_mod("engine")
    synth = pkg_mod("synth")
    C = pkg_mod("_capi")
    mols = _mols(20, 10, 100, seed=57)
    Z, P, mol_ptr = synth.flatten(mols)
    sizes = np.diff(mol_ptr)
    dev = torch.device(gpu)
    lay = engine.make_layout(sizes, dev)
    Zt = torch.tensor(Z, dtype=torch.int64, device=dev)
    Pt = torch.tensor(P.astype(np.float32), device=dev)
    model = _model(H, L, R_, seed=0, bias_scale=0.2)
    w = model.packed_weights(dev)
    cap = engine.edge_capacity_bound(sizes, False)
    plan = engine.Plan(H, L, R_, model.max_z, 5.0, lay, Zt, w, cap)
    shapes = engine.launch_shapes(H)
    assert all(len(v) >= 2 for v in shapes)

    def run():
        e = torch.empty(len(sizes), dtype=torch.float64, device=dev)
        f = torch.empty(len(Z), 3, dtype=torch.float32, device=dev)
        vir = torch.empty(len(sizes), 9, dtype=torch.float64, device=dev)
        plan.energy_forces(Pt, e, f, vir)
        torch.cuda.synchronize()
        return e, f, vir

    ref = run()
    import itertools
    for combo in itertools.product(*shapes):
        if combo[4] and combo[1] != shapes[1][0]:
            with pytest.raises(C.MlffError):       # variants exist for the default tile only
                plan.set_tiles(*combo)
            continue
        plan.set_tiles(*combo)
        got = run()
        for a, b in zip(ref, got):
            assert torch.equal(a, b), combo
    with pytest.raises(C.MlffError):
        plan.set_tiles(48, 0, 0, 0)


@pytest.mark.parametrize("H,L,R_", [(128, 3, 20), (32, 2, 16), (64, 1, 32)])
def test_device_weight_pack_equals_host_pack(gpu, H, L, R_):
    """mlff_pack_weights writes the blob from the parameters on the device in
    one launch: bit-identical to the host assembly (transposed slots, biases,
    padding and gaps zero), and an in-place update of a parameter shows in
    the next pack."""
    note(f"device weight pack H{H} L{L} R{R_}")
    engine = pkg_mod("engine")
    model = _model(H, L, R_, seed=3, bias_scale=0.1)
    sd = model.state_dict()
    dev_blob = engine.pack_weights(sd, gpu)
    host_blob = engine.pack_weights({k: v.cpu() for k, v in sd.items()}, "cpu")
    assert dev_blob.device.type == "cuda"
    assert torch.equal(dev_blob.cpu(), host_blob)
    with torch.no_grad():
        model.interactions[0].message.rbf_to_scalar[2].weight.mul_(-1.5)
        model.energy_head[4].bias.add_(0.25)
    sd = model.state_dict()
    again = engine.pack_weights(sd, gpu)
    assert torch.equal(again.cpu(), engine.pack_weights({k: v.cpu() for k, v in sd.items()},
                                                        "cpu"))
    assert not torch.equal(again.cpu(), host_blob)
