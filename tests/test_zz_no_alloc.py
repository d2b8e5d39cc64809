"""mlff_energy_forces_static (no device allocation, static scratch) against
mlff_energy_forces on a bound plan with the same config: bit-identical
energies, forces, virials, atomic energies and counts, on two calls in a
row, the second with a smaller batch, and on a periodic box (needs an
MI355X).  Called through the C ABI."""

import ctypes
import sys

import numpy as np
import pytest
import torch

from conftest import pkg, pkg_mod

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _model(H, L, R_, seed=0):
    torch.manual_seed(seed)
    return pkg().StudentForceField(hidden_dim=H, num_interactions=L, num_rbf=R_).to("cuda:0")


def _both(model, structures, cells=None, pbc=None):
    """(plan results, static results) for one batch."""
    C = pkg_mod("_capi")
    engine = pkg_mod("engine")
    synth = pkg_mod("synth")
    dev = torch.device("cuda:0")
    Z, P, mol_ptr = synth.flatten(structures)
    sizes = np.diff(mol_ptr)
    periodic = cells is not None
    layout = engine.make_layout(sizes, dev, cells, pbc)
    numbers = torch.tensor(Z, dtype=torch.int64, device=dev)
    pos = torch.tensor(P.astype(np.float32), device=dev)
    weights = model.packed_weights(dev)
    ecap = int(engine.count_edges(pos, layout, model.cutoff) * 1.25) + 64
    plan = engine.Plan(model.hidden_dim, model.num_interactions, model.num_rbf, model.max_z,
                       model.cutoff, layout, numbers, weights, ecap, streams=1)
    N, B = len(Z), len(sizes)

    def outs():
        return (torch.full((B,), float("nan"), dtype=torch.float64, device=dev),
                torch.full((N, 3), float("nan"), dtype=torch.float32, device=dev),
                torch.full((B, 9), float("nan"), dtype=torch.float64, device=dev),
                torch.full((N,), float("nan"), dtype=torch.float32, device=dev))

    ref = outs()
    plan.energy_forces(pos, *ref)
    torch.cuda.synchronize()
    ref_counts = plan.read_counts()
    got = outs()
    counts = torch.full((C.MLFF_COUNT_SLOTS,), 7, dtype=torch.int32, device=dev)
    cfg = C.Config(model.hidden_dim, model.num_interactions, model.num_rbf, model.max_z,
                   float(model.cutoff), N, B, plan.ecap, plan.pcap, int(periodic), 0, 0)
    note(f"static call: {B} structures, {N} atoms, periodic={periodic}")
    C.check("mlff_energy_forces_static", C.lib().mlff_energy_forces_static(
        C.stream_handle(), ctypes.byref(cfg), C.ptr(weights), C.ptr(numbers),
        C.ptr(layout.mol_ptr), C.ptr(layout.atom_mol), C.ptr(layout.cells), C.ptr(layout.pbc),
        C.ptr(pos), *[C.ptr(t) for t in got], C.ptr(counts)))
    torch.cuda.synchronize()
    return ref, ref_counts, got, counts.cpu().numpy()


def _same(ref, ref_counts, got, counts):
    for a, b in zip(ref, got):
        assert torch.isfinite(a).all()
        assert torch.equal(a, b)
    assert counts[0] == ref_counts[0] and counts[1] == ref_counts[1]
    assert counts[2] == 0 and counts[3] == 0


def test_static_entry_equals_plan_two_calls_second_smaller(gpu):
    synth = pkg_mod("synth")
    model = _model(128, 3, 20)
    big = synth.molecule_set(9, 10, 100, seed=2)
    small = big[:3]
    note("no-alloc: first call, 9 molecules")
    _same(*_both(model, big))
    note("no-alloc: second call, 3 molecules")
    _same(*_both(model, small))


def test_static_entry_periodic_and_compact(gpu):
    synth = pkg_mod("synth")
    box = synth.water_box(4, seed=3)
    model = _model(64, 2, 20, seed=1)
    note("no-alloc: periodic water box, H64/L2")
    _same(*_both(model, [box], box.cell[None], box.pbc[None]))
    compact = _model(32, 2, 16, seed=3)
    note("no-alloc: compact H32 student, 40 molecules")
    _same(*_both(compact, synth.molecule_set(40, 10, 60, seed=5)))


def test_static_entry_rejects_oversized_batch(gpu):
    C = pkg_mod("_capi")
    cfg = C.Config(128, 3, 20, 118, 5.0, 20000, 200, 600000, 300000, 0, 0, 0)
    fake = ctypes.c_void_p(1 << 20)
    rc = C.lib().mlff_energy_forces_static(C.stream_handle(), ctypes.byref(cfg), *([fake] * 12))
    assert rc == -4
    assert b"static scratch" in C.lib().mlff_last_error()
