"""Device Velocity Verlet and the captured-graph MD loop (needs an MI355X)."""

import sys

import numpy as np
import pytest
import torch

from conftest import pkg, pkg_mod
from oracle import painn_ref as R

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)


def _model(H=128, L=3, R_=20, seed=0):
    torch.manual_seed(seed)
    return pkg().StudentForceField(hidden_dim=H, num_interactions=L, num_rbf=R_).to("cuda:0")


def test_velocity_verlet_matches_host_integrator(gpu):
    """10 device steps == ASE-form VV on the host driven by float64 oracle forces."""
    note("VV vs host VV with oracle forces")
    synth = pkg_mod("synth")
    md_mod = pkg_mod("md")
    model = _model()
    mol = synth.single_molecule(20, seed=5)
    md = md_mod.BatchedNVE(model, [mol], timestep_fs=0.5, seed=1, max_steps=20)
    x = md.x.cpu().numpy().copy()
    p = md.p.cpu().numpy().copy()
    m = synth.masses_of(mol.numbers)[:, None]
    sd = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    dt = 0.5 * synth.FS
    F = R.energy_and_forces(sd, mol.numbers, x.astype(np.float32).astype(np.float64))[1].numpy()
    for _ in range(10):
        p += 0.5 * dt * F
        x += dt * p / m
        F = R.energy_and_forces(sd, mol.numbers, x.astype(np.float32).astype(np.float64))[1].numpy()
        p += 0.5 * dt * F
    md.run(10, graph=False)
    assert np.abs(md.positions() - x).max() < 1e-6
    assert np.abs(md.p.cpu().numpy() - p).max() < 1e-4


def test_graph_replay_is_bit_identical_to_eager(gpu):
    note("graph vs eager")
    synth = pkg_mod("synth")
    md_mod = pkg_mod("md")
    model = _model()
    mols = synth.molecule_set(16, 10, 60, seed=7)
    a = md_mod.BatchedNVE(model, mols, seed=3, max_steps=60)
    b = md_mod.BatchedNVE(model, mols, seed=3, max_steps=60)
    a.run(40, graph=True, steps_per_graph=10)
    b.run(40, graph=False)
    assert torch.equal(a.x, b.x) and torch.equal(a.p, b.p)
    ea, ka = a.energies()
    eb, kb = b.energies()
    assert np.array_equal(ea, eb) and np.array_equal(ka, kb)


def test_capacity_overflow_is_replayed(gpu):
    note("capacity overflow replay")
    synth = pkg_mod("synth")
    md_mod = pkg_mod("md")
    model = _model(64, 2, 20)
    mols = synth.molecule_set(8, 30, 60, seed=9)
    ref = md_mod.BatchedNVE(model, mols, seed=4, max_steps=30)
    tight = md_mod.BatchedNVE(model, mols, seed=4, max_steps=30, headroom=0.6)
    ref.run(20, graph=True)
    tight.run(20, graph=True)
    assert tight.plan.ecap > 0.6 * ref.plan.ecap
    assert torch.equal(ref.x, tight.x)


def test_nve_energy_error_is_second_order(gpu):
    """Velocity Verlet's energy error scales as dt^2: halving dt over the same
    simulated time cuts the largest |E(t) - E(0)| about fourfold."""
    note("NVE energy error vs dt")
    synth = pkg_mod("synth")
    md_mod = pkg_mod("md")
    model = _model()
    mols = synth.molecule_set(32, 10, 100, seed=2)
    err = {}
    for dt, n in ((0.5, 100), (0.25, 200)):
        md = md_mod.BatchedNVE(model, mols, timestep_fs=dt, seed=2, max_steps=n)
        md.run(n, graph=True)
        epot, ekin = md.energies()
        etot = epot + ekin
        err[dt] = np.abs(etot - etot[0]).max(axis=0) / md.sizes
        assert np.all(np.isfinite(etot)) and np.all(ekin[1:] > 0)
    ratio = np.median(err[0.5] / err[0.25])
    note(f"median error ratio dt/(dt/2) = {ratio:.2f}; max err at 0.5 fs {err[0.5].max():.2e} eV/atom")
    assert 3.0 < ratio < 5.5
    # random-weight molecules heat to ~2000 K within 0.05 ps; the bound is loose
    assert err[0.5].max() < 5e-3


def test_periodic_water_md_runs(gpu):
    note("periodic water MD 50 steps")
    synth = pkg_mod("synth")
    md_mod = pkg_mod("md")
    model = _model(64, 2, 20)
    box = synth.water_box(6, seed=3)
    md = md_mod.BatchedNVE(model, [box], seed=3, max_steps=60)
    md.run(50, graph=True)
    s = md.summary()
    assert np.isfinite(s["final_energy"]).all()
    assert abs(s["energy_drift_per_atom"][0]) < 1e-3


def test_graph_with_filter_stream_matches_one_stream(gpu):
    """Captured MD chunks with the filter kernels on the plan's second stream
    replay to the same bits as the one-stream graph and as eager steps."""
    note("MD graph: two streams vs one stream")
    synth = pkg_mod("synth")
    md_mod = pkg_mod("md")
    model = _model()
    mols = synth.molecule_set(16, 10, 60, seed=8)
    two = md_mod.BatchedNVE(model, mols, seed=5, max_steps=40, streams=2)
    one = md_mod.BatchedNVE(model, mols, seed=5, max_steps=40, streams=1)
    eager = md_mod.BatchedNVE(model, mols, seed=5, max_steps=40, streams=2)
    two.run(30, graph=True, steps_per_graph=10)
    one.run(30, graph=True, steps_per_graph=10)
    eager.run(30, graph=False)
    for other in (one, eager):
        assert torch.equal(two.x, other.x) and torch.equal(two.p, other.p)
        assert np.array_equal(two.energies()[0], other.energies()[0])


def test_frame_ring_drains_every_step(gpu):
    """Frames recorded into a 7-row device ring, drained to the host after
    every chunk (graph chunks of 3 steps and eager remainders), equal the
    frames of a run that keeps one device row per step."""
    note("frame ring of 7 rows vs one row per step")
    synth = pkg_mod("synth")
    md_mod = pkg_mod("md")
    model = _model(64, 2, 20)
    mols = synth.molecule_set(3, 10, 30, seed=8)
    full = md_mod.BatchedNVE(model, mols, seed=3, max_steps=30, record_frames=True)
    ring = md_mod.BatchedNVE(model, mols, seed=3, max_steps=30, record_frames=True,
                             frame_window=7)
    assert ring.frame_x.shape[0] == 7
    for md in (full, ring):
        md.run(20, graph=True, steps_per_graph=3)
        md.run(10, graph=False)
    a, b = full.frames(), ring.frames()
    for u, w in zip(a, b):
        assert u.shape == (31, full.n_atoms, 3)
        assert np.array_equal(u, w)
    with pytest.raises(ValueError):
        md_mod.BatchedNVE(model, mols, seed=3, max_steps=5, record_frames=True, frame_window=1)


def test_edge_count_maximum_is_kept_across_lists(gpu):
    """MLFF_COUNT_EDGES holds the last neighbour list's count and
    MLFF_COUNT_EDGES_MAX the largest since the plan was bound: an MD chunk
    whose overflow happened before its last step regrows from the peak."""
    note("sticky maximum edge count")
    synth = pkg_mod("synth")
    engine = pkg_mod("engine")
    C = pkg_mod("_capi")
    model = _model(32, 2, 16)
    mols = synth.molecule_set(4, 30, 60, seed=9)
    Z, P, mp = synth.flatten(mols)
    sizes = np.diff(mp)
    layout = engine.make_layout(sizes, gpu)
    numbers = torch.tensor(Z, dtype=torch.int64, device=gpu)
    plan = engine.Plan(32, 2, 16, model.max_z, model.cutoff, layout, numbers,
                       model.packed_weights(gpu), engine.edge_capacity_bound(sizes, False))
    e = torch.empty(len(sizes), dtype=torch.float64, device=gpu)
    f = torch.empty(len(Z), 3, dtype=torch.float32, device=gpu)
    counts = []
    for scale in (0.6, 1.0):
        plan.energy_forces(torch.tensor(P * scale, dtype=torch.float32, device=gpu), e, f)
        c = plan.read_counts()
        counts.append(int(c[C.MLFF_COUNT_EDGES]))
        assert c[C.MLFF_COUNT_OVERFLOW] == 0
    assert counts[0] > counts[1]
    c = plan.read_counts()
    assert int(c[C.MLFF_COUNT_EDGES]) == counts[1]
    assert int(c[C.MLFF_COUNT_EDGES_MAX]) == counts[0]
