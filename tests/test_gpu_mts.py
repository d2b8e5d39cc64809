"""Multiple-time-step integrator (needs an MI355X)."""

import numpy as np
import pytest
import torch

from conftest import pkg, pkg_mod

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)


def _model(H=128, L=3, R_=20, seed=0):
    torch.manual_seed(seed)
    return pkg().StudentForceField(hidden_dim=H, num_interactions=L, num_rbf=R_).to("cuda:0").eval()


def test_mts_with_identical_models_is_velocity_verlet(gpu):
    """fast == full: the slow force is exactly zero and r-RESPA reduces to VV
    at the inner step."""
    note("MTS(full, full) == VV")
    synth = pkg_mod("synth")
    mts = pkg_mod("mts")
    md_mod = pkg_mod("md")
    full = _model()
    mols = synth.molecule_set(8, 10, 50, seed=12)
    a = mts.MTSNVE(full, full, mols, seed=5, max_outer=20)
    b = md_mod.BatchedNVE(full, mols, seed=5, max_steps=80)
    a.run(10, chunk=5)
    b.run(40, graph=True)
    assert torch.allclose(a.x, b.x, rtol=0, atol=1e-9)
    assert torch.allclose(a.p, b.p, rtol=0, atol=1e-7)
    ea, ka = a.energies()
    eb, kb = b.energies()
    assert np.allclose(ea[-1], eb[-1], atol=1e-6) and np.allclose(ka[-1], kb[-1], atol=1e-6)


def test_full_fallback_is_velocity_verlet(gpu):
    """Threshold 0: every structure falls back to the full model inside."""
    note("fallback everywhere == VV")
    synth = pkg_mod("synth")
    mts = pkg_mod("mts")
    md_mod = pkg_mod("md")
    full = _model()
    fast = _model(32, 2, 16, seed=3)
    mols = synth.molecule_set(6, 10, 40, seed=13)
    a = mts.MTSNVE(full, fast, mols, seed=6, threshold=0.0, max_outer=20)
    b = md_mod.BatchedNVE(full, mols, seed=6, max_steps=80)
    a.run(10, chunk=5)
    b.run(40, graph=True)
    assert a.replays >= 1 and a.fallback_steps > 0
    assert torch.allclose(a.x, b.x, rtol=0, atol=1e-9)


def test_graph_replay_matches_eager(gpu):
    note("MTS graph vs eager")
    synth = pkg_mod("synth")
    mts = pkg_mod("mts")
    full = _model()
    fast = _model(32, 2, 16, seed=3)
    mols = synth.molecule_set(8, 10, 60, seed=14)
    a = mts.MTSNVE(full, fast, mols, seed=7, threshold=1e9, max_outer=40)
    b = mts.MTSNVE(full, fast, mols, seed=7, threshold=1e9, max_outer=40)
    a.run(30, chunk=10, graph=True)
    b.run(30, chunk=10, graph=False)
    assert torch.equal(a.x, b.x) and torch.equal(a.p, b.p)
    assert np.array_equal(a.energies()[0], b.energies()[0])


def test_train_student_on_md_frames(gpu):
    note("frames + compact student")
    synth = pkg_mod("synth")
    mts = pkg_mod("mts")
    full = _model()
    mols = synth.molecule_set(16, 10, 40, seed=15)
    frames = mts.md_frames(full, mols, n_frames=6, stride=4, seed=1)
    assert frames.n_frames == 96
    torch.manual_seed(0)
    student = pkg().StudentForceField(hidden_dim=32, num_interactions=2, num_rbf=16).to("cuda:0")
    hist = mts.train_student(student, frames, steps=30, batch_molecules=8, lr=3e-3)
    note(f"loss {hist[0]:.4f} -> {np.mean(hist[-5:]):.4f}")
    assert np.mean(hist[-5:]) < 0.5 * hist[0]


def _oracle_forces(sd, Z, x, batch):
    """FP64 oracle forces at the float32 positions the device model sees."""
    from oracle import painn_ref as R
    x32 = x.astype(np.float32).astype(np.float64)
    _, F = R.energy_and_forces(sd, Z, x32, cutoff=5.0, batch=batch, n_structures=int(batch.max()) + 1)
    return F.numpy()


def test_mts_matches_host_respa_on_oracle_forces(gpu):
    """r-RESPA with a distinct compact inner model (H32/L2/R16) and the full
    model (H128/L3/R20) outside, on the first molecules of the 512-molecule
    seed-5 workload: 3 outer x 4 inner steps on the device equal an r-RESPA
    written here on the host, driven by FP64 oracle forces of both models
    (p += Dt/2 (F_full - F_fast); 4 x [p += dt/2 F_fast; x += dt p/m;
    p += dt/2 F_fast]; p += Dt/2 (F_full - F_fast))."""
    note("MTS vs host r-RESPA on FP64 oracle forces (seed-5 molecules)")
    synth = pkg_mod("synth")
    mts = pkg_mod("mts")
    full = _model()
    fast = _model(32, 2, 16, seed=3)
    mols = synth.molecule_set(512, 10, 100, seed=5)[:3]
    Z, X, mol_ptr = synth.flatten(mols)
    batch = np.repeat(np.arange(len(mols)), np.diff(mol_ptr))
    m = synth.masses_of(Z)[:, None]
    rng = np.random.Generator(np.random.PCG64(9))
    P = np.concatenate([synth.maxwell_boltzmann(s.numbers, s.positions, 300.0, rng) for s in mols])
    a = mts.MTSNVE(full, fast, mols, momenta=P, threshold=1e9, max_outer=4)
    a.run(3, chunk=3, graph=False)
    sd_full = {k: v.detach().cpu() for k, v in full.state_dict().items()}
    sd_fast = {k: v.detach().cpu() for k, v in fast.state_dict().items()}
    dt = 0.5 * synth.FS
    n_inner = 4
    Dt = n_inner * dt
    x, p = X.copy(), P.copy()
    for _ in range(3):
        slow = _oracle_forces(sd_full, Z, x, batch) - _oracle_forces(sd_fast, Z, x, batch)
        p = p + 0.5 * Dt * slow
        f = _oracle_forces(sd_fast, Z, x, batch)
        for _ in range(n_inner):
            p = p + 0.5 * dt * f
            x = x + dt * p / m
            f = _oracle_forces(sd_fast, Z, x, batch)
            p = p + 0.5 * dt * f
        slow = _oracle_forces(sd_full, Z, x, batch) - f
        p = p + 0.5 * Dt * slow
    dx = np.abs(a.x.cpu().numpy() - x).max()
    dp = np.abs(a.p.cpu().numpy() - p).max()
    note(f"|dx|={dx:.2e} A  |dp|={dp:.2e}")
    assert dx <= 1e-6 and dp <= 1e-4
    # the inner model really differs from the outer one
    assert np.abs(_oracle_forces(sd_full, Z, X, batch) - _oracle_forces(sd_fast, Z, X, batch)).max() > 0.1


def test_mts_full_seed5_workload_graph_equals_eager(gpu):
    """The 512-molecule seed-5 MTS workload itself: two chunks of 5 outer
    steps replayed from HIP graphs equal the eager loop bit for bit, and the
    ledger stays finite."""
    note("MTS 512 molecules seed 5: graph vs eager, 10 outer steps")
    synth = pkg_mod("synth")
    mts = pkg_mod("mts")
    full = _model()
    fast = _model(32, 2, 16, seed=3)
    mols = synth.molecule_set(512, 10, 100, seed=5)
    a = mts.MTSNVE(full, fast, mols, seed=5, threshold=1e9, max_outer=12)
    b = mts.MTSNVE(full, fast, mols, seed=5, threshold=1e9, max_outer=12)
    a.run(10, chunk=5, graph=True)
    b.run(10, chunk=5, graph=False)
    assert a.N > 25000
    assert torch.equal(a.x, b.x) and torch.equal(a.p, b.p)
    ea, ka = a.energies()
    assert np.array_equal(ea, b.energies()[0]) and np.isfinite(ea).all() and np.isfinite(ka).all()
