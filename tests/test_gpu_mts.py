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
