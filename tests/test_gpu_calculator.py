"""Calculator and NVE-harness drop-ins on the GPU (needs an MI355X)."""

import numpy as np
import pytest
import torch

from atoms_standin import Atoms
from conftest import pkg, pkg_mod
from oracle import nl_ref
from oracle import painn_ref as R

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def ckpt(tmp_path_factory):
    torch.manual_seed(0)
    m = pkg().StudentForceField()
    path = tmp_path_factory.mktemp("ck") / "student.pt"
    m.save(path)
    return path, {k: v.detach().cpu() for k, v in m.state_dict().items()}


def _mol(n, seed):
    s = pkg_mod("synth").single_molecule(n, seed)
    return Atoms(s.numbers, s.positions)


def test_calculate_matches_oracle_and_caches(gpu, ckpt):
    note("calculator single molecule")
    path, sd = ckpt
    calc = pkg().StudentForceFieldCalculator(path, device="cuda", enable_timing=True)
    atoms = _mol(30, 4)
    atoms.calc = calc
    e = atoms.get_potential_energy()
    f = atoms.get_forces()
    E, F = R.energy_and_forces(sd, atoms.numbers, atoms.positions.astype(np.float32).astype(np.float64))
    assert abs(e - E.item()) / len(atoms) < 1e-5
    assert np.abs(f - F.numpy()).max() < 1e-4
    assert calc.n_calls == 1                    # forces came from the cached results
    atoms.positions[0, 0] += 0.01
    atoms.get_forces()
    assert calc.n_calls == 2
    stats = calc.get_timing_stats()
    assert stats["n_calls"] == 2 and stats["avg_time"] > 0


def test_invalid_atoms_raise_value_error(gpu, ckpt):
    path, _ = ckpt
    calc = pkg().StudentForceFieldCalculator(path)
    bad = Atoms([0, 1], [[0, 0, 0], [1, 0, 0]])
    with pytest.raises(ValueError):
        calc.calculate(bad)


def test_calculate_batch_equals_single(gpu, ckpt):
    note("calculate_batch")
    path, _ = ckpt
    calc = pkg().StudentForceFieldCalculator(path)
    mols = [_mol(n, 10 + n) for n in (12, 40, 25, 60)]
    res = calc.calculate_batch(mols)
    for a, r in zip(mols, res):
        single = pkg().StudentForceFieldCalculator(path)
        a.calc = single
        assert abs(a.get_potential_energy() - r["energy"]) < 1e-5 * len(a)
        assert np.abs(a.get_forces() - r["forces"]).max() < 1e-5
    assert calc.calculate_batch([]) == []


def test_periodic_stress_is_virial_over_volume(gpu, ckpt):
    note("stress periodic water")
    path, sd = ckpt
    synth = pkg_mod("synth")
    box = synth.water_box(5, seed=3)
    atoms = Atoms(box.numbers, box.positions, box.cell, True)
    calc = pkg().StudentForceFieldCalculator(path, enable_stress=True)
    atoms.calc = calc
    stress = atoms.get_stress()
    P32 = box.positions.astype(np.float32)
    rp, cl, sh = nl_ref.neighbor_csr(P32, np.array([0, len(P32)]), 5.0, box.cell[None], box.pbc[None])
    ei = np.stack([cl.astype(np.int64), np.repeat(np.arange(len(P32)), np.diff(rp))])
    _, _, W = R.energy_and_forces(sd, box.numbers, P32.astype(np.float64), edge_index=ei,
                                  shifts=sh, cell=box.cell.astype(np.float32).astype(np.float64),
                                  virial=True)
    W = W.numpy()[0]
    s = 0.5 * (W + W.T) / abs(np.linalg.det(box.cell))
    want = np.array([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])
    assert np.abs(stress - want).max() < 1e-5 + 1e-3 * np.abs(want).max()
    mol = _mol(10, 1)
    mol.calc = calc
    assert np.all(mol.get_stress() == 0)     # non-periodic: zeros, as the reference


def test_nve_harness_api(gpu, ckpt):
    note("NVEMDHarness 40 steps")
    path, _ = ckpt
    calc = pkg().StudentForceFieldCalculator(path)
    h = pkg().NVEMDHarness(_mol(20, 2), calc, temperature=300.0, timestep=0.5, log_interval=10,
                           seed=0)
    out = h.run_simulation(40)
    for k in ("energy_drift_pct", "energy_drift_abs", "avg_temperature", "wall_time_s",
              "initial_energy", "final_energy", "trajectory_data", "n_steps"):
        assert k in out
    td = out["trajectory_data"]
    assert len(td["total_energy"]) == 41 and len(td["positions"]) == 41
    assert abs(out["energy_drift_abs"]) / 20 < 1e-3
    with pytest.raises(ValueError):
        h.run_simulation(0)


def test_nve_harness_collects_every_step(gpu, ckpt, tmp_path):
    """The reference harness collects positions, velocities and forces every
    step (test_run_simulation_basic, test_trajectory_data_shapes,
    test_get_trajectory_array, test_multiple_runs_without_reset, test_reset,
    test_trajectory_file_writing).  Here the frames come from the device, 40
    steps through replayed graphs plus 3 eager ones; consecutive frames must
    satisfy the Velocity Verlet update itself, and each frame's forces must
    be the calculator's forces at that frame's positions."""
    note("NVEMDHarness per-step frames")
    path, _ = ckpt
    calc = pkg().StudentForceFieldCalculator(path)
    atoms = _mol(20, 2)
    xyz = tmp_path / "auto.xyz"
    h = pkg().NVEMDHarness(atoms, calc, temperature=300.0, timestep=0.5, seed=0,
                           trajectory_file=xyz, log_interval=10)
    h.run_simulation(43, save_interval=10)
    assert h.is_complete and h.n_atoms == 20
    x = h.get_trajectory_array("positions")
    v = h.get_trajectory_array("velocities")
    f = h.get_trajectory_array("forces")
    t = h.get_trajectory_array("time")
    assert x.shape == v.shape == f.shape == (44, 20, 3) and t.shape == (44,)
    # trajectory_file holds steps 0, 10, 20, 30, 40
    assert len(xyz.read_text().splitlines()) == 5 * (2 + 20)
    synth = pkg_mod("synth")
    m = synth.masses_of(atoms.get_atomic_numbers())[None, :, None]
    dt = 0.5 * synth.FS
    np.testing.assert_allclose(x[1:], x[:-1] + dt * v[:-1] + 0.5 * dt * dt * f[:-1] / m,
                               rtol=0, atol=1e-9)
    np.testing.assert_allclose(v[1:], v[:-1] + 0.5 * dt * (f[:-1] + f[1:]) / m, rtol=0,
                               atol=1e-9)
    np.testing.assert_allclose(x[-1], h.atoms.get_positions(), rtol=0, atol=0)
    for k in (0, 7, 25, 43):
        a = atoms.copy()
        a.set_positions(x[k])
        a.calc = calc
        np.testing.assert_allclose(f[k], a.get_forces(), rtol=0, atol=1e-4)
    ekin = 0.5 * np.sum(m * v * v, axis=(1, 2))
    np.testing.assert_allclose(ekin, h.get_trajectory_array("kinetic_energy"), rtol=1e-12)
    # a second run without reset appends its own frames, time from 0 again
    h.run_simulation(5, initialize_velocities=False)
    assert len(h.trajectory_data["time"]) == 44 + 6
    assert h.trajectory_data["time"][44] == 0.0
    np.testing.assert_allclose(h.trajectory_data["positions"][44], x[-1], rtol=0, atol=0)
    out = tmp_path / "all.xyz"
    h.save_trajectory(out, format="xyz")
    assert len(out.read_text().splitlines()) == 50 * (2 + 20)
    h.reset()
    assert not h.is_complete and len(h.trajectory_data["time"]) == 0


def test_torchscript_export_runs_on_the_kernels(gpu, ckpt, tmp_path):
    """The reference's TorchScript export (SimpleWrapper traced on a 50-atom
    molecule): the loaded file evaluates other sizes on the fused kernels with
    the same bits as the model, autograd gives the adjoint forces, and the
    calculator's use_jit path reads the same file."""
    note("TorchScript export")
    jit = pkg_mod("jit")
    path, sd = ckpt
    model = pkg().StudentForceField.load(path, device="cuda")
    out = tmp_path / "student_jit.pt"
    jit.export_to_torchscript(model, out, num_atoms=50, device="cuda")
    m = jit.load_torchscript(out, "cuda")
    assert any(k.startswith("model.interactions.0.message") for k in m.state_dict())
    atoms = _mol(37, 12)
    z = torch.tensor(atoms.numbers, device=gpu)
    x = torch.tensor(atoms.positions, dtype=torch.float32, device=gpu, requires_grad=True)
    E = m(z, x)
    F = -torch.autograd.grad(E, x)[0]
    model.eval()
    E2, F2 = model.predict_energy_and_forces(z, x.detach())
    assert E.item() == E2.item()
    assert torch.equal(F, F2)
    calc = pkg().StudentForceFieldCalculator(None, device="cuda", use_jit=True, jit_path=out)
    atoms.calc = calc
    assert np.array_equal(atoms.get_forces(), F2.cpu().numpy().astype(np.float64))


def test_reset_and_repr(gpu, ckpt):
    """The reference's calculator-state tests
    (tests/integration/test_ase_calculator.py:314-334): after reset() the
    next call computes again, and repr names the class and the checkpoint."""
    note("calculator reset and repr")
    path, _ = ckpt
    calc = pkg().StudentForceFieldCalculator(path)
    atoms = _mol(12, 5)
    atoms.calc = calc
    e1 = atoms.get_potential_energy()
    calc.reset()
    e2 = atoms.get_potential_energy()
    assert np.isfinite(e2) and e1 == e2 and calc.n_calls == 2
    r = repr(calc)
    assert "StudentForceFieldCalculator" in r and path.name in r


def test_steepest_descent_lowers_the_energy(gpu, ckpt):
    """The reference's geometry-optimisation check
    (tests/integration/test_ase_calculator.py:158-175, BFGS through ASE,
    which is absent here) as plain steepest descent: small steps along the
    calculator's forces lower the energy at every step."""
    note("steepest descent with the calculator's forces")
    path, _ = ckpt
    calc = pkg().StudentForceFieldCalculator(path)
    atoms = _mol(16, 9)
    atoms.calc = calc
    energies = [atoms.get_potential_energy()]
    for _ in range(10):
        f = atoms.get_forces()
        step = 0.01 / max(np.abs(f).max(), 1e-12)     # largest displacement 0.01 A
        atoms.positions = atoms.positions + step * f
        energies.append(atoms.get_potential_energy())
    assert np.all(np.diff(energies) < 0), energies


def test_harness_overhead_and_trajectory_memory(gpu, ckpt):
    """test_md_framework.py test_harness_overhead and
    test_trajectory_memory_efficiency: 100 steps of a water molecule finish
    in under 5 s and the position and force lists stay under 1 MB."""
    import sys
    import time
    note("NVEMDHarness water 100 steps")
    path, _ = ckpt
    calc = pkg().StudentForceFieldCalculator(path)
    water = Atoms([8, 1, 1], [[0.0, 0.0, 0.119], [0.0, 0.763, -0.477], [0.0, -0.763, -0.477]])
    h = pkg().NVEMDHarness(water, calc, temperature=100.0, timestep=1.0, seed=0)
    t0 = time.perf_counter()
    out = h.run_simulation(steps=100)
    assert time.perf_counter() - t0 < 5.0
    assert out["n_steps"] == 100 and len(h.trajectory_data["time"]) == 101
    td = h.trajectory_data
    assert sys.getsizeof(td["positions"]) + sys.getsizeof(td["forces"]) < 1_000_000
    assert not np.any(np.isnan(h.get_trajectory_array("total_energy")))


def test_torchscript_export_matches_oracle(gpu, ckpt, tmp_path):
    """The exported TorchScript file's energy and forces (forces through
    autograd of the loaded module, as export_to_torchscript.py's users take
    them) against the FP64 oracle on a molecule of another size than the
    tracing one."""
    note("TorchScript export vs oracle")
    jit = pkg_mod("jit")
    path, sd = ckpt
    model = pkg().StudentForceField.load(path, device="cuda")
    out = tmp_path / "student_jit.pt"
    jit.export_to_torchscript(model, out, num_atoms=50, device="cuda")
    m = jit.load_torchscript(out, "cuda")
    atoms = _mol(64, 31)
    x32 = atoms.positions.astype(np.float32)
    z = torch.tensor(atoms.numbers, device=gpu)
    x = torch.tensor(x32, device=gpu, requires_grad=True)
    E = m(z, x)
    F = -torch.autograd.grad(E, x)[0]
    Eref, Fref = R.energy_and_forces(sd, atoms.numbers, x32.astype(np.float64))
    assert abs(E.item() - Eref.item()) / len(atoms) < 1e-5
    assert np.abs(F.cpu().numpy() - Fref.numpy()).max() < 1e-4


def test_calculate_batch_stress_matches_oracle(gpu, ckpt):
    """calculate_batch with enable_stress on two periodic boxes of different
    cells and a molecule: each periodic structure's stress is sym(virial)/V
    of the FP64 oracle's virial, the molecule's is zeros (reference
    ase_calculator.py:590-645, 810-812)."""
    note("calculate_batch stress: two boxes + molecule")
    path, sd = ckpt
    synth = pkg_mod("synth")
    b1 = synth.water_box(4, seed=3)
    b2 = synth.water_box(4, seed=8)
    cell2 = b2.cell.copy()
    cell2[1, 0] += 1.5                         # skewed second box
    atoms = [Atoms(b1.numbers, b1.positions, b1.cell, True), _mol(20, 3),
             Atoms(b2.numbers, b2.positions, cell2, True)]
    calc = pkg().StudentForceFieldCalculator(path, enable_stress=True)
    res = calc.calculate_batch(atoms, properties=["energy", "forces", "stress"])
    assert np.all(res[1]["stress"] == 0)
    for a, r in ((atoms[0], res[0]), (atoms[2], res[2])):
        P32 = a.positions.astype(np.float32)
        c32 = a.cell.astype(np.float32)
        rp, cl, sh = nl_ref.neighbor_csr(P32, np.array([0, len(P32)]), 5.0, c32[None],
                                         np.ones((1, 3), bool))
        ei = np.stack([cl.astype(np.int64), np.repeat(np.arange(len(P32)), np.diff(rp))])
        E, F, W = R.energy_and_forces(sd, a.numbers, P32.astype(np.float64), edge_index=ei,
                                      shifts=sh, cell=c32.astype(np.float64), virial=True)
        W = W.numpy()[0]
        s = 0.5 * (W + W.T) / abs(np.linalg.det(a.cell))
        want = np.array([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])
        assert abs(r["energy"] - E.item()) / len(a) < 1e-5
        assert np.abs(r["forces"] - F.numpy()).max() < 1e-4
        assert np.abs(r["stress"] - want).max() < 1e-5 + 1e-3 * np.abs(want).max()


def test_nve_harness_continues_from_the_atoms_velocities(gpu, ckpt):
    """run_simulation(initialize_velocities=False) starts from the velocities
    already on the atoms (as ASE's VelocityVerlet does), and after a run the
    atoms hold the final positions and momenta: their kinetic energy is the
    last frame's."""
    note("NVEMDHarness: continue from the atoms' velocities")
    path, _ = ckpt
    calc = pkg().StudentForceFieldCalculator(path)
    atoms = _mol(18, 6)
    rng = np.random.default_rng(2)
    atoms.set_velocities(rng.normal(size=(18, 3)) * 0.01)
    v0 = atoms.get_velocities()
    h = pkg().NVEMDHarness(atoms, calc, temperature=300.0, seed=0)
    out = h.run_simulation(20, initialize_velocities=False)
    td = out["trajectory_data"]
    np.testing.assert_allclose(td["velocities"][0], v0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(h.atoms.get_positions(), td["positions"][-1], rtol=0, atol=0)
    assert abs(h.atoms.get_kinetic_energy() - td["kinetic_energy"][-1]) < 1e-9
    out2 = h.run_simulation(10, initialize_velocities=False)
    td2 = out2["trajectory_data"]
    np.testing.assert_allclose(td2["velocities"][21], td["velocities"][20], rtol=0, atol=1e-12)
