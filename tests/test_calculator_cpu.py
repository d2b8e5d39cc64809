"""Calculator constructor and validation contract (CPU only; no evaluation)."""

import numpy as np
import pytest
import torch

from conftest import pkg, pkg_mod


def test_missing_checkpoint_raises_file_not_found(tmp_path):
    with pytest.raises(FileNotFoundError):
        pkg().StudentForceFieldCalculator(tmp_path / "nope.pt", device="cpu")


def test_jit_without_path_raises_value_error(tmp_path):
    with pytest.raises(ValueError):
        pkg().StudentForceFieldCalculator(tmp_path / "x.pt", device="cpu", use_jit=True)


def test_validation_errors():
    calc_mod = pkg_mod("calculator")
    v = calc_mod.StudentForceFieldCalculator._validate_inputs
    ok_cell = np.eye(3) * 10
    with pytest.raises(ValueError, match="empty"):
        v(np.zeros((0, 3)), np.zeros(0, dtype=int), ok_cell, np.zeros(3, bool))
    with pytest.raises(ValueError, match="atomic numbers"):
        v(np.zeros((2, 3)), np.array([0, 1]), ok_cell, np.zeros(3, bool))
    with pytest.raises(ValueError, match="atomic numbers"):
        v(np.zeros((2, 3)), np.array([119, 1]), ok_cell, np.zeros(3, bool))
    with pytest.raises(ValueError, match="NaN"):
        v(np.array([[0, 0, np.nan]]), np.array([1]), ok_cell, np.zeros(3, bool))
    with pytest.raises(ValueError, match="Cell"):
        v(np.zeros((1, 3)), np.array([1]), np.full((3, 3), np.inf), np.ones(3, bool))
    with pytest.warns(UserWarning):
        v(np.zeros((1, 3)), np.array([1]), np.zeros((3, 3)), np.ones(3, bool))


def test_checkpoint_roundtrip_keeps_config_and_weights(tmp_path):
    model = pkg_mod("model")
    torch.manual_seed(3)
    m = model.StudentForceField(hidden_dim=64, num_interactions=2, num_rbf=16)
    m.save(tmp_path / "m.pt")
    m2 = model.StudentForceField.load(tmp_path / "m.pt")
    assert (m2.hidden_dim, m2.num_interactions, m2.num_rbf) == (64, 2, 16)
    for (k, a), (k2, b) in zip(m.state_dict().items(), m2.state_dict().items()):
        assert k == k2 and torch.equal(a, b)


def test_unsupported_hidden_fails_loudly_on_use():
    model = pkg_mod("model")
    m = model.StudentForceField(hidden_dim=48)
    with pytest.raises(RuntimeError):
        m.evaluate(torch.zeros(2, dtype=torch.long), torch.zeros(2, 3), np.array([2]))


def test_cell_grid_choice():
    engine = pkg_mod("engine")
    lay = engine.make_layout([3000, 60, 400], "cpu",
                             cells=np.stack([np.eye(3) * 31.0, np.eye(3) * 20.0,
                                             np.diag([16.0, 30.0, 12.0])]),
                             pbc=np.array([[1, 1, 1], [1, 1, 1], [1, 1, 0]], bool))
    grid, n = engine.cell_grid(lay, 5.0)
    assert tuple(grid[0]) == (6, 6, 6, 0)     # floor(31 / 5.00005) = 6
    assert grid[1, 0] == 0                    # below the atom threshold
    assert grid[2, 0] == 0                    # not periodic along z
    assert n == 216
    lay2 = engine.make_layout([500], "cpu", cells=np.eye(3)[None] * 14.0,
                              pbc=np.ones((1, 3), bool))
    assert engine.cell_grid(lay2, 5.0)[0][0, 0] == 0   # only 2 cells per axis


def test_cells_narrower_than_twice_cutoff_are_rejected():
    engine = pkg_mod("engine")
    with pytest.raises(ValueError, match="half the periodic cell"):
        engine.check_cells(np.eye(3) * 9.0, np.ones(3, bool), 5.0)
    engine.check_cells(np.eye(3) * 10.5, np.ones(3, bool), 5.0)


class _CalcStub:
    """Just what the harness constructor reads: a calculator with a model."""
    model = None
    device = "cpu"


def test_harness_state_before_any_run():
    """nve_harness tests: test_initialization, test_get_trajectory_array_invalid_key,
    test_get_trajectory_array_before_simulation; save before a run raises too."""
    from atoms_standin import Atoms
    h = pkg().NVEMDHarness(Atoms([8, 1, 1], np.eye(3)), _CalcStub(), temperature=300.0,
                           timestep=0.5)
    assert h.n_atoms == 3 and not h.is_complete
    assert h.temperature == 300.0 and h.timestep_fs == 0.5
    assert len(h.trajectory_data["time"]) == 0
    with pytest.raises(KeyError, match="Unknown trajectory key"):
        h.get_trajectory_array("nope")
    with pytest.raises(RuntimeError, match="No trajectory data collected"):
        h.get_trajectory_array("positions")
    with pytest.warns(UserWarning):
        with pytest.raises(RuntimeError, match="No trajectory data to save"):
            h.save_trajectory("unused.xyz", format="xyz")
    try:
        import ase  # noqa: F401
    except ImportError:      # a format only ASE writes fails before any step runs
        h2 = pkg().NVEMDHarness(Atoms([8, 1, 1], np.eye(3)), _CalcStub(),
                                trajectory_file="run.traj")
        with pytest.raises(ValueError, match="needs ASE"):
            h2.run_simulation(10)


def test_velocity_initialization():
    """nve_harness test_velocity_initialization: zero velocities before, then
    Maxwell-Boltzmann velocities set on the atoms, with a positive
    temperature, no centre-of-mass momentum, the atoms' own masses, and the
    calculator attached (nve_harness.py:117, 141-171)."""
    from atoms_standin import Atoms
    rng = np.random.default_rng(3)
    atoms = Atoms([8, 1, 1, 6, 1], rng.normal(size=(5, 3)) * 1.2)
    calc = _CalcStub()
    h = pkg().NVEMDHarness(atoms, calc, temperature=300.0, seed=5)
    assert h.atoms.calc is calc
    assert np.allclose(h.atoms.get_velocities(), 0.0)
    h.initialize_velocities()
    v = h.atoms.get_velocities()
    assert v.shape == (5, 3) and np.abs(v).max() > 0
    assert h.atoms.get_temperature() > 0
    p = h.atoms.get_momenta()
    assert np.abs(p.sum(0)).max() < 1e-12
    # heavier hydrogens (deuterium): the momenta follow the atoms' masses
    heavy = atoms.copy()
    m = heavy.get_masses()
    m[heavy.get_atomic_numbers() == 1] = 2.014
    heavy.set_masses(m)
    h2 = pkg().NVEMDHarness(heavy, calc, temperature=300.0, seed=5)
    h2.initialize_velocities(remove_com_translation=False, remove_com_rotation=False)
    h1 = pkg().NVEMDHarness(atoms, calc, temperature=300.0, seed=5)
    h1.initialize_velocities(remove_com_translation=False, remove_com_rotation=False)
    ratio = h2.atoms.get_momenta() / h1.atoms.get_momenta()
    np.testing.assert_allclose(ratio, np.sqrt(m / atoms.get_masses())[:, None] * np.ones(3),
                               rtol=1e-12)
    assert h._masses().tolist() == atoms.get_masses().tolist()


def test_extxyz_writer_reads_back(tmp_path):
    """The extended-XYZ writer, parsed back by hand: species, positions,
    velocities, forces and per-frame info for two frames of a periodic cell."""
    md = pkg_mod("md")
    rng = np.random.default_rng(0)
    z = np.array([8, 1, 1, 79])
    x, v, f = (rng.normal(size=(2, 4, 3)) for _ in range(3))
    info = {"time": [0.0, 0.0005], "total_energy": [-1.25, -1.5]}
    cell = np.diag([10.0, 11.0, 12.0])
    path = tmp_path / "t.xyz"
    md.write_extxyz(path, z, x, v, f, info, cell, np.array([True, True, False]))
    lines = path.read_text().splitlines()
    assert len(lines) == 2 * (2 + 4)
    for t in range(2):
        head = lines[t * 6 + 1]
        assert int(lines[t * 6]) == 4
        assert 'pbc="T T F"' in head and "Lattice=" in head
        assert f"total_energy={info['total_energy'][t]:.12g}" in head
        rows = [ln.split() for ln in lines[t * 6 + 2:t * 6 + 6]]
        assert [r[0] for r in rows] == ["O", "H", "H", "Au"]
        vals = np.array([[float(c) for c in r[1:]] for r in rows])
        np.testing.assert_allclose(vals[:, 0:3], x[t], atol=1e-9)
        np.testing.assert_allclose(vals[:, 3:6], v[t], atol=1e-9)
        np.testing.assert_allclose(vals[:, 6:9], f[t], atol=1e-9)
