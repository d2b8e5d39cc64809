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
