"""Frame datasets in the reference's ragged layout (npz container here; h5py
is not installed) and the distillation collate, on the CPU."""

import os

import numpy as np
import pytest
import torch

from atoms_standin import Atoms
from conftest import pkg_mod


def _structures(rng):
    mol = Atoms(rng.choice([1, 6, 8], size=5), rng.normal(size=(5, 3)) * 1.5)
    box = Atoms([8, 1, 1] * 4, rng.uniform(0, 7.0, size=(12, 3)), cell=np.diag([7.0, 7.5, 8.0]),
                pbc=True)
    slab = Atoms([6, 6, 1], rng.normal(size=(3, 3)), cell=np.diag([9.0, 9.0, 0.0]),
                 pbc=[True, True, False])
    return [mol, box, slab]


def _labels(structs, rng):
    E = rng.normal(size=len(structs))
    F = [rng.normal(size=(len(s), 3)).astype(np.float32) for s in structs]
    S = [None, rng.normal(size=6), None]
    return E, F, S


def test_round_trip_reference_layout(tmp_path):
    ds = pkg_mod("dataset")
    rng = np.random.default_rng(0)
    structs = _structures(rng)
    E, F, S = _labels(structs, rng)
    path = str(tmp_path / "frames.npz")
    with ds.FrameDatasetWriter(path) as w:
        w.add_batch(structs, E, F, S)
    with np.load(path, allow_pickle=False) as z:
        keys = set(z.files)
        # the reference's dataset paths and dtypes
        assert {"structures/atomic_numbers", "structures/atomic_numbers_splits",
                "structures/positions", "structures/cells", "structures/pbc", "labels/energy",
                "labels/forces", "labels/forces_splits", "labels/stress", "labels/stress_mask",
                "labels/structure_indices", "metadata"} <= keys
        assert z["structures/atomic_numbers"].dtype == np.int64
        assert z["structures/positions"].dtype == np.float64
        assert z["labels/forces"].dtype == np.float32
        np.testing.assert_array_equal(z["structures/atomic_numbers_splits"], [0, 5, 17, 20])
        # a structure is read back the way the reference's scripts slice it
        sp, fsp = z["structures/atomic_numbers_splits"], z["labels/forces_splits"]
        f1 = z["labels/forces"][fsp[1]:fsp[2]].reshape(-1, 3)
        np.testing.assert_array_equal(f1, F[1])
        np.testing.assert_array_equal(z["structures/positions"][sp[1]:sp[2]], structs[1].positions)
    r = ds.FrameDataset(path)
    assert len(r) == 3 and r.metadata["n_structures"] == 3 and r.metadata["n_atoms"] == 20
    for i, s in enumerate(structs):
        it = r[i]
        np.testing.assert_array_equal(it["atomic_numbers"], s.numbers)
        np.testing.assert_array_equal(it["positions"], s.positions)
        np.testing.assert_array_equal(it["cell"], s.cell)
        np.testing.assert_array_equal(it["pbc"], s.pbc)
        assert it["energy"] == E[i]
        np.testing.assert_array_equal(it["forces"], F[i])
        if S[i] is None:
            assert it["stress"] is None
        else:
            np.testing.assert_array_equal(it["stress"], S[i])
    with pytest.raises(IndexError):
        r[3]


def test_append_mode_and_metadata(tmp_path):
    ds = pkg_mod("dataset")
    rng = np.random.default_rng(1)
    structs = _structures(rng)
    E, F, S = _labels(structs, rng)
    path = str(tmp_path / "a.npz")
    with ds.FrameDatasetWriter(path) as w:
        w.add_batch(structs[:2], E[:2], F[:2], S[:2])
        w.finalize(extra_metadata={"teacher_model": "full", "config": {"seed": 1}})
    with ds.FrameDatasetWriter(path, mode="a") as w:
        w.add_structure(structs[2], E[2], F[2])
    r = ds.FrameDataset(path)
    assert len(r) == 3
    assert r.metadata["teacher_model"] == "full" and r.metadata["config"] == {"seed": 1}
    np.testing.assert_array_equal(r[2]["forces"], F[2])
    np.testing.assert_array_equal(r.arrays["labels/structure_indices"], [0, 1, 2])


def test_validation_rejects_bad_labels(tmp_path):
    ds = pkg_mod("dataset")
    rng = np.random.default_rng(2)
    s = _structures(rng)[0]
    w = ds.FrameDatasetWriter(str(tmp_path / "v.npz"))
    with pytest.raises(ValueError, match="non-finite energy"):
        w.add_structure(s, float("nan"), np.zeros((5, 3)))
    with pytest.raises(ValueError, match="forces for 4 atoms"):
        w.add_structure(s, 0.0, np.zeros((4, 3)))
    f = np.zeros((5, 3))
    f[2, 1] = np.inf
    with pytest.raises(ValueError, match="non-finite forces"):
        w.add_structure(s, 0.0, f)
    with pytest.raises(ValueError, match="6 Voigt"):
        w.add_structure(s, 0.0, np.zeros((5, 3)), stress=np.zeros(3))
    with pytest.raises(ValueError, match="backend"):
        ds.FrameDatasetWriter(str(tmp_path / "x.bin"), backend="csv")
    assert len(w) == 0


def test_collate_matches_reference_batch_keys(tmp_path):
    ds = pkg_mod("dataset")
    rng = np.random.default_rng(3)
    structs = _structures(rng)
    E, F, S = _labels(structs, rng)
    path = str(tmp_path / "c.npz")
    with ds.FrameDatasetWriter(path) as w:
        w.add_batch(structs, E, F, S)
    r = ds.FrameDataset(path)
    b = ds.distillation_collate([r[2], r[0]])
    assert b["atomic_numbers"].dtype == torch.int64 and b["positions"].dtype == torch.float32
    assert b["cell"].shape == (2, 3, 3) and b["pbc"].dtype == torch.bool
    assert b["batch"].tolist() == [0] * 3 + [1] * 5
    assert b["n_atoms"].tolist() == [3, 5] and b["atom_splits"].tolist() == [0, 3, 8]
    assert b["energy"].dtype == torch.float64 and b["energy"].tolist() == [E[2], E[0]]
    assert torch.equal(b["forces"], torch.tensor(np.concatenate([F[2], F[0]])))
    assert b["pbc"][0].tolist() == [True, True, False] and not b["pbc"][1].any()
    with pytest.raises(ValueError):
        ds.distillation_collate([])


def test_save_frames_from_md_frameset(tmp_path):
    ds = pkg_mod("dataset")
    mts = pkg_mod("mts")
    rng = np.random.default_rng(4)
    sizes = np.array([3, 4])
    fs = mts.FrameSet(rng.choice([1, 6], size=7), sizes, torch.device("cpu"))
    for _ in range(2):
        fs.add(torch.tensor(rng.normal(size=(7, 3)), dtype=torch.float32),
               torch.tensor(rng.normal(size=(7, 3)), dtype=torch.float32),
               torch.tensor(rng.normal(size=2), dtype=torch.float64))
    path = str(tmp_path / "f.npz")
    assert ds.save_frames(fs, path) == 4
    r = ds.FrameDataset(path)
    assert len(r) == 4 and r.metadata["source"] == "full-model MD frames"
    it = r[3]          # frame 1, molecule 1
    np.testing.assert_array_equal(it["atomic_numbers"], fs.numbers[3:7].numpy())
    np.testing.assert_array_equal(it["positions"], fs.pos[1][3:7].double().numpy())
    np.testing.assert_array_equal(it["forces"], fs.forces[1][3:7].numpy())
    assert it["energy"] == float(fs.energy[1][1])
    assert not it["pbc"].any()
