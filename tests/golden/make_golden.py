"""Write the golden vectors for the PaiNN student from the reference itself.

Run in the build container only (the reference tree is not present
elsewhere):  python tests/golden/make_golden.py

It loads /root/reference/src/mlff_distiller/models/student_model.py by file
path (the module needs only torch and numpy), evaluates it, and stores
inputs and outputs as small fixtures next to this script:

* ``tiny_model.npz``   -- a tiny configuration (hidden 8, 2 blocks, 4 RBF,
  max_z 16) with every parameter (biases randomised so bias conventions are
  visible), three molecules, their reference edge lists, float64 energies
  and autograd forces, single and batched;
* ``full_seed0.npz``   -- the seed-0 initialisation of the full (H128/L3/R20,
  max_z 118) and compact (H32/L2/R16) students as per-tensor checksums, and
  the full student's float64 energy/forces on one 12-atom molecule;
* ``nl_fp32.npz``      -- float32 positions with pairs placed within a few
  ulp of the 5 A cutoff and the reference's radius_graph_native edge list.

These fix the conventions no invariant can see: the split order of the
filter and update outputs, the einsum orientation of the mixing matrix, the
edge-vector sign and the +1e-8 in its normalisation, the missing vector
residual in the update block, and the order in which the constructor
consumes the random stream.
"""

import importlib.util
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/src/mlff_distiller/models/student_model.py"


def load_reference():
    spec = importlib.util.spec_from_file_location("ref_student_model", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def molecule(n, seed, spread=1.6):
    g = np.random.Generator(np.random.PCG64(seed))
    z = g.choice(np.array([1, 6, 7, 8, 16]), size=n)
    pos = np.zeros((n, 3))
    for k in range(1, n):
        while True:
            a = g.integers(0, k)
            d = g.normal(size=3)
            cand = pos[a] + d / np.linalg.norm(d) * g.uniform(1.0, spread)
            if np.min(np.linalg.norm(pos[:k] - cand, axis=1)) > 0.85:
                pos[k] = cand
                break
    return z.astype(np.int64), pos


def ef(model, z, pos, batch=None):
    p = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
    zt = torch.tensor(z, dtype=torch.long)
    bt = None if batch is None else torch.tensor(batch, dtype=torch.long)
    e = model(zt, p, batch=bt)
    g = torch.autograd.grad(e.sum(), p)[0]
    return e.detach().numpy(), (-g).numpy()


def checksums(model):
    out = {}
    for k, t in model.state_dict().items():
        t = t.detach().double().flatten()
        out[k] = np.array([t.sum().item(), (t * t).sum().item(),
                           *(t[:4].tolist() + [0.0] * (4 - min(4, t.numel())))])
    return out


def main():
    ref = load_reference()
    torch.set_default_dtype(torch.float32)

    # ---- tiny configuration -------------------------------------------------
    torch.manual_seed(1234)
    tiny = ref.StudentForceField(hidden_dim=8, num_interactions=2, num_rbf=4, cutoff=3.0,
                                 max_z=16, use_torch_cluster=False)
    with torch.no_grad():
        for name, prm in tiny.named_parameters():
            if name.endswith(".bias"):
                prm.copy_(torch.randn_like(prm) * 0.3)
    tiny = tiny.double()
    mols = [molecule(5, 11), molecule(9, 12), molecule(12, 13, spread=1.9)]
    fx = {}
    for k, v in tiny.state_dict().items():
        fx["param/" + k] = v.numpy()
    for i, (z, pos) in enumerate(mols):
        e, f = ef(tiny, z, pos)
        ei = ref.radius_graph_native(torch.tensor(pos), 3.0).numpy()
        fx[f"mol{i}/numbers"] = z
        fx[f"mol{i}/positions"] = pos
        fx[f"mol{i}/energy"] = np.asarray(e)
        fx[f"mol{i}/forces"] = f
        fx[f"mol{i}/edge_index"] = ei
    zb = np.concatenate([m[0] for m in mols])
    pb = np.concatenate([m[1] + np.array([20.0 * i, 0, 0]) for i, m in enumerate(mols)])
    bb = np.concatenate([np.full(len(m[0]), i) for i, m in enumerate(mols)])
    e, f = ef(tiny, zb, pb, bb)
    fx["batch/numbers"], fx["batch/positions"], fx["batch/batch"] = zb, pb, bb
    fx["batch/energy"], fx["batch/forces"] = e, f
    fx["hyper"] = np.array([8, 2, 4, 16], dtype=np.int64)
    fx["cutoff"] = np.array(3.0)
    np.savez_compressed(os.path.join(HERE, "tiny_model.npz"), **fx)

    # ---- seed-0 full and compact students --------------------------------------
    fx = {}
    torch.manual_seed(0)
    full = ref.StudentForceField(hidden_dim=128, num_interactions=3, num_rbf=20, cutoff=5.0,
                                 max_z=118, use_torch_cluster=False)
    for k, v in checksums(full).items():
        fx["full/" + k] = v
    fx["full/num_parameters"] = np.array(full.num_parameters())
    z, pos = molecule(12, 21)
    e, f = ef(full.double(), z, pos)
    fx["full_eval/numbers"], fx["full_eval/positions"] = z, pos
    fx["full_eval/energy"], fx["full_eval/forces"] = np.asarray(e), f
    torch.manual_seed(0)
    compact = ref.StudentForceField(hidden_dim=32, num_interactions=2, num_rbf=16, cutoff=5.0,
                                    max_z=118, use_torch_cluster=False)
    for k, v in checksums(compact).items():
        fx["compact/" + k] = v
    fx["compact/num_parameters"] = np.array(compact.num_parameters())
    np.savez_compressed(os.path.join(HERE, "full_seed0.npz"), **fx)

    # ---- float32 neighbour list with near-cutoff pairs ---------------------------
    g = np.random.Generator(np.random.PCG64(7))
    pts = (g.uniform(0, 12, size=(160, 3))).astype(np.float32)
    # plant pairs at 5 A +- a few ulp along random directions
    planted = []
    for k in range(40):
        a = pts[k]
        d = g.normal(size=3)
        d /= np.linalg.norm(d)
        base = (a.astype(np.float64) + 5.0 * d).astype(np.float32)
        for t in range(-3, 4):
            cand = base.copy()
            cand[0] = np.nextafter(cand[0], np.float32(np.inf) if t > 0 else np.float32(-np.inf)) \
                if t != 0 else cand[0]
            for _ in range(abs(t) - 1):
                cand[0] = np.nextafter(cand[0], np.float32(np.inf) if t > 0 else np.float32(-np.inf))
            planted.append(cand)
    pos32 = np.concatenate([pts, np.array(planted, dtype=np.float32)])
    batch = np.zeros(len(pos32), dtype=np.int64)
    batch[len(pos32) // 2:] = 1
    ei = ref.radius_graph_native(torch.tensor(pos32), 5.0).numpy()
    eib = ref.radius_graph_native(torch.tensor(pos32), 5.0, torch.tensor(batch)).numpy()
    np.savez_compressed(os.path.join(HERE, "nl_fp32.npz"), positions=pos32, edge_index=ei,
                        batch=batch, edge_index_batched=eib)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    sys.exit(main())
