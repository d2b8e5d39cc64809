"""Benchmark: batched PaiNN E+F + Velocity Verlet MD on one MI355X.

Workload (default): 256 synthetic molecules, atom counts uniform in
[10, 100] (H/C/N/O/S 50/30/8/10/2 %, bonded random walk, seed 2), PaiNN
student H128 / 3 blocks / 20 RBF / 5 A cutoff with seed-0 random weights,
float32 with float64 energy accumulation; NVE Velocity Verlet at 0.5 fs from
300 K, every step (neighbour list + energy + forces + integrator) in one
captured HIP graph.  A step is one MD step of all 256 molecules.

Printed JSON line: value = atom-steps/s of the batched path over the timed
steps, plus
  * plain: the per-molecule PyTorch-eager restatement on the same GPU
    (dense neighbour search, autograd forces, host Velocity Verlet with the
    positions/forces transfers of the reference's MD loop), on a sample;
  * accuracy: max |dF| and max |dE|/N against the float64 oracle at the GPU
    trajectory's final positions (a sample of molecules);
  * roofline: the dominant kernel's algorithmic FLOP per launch over its
    event-timed average duration, against the published peak and against the
    peak measured on the device (register-only MFMA probe, device copy);
  * cpu_baseline: the oracle (float32 PyTorch-eager restatement of the
    reference, dense neighbour search, autograd forces, VV) on the host cores.
"""

from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("atom-steps/s (E+F) batched vs plain on 1 GPU; max |ΔF| vs FP64 oracle")
PEAK_FP32_MFMA = 157.3e12
PEAK_HBM = 8.0e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_fields():
    return {"cores": torch.get_num_threads(), "cpu_count": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)), "cpu_model": cpu_model()}


# ---------------------------------------------------------------------------
# plain path: eager PyTorch, one molecule at a time (student_model.py:634-795)
# ---------------------------------------------------------------------------

def _lin(x, sd, name):
    return torch.nn.functional.linear(x, sd[name + ".weight"], sd[name + ".bias"])


def plain_energy(sd, H, L, Z, pos, rc):
    n = pos.shape[0]
    diff = pos.unsqueeze(0) - pos.unsqueeze(1)
    dist = torch.norm(diff, dim=2)
    mask = (dist <= rc) & ~torch.eye(n, dtype=torch.bool, device=pos.device)
    src, dst = torch.where(mask)
    vec = pos[src] - pos[dst]
    d = torch.norm(vec, dim=1)
    unit = vec / (d.unsqueeze(1) + 1e-8)
    rbf = torch.exp(-(1.0 / sd["rbf.widths"] ** 2) * (d.unsqueeze(-1) - sd["rbf.centers"]) ** 2)
    fc = 0.5 * (torch.cos(math.pi * d / rc) + 1.0) * (d < rc).float()
    erbf = rbf * fc.unsqueeze(-1)
    s = sd["embedding.weight"][Z]
    v = torch.zeros(n, 3, H, device=pos.device)
    silu = torch.nn.functional.silu
    for k in range(L):
        pre = f"interactions.{k}.message.rbf_to_scalar"
        w = _lin(silu(_lin(erbf, sd, pre + ".0")), sd, pre + ".2")
        ws, wv1, wv2 = torch.split(w, H, dim=-1)
        s = s + torch.zeros_like(s).index_add_(0, dst, s[src] * ws)
        vm = v[src] * wv1.unsqueeze(1) + unit.unsqueeze(-1) * wv2.unsqueeze(1)
        v = v + torch.zeros_like(v).index_add_(0, dst, vm)
        pre = f"interactions.{k}.update"
        a = _lin(silu(_lin(torch.cat([s, torch.norm(v, dim=1)], -1), sd, pre + ".update_mlp.0")),
                 sd, pre + ".update_mlp.2")
        ds, g1, g2 = torch.split(a, H, dim=-1)
        mixed = torch.einsum("ij,njk->nik", sd[pre + ".mixing_matrix"], v)
        s = s + ds
        v = v * g1.unsqueeze(1) + mixed * g2.unsqueeze(1)
    h = silu(_lin(silu(_lin(s, sd, "energy_head.0")), sd, "energy_head.2"))
    return _lin(h, sd, "energy_head.4").sum()


def plain_md_rate(sd, H, L, mols, masses_of, dt, steps, device, rc=5.0):
    """atom-steps/s of the plain path (ASE-style host VV, one molecule at a time)."""
    state = []
    for m in mols:
        x = m.positions.copy()
        p = np.zeros_like(x)
        state.append([m, x, p, masses_of(m.numbers)[:, None], None])
    Zs = [torch.tensor(m.numbers, device=device) for m in mols]

    def forces(k, x):
        pos = torch.tensor(x, dtype=torch.float32, device=device, requires_grad=True)
        e = plain_energy(sd, H, L, Zs[k], pos, rc)
        f = -torch.autograd.grad(e, pos)[0]
        return f.detach().cpu().numpy().astype(np.float64), float(e.detach().cpu())

    for k, st in enumerate(state):
        st[4] = forces(k, st[1])[0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    atoms = 0
    for _ in range(steps):
        for k, st in enumerate(state):
            m, x, p, mass, F = st
            p += 0.5 * dt * F
            x += dt * p / mass
            F, _ = forces(k, x)
            p += 0.5 * dt * F
            st[4] = F
            atoms += len(m)
    torch.cuda.synchronize()
    return atoms / (time.perf_counter() - t0)


# ---------------------------------------------------------------------------
# CPU baseline: the oracle on the host cores
# ---------------------------------------------------------------------------

def cpu_baseline(sd_cpu, mols, masses_of, dt, steps):
    from oracle import painn_ref as R
    sd32 = {k: v.float() for k, v in sd_cpu.items()}
    t0 = time.perf_counter()
    atoms = 0
    for m in mols:
        x = m.positions.astype(np.float32).astype(np.float64)
        mass = masses_of(m.numbers)[:, None]
        p = np.zeros_like(x)
        _, F = R.energy_and_forces(sd32, m.numbers, x, dtype=torch.float32)
        F = F.numpy().astype(np.float64)
        for _ in range(steps):
            p += 0.5 * dt * F
            x += dt * p / mass
            _, F = R.energy_and_forces(sd32, m.numbers, x, dtype=torch.float32)
            F = F.numpy().astype(np.float64)
            p += 0.5 * dt * F
            atoms += len(m)
    return atoms / (time.perf_counter() - t0)


def accuracy(sd_cpu, md, sample=None):
    """max|dF| (eV/A) and max|dE|/N (eV/atom) vs the float64 oracle at the
    trajectory's current positions, for the first `sample` molecules (all
    when None)."""
    from oracle import painn_ref as R
    lib = md.plan
    x32 = md.pos32.cpu().numpy().astype(np.float64)
    F = md.forces.cpu().numpy()
    E = md.energy.cpu().numpy()
    mp = np.concatenate([[0], np.cumsum(md.sizes)])
    df = de = 0.0
    for b in range(md.B if sample is None else min(sample, md.B)):
        sl = slice(mp[b], mp[b + 1])
        Eref, Fref = R.energy_and_forces(sd_cpu, md.structures[b].numbers, x32[sl])
        df = max(df, float(np.abs(F[sl] - Fref.numpy()).max()))
        de = max(de, abs(float(E[b]) - Eref.item()) / md.sizes[b])
    del lib
    return df, de


def window_edges(model, mols, args, md_timed):
    """Mean directed-edge count over the timed MD steps.  The step loop is
    bit-reproducible, so a second run from the same state, one eager step at
    a time with the edge count read after each, visits the same trajectory;
    returns (mean, whether the replay ended on the timed run's positions)."""
    mdmod = importlib.import_module("mlff-distiller_amd.md")
    md = mdmod.BatchedNVE(model, mols, timestep_fs=0.5, temperature_K=300.0, seed=args.seed,
                          max_steps=args.warmup + args.steps + 1, streams=args.streams,
                          persistent_blocks=md_timed._persistent_blocks)
    md.run(args.warmup, graph=False)
    edges = []
    for _ in range(args.steps):
        md.run(1, graph=False)
        edges.append(int(md.plan.read_counts()[0]))
    same = bool(torch.equal(md.x, md_timed.x) and torch.equal(md.p, md_timed.p))
    return float(np.mean(edges)), same


# ---------------------------------------------------------------------------
# per-kernel timing
# ---------------------------------------------------------------------------

STAGES = ["neighbor", "filter_fwd", "message_fwd", "update_fwd", "readout", "update_bwd",
          "filter_bwd", "message_bwd", "forces"]


def stage_times(md, reps=20):
    """Average duration (s) of every stage per launch, HIP events on the
    stream the stage is launched on."""
    C = importlib.import_module("mlff-distiller_amd._capi")
    lib = C.lib()
    L = md.model.num_interactions
    out = {}
    stream = torch.cuda.current_stream()
    for sid, name in enumerate(STAGES):
        per_layer = sid in (1, 2, 3, 5, 6, 7)
        layers = range(L) if per_layer else [0]
        tot = 0.0
        for l in layers:
            C.check("stage", lib.mlff_launch_stage(md.plan.handle, stream.cuda_stream, sid, l))
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(reps):
                C.check("stage", lib.mlff_launch_stage(md.plan.handle, stream.cuda_stream, sid, l))
            b.record(stream)
            b.synchronize()
            tot += a.elapsed_time(b) / 1e3 / reps
        out[name] = {"avg_s": tot / len(layers), "launches_per_step": len(layers),
                     "per_step_s": tot}
    return out


def measured_peaks(dev, iters=4000, copy_mib=1024):
    """Sustained peaks on this device: dense FP32 (16x16x4) and bf16
    (16x16x32) MFMA from a register-only probe (four waves per SIMD, eight
    independent chains each), and HBM bandwidth from a device-to-device copy
    (bytes read + written per second)."""
    C = importlib.import_module("mlff-distiller_amd._capi")
    lib = C.lib()
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    blocks = 4 * cus
    out = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    res = {}
    for kind, name in ((0, "fp32_mfma_tflops"), (1, "bf16_mfma_tflops")):
        C.check("probe", lib.mlff_probe_mfma(stream.cuda_stream, kind, blocks, iters, C.ptr(out)))
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(3):
            C.check("probe", lib.mlff_probe_mfma(stream.cuda_stream, kind, blocks, iters,
                                                 C.ptr(out)))
        b.record(stream)
        b.synchronize()
        flops = 3 * lib.mlff_probe_mfma_flops(kind, blocks, iters)
        res[name] = flops / (a.elapsed_time(b) / 1e3) / 1e12
    n = copy_mib << 18
    x = torch.empty(n, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    y.copy_(x)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(5):
        y.copy_(x)
    b.record(stream)
    b.synchronize()
    res["hbm_copy_gbps"] = 5 * 2 * 4 * n / (a.elapsed_time(b) / 1e3) / 1e9
    res["method"] = (f"MFMA: {blocks} workgroups x 4 waves, {iters} x 8 independent MFMAs per "
                     f"wave, 3 launches; copy: {copy_mib} MiB device to device, 5 times")
    del x, y
    return res


# ---------------------------------------------------------------------------
# the other configurations (one JSON line each)
# ---------------------------------------------------------------------------

def _common(metric, value, unit, steps, warmup, seconds, dtype, workload, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": seconds / max(steps, 1) * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic (seeded structures, random-init weights)",
           "config": {"workload": workload}}
    out.update(extra)
    return out


def config1(args, pkg, synth, mdmod, dev):
    """One 60-atom molecule, 1000 NVE steps at 0.5 fs (seed 1, weights seed 0)."""
    torch.manual_seed(0)
    model = pkg.StudentForceField().to(dev).eval()
    mol = synth.single_molecule(60, seed=1)
    steps = args.steps if args.steps is not None else 1000
    warmup = args.warmup if args.warmup is not None else 20
    args.warmup = warmup
    md = mdmod.BatchedNVE(model, [mol], seed=1, max_steps=steps + args.warmup + 1,
                          streams=args.streams)
    md.run(args.warmup, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    md.run(steps, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = 60 * steps / wall
    sd_cpu = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    df, de = accuracy(sd_cpu, md, 1)
    s = md.summary()
    sd_dev = {k: v.detach() for k, v in model.state_dict().items()}
    plain = plain_md_rate(sd_dev, 128, 3, [mol], synth.masses_of, 0.5 * synth.FS, 50, dev)
    cpu_steps = args.cpu_steps if args.cpu_steps is not None else 1000
    cpu = cpu_baseline(sd_cpu, [mol], synth.masses_of, 0.5 * synth.FS, cpu_steps)
    return _common("atom-steps/s (E+F), single 60-atom molecule NVE", value, "atom-steps/s",
                   steps, args.warmup, wall, "f32",
                   "1 molecule, 60 atoms (seed 1), PaiNN H128/L3/R20 seed 0, hipGraph NVE dt 0.5 fs",
                   accuracy={"max_abs_dF_eV_per_A": df, "max_abs_dE_per_atom_eV": de},
                   plain={"value": plain, "unit": "atom-steps/s", "speedup": value / plain},
                   nve={"drift_per_atom_eV": float(s["energy_drift_per_atom"][0]),
                        "drift_pct": float(s["energy_drift_pct"][0])},
                   cpu_baseline={"value": cpu, "unit": "atom-steps/s", "kind": "port",
                                 "sample": f"the same molecule, {cpu_steps} MD steps",
                                 **cpu_fields()})


def config3(args, pkg, synth, mdmod, dev):
    """1000 waters (3000 atoms) in a periodic box, 10 ps NVE with drift and stress."""
    from oracle import nl_ref
    from oracle import painn_ref as R
    torch.manual_seed(0)
    model = pkg.StudentForceField().to(dev).eval()
    box = synth.water_box(10, seed=3)
    steps = args.steps if args.steps is not None else 20000
    args.warmup = args.warmup if args.warmup is not None else 20
    md = mdmod.BatchedNVE(model, [box], seed=3, max_steps=steps + args.warmup + 1,
                          streams=args.streams)
    sd = {k: v.detach().cpu() for k, v in model.state_dict().items()}

    def accuracy_at(P32):
        """GPU E/F/virial vs the float64 oracle on the same edge set, and the
        error of the same algorithm run in float32 on the host."""
        rp, cl, sh = nl_ref.neighbor_csr(P32, np.array([0, 3000]), 5.0, box.cell[None],
                                         box.pbc[None])
        ei = np.stack([cl.astype(np.int64), np.repeat(np.arange(3000), np.diff(rp))])
        E, F, W = R.energy_and_forces(sd, box.numbers, P32.astype(np.float64), edge_index=ei,
                                      shifts=sh, cell=box.cell.astype(np.float32).astype(np.float64),
                                      virial=True)
        E32, F32 = R.energy_and_forces({k: v.float() for k, v in sd.items()}, box.numbers,
                                       P32.astype(np.float64), edge_index=ei, shifts=sh,
                                       cell=box.cell, dtype=torch.float32)
        e64, f, vir = model.energy_forces_virial(torch.tensor(box.numbers, device=dev),
                                                 torch.tensor(P32, device=dev),
                                                 torch.tensor(box.cell), torch.tensor(box.pbc))
        return ({"max_abs_dF_eV_per_A": float(np.abs(f.cpu().numpy() - F.numpy()).max()),
                 "max_abs_dE_per_atom_eV": abs(e64.item() - E.item()) / 3000,
                 "max_abs_dvirial_eV": float(np.abs(vir.cpu().numpy()[0] - W.numpy()[0]).max()),
                 "max_abs_F_eV_per_A": float(np.abs(F.numpy()).max()),
                 "fp32_reference_max_abs_dF": float(np.abs(F32.numpy() - F.numpy()).max()),
                 "edges": int(cl.size)}, W, ei, sh)

    acc0 = accuracy_at(md.pos32.cpu().numpy())[0]
    tuning = md.autotune()
    md.run(args.warmup, graph=True)
    torch.cuda.synchronize()
    c = md.plan.read_counts()
    t0 = time.perf_counter()
    md.run(steps, graph=True, steps_per_graph=20)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = 3000 * steps / wall
    s = md.summary()
    # accuracy and virial at the final positions (the box has heated far above
    # room temperature under random-init forces, so forces are large there)
    P32 = md.pos32.cpu().numpy()
    acc1, W, ei, sh = accuracy_at(P32)
    # CPU baseline: float32 oracle E+F on the host (same edge set), a few evaluations
    sd32 = {k: v.float() for k, v in sd.items()}
    t2 = time.perf_counter()
    nrep = 3
    for _ in range(nrep):
        R.energy_and_forces(sd32, box.numbers, P32.astype(np.float64), edge_index=ei, shifts=sh,
                            cell=box.cell, dtype=torch.float32)
    cpu_rate = 3000 * nrep / (time.perf_counter() - t2)
    vol = abs(np.linalg.det(box.cell))
    return _common("atom-steps/s (E+F), periodic water box NVE", value, "atom-steps/s", steps,
                   args.warmup, wall, "f32",
                   "1000 waters (3000 atoms, seed 3) in a 31 A periodic box, PaiNN H128/L3/R20 "
                   "seed 0, cell-free minimum image, hipGraph NVE dt 0.5 fs",
                   edges_start=int(c[0]),
                   autotune={"persistent_blocks": int(tuning["blocks"]),
                             "launch_shapes": list(tuning["tiles"]),
                             "evaluation_us": {str(k): round(v, 1)
                                               for k, v in tuning["timings_us"].items()}},
                   accuracy={"t0": acc0, "final": acc1},
                   nve={"simulated_ps": steps * 0.5e-3,
                        "drift_per_atom_eV": float(s["energy_drift_per_atom"][0]),
                        "drift_rate_eV_per_atom_ps": float(s["drift_rate_eV_per_atom_ps"][0]),
                        "drift_pct": float(s["energy_drift_pct"][0]),
                        "avg_temperature_K": float(s["avg_temperature"][0])},
                   pressure_eV_per_A3=float(-np.trace(W.numpy()[0]) / (3 * vol)),
                   cpu_baseline={"value": cpu_rate, "unit": "atom-evals/s (E+F only)",
                                 "kind": "port",
                                 "sample": f"{nrep} float32 E+F evaluations of the final frame",
                                 **cpu_fields()})


def config4(args, pkg, synth, dev):
    """Force-matching training step: 64 molecules of 20-60 atoms (seed 4),
    labels from the float64 model with seed-1 weights, loss E-MSE + 100 F-MSE."""
    from oracle import painn_ref as R
    train = importlib.import_module("mlff-distiller_amd.train")
    torch.manual_seed(1)
    teacher = pkg.StudentForceField()
    torch.manual_seed(0)
    student = pkg.StudentForceField().to(dev)
    mols = synth.molecule_set(64, 20, 60, seed=4)
    Z, P, mp = synth.flatten(mols)
    sizes = np.diff(mp)
    b = np.repeat(np.arange(64), sizes)
    P32 = P.astype(np.float32)
    sdt = {k: v.double() for k, v in teacher.state_dict().items()}
    El, Fl = R.energy_and_forces(sdt, Z, P32.astype(np.float64), batch=b)
    Zt = torch.tensor(Z, device=dev)
    Pt = torch.tensor(P32, device=dev)
    Elt, Flt = El.float().to(dev), Fl.float().to(dev)
    tr = train.ForceMatchingTrainer(student, lr=1e-4)
    # gradient check against float64 autograd on the starting weights
    loss, _, _, grads = tr.loss_and_grads(Zt, Pt, sizes, Elt, Flt)
    sd = {k: v.detach().cpu().double().requires_grad_(True) for k, v in student.state_dict().items()}
    pos = torch.tensor(P32, dtype=torch.float64, requires_grad=True)
    E = R.energy(sd, torch.tensor(Z), pos, 5.0, batch=torch.tensor(b), n_structures=64)
    F = -torch.autograd.grad(E.sum(), pos, create_graph=True)[0]
    Lref = ((E - El) ** 2).mean() + 100.0 * ((F - Fl) ** 2).mean()
    names = [k for k in sd if not k.startswith("rbf.")]
    gref = torch.autograd.grad(Lref, [sd[k] for k in names], allow_unused=True)
    rel = max(((grads[k].double().cpu() - g).norm() / (g.norm() + 1e-30)).item()
              for k, g in zip(names, gref) if g is not None)
    steps = args.steps if args.steps is not None else 50
    for _ in range(args.warmup if args.warmup is not None else 3):
        tr.step(Zt, Pt, sizes, Elt, Flt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tr.step(Zt, Pt, sizes, Elt, Flt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n = len(Z)
    # CPU baseline: float32 autograd double backward of the oracle, one step
    sd32 = {k: v.detach().cpu().float().requires_grad_(True)
            for k, v in student.state_dict().items()}
    t1 = time.perf_counter()
    pos32 = torch.tensor(P32, requires_grad=True)
    E = R.energy(sd32, torch.tensor(Z), pos32, 5.0, batch=torch.tensor(b), n_structures=64)
    F = -torch.autograd.grad(E.sum(), pos32, create_graph=True)[0]
    L = ((E - El.float()) ** 2).mean() + 100.0 * ((F - Fl.float()) ** 2).mean()
    L.backward()
    cpu_s = time.perf_counter() - t1
    return _common("force-matching training atoms/s (E+F loss, exact weight gradient)",
                   n * steps / wall, "atoms/s", steps, 3, wall, "f32",
                   "64 molecules N~U{20..60} (seed 4), PaiNN H128/L3/R20 student seed 0, "
                   "labels from the float64 seed-1 model, loss MSE(E) + 100 MSE(F), AdamW",
                   n_atoms=n, train_steps_per_s=steps / wall,
                   accuracy={"max_rel_grad_err_vs_fp64_autograd": rel,
                             "loss_fp32": float(loss), "loss_fp64": Lref.item()},
                   cpu_baseline={"value": n / cpu_s, "unit": "atoms/s", "kind": "port",
                                 "sample": "one float32 autograd double-backward step of the "
                                           "oracle on the same batch",
                                 **cpu_fields()})


def config5(args, pkg, synth, mdmod, dev):
    """Compact student trained on full-model MD frames, then MTS vs full MD."""
    mts = importlib.import_module("mlff-distiller_amd.mts")
    torch.manual_seed(0)
    full = pkg.StudentForceField().to(dev).eval()
    t0 = time.perf_counter()
    frames = mts.md_frames(full, synth.molecule_set(256, 10, 100, seed=2), n_frames=args.frames,
                           stride=args.frame_stride, seed=2)
    t_frames = time.perf_counter() - t0
    torch.manual_seed(0)
    student = pkg.StudentForceField(hidden_dim=32, num_interactions=2, num_rbf=16).to(dev)
    t0 = time.perf_counter()
    hist = mts.train_student(student, frames, steps=args.train_steps, batch_molecules=64,
                             lr=2e-3, seed=0)
    t_train = time.perf_counter() - t0
    student.eval()
    mols = synth.molecule_set(512, 10, 100, seed=5)
    n_atoms = sum(len(m) for m in mols)
    n_in = args.mts_inner
    outer = args.mts_outer * 4 // n_in          # the same simulated time for every n_inner
    steps = n_in * outer
    ref = mdmod.BatchedNVE(full, mols, seed=5, max_steps=steps + 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref.run(steps, graph=True, steps_per_graph=20)
    torch.cuda.synchronize()
    t_full = time.perf_counter() - t0
    m = mts.MTSNVE(full, student, mols, seed=5, threshold=args.mts_threshold,
                   n_inner=n_in, max_outer=outer + 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.run(outer, chunk=10)
    torch.cuda.synchronize()
    t_mts = time.perf_counter() - t0
    sf, sm = ref.summary(), m.summary()
    full_rate = n_atoms * steps / t_full
    mts_rate = n_atoms * steps / t_mts
    return _common("MTS atom-steps/s (0.5 fs equivalent) vs full-model MD", mts_rate,
                   "atom-steps/s", steps, 0, t_mts, "f32",
                   f"512 molecules N~U{{10..100}} (seed 5), {steps * 0.5e-3:.2f} ps NVE; inner "
                   f"compact student H32/L2/R16 at 0.5 fs, outer full H128/L3/R20 at "
                   f"{0.5 * n_in:g} fs",
                   n_atoms=n_atoms,
                   full_md={"value": full_rate, "unit": "atom-steps/s",
                            "max_abs_drift_eV_per_atom": float(np.abs(sf["energy_drift_per_atom"]).max()),
                            "median_abs_drift_eV_per_atom": float(np.median(np.abs(sf["energy_drift_per_atom"])))},
                   mts={"speedup_vs_full": mts_rate / full_rate,
                        "max_abs_drift_eV_per_atom": float(np.abs(sm["energy_drift_per_atom"]).max()),
                        "median_abs_drift_eV_per_atom": float(np.median(np.abs(sm["energy_drift_per_atom"]))),
                        "fallback_structure_steps": sm["fallback_structure_steps"],
                        "replayed_chunks": sm["replays"], "threshold_eV_per_A": args.mts_threshold},
                   student={"frames": frames.n_frames, "frame_stride_steps": args.frame_stride,
                            "frame_seconds": t_frames,
                            "train_steps": args.train_steps, "train_seconds": t_train,
                            "loss_first": hist[0], "loss_last_mean10": float(np.mean(hist[-10:]))})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--frames", type=int, default=80)
    ap.add_argument("--frame-stride", type=int, default=250,
                    help="MD steps between training frames (config 5; 80 x 250 steps = 10 ps)")
    ap.add_argument("--train-steps", type=int, default=5000)
    ap.add_argument("--mts-outer", type=int, default=5000)
    ap.add_argument("--mts-threshold", type=float, default=5.0)
    ap.add_argument("--mts-inner", type=int, default=4,
                    help="inner 0.5 fs steps per outer step (config 5)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 20 for the batched workload, the workload's "
                         "own length for the others)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 5)")
    ap.add_argument("--molecules", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--steps-per-graph", type=int, default=10)
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="MD steps per molecule of the CPU baseline (default 10; 1000 for the "
                         "single molecule)")
    ap.add_argument("--accuracy-sample", type=int, default=0,
                    help="molecules checked against the float64 oracle (0: all)")
    ap.add_argument("--skip-baselines", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--stage-reps", type=int, default=20)
    ap.add_argument("--blocks", type=int, default=0,
                    help="persistent grid of the pair kernels (0: autotune)")
    ap.add_argument("--streams", type=int, default=2, choices=[1, 2],
                    help="2: pair filter kernels on a second stream beside the atom-side kernels")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB copied by a plain device copy at the end (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); every box here has one "
              "GPU and no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("mlff-distiller_amd")
    synth = importlib.import_module("mlff-distiller_amd.synth")
    mdmod = importlib.import_module("mlff-distiller_amd.md")
    dev = torch.device("cuda:0")
    if args.config != 2:
        fn = {1: config1, 3: config3, 5: config5}.get(args.config)
        out = config4(args, pkg, synth, dev) if args.config == 4 else fn(args, pkg, synth, mdmod,
                                                                          dev)
        print(json.dumps(out), flush=True)
        return

    args.steps = args.steps if args.steps is not None else 20
    args.warmup = args.warmup if args.warmup is not None else 5
    args.cpu_steps = args.cpu_steps if args.cpu_steps is not None else 10
    torch.manual_seed(0)
    model = pkg.StudentForceField(hidden_dim=128, num_interactions=3, num_rbf=20, cutoff=5.0)
    model = model.to(dev)
    H, L, R = 128, 3, 20
    mols = synth.molecule_set(args.molecules, 10, 100, args.seed)
    n_atoms = sum(len(m) for m in mols)
    log(f"{len(mols)} molecules, {n_atoms} atoms")
    total_steps = args.warmup + args.steps
    md = mdmod.BatchedNVE(model, mols, timestep_fs=0.5, temperature_K=300.0, seed=args.seed,
                          max_steps=total_steps + 1, streams=args.streams,
                          persistent_blocks=args.blocks)
    tuning = md.autotune() if not (args.no_autotune or args.blocks) else None
    if tuning:
        log(f"autotune: persistent grid {tuning['blocks']} workgroups "
            f"{ {k: round(v, 1) for k, v in tuning['timings_us'].items()} }")
        log(f"autotune: launch shapes {tuning['tiles']} "
            f"{ {k: round(v, 1) for k, v in tuning['shape_timings_us'].items()} }")
    # E+F alone on the starting (compact) structures, and per-kernel timings there
    c0 = md.plan.read_counts()
    E0, P0 = int(c0[0]), int(c0[1])
    e_tmp = torch.empty(md.B, dtype=torch.float64, device=dev)
    f_tmp = torch.empty(n_atoms, 3, dtype=torch.float32, device=dev)
    autotune = importlib.import_module("mlff-distiller_amd.autotune")
    ef_s = autotune.time_evaluation(md.plan, md.pos32, e_tmp, f_tmp, reps=20)
    log(f"static E+F at t=0: {ef_s * 1e3:.3f} ms ({n_atoms / ef_s:.4g} atom-evals/s), "
        f"E={E0}, P={P0}, {args.streams} stream(s)")
    ef_other = None
    if args.streams == 2:
        md.plan.set_streams(1)
        ef_other = autotune.time_evaluation(md.plan, md.pos32, e_tmp, f_tmp, reps=20)
        md.plan.set_streams(2)
        log(f"  same evaluation on one stream: {ef_other * 1e3:.3f} ms")
    st = stage_times(md, reps=args.stage_reps)
    for k, v in st.items():
        log(f"  {k:12s} {v['avg_s'] * 1e6:9.1f} us/launch x{v['launches_per_step']}")
    log("warmup")
    md.run(args.warmup, graph=True, steps_per_graph=args.steps_per_graph)
    torch.cuda.synchronize()
    cw = md.plan.read_counts()
    E_start = int(cw[0])
    log("timed steps")
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(md.stream)
    md.run(args.steps, graph=True, steps_per_graph=args.steps_per_graph)
    stop.record(md.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = start.elapsed_time(stop) / 1e3
    counts = md.plan.read_counts()
    E_edges, P_pairs = int(counts[0]), int(counts[1])
    elapsed = max(wall, gpu_s)
    value = n_atoms * args.steps / elapsed
    log(f"batched: {value:.4g} atom-steps/s ({elapsed / args.steps * 1e3:.3f} ms/step, "
        f"E={E_edges}, P={P_pairs})")

    summ = md.summary()
    edges_mean, replay_same = window_edges(model, mols, args, md)
    log(f"edges over the timed window: mean {edges_mean:.0f} (deterministic replay, "
        f"bit-identical trajectory: {replay_same})")
    sd_cpu = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    acc_n = args.accuracy_sample or len(mols)
    df, de = accuracy(sd_cpu, md, sample=acc_n)
    log(f"accuracy vs FP64 oracle ({acc_n} molecules): max|dF|={df:.3e} eV/A, "
        f"max|dE|/N={de:.3e} eV/atom")

    flops = {
        "filter_fwd": P0 * 2 * (R * H + 3 * H * H),
        "filter_bwd": P0 * 2 * (3 * H * H + H * R),
        "update_fwd": n_atoms * 2 * (2 * H * H + 3 * H * H),
        "update_bwd": n_atoms * 2 * (3 * H * H + 2 * H * H),
    }
    dom = max(st, key=lambda k: st[k]["per_step_s"])
    if dom in flops:
        ach = flops[dom] / st[dom]["avg_s"]
        roof = {"kernel": dom, "bound": "mfma", "achieved": ach / 1e12, "peak": PEAK_FP32_MFMA / 1e12,
                "unit": "TFLOP/s", "frac": ach / PEAK_FP32_MFMA, "traffic": None,
                "algorithmic_per_launch": flops[dom], "avg_launch_us": st[dom]["avg_s"] * 1e6}
    else:
        nbytes = (n_atoms * 4 * H * 4 * 2 + P0 * 3 * H * 4 + E0 * 8 + P0 * 16)
        ach = nbytes / st[dom]["avg_s"]
        roof = {"kernel": dom, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": ach / PEAK_HBM, "traffic": None,
                "algorithmic_per_launch": nbytes, "avg_launch_us": st[dom]["avg_s"] * 1e6}

    peaks = measured_peaks(dev)
    log(f"measured peaks: {peaks}")
    if roof["bound"] == "mfma":
        roof["peak_measured"] = peaks["fp32_mfma_tflops"]
    else:
        roof["peak_measured"] = peaks["hbm_copy_gbps"]
    roof["frac_of_measured"] = roof["achieved"] / roof["peak_measured"]

    traffic_file = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            tr = json.load(fh)
        if tr.get("kernel") == dom and abs(tr.get("pairs_t0", -1) - P0) <= 0.01 * P0:
            roof["traffic"] = tr["traffic_bytes"]
            roof["traffic_source"] = "profiles/pmc_traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"

    if args.copy_probe:
        nb = args.copy_probe * (1 << 20)
        src = torch.empty(nb, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {nb} bytes read + {nb} bytes written")

    plain = None
    cpu = None
    if not args.skip_baselines:
        sd_dev = {k: v.detach() for k, v in model.state_dict().items()}
        sample = mols[:args.plain_sample]
        log(f"plain eager GPU path on {len(sample)} molecules x {args.plain_steps} steps")
        plain_rate = plain_md_rate(sd_dev, H, L, sample, synth.masses_of, 0.5 * synth.FS,
                                   args.plain_steps, dev)
        plain = {"value": plain_rate, "unit": "atom-steps/s",
                 "sample": f"{len(sample)} molecules, {args.plain_steps} MD steps each",
                 "speedup": value / plain_rate}
        log(f"plain: {plain_rate:.4g} atom-steps/s -> batched/plain = {value / plain_rate:.1f}x")
        csample = mols[:args.cpu_sample]
        log(f"cpu baseline on {len(csample)} molecules x {args.cpu_steps} steps")
        cpu_rate = cpu_baseline(sd_cpu, csample, synth.masses_of, 0.5 * synth.FS, args.cpu_steps)
        cpu = {"value": cpu_rate, "unit": "atom-steps/s", "kind": "port",
               "sample": f"{len(csample)} molecules of the workload, {args.cpu_steps} MD "
                         "steps each, one molecule at a time, float32 PyTorch-eager "
                         "restatement of the reference (dense neighbour search, autograd "
                         "forces, Velocity Verlet)",
               **cpu_fields()}
        log(f"cpu: {cpu_rate:.4g} atom-steps/s on {torch.get_num_threads()} threads")

    drift = summ["energy_drift_per_atom"]
    out = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded molecules, seed-0 random-init weights)",
        "config": {"workload": "256 molecules N~U{10..100} seed 2, PaiNN H128/L3/R20 rc 5 A, "
                               "hipGraph NVE Velocity Verlet dt 0.5 fs",
                   "n_atoms": n_atoms, "steps_per_graph": args.steps_per_graph,
                   "edges_t0": E0, "pairs_t0": P0, "edges_window_start": E_start,
                   "edges_window_end": E_edges, "edges_window_mean": edges_mean,
                   "note": "random-init forces are not restoring: molecules expand during "
                           "the run and the edge count falls over the window; the mean "
                           "comes from a deterministic replay of the same steps"},
        "static_ef_t0": {"ms": ef_s * 1e3, "atom_evals_per_s": n_atoms / ef_s,
                         "streams": args.streams,
                         "one_stream_ms": ef_other * 1e3 if ef_other else None},
        "autotune": ({"persistent_blocks": tuning["blocks"],
                      "launch_shapes": dict(zip(("filter_fwd_pairs", "filter_bwd_pairs",
                                                 "message_fwd_unroll", "message_bwd_unroll",
                                                 "filter_bwd_variant"), tuning["tiles"])),
                      "evaluation_us": {str(k): round(v, 1)
                                          for k, v in tuning["timings_us"].items()},
                      "shape_evaluation_us": {k: round(v, 1)
                                              for k, v in tuning["shape_timings_us"].items()}}
                     if tuning else None),
        "accuracy": {"max_abs_dF_eV_per_A": df, "max_abs_dE_per_atom_eV": de,
                     "sample_molecules": acc_n},
        "replay_bit_identical": replay_same,
        "plain": plain,
        "nve": {"max_abs_drift_eV_per_atom": float(np.abs(drift).max()),
                "steps": int(summ["n_steps"])},
        "stages_us": {k: round(v["avg_s"] * 1e6, 2) for k, v in st.items()},
        "roofline": roof,
        "peaks_measured": peaks,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
