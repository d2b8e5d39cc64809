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
    event-timed average duration;
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


def accuracy(sd_cpu, md, sample):
    """max|dF| (eV/A) and max|dE|/N (eV/atom) vs the float64 oracle at the
    trajectory's current positions, for the first `sample` molecules."""
    from oracle import painn_ref as R
    lib = md.plan
    x32 = md.pos32.cpu().numpy().astype(np.float64)
    F = md.forces.cpu().numpy()
    E = md.energy.cpu().numpy()
    mp = np.concatenate([[0], np.cumsum(md.sizes)])
    df = de = 0.0
    for b in range(min(sample, md.B)):
        sl = slice(mp[b], mp[b + 1])
        Eref, Fref = R.energy_and_forces(sd_cpu, md.structures[b].numbers, x32[sl])
        df = max(df, float(np.abs(F[sl] - Fref.numpy()).max()))
        de = max(de, abs(float(E[b]) - Eref.item()) / md.sizes[b])
    del lib
    return df, de


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--molecules", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--steps-per-graph", type=int, default=10)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--skip-baselines", action="store_true")
    ap.add_argument("--stage-reps", type=int, default=20)
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

    torch.manual_seed(0)
    model = pkg.StudentForceField(hidden_dim=128, num_interactions=3, num_rbf=20, cutoff=5.0)
    model = model.to(dev)
    H, L, R = 128, 3, 20
    mols = synth.molecule_set(args.molecules, 10, 100, args.seed)
    n_atoms = sum(len(m) for m in mols)
    log(f"{len(mols)} molecules, {n_atoms} atoms")
    total_steps = args.warmup + args.steps
    md = mdmod.BatchedNVE(model, mols, timestep_fs=0.5, temperature_K=300.0, seed=args.seed,
                          max_steps=total_steps + 1)
    # E+F alone on the starting (compact) structures, and per-kernel timings there
    c0 = md.plan.read_counts()
    E0, P0 = int(c0[0]), int(c0[1])
    e_tmp = torch.empty(md.B, dtype=torch.float64, device=dev)
    f_tmp = torch.empty(n_atoms, 3, dtype=torch.float32, device=dev)
    for _ in range(3):
        md.plan.energy_forces(md.pos32, e_tmp, f_tmp)
    torch.cuda.synchronize()
    a_ev = torch.cuda.Event(enable_timing=True)
    b_ev = torch.cuda.Event(enable_timing=True)
    reps = 20
    a_ev.record()
    for _ in range(reps):
        md.plan.energy_forces(md.pos32, e_tmp, f_tmp)
    b_ev.record()
    b_ev.synchronize()
    ef_s = a_ev.elapsed_time(b_ev) / 1e3 / reps
    log(f"static E+F at t=0: {ef_s * 1e3:.3f} ms ({n_atoms / ef_s:.4g} atom-evals/s), "
        f"E={E0}, P={P0}")
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
    sd_cpu = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    df, de = accuracy(sd_cpu, md, sample=16)
    log(f"accuracy vs FP64 oracle: max|dF|={df:.3e} eV/A, max|dE|/N={de:.3e} eV/atom")

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
        log(f"plain eager GPU path on {len(sample)} molecules")
        plain_rate = plain_md_rate(sd_dev, H, L, sample, synth.masses_of, 0.5 * synth.FS, 2, dev)
        plain = {"value": plain_rate, "unit": "atom-steps/s",
                 "sample": f"first {len(sample)} of the molecules, 2 MD steps each",
                 "speedup": value / plain_rate}
        log(f"plain: {plain_rate:.4g} atom-steps/s -> batched/plain = {value / plain_rate:.1f}x")
        csample = mols[:args.cpu_sample]
        log(f"cpu baseline on {len(csample)} molecules x {args.cpu_steps} steps")
        cpu_rate = cpu_baseline(sd_cpu, csample, synth.masses_of, 0.5 * synth.FS, args.cpu_steps)
        cpu = {"value": cpu_rate, "unit": "atom-steps/s", "cores": torch.get_num_threads(),
               "kind": "port",
               "sample": f"first {len(csample)} molecules of the workload, {args.cpu_steps} MD "
                         "steps each, float32 PyTorch-eager restatement of the reference "
                         "(dense neighbour search, autograd forces, Velocity Verlet)",
               "cpu_count": os.cpu_count(),
               "affinity": len(os.sched_getaffinity(0))}
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
                   "edges_window_end": E_edges,
                   "note": "random-init forces are not restoring: molecules expand during "
                           "the run and the edge count falls over the window"},
        "static_ef_t0": {"ms": ef_s * 1e3, "atom_evals_per_s": n_atoms / ef_s},
        "accuracy": {"max_abs_dF_eV_per_A": df, "max_abs_dE_per_atom_eV": de,
                     "sample_molecules": 16},
        "plain": plain,
        "nve": {"max_abs_drift_eV_per_atom": float(np.abs(drift).max()),
                "steps": int(summ["n_steps"])},
        "stages_us": {k: round(v["avg_s"] * 1e6, 2) for k, v in st.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
