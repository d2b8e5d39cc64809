"""Launch-shape autotuner for the pair and message kernels.

Two kinds of choice, both timed on the plan's own workload with HIP events
around whole evaluations (neighbour list, forward, adjoint: the unit the
caller repeats):

* the persistent grid of the pair kernels (filter forward and adjoint),
  which trades load balance against L2 locality and occupancy, and, with the
  filter kernels on their own stream, against the room left for the
  atom-side kernels running beside them;
* the launch shapes among the pre-instantiated variants
  (engine.launch_shapes): pairs per filter-forward tile, pairs per
  filter-adjoint tile, and edges in flight per message row (forward,
  adjoint).  They trade registers and LDS (blocks per CU) against work per
  tile; every variant gives bit-identical results.

The grid is searched first with the default shapes, then each shape kind in
turn is switched to its alternative and kept when the evaluation gets
faster (coordinate descent).  The choice is remembered per (device, hidden
size, streams, pair-count bucket).
"""

from __future__ import annotations

import json
import math
import os
from typing import Dict, Iterable, Optional

import torch

from . import _capi as C

_CACHE: Dict[str, int] = {}


def _key(plan) -> str:
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    counts = plan.read_counts()
    pairs = max(int(counts[C.MLFF_COUNT_PAIRS]), 1)
    streams = getattr(plan, "streams", 2)
    return f"{name}|H{plan.H}|L{plan.L}|S{streams}|P2^{int(math.log2(pairs))}"


def default_candidates(cus: Optional[int] = None) -> Iterable[int]:
    if cus is None:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
    return [k * cus for k in (2, 3, 4, 6, 8, 12, 16, 24, 32)]


def time_evaluation(plan, positions: torch.Tensor, e: torch.Tensor, f: torch.Tensor,
                    reps: int = 10) -> float:
    """Seconds per full energy + forces evaluation on the current stream."""
    s = torch.cuda.current_stream()
    plan.energy_forces(positions, e, f, stream=s)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        plan.energy_forces(positions, e, f, stream=s)
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / 1e3 / reps


SHAPE_NAMES = ("filter_fwd_pairs", "filter_bwd_pairs", "message_fwd_unroll",
               "message_bwd_unroll", "filter_bwd_variant")


def tune(plan, positions: torch.Tensor, candidates: Optional[Iterable[int]] = None,
         reps: int = 10, cache_file: Optional[str] = None,
         shapes: bool = True) -> Dict[str, object]:
    """Pick the persistent grid and the launch shapes for ``plan`` (evaluated
    at ``positions``).

    Returns {"blocks": best grid, "tiles": (fwd pairs, bwd pairs, message
    fwd unroll, message bwd unroll), "timings_us": {blocks: us},
    "shape_timings_us": {shape label: us}}; the plan keeps the choice.  With
    ``cache_file`` the choice is read from / written to a JSON file so later
    runs skip the search; ``shapes=False`` tunes the grid only.
    """
    from .engine import launch_shapes
    lib = C.lib()
    B = plan.layout.n_structures
    e = torch.empty(B, dtype=torch.float64, device=positions.device)
    f = torch.empty(plan.layout.n_atoms, 3, dtype=torch.float32, device=positions.device)
    plan.energy_forces(positions, e, f)
    key = _key(plan) + ("" if shapes else "|grid")
    if cache_file and os.path.exists(cache_file):
        with open(cache_file) as fh:
            _CACHE.update(json.load(fh))
    if key in _CACHE:
        hit = _CACHE[key]
        blocks = int(hit["blocks"] if isinstance(hit, dict) else hit)
        tiles = tuple(hit.get("tiles", plan.tiles)) if isinstance(hit, dict) else plan.tiles
        C.check("set_blocks", lib.mlff_plan_set_blocks(plan.handle, blocks))
        plan.set_tiles(*tiles)
        return {"blocks": blocks, "tiles": tuple(plan.tiles), "timings_us": {},
                "shape_timings_us": {}, "cached": True}
    cands = sorted({(int(c) + 7) // 8 * 8 for c in (candidates or default_candidates())})
    timings = {}
    for blocks in cands:
        C.check("set_blocks", lib.mlff_plan_set_blocks(plan.handle, blocks))
        timings[blocks] = time_evaluation(plan, positions, e, f, reps) * 1e6
    best = min(timings, key=timings.get)
    C.check("set_blocks", lib.mlff_plan_set_blocks(plan.handle, best))
    shape_timings = {}
    tiles = list(plan.tiles)
    if shapes:
        best_us = time_evaluation(plan, positions, e, f, reps) * 1e6
        shape_timings["default"] = best_us
        shapes_h = launch_shapes(plan.H)
        for k, values in enumerate(shapes_h):
            best_k = list(tiles)
            for v in values:
                if v == tiles[k]:
                    continue
                trial = list(tiles)
                trial[k] = v
                if trial[4] and trial[1] != shapes_h[1][0]:
                    continue             # adjoint variants exist for the default tile only
                plan.set_tiles(*trial)
                us = time_evaluation(plan, positions, e, f, reps) * 1e6
                shape_timings[f"{SHAPE_NAMES[k]}={v}"] = us
                if us < best_us:
                    best_us, best_k = us, trial
            tiles = best_k
        plan.set_tiles(*tiles)
    _CACHE[key] = {"blocks": best, "tiles": list(plan.tiles)}
    if cache_file:
        os.makedirs(os.path.dirname(os.path.abspath(cache_file)), exist_ok=True)
        with open(cache_file, "w") as fh:
            json.dump(_CACHE, fh, indent=1)
    return {"blocks": best, "tiles": tuple(plan.tiles), "timings_us": timings,
            "shape_timings_us": shape_timings, "cached": False}
