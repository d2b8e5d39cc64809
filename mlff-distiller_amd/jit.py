"""TorchScript export of the student force field (the reference's
scripts/export_to_torchscript.py:44-160 and the calculator's use_jit path,
ase_calculator.py:210-237, 319-335).

The fused evaluation is registered as a PyTorch operator,
``torch.ops.mlff_amd.energy_forces(atomic_numbers, positions, weights, hyper,
cutoff) -> (energy, forces)``, with an autograd rule (dE/dx = -forces from the
hand-written adjoint).  ``TorchScriptEnergy`` is the reference's SimpleWrapper,
``forward(atomic_numbers, positions) -> E``, calling that operator with the
model's own parameters, so ``torch.jit.trace`` records one operator call and
the saved file keeps the weights under the reference's state-dict keys
(``model.<key>``).  Loading a saved file needs this module imported first
(it registers the operator); the operator runs on the ROCm device only.
Forces of the exported module come from autograd.grad of its energy; second
derivatives (a force loss through the exported module) are not supported.
"""

from __future__ import annotations

from collections import OrderedDict
from typing import List, Tuple

import numpy as np
import torch
from torch import Tensor, nn

_EVALUATORS: "OrderedDict[tuple, nn.Module]" = OrderedDict()


def _evaluator(weights: List[Tensor], hyper: List[int], cutoff: float):
    """A StudentForceField holding ``weights`` (state-dict order), cached per
    (architecture, device, weight storage and version)."""
    from .model import StudentForceField
    H, L, R, max_z = (int(x) for x in hyper)
    sig = (H, L, R, max_z, float(cutoff), str(weights[0].device)) + tuple(
        (t.data_ptr(), t._version) for t in weights)
    ev = _EVALUATORS.get(sig)
    if ev is None:
        ev = StudentForceField(hidden_dim=H, num_interactions=L, num_rbf=R, cutoff=float(cutoff),
                               max_z=max_z)
        keys = list(ev.state_dict().keys())
        if len(keys) != len(weights):
            raise ValueError(f"expected {len(keys)} weight tensors, got {len(weights)}")
        ev.load_state_dict({k: w.detach() for k, w in zip(keys, weights)})
        ev = ev.to(weights[0].device).eval()
        for p in ev.parameters():
            p.requires_grad_(False)
        _EVALUATORS[sig] = ev
        while len(_EVALUATORS) > 4:
            _EVALUATORS.popitem(last=False)
    else:
        _EVALUATORS.move_to_end(sig)
    return ev


@torch.library.custom_op("mlff_amd::energy_forces", mutates_args=())
def energy_forces(atomic_numbers: Tensor, positions: Tensor, weights: List[Tensor],
                  hyper: List[int], cutoff: float) -> Tuple[Tensor, Tensor]:
    """Energy (0-d, the positions' dtype) and forces [N, 3] of one structure."""
    ev = _evaluator(weights, hyper, cutoff)
    with torch.no_grad():
        e64, f, _ = ev.evaluate(atomic_numbers, positions.detach(),
                                np.array([positions.shape[0]], dtype=np.int64))
    return e64.to(positions.dtype).reshape(()), f.to(positions.dtype)


@energy_forces.register_fake
def _(atomic_numbers, positions, weights, hyper, cutoff):
    return positions.new_empty(()), positions.new_empty(positions.shape)


def _setup_context(ctx, inputs, output):
    ctx.n_weights = len(inputs[2])
    ctx.save_for_backward(output[1])


def _backward(ctx, grad_energy, grad_forces):
    (forces,) = ctx.saved_tensors
    if grad_forces is not None and bool(grad_forces.ne(0).any()):
        raise NotImplementedError("the exported module gives first derivatives only")
    return None, -grad_energy * forces, [None] * ctx.n_weights, None, None


energy_forces.register_autograd(_backward, setup_context=_setup_context)


def _tensor_tree(model: nn.Module) -> nn.Module:
    """Plain nn.Modules nested like ``model``, holding its parameters and
    buffers under the same names (the saved file then needs no class of this
    package, whose directory name is not a Python identifier)."""
    root = nn.Module()
    for name, t in model.state_dict(keep_vars=True).items():
        *path, leaf = name.split(".")
        node = root
        for part in path:
            child = node._modules.get(part)
            if child is None:
                child = nn.Module()
                node.add_module(part, child)
            node = child
        if isinstance(t, nn.Parameter):
            node.register_parameter(leaf, t)
        else:
            node.register_buffer(leaf, t)
    return root


class TorchScriptEnergy(nn.Module):
    """The reference export's SimpleWrapper: ``forward(atomic_numbers,
    positions) -> E`` for one structure, on the fused kernels."""

    def __init__(self, model: nn.Module):
        super().__init__()
        self.model = _tensor_tree(model)
        self.hyper = [int(model.hidden_dim), int(model.num_interactions), int(model.num_rbf),
                      int(model.max_z)]
        self.cutoff = float(model.cutoff)
        self._names = list(model.state_dict().keys())

    def forward(self, atomic_numbers: Tensor, positions: Tensor) -> Tensor:
        tensors = dict(self.model.named_parameters())
        tensors.update(dict(self.model.named_buffers()))
        weights = [tensors[k] for k in self._names]
        return torch.ops.mlff_amd.energy_forces(atomic_numbers, positions, weights, self.hyper,
                                                self.cutoff)[0]


# the traced module's type is saved under this qualified name
TorchScriptEnergy.__module__ = "mlff_amd.jit"


def export_to_torchscript(model: nn.Module, output_path: str, num_atoms: int = 50,
                          device="cuda", seed: int = 0) -> torch.jit.ScriptModule:
    """Trace ``TorchScriptEnergy(model)`` on a seeded dummy molecule and save
    it (scripts/export_to_torchscript.py:70-106, 149-158)."""
    from . import synth
    model = model.to(device).eval()
    wrapped = TorchScriptEnergy(model).eval()
    mol = synth.single_molecule(num_atoms, seed=seed)
    z = torch.tensor(mol.numbers, dtype=torch.int64, device=device)
    x = torch.tensor(mol.positions, dtype=torch.float32, device=device)
    with torch.no_grad():
        traced = torch.jit.trace(wrapped, (z, x), check_trace=False)
    torch.jit.save(traced, str(output_path))
    return traced


def load_torchscript(path: str, device="cuda") -> torch.jit.ScriptModule:
    """Load an export of this package (the operator is registered by this module)."""
    m = torch.jit.load(str(path), map_location=device)
    return m.eval()
