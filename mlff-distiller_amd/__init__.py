"""MI355X-native PaiNN student force field (drop-in for atfrank/MLFF-Distiller's
StudentForceField / StudentForceFieldCalculator hot path).

The directory name contains a hyphen, so import it with importlib::

    mlff = importlib.import_module("mlff-distiller_amd")
    model = mlff.StudentForceField()

Evaluation runs in libmlff_amd.so (hand-written gfx950 HIP kernels behind a
C ABI, include/mlff_amd.h), reached through ctypes.  There is no CPU path:
building the library is required (``build_lib.build()``).
"""

from .model import (CosineCutoff, GaussianRBF, PaiNNInteraction, PaiNNMessage, PaiNNUpdate,
                    StudentForceField)
from .graph import radius_graph, radius_graph_native, radius_graph_csr
from .calculator import StudentForceFieldCalculator
from .md import BatchedNVE, NVEMDHarness
from .dataset import (DistillationWrapper, FrameDataset, FrameDatasetWriter,
                      distillation_collate)

__all__ = ["StudentForceField", "PaiNNInteraction", "PaiNNMessage", "PaiNNUpdate",
           "GaussianRBF", "CosineCutoff", "radius_graph", "radius_graph_native",
           "radius_graph_csr", "StudentForceFieldCalculator", "BatchedNVE", "NVEMDHarness",
           "DistillationWrapper", "FrameDataset", "FrameDatasetWriter", "distillation_collate"]
