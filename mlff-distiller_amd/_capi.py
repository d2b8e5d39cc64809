"""ctypes binding of libmlff_amd.so (declarations in include/mlff_amd.h).

torch is imported first on purpose: the library links libamdhip64.so.7,
and loading it after torch makes the dynamic linker reuse the HIP runtime
torch already loaded (same SONAME), so there is one runtime per process.

There is no fallback: if the library is missing, every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (must precede the library load, see above)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmlff_amd.so")

MLFF_COUNT_EDGES = 0
MLFF_COUNT_PAIRS = 1
MLFF_COUNT_OVERFLOW = 2
MLFF_COUNT_CELL_SMALL = 3
MLFF_COUNT_STEP = 4
MLFF_COUNT_EDGES_MAX = 5
MLFF_COUNT_SLOTS = 8

T_EMBEDDING, T_RBF_CENTERS, T_RBF_WIDTHS = 0, 1, 2
T_F1_W, T_F1_B, T_F2_W, T_F2_B = 3, 4, 5, 6
T_U1_W, T_U1_B, T_U2_W, T_U2_B, T_MIX = 7, 8, 9, 10, 11
T_R1_W, T_R1_B, T_R2_W, T_R2_B, T_R3_W, T_R3_B = 12, 13, 14, 15, 16, 17

STATUS = {0: "ok", -1: "invalid argument", -2: "unsupported", -3: "HIP error",
          -4: "workspace", -5: "call order"}

# every symbol include/mlff_amd.h declares
EXPORTS = (
    "mlff_version", "mlff_last_error", "mlff_weights_layout", "mlff_neighbor_workspace_bytes",
    "mlff_neighbor_list", "mlff_neighbor_pairs", "mlff_plan_create", "mlff_plan_destroy", "mlff_plan_workspace_bytes",
    "mlff_plan_bind", "mlff_plan_counts", "mlff_energy_forces", "mlff_md_bind",
    "mlff_md_bind_frames", "mlff_md_init_forces", "mlff_md_run", "mlff_md_graph_run",
    "mlff_launch_stage", "mlff_plan_set_blocks", "mlff_plan_set_cell_grid", "mlff_mts_bind", "mlff_mts_init_forces",
    "mlff_mts_run", "mlff_mts_graph_run", "mlff_plan_set_streams", "mlff_energy_forces_static",
    "mlff_train_message", "mlff_train_message_adjoint", "mlff_train_pair_grad",
    "mlff_segment_sum", "mlff_plan_set_tiles", "mlff_edge_message", "mlff_edge_grad",
    "mlff_weight_grad_workspace_bytes", "mlff_weight_grad", "mlff_weight_grad_partial_count",
    "mlff_weight_grad_partials", "mlff_weight_grad_reduce_group", "mlff_xyz_outer_partial_count",
    "mlff_xyz_outer_partials", "mlff_train_geometry", "mlff_silu_chain",
    "mlff_update_norm", "mlff_update_apply", "mlff_update_norm_tangent",
    "mlff_update_apply_tangent", "mlff_update_gate_adjoint", "mlff_update_input_adjoint",
    "mlff_probe_mfma_flops", "mlff_probe_mfma", "mlff_plan_set_weights",
    "mlff_train_filter", "mlff_train_filter_adjoint", "mlff_split_bf16x3", "mlff_pack_weights",
)

# bytes of the static device scratch behind mlff_energy_forces_static
STATIC_SCRATCH_BYTES = 64 << 20

STAGE_NEIGHBOR, STAGE_FILTER_FWD, STAGE_MESSAGE_FWD, STAGE_UPDATE_FWD = 0, 1, 2, 3
STAGE_READOUT, STAGE_UPDATE_BWD, STAGE_FILTER_BWD, STAGE_MESSAGE_BWD, STAGE_FORCES = 4, 5, 6, 7, 8


class Slot(ctypes.Structure):
    _fields_ = [("tensor", ctypes.c_int32), ("layer", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("rows", ctypes.c_int32), ("cols", ctypes.c_int32), ("ld", ctypes.c_int32),
                ("offset", ctypes.c_int64)]


class Config(ctypes.Structure):
    _fields_ = [("hidden_dim", ctypes.c_int32), ("num_interactions", ctypes.c_int32),
                ("num_rbf", ctypes.c_int32), ("max_z", ctypes.c_int32),
                ("cutoff", ctypes.c_float),
                ("n_atoms", ctypes.c_int32), ("n_structures", ctypes.c_int32),
                ("edge_capacity", ctypes.c_int32), ("pair_capacity", ctypes.c_int32),
                ("any_periodic", ctypes.c_int32), ("persistent_blocks", ctypes.c_int32),
                ("cell_capacity", ctypes.c_int32)]


class MlffError(RuntimeError):
    """A C entry point returned a non-zero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({STATUS.get(code, code)}): {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()
vp = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64


def _declare(lib):
    lib.mlff_version.restype = ctypes.c_int
    lib.mlff_version.argtypes = []
    lib.mlff_last_error.restype = ctypes.c_char_p
    lib.mlff_last_error.argtypes = []
    lib.mlff_weights_layout.restype = ctypes.c_int
    lib.mlff_weights_layout.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(Slot), ctypes.c_int,
                                                             ctypes.POINTER(ctypes.c_int),
                                                             ctypes.POINTER(i64)]
    lib.mlff_neighbor_workspace_bytes.restype = i64
    lib.mlff_neighbor_workspace_bytes.argtypes = [ctypes.c_int]
    lib.mlff_neighbor_list.restype = ctypes.c_int
    lib.mlff_neighbor_list.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp,
                                       ctypes.c_float, vp, vp, vp, i32, vp, vp]
    lib.mlff_neighbor_pairs.restype = ctypes.c_int
    lib.mlff_neighbor_pairs.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp, vp,
                                        ctypes.c_float, vp, vp, i32, vp, vp, vp, vp, i32, vp, vp]
    lib.mlff_plan_create.restype = ctypes.c_int
    lib.mlff_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Config)]
    lib.mlff_plan_destroy.restype = ctypes.c_int
    lib.mlff_plan_destroy.argtypes = [vp]
    lib.mlff_plan_workspace_bytes.restype = i64
    lib.mlff_plan_workspace_bytes.argtypes = [vp]
    lib.mlff_plan_bind.restype = ctypes.c_int
    lib.mlff_plan_bind.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp, vp]
    lib.mlff_plan_counts.restype = vp
    lib.mlff_plan_counts.argtypes = [vp]
    lib.mlff_energy_forces.restype = ctypes.c_int
    lib.mlff_energy_forces.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.mlff_md_bind.restype = ctypes.c_int
    lib.mlff_md_bind.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i32, ctypes.c_double]
    lib.mlff_md_bind_frames.restype = ctypes.c_int
    lib.mlff_md_bind_frames.argtypes = [vp, vp, vp, vp, i32]
    lib.mlff_md_init_forces.restype = ctypes.c_int
    lib.mlff_md_init_forces.argtypes = [vp, vp]
    lib.mlff_md_run.restype = ctypes.c_int
    lib.mlff_md_run.argtypes = [vp, vp, ctypes.c_int]
    lib.mlff_md_graph_run.restype = ctypes.c_int
    lib.mlff_md_graph_run.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    lib.mlff_launch_stage.restype = ctypes.c_int
    lib.mlff_launch_stage.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    lib.mlff_plan_set_blocks.restype = ctypes.c_int
    lib.mlff_plan_set_blocks.argtypes = [vp, ctypes.c_int]
    lib.mlff_plan_set_tiles.restype = ctypes.c_int
    lib.mlff_plan_set_tiles.argtypes = [vp] + [ctypes.c_int] * 5
    lib.mlff_plan_set_streams.restype = ctypes.c_int
    lib.mlff_plan_set_streams.argtypes = [vp, ctypes.c_int]
    lib.mlff_plan_set_cell_grid.restype = ctypes.c_int
    lib.mlff_plan_set_cell_grid.argtypes = [vp, vp, ctypes.c_int]
    lib.mlff_mts_bind.restype = ctypes.c_int
    lib.mlff_mts_bind.argtypes = [vp] * 14 + [i32, ctypes.c_double, i32]
    lib.mlff_mts_init_forces.restype = ctypes.c_int
    lib.mlff_mts_init_forces.argtypes = [vp, vp]
    lib.mlff_mts_run.restype = ctypes.c_int
    lib.mlff_mts_run.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    lib.mlff_mts_graph_run.restype = ctypes.c_int
    lib.mlff_mts_graph_run.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    lib.mlff_train_message.restype = ctypes.c_int
    lib.mlff_train_message.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * 15
    lib.mlff_train_message_adjoint.restype = ctypes.c_int
    lib.mlff_train_message_adjoint.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * 13
    lib.mlff_train_pair_grad.restype = ctypes.c_int
    lib.mlff_train_pair_grad.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * 14
    lib.mlff_edge_message.restype = ctypes.c_int
    lib.mlff_edge_message.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * 9
    lib.mlff_edge_grad.restype = ctypes.c_int
    lib.mlff_edge_grad.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * 10
    lib.mlff_segment_sum.restype = ctypes.c_int
    lib.mlff_segment_sum.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * 5
    lib.mlff_weight_grad_workspace_bytes.restype = ctypes.c_size_t
    lib.mlff_weight_grad_workspace_bytes.argtypes = [ctypes.c_int] * 3
    lib.mlff_weight_grad.restype = ctypes.c_int
    lib.mlff_weight_grad.argtypes = ([vp] + [ctypes.c_int] * 3 + [vp, ctypes.c_int] * 4
                                     + [vp, vp, ctypes.c_size_t, vp, vp])
    lib.mlff_weight_grad_partial_count.restype = ctypes.c_int
    lib.mlff_weight_grad_partial_count.argtypes = [ctypes.c_int] * 3
    lib.mlff_weight_grad_partials.restype = ctypes.c_int
    lib.mlff_weight_grad_partials.argtypes = ([vp] + [ctypes.c_int] * 3 + [vp, ctypes.c_int] * 4
                                              + [vp, ctypes.c_int, vp, ctypes.c_size_t])
    lib.mlff_train_geometry.restype = ctypes.c_int
    lib.mlff_train_geometry.argtypes = ([vp, ctypes.c_int, ctypes.c_int] + [vp] * 7
                                        + [ctypes.c_float, ctypes.c_float] + [vp] * 9)
    lib.mlff_xyz_outer_partial_count.restype = ctypes.c_int
    lib.mlff_xyz_outer_partial_count.argtypes = [ctypes.c_int]
    lib.mlff_xyz_outer_partials.restype = ctypes.c_int
    lib.mlff_xyz_outer_partials.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * 6
    lib.mlff_weight_grad_reduce_group.restype = ctypes.c_int
    lib.mlff_weight_grad_reduce_group.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
    lib.mlff_silu_chain.restype = ctypes.c_int
    lib.mlff_silu_chain.argtypes = [vp, ctypes.c_int64] + [vp] * 8
    for name, n_ptrs in (("mlff_update_norm", 3), ("mlff_update_apply", 7),
                         ("mlff_update_norm_tangent", 5), ("mlff_update_apply_tangent", 10),
                         ("mlff_update_gate_adjoint", 10), ("mlff_update_input_adjoint", 19)):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * n_ptrs
    lib.mlff_plan_set_weights.restype = ctypes.c_int
    lib.mlff_plan_set_weights.argtypes = [vp, vp, vp]
    lib.mlff_train_filter.restype = ctypes.c_int
    lib.mlff_train_filter.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int] + [vp] * 9
    lib.mlff_train_filter_adjoint.restype = ctypes.c_int
    lib.mlff_train_filter_adjoint.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [vp] * 7
    lib.mlff_split_bf16x3.restype = ctypes.c_int
    lib.mlff_split_bf16x3.argtypes = [vp, ctypes.c_int64, vp, vp]
    lib.mlff_pack_weights.restype = ctypes.c_int
    lib.mlff_pack_weights.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int64, vp]
    lib.mlff_probe_mfma_flops.restype = ctypes.c_int64
    lib.mlff_probe_mfma_flops.argtypes = [ctypes.c_int] * 3
    lib.mlff_probe_mfma.restype = ctypes.c_int
    lib.mlff_probe_mfma.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    lib.mlff_energy_forces_static.restype = ctypes.c_int
    lib.mlff_energy_forces_static.argtypes = [vp, ctypes.POINTER(Config)] + [vp] * 12


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                        "(the HIP library is required; there is no CPU fallback)")
                handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
                _declare(handle)
                _lib = handle
    return _lib


def check(fn: str, rc: int):
    if rc != 0:
        msg = lib().mlff_last_error().decode(errors="replace")
        raise MlffError(fn, rc, msg)


def ptr(t) -> int:
    """Device (or host) address of a tensor, or 0 for None."""
    return 0 if t is None else t.data_ptr()


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_cur_device = getattr(torch._C, "_cuda_getDevice", None)


def stream_handle(stream=None) -> int:
    """hipStream_t of ``stream``, or of the current stream of the current
    device (read directly: torch.cuda.current_stream() costs several
    microseconds of Python per call, and the training pass makes thousands)."""
    if stream is not None:
        return stream.cuda_stream
    if _raw_stream is not None and _cur_device is not None:
        return _raw_stream(_cur_device())
    return torch.cuda.current_stream().cuda_stream


def weights_layout(H: int, L: int, R: int, max_z: int):
    n = ctypes.c_int(0)
    total = i64(0)
    check("mlff_weights_layout", lib().mlff_weights_layout(H, L, R, max_z, None, 0,
                                                           ctypes.byref(n), ctypes.byref(total)))
    slots = (Slot * n.value)()
    check("mlff_weights_layout", lib().mlff_weights_layout(H, L, R, max_z, slots, n.value,
                                                           ctypes.byref(n), ctypes.byref(total)))
    return list(slots), int(total.value)
