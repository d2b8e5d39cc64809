"""The C ABI library loads, exports every declared symbol, and its host-only
entry points (layout, plan bookkeeping, argument checks) behave (CPU only)."""

import ctypes
import os
import re

import pytest
import torch

from conftest import ROOT, pkg_mod

C = pkg_mod("_capi")


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "mlff_amd.h")).read()
    return set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(mlff_\w+)\s*\(", text, flags=re.M))


def test_header_and_binding_agree():
    assert _declared_symbols() == set(C.EXPORTS)


def _declared_arity():
    """Parameter count of every entry point in the header."""
    text = open(os.path.join(ROOT, "include", "mlff_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"(mlff_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_binding_arity_matches_header():
    """Every ctypes declaration lists as many arguments as the C prototype.
    ctypes passes surplus arguments with default int conversion, which
    truncates a 64-bit device pointer: a missing argtype is a wrong address."""
    lib = C.lib()
    arity = _declared_arity()
    assert set(arity) == set(C.EXPORTS)
    for name, n in arity.items():
        assert len(getattr(lib, name).argtypes) == n, (name, n)


def test_library_exports_every_symbol():
    lib = C.lib()
    for name in C.EXPORTS:
        assert hasattr(lib, name), name
    assert lib.mlff_version() >= 100


def test_weights_layout_full_model():
    slots, total = C.weights_layout(128, 3, 20, 118)
    # 3 model tensors, 13 slots per block, 8 readout slots
    assert len(slots) == 3 + 13 * 3 + 8
    offs = sorted(s.offset for s in slots)
    assert offs[0] == 0 and len(set(offs)) == len(offs)
    for s in slots:
        assert s.offset % 64 == 0
        # MFMA operands need 16-byte rows; the 3x3 mixing matrix is read as scalars
        assert s.ld % 4 == 0 or s.rows == 1 or s.tensor == C.T_MIX
        assert s.offset + s.rows * s.ld <= total
    f1 = [s for s in slots if s.tensor == C.T_F1_W and s.kind == 0][0]
    assert (f1.rows, f1.cols, f1.ld) == (128, 20, 32)


def test_layout_rejects_unsupported_hidden():
    n = ctypes.c_int(0)
    tot = ctypes.c_int64(0)
    rc = C.lib().mlff_weights_layout(48, 3, 20, 118, None, 0, ctypes.byref(n), ctypes.byref(tot))
    assert rc == -2
    assert b"hidden_dim 48" in C.lib().mlff_last_error()


def test_plan_workspace_accounting():
    cfg = C.Config(128, 3, 20, 118, 5.0, 1000, 10, 40000, 20000, 0, 64)
    h = C.vp()
    C.check("create", C.lib().mlff_plan_create(ctypes.byref(h), ctypes.byref(cfg)))
    nbytes = C.lib().mlff_plan_workspace_bytes(h)
    # W per block (3H floats per pair) dominates: 3 * 20000 * 384 * 4 bytes
    assert nbytes > 3 * 20000 * 384 * 4
    rc = C.lib().mlff_plan_bind(h, None, 0, None, None, None, None, None, None)
    assert rc == -4
    assert C.lib().mlff_energy_forces(h, None, None, None, None, None, None) == -5
    # a weight rebind needs a bound plan; a null plan is refused the same way
    assert C.lib().mlff_plan_set_weights(h, None, ctypes.c_void_p(1 << 20)) == -5
    assert b"plan not bound" in C.lib().mlff_last_error()
    assert C.lib().mlff_plan_set_weights(None, None, ctypes.c_void_p(1 << 20)) == -5
    C.lib().mlff_plan_destroy(h)


def test_plan_bind_rejects_misaligned_buffers():
    """The kernels access the workspace and the weights with 16-byte vector
    loads; bind refuses other alignments before it touches the device (the
    pointers below are never dereferenced)."""
    cfg = C.Config(32, 2, 16, 118, 5.0, 100, 2, 4000, 2000, 0, 64)
    h = C.vp()
    C.check("create", C.lib().mlff_plan_create(ctypes.byref(h), ctypes.byref(cfg)))
    nbytes = C.lib().mlff_plan_workspace_bytes(h)
    ok, odd = 1 << 20, (1 << 20) + 4
    args = (ctypes.c_void_p(ok),) * 4                       # numbers, mol_ptr, atom_mol
    for ws, w in ((odd, ok), (ok, odd)):
        rc = C.lib().mlff_plan_bind(h, ctypes.c_void_p(ws), nbytes, ctypes.c_void_p(w),
                                    *args[:3], None, None)
        assert rc == -1
        assert b"16-byte aligned" in C.lib().mlff_last_error()
    C.lib().mlff_plan_destroy(h)


def test_plan_create_rejects_bad_sizes():
    cfg = C.Config(128, 3, 20, 118, 5.0, 0, 1, 10, 10, 0, 0)
    h = C.vp()
    assert C.lib().mlff_plan_create(ctypes.byref(h), ctypes.byref(cfg)) == -1


def test_weight_packing_places_tensors():
    engine = pkg_mod("engine")
    model = pkg_mod("model")
    torch.manual_seed(0)
    m = model.StudentForceField(hidden_dim=32, num_interactions=2, num_rbf=16)
    blob = engine.pack_weights(m.state_dict(), "cpu")
    slots, total = C.weights_layout(32, 2, 16, 118)
    assert blob.numel() == total
    sd = m.state_dict()
    for s in slots:
        if s.tensor == C.T_F2_W and s.layer == 1:
            view = blob[s.offset:s.offset + s.rows * s.ld].view(s.rows, s.ld)
            ref = sd["interactions.1.message.rbf_to_scalar.2.weight"]
            if s.kind == 0:
                assert torch.equal(view[:, :32], ref)
            else:
                assert torch.equal(view[:32, :96], ref.t())
        if s.tensor == C.T_R3_W:
            view = blob[s.offset:s.offset + s.ld]
            assert torch.equal(view[:8], sd["energy_head.4.weight"][0])
            assert torch.all(view[8:] == 0)


def test_static_entry_rejects_before_touching_the_device():
    """mlff_energy_forces_static checks its arguments and the scratch bound
    on the host: null arguments and a batch whose workspace exceeds the
    static scratch return errors without any device call (the non-null
    pointers below are never dereferenced)."""
    lib = C.lib()
    fake = ctypes.c_void_p(1 << 20)
    cfg = C.Config(128, 3, 20, 118, 5.0, 1000, 10, 400000, 200000, 0, 0, 0)
    assert lib.mlff_energy_forces_static(None, None, *([fake] * 12)) == -1
    rc = lib.mlff_energy_forces_static(None, ctypes.byref(cfg), *([fake] * 12))
    assert rc == -4
    msg = lib.mlff_last_error().decode()
    assert "static scratch holds %d" % C.STATIC_SCRATCH_BYTES in msg
    h = C.vp()
    C.check("create", lib.mlff_plan_create(ctypes.byref(h), ctypes.byref(cfg)))
    need = lib.mlff_plan_workspace_bytes(h)
    lib.mlff_plan_destroy(h)
    assert "needs %d bytes" % need in msg and need > C.STATIC_SCRATCH_BYTES
    bad = C.Config(128, 3, 20, 118, 5.0, 0, 1, 10, 10, 0, 0, 0)
    assert lib.mlff_energy_forces_static(None, ctypes.byref(bad), *([fake] * 12)) == -1


@pytest.mark.parametrize("H", [32, 64, 128])
def test_launch_shape_defaults_are_the_first_values(H):
    """Each launch-shape kind lists its default first; a plan starts on the
    defaults (five kinds: filter pairs forward and adjoint, message edges in
    flight forward and adjoint, filter adjoint variant 0)."""
    engine = pkg_mod("engine")
    shapes = engine.launch_shapes(H)
    assert len(shapes) == 5 and all(len(v) >= 2 for v in shapes)
    tiles = engine.default_tiles(H)
    assert tiles == tuple(v[0] for v in shapes) and tiles[4] == 0


def test_weight_grad_checks_arguments_on_the_host():
    """mlff_weight_grad validates shapes, strides, alignment and the
    workspace size before any device call (the fake pointers below are never
    dereferenced); the workspace size follows the row split."""
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    odd = ctypes.c_void_p((1 << 20) + 4)
    big = 1 << 30

    def call(K, M, N, a0=f, lda0=None, b0=f, ldb0=None, a1=None, b1=None, ws=f, nbytes=big,
             out=f):
        lda0 = M if lda0 is None else lda0
        ldb0 = N if ldb0 is None else ldb0
        return lib.mlff_weight_grad(None, K, M, N, a0, lda0, b0, ldb0, a1, M, b1, N, None, ws,
                                    nbytes, out, None)

    assert call(100, 6, 8) == -1          # M not a multiple of 4
    assert call(100, 8, 10) == -1         # N not a multiple of 4
    assert call(-1, 8, 8) == -1
    assert call(100, 8, 8, lda0=6) == -1  # stride below M
    assert call(100, 8, 8, a1=f) == -1    # a1 without b1
    assert call(100, 8, 8, b0=odd) == -1  # misaligned
    assert call(100, 8, 8, out=None) == -1
    need = lib.mlff_weight_grad_workspace_bytes(100, 8, 8)
    assert need > 0
    assert call(100, 8, 8, nbytes=need - 4) == -1
    assert "workspace" in lib.mlff_last_error().decode()
    # the partial count grows with K up to its cap, then the rows per wave grow
    small = lib.mlff_weight_grad_workspace_bytes(1000, 384, 128)
    large = lib.mlff_weight_grad_workspace_bytes(1_000_000, 384, 128)
    assert small < large == lib.mlff_weight_grad_workspace_bytes(2_000_000, 384, 128)
    assert lib.mlff_weight_grad_workspace_bytes(0, 8, 8) == (8 * 8 + 8) * 4


def test_grouped_weight_grad_checks_arguments_on_the_host():
    """mlff_weight_grad_partials checks what mlff_weight_grad checks, its
    partial count matches the workspace size, and the grouped reduce rejects
    bad tables before any device call."""
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    odd = ctypes.c_void_p((1 << 20) + 4)
    big = 1 << 30

    def call(K, M, N, a0=f, b0=f, a1=None, b1=None, ws=f, nbytes=big):
        return lib.mlff_weight_grad_partials(None, K, M, N, a0, M, b0, N, a1, M, b1, N, None, 1,
                                             ws, nbytes)

    assert call(100, 6, 8) == -1
    assert call(100, 8, 8, a1=f) == -1
    assert call(100, 8, 8, b0=odd) == -1
    assert call(100, 8, 8, ws=None) == -1
    need = lib.mlff_weight_grad_workspace_bytes(100, 8, 8)
    assert call(100, 8, 8, nbytes=need - 4) == -1
    for K, M, N in ((0, 8, 8), (100, 8, 8), (50132, 384, 128), (2736, 128, 20)):
        S = lib.mlff_weight_grad_partial_count(K, M, N)
        assert S >= 1
        assert lib.mlff_weight_grad_workspace_bytes(K, M, N) == S * (M * N + M) * 4
    assert lib.mlff_weight_grad_reduce_group(None, 0, None, 0) == 0      # nothing to add
    assert lib.mlff_weight_grad_reduce_group(None, -1, f, 1) == -1
    assert lib.mlff_weight_grad_reduce_group(None, 257, f, 1) == -1
    assert lib.mlff_weight_grad_reduce_group(None, 2, None, 1) == -1
    assert lib.mlff_weight_grad_reduce_group(None, 2, f, 0) == -1
    assert lib.mlff_weight_grad_reduce_group(None, 2, odd, 1) == -1
    assert "n_rows" in lib.mlff_last_error().decode()


def test_mixing_gradient_checks_arguments_on_the_host():
    """mlff_xyz_outer_partials refuses bad sizes, missing or unpaired
    operands and misaligned buffers before any device call; the partial count
    is at least one and at most 256."""
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    odd = ctypes.c_void_p((1 << 20) + 4)
    call = lib.mlff_xyz_outer_partials
    assert call(None, -1, 128, f, f, None, None, None, f) == -1
    assert call(None, 10, 6, f, f, None, None, None, f) == -1        # hidden not a multiple of 4
    assert call(None, 10, 128, None, f, None, None, None, f) == -1
    assert call(None, 10, 128, f, f, f, None, None, f) == -1         # a1 without b1
    assert call(None, 10, 128, f, odd, None, None, None, f) == -1
    assert call(None, 10, 128, f, f, None, None, None, odd) == -1
    assert lib.mlff_xyz_outer_partial_count(0) == 1
    assert lib.mlff_xyz_outer_partial_count(2736) == 86
    assert lib.mlff_xyz_outer_partial_count(10 ** 7) <= 256
    assert lib.mlff_xyz_outer_partial_count(-1) == 0


def test_train_geometry_checks_arguments_on_the_host():
    """mlff_train_geometry refuses bad sizes and missing outputs (the forward
    writes six arrays, the tangent three) before any device call; no pairs is
    a no-op."""
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    call = lib.mlff_train_geometry
    fwd = (f, f, f, f, f, f, None, None, None)
    tan = (None,) * 6 + (f, f, f)
    assert call(None, -1, 20, f, None, f, f, None, f, f, 5.0, 1e-8, *fwd) == -1
    assert call(None, 10, 0, f, None, f, f, None, f, f, 5.0, 1e-8, *fwd) == -1
    assert call(None, 10, 20, f, None, f, f, None, f, f, 0.0, 1e-8, *fwd) == -1
    assert call(None, 10, 20, f, None, f, f, None, f, f, 5.0, 1e-8, *tan) == -1  # forward, no e
    assert call(None, 10, 20, f, f, f, f, None, f, f, 5.0, 1e-8, *fwd) == -1     # tangent, no dd
    assert call(None, 10, 20, None, None, f, f, None, f, f, 5.0, 1e-8, *fwd) == -1
    assert call(None, 0, 20, f, None, f, f, None, f, f, 5.0, 1e-8, *fwd) == 0


def test_neighbor_pairs_checks_arguments_on_the_host():
    """mlff_neighbor_pairs refuses bad sizes, missing outputs and a
    misaligned pair geometry before any device call; no atoms is a no-op."""
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    odd = ctypes.c_void_p((1 << 20) + 4)
    call = lib.mlff_neighbor_pairs
    ok = dict(edge_pair=f, lo=f, hi=f, geom=f)

    def run(n=10, cap=100, pcap=50, rc=5.0, **kw):
        a = {**ok, **kw}
        return call(None, f, n, 2, f, f, rc, f, f, cap, a["edge_pair"], a["lo"], a["hi"],
                    a["geom"], pcap, f, f)

    assert run(n=-1) == -1
    assert run(rc=0.0) == -1
    assert run(pcap=-1) == -1
    assert run(edge_pair=None) == -1
    assert run(lo=None) == -1
    assert run(geom=None) == -1
    assert run(geom=odd) == -1
    assert run(n=0) == 0


def test_silu_chain_checks_arguments_on_the_host():
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    odd = ctypes.c_void_p((1 << 20) + 8)
    assert lib.mlff_silu_chain(None, 6, f, None, None, None, f, None, None, None) == -1
    assert lib.mlff_silu_chain(None, 8, None, None, None, None, f, None, None, None) == -1
    assert lib.mlff_silu_chain(None, 8, f, None, None, None, None, f, None, None) == -1  # dh, no da
    assert lib.mlff_silu_chain(None, 8, f, f, None, None, None, None, None, f) == -1     # dga, no g
    assert lib.mlff_silu_chain(None, 8, odd, None, None, None, f, None, None, None) == -1
    assert lib.mlff_silu_chain(None, 0, f, None, None, None, f, None, None, None) == 0


def test_train_filter_checks_arguments_on_the_host():
    """The fused training filter kernel refuses unsupported hidden sizes, a
    basis wider than 32, null buffers (the tangent needs a, the forward the
    biases) and misaligned weights before any device call; no pairs is a
    no-op.  The bf16 split refuses negative sizes and null buffers."""
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    odd = ctypes.c_void_p((1 << 20) + 4)
    ok = (f, f, f, f, f, None, f, f, f)               # forward: x w1 b1 w2 b2 a out...
    assert lib.mlff_train_filter(None, 0, 48, 10, 20, *ok) == -2
    assert lib.mlff_train_filter(None, 0, 128, 10, 33, *ok) == -1
    assert lib.mlff_train_filter(None, 0, 128, -1, 20, *ok) == -1
    assert lib.mlff_train_filter(None, 0, 128, 10, 20, f, f, None, f, f, None, f, f, f) == -1
    assert lib.mlff_train_filter(None, 1, 128, 10, 20, f, f, None, f, None, None, f, f,
                                 f) == -1          # tangent without a
    assert lib.mlff_train_filter(None, 0, 128, 10, 20, f, odd, f, f, f, None, f, f, f) == -1
    assert lib.mlff_train_filter(None, 0, 128, 0, 20, *ok) == 0
    # the fused filter adjoint: hidden size, counts, buffers, alignment
    ta = lib.mlff_train_filter_adjoint
    assert ta(None, 1, 48, 10, f, f, f, f, f, f, f) == -2
    assert ta(None, 1, 128, -1, f, f, f, f, f, f, f) == -1
    assert ta(None, 0, 128, 10, None, None, f, f, None, f, None) == -1
    assert ta(None, 1, 128, 10, f, None, f, f, f, f, f) == -1    # tangent without dgw
    assert ta(None, 1, 128, 10, f, f, f, f, f, f, None) == -1    # tangent without dga
    assert ta(None, 0, 128, 10, odd, None, f, f, None, f, None) == -1
    assert ta(None, 0, 128, 10, f, None, odd, f, None, f, None) == -1
    assert ta(None, 1, 128, 0, f, f, f, f, f, f, f) == 0
    assert lib.mlff_split_bf16x3(None, -1, f, f) == -1
    assert lib.mlff_split_bf16x3(None, 4, None, f) == -1
    assert lib.mlff_split_bf16x3(None, 0, None, None) == 0


def test_pack_weights_checks_arguments_on_the_host():
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    assert lib.mlff_pack_weights(None, -1, f, 10, f) == -1
    assert lib.mlff_pack_weights(None, 257, f, 10, f) == -1
    assert lib.mlff_pack_weights(None, 3, None, 10, f) == -1
    assert lib.mlff_pack_weights(None, 3, f, 10, None) == -1
    assert lib.mlff_pack_weights(None, 3, f, 0, None) == 0


def test_update_glue_checks_arguments_on_the_host():
    """The update-block kernels refuse bad sizes, null and misaligned
    buffers, and a tangent without its inputs, before any device call."""
    lib = C.lib()
    f = ctypes.c_void_p(1 << 20)
    odd = ctypes.c_void_p((1 << 20) + 4)
    assert lib.mlff_update_norm(None, 10, 6, f, f, f) == -1
    assert lib.mlff_update_norm(None, 10, 8, f, None, f) == -1
    assert lib.mlff_update_norm(None, 10, 8, f, odd, f) == -1
    assert lib.mlff_update_norm(None, 0, 8, f, f, f) == 0
    assert lib.mlff_update_apply(None, 10, 8, None, f, f, f, f, f, f) == -1
    assert lib.mlff_update_gate_adjoint(None, 10, 8, f, f, f, f, None, None, None, None, f,
                                        f) == -1        # tangent output without inputs
    assert lib.mlff_update_input_adjoint(None, 10, 8, f, f, f, f, f, f, f, None, None, None,
                                         None, None, None, f, f, f, f, None, None) == -1


def test_probe_flop_count_and_argument_checks():
    """The matrix-core probe's FLOP count (blocks x 4 waves x iters x 8
    MFMAs x FLOP per MFMA) and its host-side argument checks."""
    lib = C.lib()
    assert lib.mlff_probe_mfma_flops(0, 2, 10) == 2 * 4 * 10 * 8 * (16 * 16 * 4 * 2)
    assert lib.mlff_probe_mfma_flops(1, 1, 1) == 4 * 8 * (16 * 16 * 32 * 2)
    assert lib.mlff_probe_mfma_flops(2, 1, 1) == -1
    f = ctypes.c_void_p(1 << 20)
    assert lib.mlff_probe_mfma(None, 2, 1, 1, f) == -1
    assert lib.mlff_probe_mfma(None, 0, 0, 1, f) == -1
    assert lib.mlff_probe_mfma(None, 0, 1, 1, None) == -1
