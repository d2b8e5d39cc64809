"""ORACLE -- test infrastructure only.

Hand-derived force-matching gradients in plain torch ops (any device, any
float dtype), the checker for the HIP training pass
(mlff-distiller_amd/train.py).  Only ``tests/`` may import it.

The reference trains with torch autograd double backward
(models/distillation_wrapper.py:86-91, training/trainer.py:325-328, and the
compact recipe scripts/train_compact_models.py:229-372 with loss
wE * MSE(E) + wF * MSE(F), :340-342).  Here the gradient of the loss with
respect to the weights is formed explicitly:

  dL/dtheta = sum_b dL/dE_b dE_b/dtheta  -  d/dtheta < lambda, grad_x E >,
  lambda = dL/dF.

The first term is the reverse (adjoint) pass with parameter gradients.  The
second is the directional derivative, along the position perturbation
lambda, of the parameter gradient of the total energy: the tangent-linear
version of the same reverse pass (forward-over-reverse).  Every line below
is an explicit derivative of the corresponding line of
student_model.py:634-757.

Pinned (tests/test_train_cpu.py, float64): the gradients equal the
oracle's autograd double backward (painn_ref.energy) to 1e-9, including
learnable basis centres and widths, and central finite differences of the
loss; forward energies and forces equal painn_ref's.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

Tensor = torch.Tensor
EPS = 1e-8   # edge_vector / (d + 1e-8), student_model.py:715


@dataclass
class PairGraph:
    lo: Tensor          # [P] int64
    hi: Tensor          # [P] int64
    shift: Optional[Tensor]  # [P,3] float (already multiplied by the cell) or None
    batch: Tensor       # [N] int64 structure of each atom
    n_structures: int


def pair_graph_from_csr(row_ptr: Tensor, col: Tensor, batch: Tensor, n_structures: int,
                        shift_vec: Optional[Tensor] = None) -> PairGraph:
    """Unordered pairs (row < col) from a symmetric receiver CSR."""
    n = row_ptr.numel() - 1
    deg = (row_ptr[1:] - row_ptr[:-1]).long()
    row = torch.repeat_interleave(torch.arange(n, device=col.device), deg)
    colL = col.long()
    keep = colL > row
    sv = None if shift_vec is None else shift_vec[keep]
    return PairGraph(row[keep], colL[keep], sv, batch.long(), n_structures)


# ---------------------------------------------------------------------------
# elementwise pieces
# ---------------------------------------------------------------------------

def _silu(x):
    return x * torch.sigmoid(x)


def _silu1(x):
    s = torch.sigmoid(x)
    return s * (1 + x * (1 - s))


def _silu2(x):
    s = torch.sigmoid(x)
    return s * (1 - s) * (2 + x * (1 - 2 * s))


def _edge_features(d, centers, widths, rc):
    """e_k(d) = exp(-(d-mu_k)^2 / w_k^2) * 0.5 (cos(pi d/rc) + 1) [d < rc] and de/dd."""
    diff = d.unsqueeze(-1) - centers
    gamma = 1.0 / widths ** 2
    phi = torch.exp(-gamma * diff ** 2)
    inside = (d < rc).to(d.dtype).unsqueeze(-1)
    arg = math.pi * d.unsqueeze(-1) / rc
    fc = 0.5 * (torch.cos(arg) + 1.0) * inside
    dfc = -0.5 * (math.pi / rc) * torch.sin(arg) * inside
    return phi * fc, (-2.0 * gamma * diff * phi) * fc + phi * dfc


def _edge_features_d2(d, centers, widths, rc):
    """d^2 e_k / dd^2 (used by the tangent of de/dd)."""
    diff = d.unsqueeze(-1) - centers
    gamma = 1.0 / widths ** 2
    phi = torch.exp(-gamma * diff ** 2)
    dphi = -2.0 * gamma * diff * phi
    d2phi = (-2.0 * gamma + 4.0 * gamma ** 2 * diff ** 2) * phi
    inside = (d < rc).to(d.dtype).unsqueeze(-1)
    arg = math.pi * d.unsqueeze(-1) / rc
    k = math.pi / rc
    fc = 0.5 * (torch.cos(arg) + 1.0) * inside
    dfc = -0.5 * k * torch.sin(arg) * inside
    d2fc = -0.5 * k * k * torch.cos(arg) * inside
    return d2phi * fc + 2 * dphi * dfc + phi * d2fc


def _edge_feature_param_derivs(d, centers, widths, rc):
    """Derivatives of e_k(d) = phi_k(d) fc(d) in the basis parameters
    (learnable_rbf, student_model.py:226-236): de_k/dmu_k, de_k/dw_k, and
    the d-derivatives of both (for the tangent pass), each [P, R].
    phi = exp(-(d-mu)^2 / w^2): dphi/dmu = 2 gamma (d-mu) phi,
    dphi/dw = 2 (d-mu)^2 phi / w^3, gamma = 1/w^2."""
    diff = d.unsqueeze(-1) - centers
    gamma = 1.0 / widths ** 2
    phi = torch.exp(-gamma * diff ** 2)
    inside = (d < rc).to(d.dtype).unsqueeze(-1)
    arg = math.pi * d.unsqueeze(-1) / rc
    fc = 0.5 * (torch.cos(arg) + 1.0) * inside
    dfc = -0.5 * (math.pi / rc) * torch.sin(arg) * inside
    w3 = widths ** 3
    A = 2.0 * gamma * diff * phi
    Bw = 2.0 * diff ** 2 * phi / w3
    dA = 2.0 * gamma * phi * (1.0 - 2.0 * gamma * diff ** 2)
    dBw = (2.0 * diff * phi / w3) * (2.0 - 2.0 * gamma * diff ** 2)
    return fc * A, fc * Bw, dfc * A + fc * dA, dfc * Bw + fc * dBw


def _tmm(a: Tensor, b: Tensor, chunk: int = 2048) -> Tensor:
    """a.T @ b for tall a [K, M], b [K, N] (a weight gradient, K = pairs or
    atoms): the K reduction is split into chunks computed as one batched
    product and then summed, because a single long reduction leaves most
    of the GPU idle in the library's kernels."""
    K = a.shape[0]
    if K < 4 * chunk:
        return a.T @ b
    nc = K // chunk
    head = nc * chunk
    at = a[:head].reshape(nc, chunk, a.shape[1]).transpose(1, 2).contiguous()
    out = torch.bmm(at, b[:head].reshape(nc, chunk, b.shape[1]).contiguous()).sum(0)
    if head < K:
        out = out + a[head:].T @ b[head:]
    return out


def _lin(x, p, name):
    return x @ p[name + ".weight"].T + p[name + ".bias"]


def _add_pairs(n, lo, hi, to_lo, to_hi):
    """Segment sum onto atoms of per-pair contributions to lo and to hi."""
    out = torch.zeros((n,) + to_lo.shape[1:], dtype=to_lo.dtype, device=to_lo.device)
    out.index_add_(0, lo, to_lo)
    out.index_add_(0, hi, to_hi)
    return out


def _mixT(M, v):
    """(M^T v)_z = sum_x M[x][z] v_x over the xyz axis of [N,3,H] (three
    weighted sums: a 3x3 contraction as a batched product would launch one
    tiny GEMM per atom)."""
    v0, v1, v2 = v[:, 0], v[:, 1], v[:, 2]
    return torch.stack([M[0, z] * v0 + M[1, z] * v1 + M[2, z] * v2 for z in range(3)], dim=1)


def _xyz_outer(a, b):
    """sum_{n,h} a[n,x,h] b[n,z,h] -> [3,3] (gradient of a 3x3 mixing matrix)."""
    return _tmm(a.permute(0, 2, 1).reshape(-1, 3), b.permute(0, 2, 1).reshape(-1, 3))


def _mix(M, v):
    """(M v)_x = sum_z M[x][z] v_z over the xyz axis of [N,3,H]."""
    v0, v1, v2 = v[:, 0], v[:, 1], v[:, 2]
    return torch.stack([M[x, 0] * v0 + M[x, 1] * v1 + M[x, 2] * v2 for x in range(3)], dim=1)


# ---------------------------------------------------------------------------
# forward with saved activations
# ---------------------------------------------------------------------------

def forward(p: Dict[str, Tensor], Z: Tensor, x: Tensor, g: PairGraph, rc: float):
    """Primal pass; returns (E [B], cache)."""
    H = p["embedding.weight"].shape[1]
    L = len({int(k.split(".")[1]) for k in p if k.startswith("interactions.")})
    n = Z.shape[0]
    dt = x.dtype
    vec = x[g.hi] - x[g.lo]
    if g.shift is not None:
        vec = vec + g.shift.to(dt)
    d = torch.linalg.vector_norm(vec, dim=1)
    den = d + EPS
    u = vec / den.unsqueeze(1)
    e, de = _edge_features(d, p["rbf.centers"].to(dt), p["rbf.widths"].to(dt), rc)
    s = p["embedding.weight"][Z].to(dt)
    v = torch.zeros(n, 3, H, dtype=dt, device=x.device)
    layers = []
    for k in range(L):
        pre = f"interactions.{k}.message.rbf_to_scalar"
        a1 = _lin(e, p, pre + ".0")
        h = _silu(a1)
        W = _lin(h, p, pre + ".2")
        Ws, Wv1, Wv2 = W[:, :H], W[:, H:2 * H], W[:, 2 * H:]
        s_m = s + _add_pairs(n, g.lo, g.hi, s[g.hi] * Ws, s[g.lo] * Ws)
        uW = u.unsqueeze(-1) * Wv2.unsqueeze(1)
        v_m = v + _add_pairs(n, g.lo, g.hi, v[g.hi] * Wv1.unsqueeze(1) + uW,
                             v[g.lo] * Wv1.unsqueeze(1) - uW)
        upre = f"interactions.{k}.update"
        nrm = torch.linalg.vector_norm(v_m, dim=1)
        xin = torch.cat([s_m, nrm], dim=-1)
        a2 = _lin(xin, p, upre + ".update_mlp.0")
        h2 = _silu(a2)
        y = _lin(h2, p, upre + ".update_mlp.2")
        ds, g1, g2 = y[:, :H], y[:, H:2 * H], y[:, 2 * H:]
        M = p[upre + ".mixing_matrix"].to(dt)
        Mv = _mix(M, v_m)
        s_new = s_m + ds
        v_new = v_m * g1.unsqueeze(1) + Mv * g2.unsqueeze(1)
        layers.append(dict(s=s, v=v, a1=a1, h=h, W=W, s_m=s_m, v_m=v_m, nrm=nrm, xin=xin,
                           a2=a2, h2=h2, g1=g1, g2=g2, Mv=Mv, M=M))
        s, v = s_new, v_new
    r1 = _lin(s, p, "energy_head.0")
    h3 = _silu(r1)
    r2 = _lin(h3, p, "energy_head.2")
    h4 = _silu(r2)
    eps = _lin(h4, p, "energy_head.4")[:, 0]
    E = torch.zeros(g.n_structures, dtype=dt, device=x.device).index_add_(0, g.batch, eps)
    cache = dict(H=H, L=L, n=n, vec=vec, d=d, den=den, u=u, e=e, de=de, layers=layers,
                 sL=s, r1=r1, h3=h3, r2=r2, h4=h4, Z=Z, rc=rc)
    return E, cache


# ---------------------------------------------------------------------------
# reverse pass (first order), optionally with its tangent along lambda
# ---------------------------------------------------------------------------

def reverse(p: Dict[str, Tensor], c: dict, g: PairGraph, gE: Tensor,
            lam: Optional[Tensor] = None, want_forces: bool = True, want_grads: bool = True,
            gw: Optional[Tensor] = None, rbf_grads: bool = False):
    """Adjoint of E weighted by gE [B].

    Returns (grads, forces, tgrads): parameter gradients of sum_b gw_b E_b
    (gw defaults to gE), forces -d(sum gE E)/dx, and -- when ``lam`` [N,3] is
    given -- the parameter gradients of d/deps [sum_b gE_b dE_b/dx] . lam,
    i.e. of < lam, grad_x (sum gE E) >.

    The adjoint is linear in the structure weights and structures do not
    interact, so the parameter gradients for weights gw are the terms of
    the gE adjoint with each atom or pair row scaled by gw_b / gE_b of its
    structure; ``gw`` is applied that way (gE must then be all ones), which
    lets one pass give the unit-weight tangent and the gw-weighted
    gradients together.  ``want_grads=False`` skips the parameter gradients.
    ``rbf_grads`` adds the gradients of the basis centres and widths
    (learnable_rbf); they are buffers otherwise and get none.
    """
    H, L, n = c["H"], c["L"], c["n"]
    T = lam is not None
    dt = c["d"].dtype
    grads: Dict[str, Tensor] = {}
    tg: Dict[str, Tensor] = {}

    def acc(store, name, val):
        store[name] = store[name] + val if name in store else val

    def accg(name, fn):
        """parameter gradient term (evaluated only when wanted)"""
        if want_grads:
            acc(grads, name, fn())

    if gw is not None:
        w_atom = gw[g.batch]
        w_pair = gw[g.batch[g.lo]]
        wa = lambda t: t * w_atom.view(-1, *([1] * (t.dim() - 1)))
        wp = lambda t: t * w_pair.view(-1, *([1] * (t.dim() - 1)))
    else:
        wa = wp = lambda t: t

    # ---- tangents of the forward pass along lam --------------------------------
    if T:
        vec, d, den, u = c["vec"], c["d"], c["den"], c["u"]
        dvec = lam[g.hi] - lam[g.lo]
        dd = (vec * dvec).sum(1) / d
        du = dvec / den.unsqueeze(1) - vec * (dd / den ** 2).unsqueeze(1)
        de_t = c["de"] * dd.unsqueeze(-1)
        ds_ = torch.zeros(n, H, dtype=dt, device=vec.device)
        dv_ = torch.zeros(n, 3, H, dtype=dt, device=vec.device)
        for k, lay in enumerate(c["layers"]):
            pre = f"interactions.{k}.message.rbf_to_scalar"
            lay["ds"], lay["dv"] = ds_, dv_
            da1 = de_t @ p[pre + ".0.weight"].T
            dh = _silu1(lay["a1"]) * da1
            dW = dh @ p[pre + ".2.weight"].T
            lay["da1"], lay["dh"], lay["dW"] = da1, dh, dW
            W = lay["W"]
            Ws, Wv1, Wv2 = W[:, :H], W[:, H:2 * H], W[:, 2 * H:]
            dWs, dWv1, dWv2 = dW[:, :H], dW[:, H:2 * H], dW[:, 2 * H:]
            s, v = lay["s"], lay["v"]
            dsm = ds_ + _add_pairs(n, g.lo, g.hi, ds_[g.hi] * Ws + s[g.hi] * dWs,
                                   ds_[g.lo] * Ws + s[g.lo] * dWs)
            duW = du.unsqueeze(-1) * Wv2.unsqueeze(1) + u.unsqueeze(-1) * dWv2.unsqueeze(1)
            dvm = dv_ + _add_pairs(
                n, g.lo, g.hi,
                dv_[g.hi] * Wv1.unsqueeze(1) + v[g.hi] * dWv1.unsqueeze(1) + duW,
                dv_[g.lo] * Wv1.unsqueeze(1) + v[g.lo] * dWv1.unsqueeze(1) - duW)
            nrm, v_m = lay["nrm"], lay["v_m"]
            safe = torch.where(nrm > 0, nrm, torch.ones_like(nrm))
            dnrm = torch.where(nrm > 0, (v_m * dvm).sum(1) / safe, torch.zeros_like(nrm))
            dxin = torch.cat([dsm, dnrm], dim=-1)
            upre = f"interactions.{k}.update"
            da2 = dxin @ p[upre + ".update_mlp.0.weight"].T
            dh2 = _silu1(lay["a2"]) * da2
            dy = dh2 @ p[upre + ".update_mlp.2.weight"].T
            M = lay["M"]
            dMv = _mix(M, dvm)
            lay.update(dsm=dsm, dvm=dvm, dnrm=dnrm, dxin=dxin, da2=da2, dh2=dh2, dy=dy, dMv=dMv)
            g1, g2 = lay["g1"], lay["g2"]
            dg1, dg2 = dy[:, H:2 * H], dy[:, 2 * H:]
            ds_ = dsm + dy[:, :H]
            dv_ = (dvm * g1.unsqueeze(1) + v_m * dg1.unsqueeze(1) + dMv * g2.unsqueeze(1)
                   + lay["Mv"] * dg2.unsqueeze(1))
        dsL = ds_
        dr1 = dsL @ p["energy_head.0.weight"].T
        dh3 = _silu1(c["r1"]) * dr1
        dr2 = dh3 @ p["energy_head.2.weight"].T
        dh4 = _silu1(c["r2"]) * dr2

    # ---- readout reverse ------------------------------------------------------
    geps = gE[g.batch]
    r1, h3, r2, h4, sL = c["r1"], c["h3"], c["r2"], c["h4"], c["sL"]
    Q3 = p["energy_head.4.weight"]
    accg("energy_head.4.weight", lambda: (wa(geps).unsqueeze(1) * h4).sum(0, keepdim=True))
    accg("energy_head.4.bias", lambda: wa(geps).sum().reshape(1))
    gh4 = geps.unsqueeze(1) * Q3
    gr2 = gh4 * _silu1(r2)
    accg("energy_head.2.weight", lambda: _tmm(wa(gr2), h3))
    accg("energy_head.2.bias", lambda: wa(gr2).sum(0))
    gh3 = gr2 @ p["energy_head.2.weight"]
    gr1 = gh3 * _silu1(r1)
    accg("energy_head.0.weight", lambda: _tmm(wa(gr1), sL))
    accg("energy_head.0.bias", lambda: wa(gr1).sum(0))
    gs = gr1 @ p["energy_head.0.weight"]
    gv = torch.zeros(n, 3, H, dtype=dt, device=gs.device)
    if T:
        acc(tg, "energy_head.4.weight", (geps.unsqueeze(1) * dh4).sum(0, keepdim=True))
        acc(tg, "energy_head.4.bias", torch.zeros(1, dtype=dt, device=gs.device))
        dgr2 = gh4 * _silu2(r2) * dr2
        acc(tg, "energy_head.2.weight", _tmm(dgr2, h3) + _tmm(gr2, dh3))
        acc(tg, "energy_head.2.bias", dgr2.sum(0))
        dgh3 = dgr2 @ p["energy_head.2.weight"]
        dgr1 = dgh3 * _silu1(r1) + gh3 * _silu2(r1) * dr1
        acc(tg, "energy_head.0.weight", _tmm(dgr1, sL) + _tmm(gr1, dsL))
        acc(tg, "energy_head.0.bias", dgr1.sum(0))
        dgs = dgr1 @ p["energy_head.0.weight"]
        dgv = torch.zeros_like(gv)

    gd = torch.zeros_like(c["d"])
    gu = torch.zeros_like(c["u"])
    u = c["u"]
    if rbf_grads:
        e_mu, e_w, de_mu, de_w = _edge_feature_param_derivs(
            c["d"], p["rbf.centers"].to(dt), p["rbf.widths"].to(dt), c["rc"])
    for k in range(L - 1, -1, -1):
        lay = c["layers"][k]
        upre = f"interactions.{k}.update"
        mpre = f"interactions.{k}.message.rbf_to_scalar"
        U1 = p[upre + ".update_mlp.0.weight"]
        U2 = p[upre + ".update_mlp.2.weight"]
        M = lay["M"]
        v_m, g1, g2, Mv, nrm = lay["v_m"], lay["g1"], lay["g2"], lay["Mv"], lay["nrm"]
        # update reverse
        gg1 = (gv * v_m).sum(1)
        gg2 = (gv * Mv).sum(1)
        gy = torch.cat([gs, gg1, gg2], dim=-1)
        accg(upre + ".update_mlp.2.weight", lambda: _tmm(wa(gy), lay["h2"]))
        accg(upre + ".update_mlp.2.bias", lambda: wa(gy).sum(0))
        gh2 = gy @ U2
        ga2 = gh2 * _silu1(lay["a2"])
        accg(upre + ".update_mlp.0.weight", lambda: _tmm(wa(ga2), lay["xin"]))
        accg(upre + ".update_mlp.0.bias", lambda: wa(ga2).sum(0))
        gx = ga2 @ U1
        gs_x, gn = gx[:, :H], gx[:, H:]
        gvg2 = gv * g2.unsqueeze(1)
        accg(upre + ".mixing_matrix", lambda: _xyz_outer(wa(gvg2), v_m))
        safe = torch.where(nrm > 0, nrm, torch.ones_like(nrm))
        unit_n = torch.where((nrm > 0).unsqueeze(1), v_m / safe.unsqueeze(1), torch.zeros_like(v_m))
        gs_m = gs + gs_x
        gv_m = gv * g1.unsqueeze(1) + _mixT(M, gvg2) + gn.unsqueeze(1) * unit_n
        if T:
            dvm, dMv, dxin = lay["dvm"], lay["dMv"], lay["dxin"]
            dg1, dg2 = lay["dy"][:, H:2 * H], lay["dy"][:, 2 * H:]
            dgg1 = (dgv * v_m + gv * dvm).sum(1)
            dgg2 = (dgv * Mv + gv * dMv).sum(1)
            dgy = torch.cat([dgs, dgg1, dgg2], dim=-1)
            acc(tg, upre + ".update_mlp.2.weight", _tmm(dgy, lay["h2"]) + _tmm(gy, lay["dh2"]))
            acc(tg, upre + ".update_mlp.2.bias", dgy.sum(0))
            dgh2 = dgy @ U2
            dga2 = dgh2 * _silu1(lay["a2"]) + gh2 * _silu2(lay["a2"]) * lay["da2"]
            acc(tg, upre + ".update_mlp.0.weight", _tmm(dga2, lay["xin"]) + _tmm(ga2, dxin))
            acc(tg, upre + ".update_mlp.0.bias", dga2.sum(0))
            dgx = dga2 @ U1
            dgs_x, dgn = dgx[:, :H], dgx[:, H:]
            dgvg2 = dgv * g2.unsqueeze(1) + gv * dg2.unsqueeze(1)
            acc(tg, upre + ".mixing_matrix",
                _xyz_outer(dgvg2, v_m) + _xyz_outer(gvg2, dvm))
            dnrm = lay["dnrm"]
            dunit = torch.where((nrm > 0).unsqueeze(1),
                                dvm / safe.unsqueeze(1) - v_m * (dnrm / safe ** 2).unsqueeze(1),
                                torch.zeros_like(v_m))
            dgs_m = dgs + dgs_x
            dgv_m = (dgv * g1.unsqueeze(1) + gv * dg1.unsqueeze(1) + _mixT(M, dgvg2)
                     + dgn.unsqueeze(1) * unit_n + gn.unsqueeze(1) * dunit)
        # message reverse
        s, v, W = lay["s"], lay["v"], lay["W"]
        Ws, Wv1, Wv2 = W[:, :H], W[:, H:2 * H], W[:, 2 * H:]
        ga, gb = g.lo, g.hi
        gWs = gs_m[ga] * s[gb] + gs_m[gb] * s[ga]
        gWv1 = (gv_m[ga] * v[gb]).sum(1) + (gv_m[gb] * v[ga]).sum(1)
        dgv_ab = gv_m[ga] - gv_m[gb]
        gWv2 = (u.unsqueeze(-1) * dgv_ab).sum(1)
        gu = gu + (dgv_ab * Wv2.unsqueeze(1)).sum(-1)
        gW = torch.cat([gWs, gWv1, gWv2], dim=-1)
        W2 = p[mpre + ".2.weight"]
        W1 = p[mpre + ".0.weight"]
        accg(mpre + ".2.weight", lambda: _tmm(wp(gW), lay["h"]))
        accg(mpre + ".2.bias", lambda: wp(gW).sum(0))
        gh = gW @ W2
        ga1 = gh * _silu1(lay["a1"])
        accg(mpre + ".0.weight", lambda: _tmm(wp(ga1), c["e"]))
        accg(mpre + ".0.bias", lambda: wp(ga1).sum(0))
        ge = ga1 @ W1
        gd = gd + (ge * c["de"]).sum(-1)
        if rbf_grads:
            accg("rbf.centers", lambda: (wp(ge) * e_mu).sum(0))
            accg("rbf.widths", lambda: (wp(ge) * e_w).sum(0))
        gs_new = gs_m + _add_pairs(n, ga, gb, gs_m[gb] * Ws, gs_m[ga] * Ws)
        gv_new = gv_m + _add_pairs(n, ga, gb, gv_m[gb] * Wv1.unsqueeze(1),
                                   gv_m[ga] * Wv1.unsqueeze(1))
        if T:
            ds, dv, dW, dh, da1 = lay["ds"], lay["dv"], lay["dW"], lay["dh"], lay["da1"]
            dWs, dWv1, dWv2 = dW[:, :H], dW[:, H:2 * H], dW[:, 2 * H:]
            dgWs = (dgs_m[ga] * s[gb] + gs_m[ga] * ds[gb] + dgs_m[gb] * s[ga] + gs_m[gb] * ds[ga])
            dgWv1 = ((dgv_m[ga] * v[gb] + gv_m[ga] * dv[gb]).sum(1)
                     + (dgv_m[gb] * v[ga] + gv_m[gb] * dv[ga]).sum(1))
            ddgv_ab = dgv_m[ga] - dgv_m[gb]
            dgWv2 = (du.unsqueeze(-1) * dgv_ab).sum(1) + (u.unsqueeze(-1) * ddgv_ab).sum(1)
            dgW = torch.cat([dgWs, dgWv1, dgWv2], dim=-1)
            acc(tg, mpre + ".2.weight", _tmm(dgW, lay["h"]) + _tmm(gW, dh))
            acc(tg, mpre + ".2.bias", dgW.sum(0))
            dgh = dgW @ W2
            dga1 = dgh * _silu1(lay["a1"]) + gh * _silu2(lay["a1"]) * da1
            acc(tg, mpre + ".0.weight", _tmm(dga1, c["e"]) + _tmm(ga1, de_t))
            acc(tg, mpre + ".0.bias", dga1.sum(0))
            if rbf_grads:
                dge = dga1 @ W1
                acc(tg, "rbf.centers", (dge * e_mu + ge * de_mu * dd.unsqueeze(-1)).sum(0))
                acc(tg, "rbf.widths", (dge * e_w + ge * de_w * dd.unsqueeze(-1)).sum(0))
            dgs = dgs_m + _add_pairs(n, ga, gb, dgs_m[gb] * Ws + gs_m[gb] * dWs,
                                     dgs_m[ga] * Ws + gs_m[ga] * dWs)
            dgv = dgv_m + _add_pairs(n, ga, gb,
                                     dgv_m[gb] * Wv1.unsqueeze(1) + gv_m[gb] * dWv1.unsqueeze(1),
                                     dgv_m[ga] * Wv1.unsqueeze(1) + gv_m[ga] * dWv1.unsqueeze(1))
        gs, gv = gs_new, gv_new
    if want_grads:
        emb = torch.zeros_like(p["embedding.weight"])
        emb.index_add_(0, c["Z"], wa(gs).to(emb.dtype))
        acc(grads, "embedding.weight", emb)
    if T:
        demb = torch.zeros_like(p["embedding.weight"])
        demb.index_add_(0, c["Z"], dgs.to(demb.dtype))
        acc(tg, "embedding.weight", demb)

    forces = None
    if want_forces:
        vec, d, den = c["vec"], c["d"], c["den"]
        vdot = (vec * gu).sum(1)
        G = (gd / d).unsqueeze(1) * vec + gu / den.unsqueeze(1) \
            - vec * (vdot / (d * den * den)).unsqueeze(1)
        forces = torch.zeros(n, 3, dtype=dt, device=vec.device)
        forces.index_add_(0, g.lo, G)
        forces.index_add_(0, g.hi, -G)
    return grads, forces, (tg if T else None)


# ---------------------------------------------------------------------------
# loss
# ---------------------------------------------------------------------------

def force_matching_grads(p: Dict[str, Tensor], Z: Tensor, x: Tensor, g: PairGraph, rc: float,
                         E_label: Tensor, F_label: Tensor, w_energy: float = 1.0,
                         w_force: float = 100.0, forces: Optional[Tensor] = None,
                         rbf_grads: bool = False):
    """loss = wE MSE(E) + wF MSE(F) (train_compact_models.py:340-342) and
    its exact gradient with respect to every parameter in ``p``.
    ``forces``: the model's forces at ``x`` when already known (the fused
    evaluation computes them far faster than the unit adjoint pass here)."""
    E, cache = forward(p, Z, x, g, rc)
    B = E.shape[0]
    ones = torch.ones_like(E)
    if forces is None:
        _, F, _ = reverse(p, cache, g, ones, lam=None, want_forces=True, want_grads=False)
    else:
        F = forces.to(E.dtype)
    eE = E - E_label
    eF = F - F_label
    loss = w_energy * (eE ** 2).mean() + w_force * (eF ** 2).mean()
    dLdE = w_energy * 2.0 * eE / B
    dLdF = w_force * 2.0 * eF / eF.numel()
    # one more pass gives both parts:
    #   d/dtheta of wE MSE(E): the unit adjoint's terms weighted by dL/dE (gw)
    #   d/dtheta of wF MSE(F) = d/dtheta < dL/dF, F > = - d/dtheta < dL/dF, grad_x sum E >
    gE, _, tg = reverse(p, cache, g, ones, lam=dLdF, want_forces=False, gw=dLdE,
                        rbf_grads=rbf_grads)
    grads = {}
    for name in p:
        a = gE.get(name)
        b = tg.get(name)
        if a is None and b is None:
            continue
        val = (a if a is not None else 0) - (b if b is not None else 0)
        grads[name] = val
    return loss, E, F, grads


