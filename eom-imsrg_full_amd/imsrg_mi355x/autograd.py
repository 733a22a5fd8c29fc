"""Gradient of the Magnus-flow energy with respect to the Hamiltonian.

``flow_e0(h, ctx)`` returns E0(s) after ``nsteps`` Magnus steps (White or atan
generator) as a torch scalar whose gradient with respect to ``h`` (the bare
Hamiltonian as a vectorize-order stream: E0, one-body blocks, two-body
J-coupled matrix elements) comes from the device adjoint flow
(imsrg_flow_gradient): adjoint commutators through every BCH and Magnus
series, Omega(s) checkpointed and recomputed.  PyTorch is only the interface;
the forward and the backward both run in the HIP library.
"""
from __future__ import annotations

import numpy as np
import torch

from . import engine as E


def energy_and_gradient(ctx, h_bare, kind="white", ds=0.5, nsteps=8, ncheck=8):
    """(E0(s_N), dE0/dh) for the bare Hamiltonian stream h_bare (numpy)."""
    H = ctx.op(1, np.asarray(h_bare, dtype=np.float64)).normal_order()
    e0, Hb = ctx.flow_gradient(H, kind, ds, nsteps, ncheck)
    return e0, Hb.normal_order_adjoint().download_cotangent()


def spe_indices(basis):
    """positions of the single-particle energies f_ii in the vectorize stream"""
    nA, nP = len(basis.holes), len(basis.parts)
    hh = [1 + i * nA + i for i in range(nA)]
    pp = [1 + nA * nA + i * nP + i for i in range(nP)]
    return np.array(hh + pp)


class MagnusFlowE0(torch.autograd.Function):
    @staticmethod
    def forward(fctx, h, ctx, kind, ds, nsteps, ncheck):
        e0, g = energy_and_gradient(ctx, h.detach().cpu().numpy(), kind, ds, nsteps, ncheck)
        fctx.save_for_backward(torch.from_numpy(g))
        return torch.tensor(e0, dtype=h.dtype, device=h.device)

    @staticmethod
    def backward(fctx, gout):
        (g,) = fctx.saved_tensors
        return gout * g.to(device=gout.device, dtype=gout.dtype), None, None, None, None, None


def flow_e0(h, ctx, kind="white", ds=0.5, nsteps=8, ncheck=8):
    """E0 after nsteps Magnus steps, differentiable with respect to h."""
    return MagnusFlowE0.apply(h, ctx, kind, ds, nsteps, ncheck)
