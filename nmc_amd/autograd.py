"""torch.autograd wrappers of the layered transport (PyTorch is used only here).

    rd = LayeredRd.apply(mua, mus, runner)
    absorbed = LayeredAbsorbed.apply(mua, mus, runner)   # runner with absorbed_grad = 1

mua, mus: float64 tensors [n_stacks, n_layers]; rd: float64 tensor
[n_stacks, nr+1], Rd(r) per launched photon.  forward() uploads the
properties and runs the transport (keeping the per-photon path records in
HBM); backward() runs the perturbation-MC backward kernel over those records.
The same photons are used by every forward call unless the runner's seed or
first photon id is changed (common random numbers across optimisation steps).
"""
from __future__ import annotations

import numpy as np
import torch

from . import params as P
from .layered import LayeredMC


class LayeredRd(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mua: torch.Tensor, mus: torch.Tensor, runner: LayeredMC):
        nl = mua.shape[1]
        full_mua = np.zeros((runner.n_stacks, P.MAX_LAYERS))
        full_mus = np.zeros((runner.n_stacks, P.MAX_LAYERS))
        full_mua[:, :nl] = mua.detach().cpu().double().numpy()
        full_mus[:, :nl] = mus.detach().cpu().double().numpy()
        runner.set_properties(full_mua, full_mus)
        runner.run()
        ctx.runner = runner
        ctx.epoch = runner.epoch
        ctx.nl = nl
        return torch.from_numpy(runner.rd_per_photon()).to(mua.device)

    @staticmethod
    def backward(ctx, grad_rd: torch.Tensor):
        if ctx.runner.epoch != ctx.epoch:
            raise RuntimeError("the runner has run or changed properties since this forward; "
                               "its path records belong to a later call (use one runner per "
                               "forward that is differentiated)")
        gmua, gmus = ctx.runner.backward(grad_rd.detach().cpu().double().numpy())
        nl = ctx.nl
        dev = grad_rd.device
        return (torch.from_numpy(gmua[:, :nl]).to(dev), torch.from_numpy(gmus[:, :nl]).to(dev), None)


def layered_rd(mua: torch.Tensor, mus: torch.Tensor, runner: LayeredMC) -> torch.Tensor:
    return LayeredRd.apply(mua, mus, runner)


class LayeredAbsorbed(torch.autograd.Function):
    """Absorbed weight per layer and launched photon, [n_stacks, n_layers],
    with its perturbation-MC sensitivities as the Jacobian of backward (the
    runner prepared with absorbed_grad = 1)."""

    @staticmethod
    def forward(ctx, mua: torch.Tensor, mus: torch.Tensor, runner: LayeredMC):
        nl = mua.shape[1]
        full_mua = np.zeros((runner.n_stacks, P.MAX_LAYERS))
        full_mus = np.zeros((runner.n_stacks, P.MAX_LAYERS))
        full_mua[:, :nl] = mua.detach().cpu().double().numpy()
        full_mus[:, :nl] = mus.detach().cpu().double().numpy()
        runner.set_properties(full_mua, full_mus)
        runner.run()
        n = np.array([s.n_photons for s in runner.stacks()], np.float64)[:, None]
        absorbed = runner.results()[1][:, :nl] / np.maximum(n, 1.0)
        da, ds = runner.absorbed_grad()
        ctx.jac = (torch.from_numpy(da[:, :nl, :nl] / np.maximum(n, 1.0)[:, :, None]),
                   torch.from_numpy(ds[:, :nl, :nl] / np.maximum(n, 1.0)[:, :, None]))
        return torch.from_numpy(absorbed).to(mua.device)

    @staticmethod
    def backward(ctx, grad_a: torch.Tensor):
        ja, js = ctx.jac
        g = grad_a.detach().cpu().double()
        dev = grad_a.device
        # dLoss/dtheta_t = sum_l dLoss/dA_l dA_l/dtheta_t
        return (torch.einsum("jl,jlt->jt", g, ja).to(dev), torch.einsum("jl,jlt->jt", g, js).to(dev),
                None)


def layered_absorbed(mua: torch.Tensor, mus: torch.Tensor, runner: LayeredMC) -> torch.Tensor:
    return LayeredAbsorbed.apply(mua, mus, runner)
