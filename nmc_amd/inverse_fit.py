"""Batched inverse fit: recover per-layer mua/mus of many layered instances
from noisy Rd(r) profiles with Adam on perturbation-MC gradients.

Every optimisation step is ONE forward launch over all instances (their
uneven photon budgets share one device queue) plus ONE backward launch over
the stored path records.  The target is the forward model at the hidden
parameters, run on a disjoint photon-id range, plus 1 % Gaussian noise.

    python -m nmc_amd.inverse_fit --instances 256 --steps 50
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import params as P
from . import workloads as W
from .layered import LayeredMC

TARGET_FIRST_PHOTON = 1 << 40   # photon ids of the target run (disjoint from the fit's)


def _props(stacks, nl):
    mua = np.array([[s.layer[l].mua for l in range(nl)] for s in stacks], np.float64)
    mus = np.array([[s.layer[l].mus for l in range(nl)] for s in stacks], np.float64)
    return mua, mus


def make_target(runner: LayeredMC, seed: int, noise: float = 0.01) -> np.ndarray:
    runner.set_seed(runner.cfg.seed, TARGET_FIRST_PHOTON)
    runner.run()
    rd = runner.rd_per_photon()
    runner.set_seed(runner.cfg.seed, 0)
    rng = np.random.default_rng(seed + 1000)
    return rd * (1.0 + noise * rng.standard_normal(rd.shape))


def fit(n_inst: int = 256, steps: int = 50, budget_scale: float = 1.0, lr: float = 0.05,
        seed: int = 5, init_spread: float = 0.4, device: int = 0, verbose: bool = False,
        plain_compare: bool = True) -> dict:
    cfg, hidden = W.inverse_fit_instances(n_inst, seed, budget_scale)
    nl = hidden[0].n_layers
    true_mua, true_mus = _props(hidden, nl)
    with LayeredMC(cfg, hidden, device=device) as runner:
        return _fit(runner, cfg, hidden, nl, true_mua, true_mus, steps, lr, seed, init_spread,
                    device, verbose, plain_compare)


def _fit(runner, cfg, hidden, nl, true_mua, true_mus, steps, lr, seed, init_spread, device,
         verbose, plain_compare) -> dict:
    import torch

    from .autograd import layered_rd

    target = make_target(runner, seed)
    # identifiability at the hidden parameters, from the target run's path
    # records, with the target's own noise: 1 % of each bin's noise-free value
    # (bins no light reaches carry no information)
    rd0 = runner.rd_per_photon()
    crb = cramer_rao(runner, np.where(rd0 > 0, 0.01 * rd0, np.inf), nl)
    rng = np.random.default_rng(seed + 2000)
    start_mua = true_mua * np.exp(rng.uniform(-init_spread, init_spread, true_mua.shape))
    start_mus = true_mus * np.exp(rng.uniform(-init_spread, init_spread, true_mus.shape))
    la = torch.tensor(np.log(start_mua), requires_grad=True)
    ls = torch.tensor(np.log(start_mus), requires_grad=True)
    tgt = torch.from_numpy(target)
    sigma = torch.from_numpy(fit_sigma(target))
    opt = torch.optim.Adam([la, ls], lr=lr)
    history = []
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for step in range(steps):
        opt.zero_grad()
        rd = layered_rd(torch.exp(la), torch.exp(ls), runner)
        kernel_ms += runner.last_kernel_ms()
        loss = (((rd - tgt) / sigma) ** 2).mean()
        loss.backward()
        kernel_ms += runner.last_kernel_ms()
        opt.step()
        history.append(loss.item())
        if verbose:
            print(f"step {step} loss {history[-1]:.4g}", flush=True)
    wall = time.perf_counter() - t0
    fit_mua, fit_mus = np.exp(la.detach().numpy()), np.exp(ls.detach().numpy())
    err_a = np.abs(fit_mua - true_mua) / true_mua
    err_s = np.abs(fit_mus - true_mus) / true_mus
    start_err_a = np.abs(start_mua - true_mua) / true_mua
    start_err_s = np.abs(start_mus - true_mus) / true_mus
    out = {
        "instances": len(hidden), "layers": nl, "steps": steps,
        "photons_per_step": int(sum(s.n_photons for s in hidden)),
        "loss_first": history[0], "loss_last": history[-1],
        "median_rel_err_mua_per_layer": np.median(err_a, axis=0).tolist(),
        "median_rel_err_mus_per_layer": np.median(err_s, axis=0).tolist(),
        "median_start_err_mua_per_layer": np.median(start_err_a, axis=0).tolist(),
        "median_start_err_mus_per_layer": np.median(start_err_s, axis=0).tolist(),
        "median_crb_rel_mua_per_layer": crb["median_crb_rel_mua_per_layer"],
        "median_crb_rel_mus_per_layer": crb["median_crb_rel_mus_per_layer"],
        # share of instances whose error is within 3 times their own bound
        "share_within_3crb_mua_per_layer": np.mean(err_a <= 3 * crb["crb_rel"][:, :nl], axis=0).tolist(),
        "share_within_3crb_mus_per_layer": np.mean(err_s <= 3 * crb["crb_rel"][:, nl:], axis=0).tolist(),
        "wall_s": wall, "wall_s_per_step": wall / steps, "kernel_ms_per_step": kernel_ms / steps,
    }
    if plain_compare:
        out.update(plain_step_time(cfg, hidden, device))
        out["batched_speedup_per_step"] = out["plain_wall_s_per_step"] / out["wall_s_per_step"]
    return out


def fit_sigma(target: np.ndarray, noise: float = 0.01) -> np.ndarray:
    """Per-bin standard deviation the fit's loss assumes: the target's relative
    noise plus a floor of 1e-3 of each profile's largest bin."""
    return noise * np.abs(target) + 1e-3 * np.abs(target).max(axis=1, keepdims=True) + 1e-12


def cramer_rao(runner: LayeredMC, sigma: np.ndarray, nl: int) -> dict:
    """Identifiability of each layer's mua and mus from Rd(r) with noise sigma
    [n_stacks, nr+1] at the runner's current properties (its last run's path
    records): the Fisher information F = J^T diag(sigma^-2) J of the
    log-parameters, with the Jacobian J = dRd(r) / dlog(theta) assembled from
    one perturbation-MC backward call per Rd bin.  sqrt(diag(F^-1)) bounds
    from below the standard deviation of any unbiased estimate of
    log(theta): the relative error no fit of this data can beat."""
    ns, nb = sigma.shape
    mua = np.array([[s.layer[l].mua for l in range(nl)] for s in runner.stacks()])
    mus = np.array([[s.layer[l].mus for l in range(nl)] for s in runner.stacks()])
    J = np.zeros((ns, nb, 2 * nl))
    for k in range(nb):
        g = np.zeros((ns, nb))
        g[:, k] = 1.0
        gmua, gmus = runner.backward(g)
        J[:, k, :nl] = gmua[:, :nl] * mua
        J[:, k, nl:] = gmus[:, :nl] * mus
    Jw = J / sigma[:, :, None]
    F = np.einsum("skp,skq->spq", Jw, Jw)
    crb = np.full((ns, 2 * nl), np.inf)
    for j in range(ns):
        try:
            crb[j] = np.sqrt(np.clip(np.diag(np.linalg.inv(F[j])), 0.0, None))
        except np.linalg.LinAlgError:
            pass
    return {"median_crb_rel_mua_per_layer": np.median(crb[:, :nl], axis=0).tolist(),
            "median_crb_rel_mus_per_layer": np.median(crb[:, nl:], axis=0).tolist(),
            "crb_rel": crb}


def plain_step_time(cfg: P.LayeredConfig, stacks, device: int = 0) -> dict:
    """One optimisation step done the unbatched way: every instance its own
    forward and backward launch."""
    t0 = time.perf_counter()
    for j, s in enumerate(stacks):
        with LayeredMC(cfg, [s], device=device) as r:
            r.run()
            r.backward(np.ones((1, cfg.nr + 1)))
    return {"plain_wall_s_per_step": time.perf_counter() - t0}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--budget-scale", type=float, default=1.0)
    ap.add_argument("--lr", type=float, default=0.05)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    r = fit(a.instances, a.steps, a.budget_scale, a.lr, verbose=a.verbose,
            plain_compare=not a.no_plain)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
