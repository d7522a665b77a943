"""Batched inverse fit: recover per-layer mua/mus of many layered instances
from noisy Rd(r) profiles with Adam on perturbation-MC gradients.

Every optimisation step is ONE forward launch over all instances (their
uneven photon budgets share one device queue) plus ONE backward launch over
the stored path records.  The target is the forward model at the hidden
parameters, run as 32 batches on disjoint photon-id ranges (together the
instance's budget), plus 1 % Gaussian noise per bin.  The spread of the batch
means gives each bin's Monte Carlo variance (pooled over neighbouring bins as
a variance-to-mean ratio, so that bins a handful of photons reach still get
an estimate with many degrees of freedom); the loss weights each bin by its
noise at the hidden parameters (target noise, the target's Monte Carlo
variance and the model run's own, which has the same budget).  The start
values are the hidden ones times exp(U(-s, s)), s = 0.4, so the loss also
holds that knowledge as a Gaussian prior on each log-parameter around its
start value (standard deviation s / sqrt 3): the fit is a maximum a
posteriori estimate, and the Cramer-Rao bound of every parameter is computed
for the same likelihood and prior (and for the data alone).  After the Adam
steps, a few Levenberg-Marquardt steps on the same posterior (the Jacobian
from one perturbation-MC backward call per Rd bin) take the estimate to the
optimum, which Adam's fixed step size approaches slowly on this stiff
least-squares problem; both results are reported.

    python -m nmc_amd.inverse_fit --instances 256 --steps 50
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import time

import numpy as np

from . import params as P
from . import workloads as W
from .layered import LayeredMC

TARGET_FIRST_PHOTON = 1 << 40   # photon ids of the target run (disjoint from the fit's)
TARGET_BATCHES = 32
POOL_BINS = 3      # the variance-to-mean ratio of bin k is pooled over k +- 3


def _props(stacks, nl):
    mua = np.array([[s.layer[l].mua for l in range(nl)] for s in stacks], np.float64)
    mus = np.array([[s.layer[l].mus for l in range(nl)] for s in stacks], np.float64)
    return mua, mus


def make_target(cfg: P.LayeredConfig, hidden, seed: int, device: int = 0, noise: float = 0.01,
                batches: int = TARGET_BATCHES):
    """Rd(r) per photon of every instance at the hidden parameters, as the
    mean of `batches` runs on disjoint photon-id ranges (each with 1/batches
    of the instance's budget), times (1 + noise N(0,1)) per bin.  Returns
    (target, clean mean, Monte Carlo variance of one full-budget run per bin).
    The variance comes from the spread of the batch means, pooled: a bin's
    variance is its mean times the ratio sum(var) / sum(mean) over the bins
    within POOL_BINS of it (a photon's exit weight, which sets that ratio,
    changes slowly with r)."""
    sub = []
    for s in hidden:
        t = P.Stack()
        C.memmove(C.addressof(t), C.addressof(s), C.sizeof(P.Stack))
        t.n_photons = max(1, s.n_photons // batches)
        sub.append(t)
    runs = []
    with LayeredMC(cfg, sub, device=device) as r:
        for k in range(batches):
            r.set_seed(cfg.seed, TARGET_FIRST_PHOTON + (k << 32))
            r.run()
            runs.append(r.rd_per_photon())
    runs = np.array(runs)
    mean = runs.mean(axis=0)
    var_raw = runs.var(axis=0, ddof=1) / batches
    kern = np.ones(2 * POOL_BINS + 1)
    pool = lambda a: np.apply_along_axis(lambda r: np.convolve(r, kern, mode="same"), 1, a)
    sm, sv = pool(mean), pool(var_raw)
    var_run = np.where(sm > 0, mean * sv / np.where(sm > 0, sm, 1.0), 0.0)
    rng = np.random.default_rng(seed + 1000)
    return mean * (1.0 + noise * rng.standard_normal(mean.shape)), mean, var_run


def noise_sigma(mean: np.ndarray, var_run: np.ndarray, noise: float = 0.01) -> np.ndarray:
    """Standard deviation of (model - target) per bin at the hidden parameters:
    the target's relative noise, its Monte Carlo variance and the model run's
    (the same budget, other photons).  Bins no photon reached carry no
    information: infinite."""
    var = (noise * mean) ** 2 + 2.0 * var_run
    return np.where(var > 0, np.sqrt(np.where(var > 0, var, 1.0)), np.inf)


def fit(n_inst: int = 256, steps: int = 50, budget_scale: float = 1.0, lr: float = 0.05,
        seed: int = 5, init_spread: float = 0.4, device: int = 0, verbose: bool = False,
        plain_compare: bool = True, lm_steps: int = 8) -> dict:
    cfg, hidden = W.inverse_fit_instances(n_inst, seed, budget_scale)
    nl = hidden[0].n_layers
    true_mua, true_mus = _props(hidden, nl)
    with LayeredMC(cfg, hidden, device=device) as runner:
        return _fit(runner, cfg, hidden, nl, true_mua, true_mus, steps, lr, seed, init_spread,
                    device, verbose, plain_compare, lm_steps)


def _fit(runner, cfg, hidden, nl, true_mua, true_mus, steps, lr, seed, init_spread, device,
         verbose, plain_compare, lm_steps) -> dict:
    import torch

    from .autograd import layered_rd

    target, clean, var_run = make_target(cfg, hidden, seed, device)
    sig = noise_sigma(clean, var_run)
    # identifiability at the hidden parameters (the runner's properties at
    # construction), from the path records of the fit's own photons, for the
    # noise the loss assumes
    prior_sd = init_spread / np.sqrt(3.0)
    runner.run()
    crb = cramer_rao(runner, sig, nl, prior_sd)
    rng = np.random.default_rng(seed + 2000)
    start_mua = true_mua * np.exp(rng.uniform(-init_spread, init_spread, true_mua.shape))
    start_mus = true_mus * np.exp(rng.uniform(-init_spread, init_spread, true_mus.shape))
    la0 = torch.tensor(np.log(start_mua))
    ls0 = torch.tensor(np.log(start_mus))
    la = la0.clone().requires_grad_(True)
    ls = ls0.clone().requires_grad_(True)
    tgt = torch.from_numpy(target)
    sigma = torch.from_numpy(sig)
    n_terms = float(target.size)
    opt = torch.optim.Adam([la, ls], lr=lr)
    history = []
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for step in range(steps):
        opt.zero_grad()
        rd = layered_rd(torch.exp(la), torch.exp(ls), runner)
        kernel_ms += runner.last_kernel_ms()
        # -2 log posterior per Rd bin: the noise-weighted residuals and the prior
        loss = ((((rd - tgt) / sigma) ** 2).sum() + (((la - la0) / prior_sd) ** 2).sum() +
                (((ls - ls0) / prior_sd) ** 2).sum()) / n_terms
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
        "median_crb_data_only_rel_mua_per_layer": crb["median_crb_data_only_rel_mua_per_layer"],
        "median_crb_data_only_rel_mus_per_layer": crb["median_crb_data_only_rel_mus_per_layer"],
        # share of instances whose error is within 3 times their own bound
        "share_within_3crb_mua_per_layer": np.mean(err_a <= 3 * crb["crb_rel"][:, :nl], axis=0).tolist(),
        "share_within_3crb_mus_per_layer": np.mean(err_s <= 3 * crb["crb_rel"][:, nl:], axis=0).tolist(),
        "wall_s": wall, "wall_s_per_step": wall / steps, "kernel_ms_per_step": kernel_ms / steps,
    }
    if lm_steps > 0:
        th0 = np.concatenate([la0.numpy(), ls0.numpy()], axis=1)
        th = np.concatenate([la.detach().numpy(), ls.detach().numpy()], axis=1)
        t1 = time.perf_counter()
        th, lm_hist = levenberg_marquardt(runner, target, sig, th0, th, prior_sd, nl, lm_steps)
        lm_err_a = np.abs(np.exp(th[:, :nl]) - true_mua) / true_mua
        lm_err_s = np.abs(np.exp(th[:, nl:]) - true_mus) / true_mus
        out.update({
            "lm_steps": lm_steps, "lm_wall_s": time.perf_counter() - t1,
            "lm_loss_last": lm_hist[-1] / float(target.size),
            "lm_median_rel_err_mua_per_layer": np.median(lm_err_a, axis=0).tolist(),
            "lm_median_rel_err_mus_per_layer": np.median(lm_err_s, axis=0).tolist(),
            "lm_share_within_3crb_mua_per_layer":
                np.mean(lm_err_a <= 3 * crb["crb_rel"][:, :nl], axis=0).tolist(),
            "lm_share_within_3crb_mus_per_layer":
                np.mean(lm_err_s <= 3 * crb["crb_rel"][:, nl:], axis=0).tolist(),
        })
    if plain_compare:
        out.update(plain_step_time(cfg, hidden, device))
        out["batched_speedup_per_step"] = out["plain_wall_s_per_step"] / out["wall_s_per_step"]
    return out


def _run_at(runner: LayeredMC, th: np.ndarray, nl: int) -> np.ndarray:
    """Rd(r) per photon with every instance's log-parameters th [n, 2 nl]."""
    full_mua = np.zeros((runner.n_stacks, P.MAX_LAYERS))
    full_mus = np.zeros((runner.n_stacks, P.MAX_LAYERS))
    full_mua[:, :nl] = np.exp(th[:, :nl])
    full_mus[:, :nl] = np.exp(th[:, nl:])
    runner.set_properties(full_mua, full_mus)
    runner.run()
    return runner.rd_per_photon()


def _posterior(rd, target, w, th, th0, prior_sd):
    """-2 log posterior per instance (w = 1/sigma, 0 for bins without information)."""
    return (((rd - target) * w) ** 2).sum(axis=1) + (((th - th0) / prior_sd) ** 2).sum(axis=1)


def levenberg_marquardt(runner: LayeredMC, target, sigma, th0, th, prior_sd, nl, steps):
    """Levenberg-Marquardt on the posterior of every instance's log-parameters
    (independent 2 nl-parameter problems solved side by side): per step one
    forward run at the current values, the Jacobian from nr+1 backward calls,
    and one forward run at the proposed values; a step that does not lower an
    instance's posterior is not taken and its damping grows."""
    w = np.where(np.isfinite(sigma), 1.0 / np.where(np.isfinite(sigma), sigma, 1.0), 0.0)
    npar = 2 * nl
    lam = np.full(len(th), 1e-3)
    rd = _run_at(runner, th, nl)
    cost = _posterior(rd, target, w, th, th0, prior_sd)
    hist = [cost.sum()]
    for _ in range(steps):
        J = jacobian(runner, nl) * w[:, :, None]
        r = (rd - target) * w
        A = np.einsum("skp,skq->spq", J, J) + np.eye(npar) / prior_sd ** 2
        g = np.einsum("skp,sk->sp", J, r) + (th - th0) / prior_sd ** 2
        D = A + lam[:, None, None] * A * np.eye(npar)
        step = -np.linalg.solve(D, g[:, :, None])[:, :, 0]
        trial = th + np.clip(step, -0.5, 0.5)
        rd_t = _run_at(runner, trial, nl)
        cost_t = _posterior(rd_t, target, w, trial, th0, prior_sd)
        take = cost_t < cost
        th = np.where(take[:, None], trial, th)
        cost = np.where(take, cost_t, cost)
        lam = np.where(take, lam / 3.0, lam * 4.0)
        hist.append(cost.sum())
        # the next Jacobian needs the path records of a run at the kept values
        rd = rd_t if take.all() else _run_at(runner, th, nl)
    return th, hist


def jacobian(runner: LayeredMC, nl: int) -> np.ndarray:
    """dRd(r)/dlog(theta) [n, nr+1, 2 nl] at the runner's last run (its path
    records), one perturbation-MC backward call per Rd bin."""
    ns, nb = runner.n_stacks, runner.cfg.nr + 1
    mua = np.array([[s.layer[l].mua for l in range(nl)] for s in runner.stacks()])
    mus = np.array([[s.layer[l].mus for l in range(nl)] for s in runner.stacks()])
    J = np.zeros((ns, nb, 2 * nl))
    for k in range(nb):
        g = np.zeros((ns, nb))
        g[:, k] = 1.0
        gmua, gmus = runner.backward(g)
        J[:, k, :nl] = gmua[:, :nl] * mua
        J[:, k, nl:] = gmus[:, :nl] * mus
    return J


def cramer_rao(runner: LayeredMC, sigma: np.ndarray, nl: int, prior_sd: float | None = None) -> dict:
    """Identifiability of each layer's mua and mus from Rd(r) with noise sigma
    [n_stacks, nr+1] at the runner's current properties (its last run's path
    records): the Fisher information F = J^T diag(sigma^-2) J of the
    log-parameters, with the Jacobian J = dRd(r) / dlog(theta) assembled from
    one perturbation-MC backward call per Rd bin.  sqrt(diag(F^-1)) bounds
    from below the standard deviation of any unbiased estimate of
    log(theta): the relative error no fit of this data can beat.  With a
    Gaussian prior of standard deviation prior_sd on every log-parameter, F
    gains prior_sd^-2 on its diagonal (the Bayesian bound of the maximum a
    posteriori fit); the data-only bound is returned as well."""
    ns, nb = sigma.shape
    Jw = jacobian(runner, nl) / sigma[:, :, None]
    F = np.einsum("skp,skq->spq", Jw, Jw)

    def bound(F):
        crb = np.full((ns, 2 * nl), np.inf)
        for j in range(ns):
            try:
                crb[j] = np.sqrt(np.clip(np.diag(np.linalg.inv(F[j])), 0.0, None))
            except np.linalg.LinAlgError:
                pass
        return crb

    data = bound(F)
    crb = bound(F + np.eye(2 * nl) / prior_sd ** 2) if prior_sd else data
    return {"median_crb_rel_mua_per_layer": np.median(crb[:, :nl], axis=0).tolist(),
            "median_crb_rel_mus_per_layer": np.median(crb[:, nl:], axis=0).tolist(),
            "median_crb_data_only_rel_mua_per_layer": np.median(data[:, :nl], axis=0).tolist(),
            "median_crb_data_only_rel_mus_per_layer": np.median(data[:, nl:], axis=0).tolist(),
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
    ap.add_argument("--lm-steps", type=int, default=8)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    r = fit(a.instances, a.steps, a.budget_scale, a.lr, verbose=a.verbose,
            plain_compare=not a.no_plain, lm_steps=a.lm_steps)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
