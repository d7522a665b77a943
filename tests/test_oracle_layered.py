"""Pins for the oracle's layered transport and its perturbation-MC gradient
(CPU only): published slab values, the semi-infinite H-function limit, weight
balance, agreement with the independent mcsub estimator, and the pMC gradient
against central finite differences with common random numbers."""
from __future__ import annotations

import math

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from oracle import oracle as O

from test_oracle import chandrasekhar_R


def _balance(t):
    return (t["specular"] + t["rd"] + t["absorbed"] + t["tt"] + t["roulette_loss"] + t["capped"]
            - t["roulette_gain"]) / t["photons"]


@pytest.mark.parametrize("n_slab,mua", [(1.5, 20.0), (1.4, 5.0), (1.0, 10.0)])
def test_non_scattering_slab_multiple_reflections(n_slab, mua):
    """Non-scattering slab in air at normal incidence: with R = ((n-1)/(n+1))^2
    and a = exp(-mua d), the all-or-none interface sampling must reproduce
    Tt = (1-R)^2 a / (1 - R^2 a^2) and Rd = (1-R)^2 R a^2 / (1 - R^2 a^2)
    (the geometric series of internal reflections)."""
    d = 0.05
    cfg = P.LayeredConfig.default(nr=10, dr=1.0, records=0)
    st = W.make_stack([d], [mua], [0.0], [n_slab], 1_000_000, g=0.0, n_below=1.0)
    t = O.run_layered(cfg, [st], threads=8, records=False).totals[0]
    N = t["photons"]
    R = ((n_slab - 1) / (n_slab + 1)) ** 2
    a = math.exp(-mua * d)
    tt = (1 - R) ** 2 * a / (1 - R * R * a * a)
    rd = (1 - R) ** 2 * R * a * a / (1 - R * R * a * a)
    assert t["specular"] / N == pytest.approx(R, rel=1e-6, abs=1e-12)
    assert t["interactions"] == 0
    # each photon leaves with weight ~ (1-R) a^k: sigma <= (1-R) sqrt(p/N)
    assert abs(t["tt"] / N - tt) < 5 * math.sqrt(tt * (1 - R) / N) + 1e-6
    assert abs(t["rd"] / N - rd) < 5 * math.sqrt(max(rd, 1e-9) * (1 - R) / N) + 1e-6
    assert abs(_balance(t) - 1) < 1e-5


@pytest.mark.parametrize("albedo", [0.9, 0.99])
def test_semi_infinite_isotropic_matched(albedo):
    cfg = P.LayeredConfig.default(nr=10, dr=1.0, records=0)
    st = W.make_stack([0.0], [1.0 - albedo], [albedo], [1.0], 300_000, g=0.0)
    o = O.run_layered(cfg, [st], threads=8, records=False)
    t = o.totals[0]
    R = chandrasekhar_R(albedo)
    assert abs(t["rd"] / t["photons"] - R) < 4 * 0.5 / math.sqrt(t["photons"]) + 1e-4


def test_agrees_with_mcsub_estimator():
    """One semi-infinite layer is the mcsub medium; mcsub uses albedo weighting,
    the layered walk continuous absorption - independent estimators of the
    same Rd(r)."""
    n = 200_000
    w = W.mcsub_semi_infinite()
    w.ext.na = 1
    sub = O.run(w, n, threads=8)
    cfg = P.LayeredConfig.default(nr=100, dr=0.01, records=0)
    st = W.make_stack([0.0], [1.0], [100.0], [1.33], n, g=0.9, n_above=1.0)
    lay = O.run_layered(cfg, [st], threads=8, records=False)
    a, b = sub.totals["rd"] / n, lay.totals[0]["rd"] / n
    assert abs(a - b) < 4 * 0.5 * math.sqrt(2.0 / n), (a, b)
    assert sub.totals["specular"] / n == pytest.approx(lay.totals[0]["specular"] / n, rel=1e-6)
    # radial profile: compare the first 20 bins with a chi-square-like check
    ra, rb = sub.rd_ra[:20] / n, lay.rd[0, :20] / n
    sig = np.sqrt((sub.rd_ra[:20] + lay.rd[0, :20]) / n / n) + 1e-12
    assert np.mean(((ra - rb) / sig) ** 2) < 3.0


def test_multilayer_balance_and_records():
    cfg, stacks = W.multilayer_5(n_photons=20_000)
    o = O.run_layered(cfg, stacks, threads=8)
    t = o.totals[0]
    assert abs(_balance(t) - 1) < 1e-5
    assert o.n_records == t["detections"]
    # record weights sum to the Rd tally bin by bin
    per_bin = np.bincount(o.rec_bin[:o.n_records], weights=o.rec_w[:o.n_records].astype(np.float64),
                          minlength=cfg.nr + 1)
    np.testing.assert_allclose(per_bin, o.rd[0], rtol=1e-6)
    # recorded path lengths only in layers the stack has
    assert (o.rec_L[5:, :o.n_records] == 0).all()


def test_pmc_matches_central_difference_for_mua():
    """With continuous absorption, changing mua leaves trajectories unchanged
    (common random numbers), so FD of Rd and the pMC estimate must agree to
    far better than the 1% bar; roulette decisions are the only difference."""
    cfg, stacks = W.multilayer_5(n_photons=100_000)
    base = O.run_layered(cfg, stacks, threads=8)
    grad_rd = np.zeros((1, cfg.nr + 1))
    grad_rd[0, :40] = 1.0                      # d(sum of the first 40 bins)
    gmua, gmus = O.layered_backward(base, grad_rd)
    scale = np.abs(gmua[0, :5]).max()
    for l in range(5):
        h = 0.02 * stacks[0].layer[l].mua
        vals = []
        for sgn in (+1, -1):
            s = P.Stack.from_buffer_copy(stacks[0])
            s.layer[l].mua = stacks[0].layer[l].mua + sgn * h
            o = O.run_layered(cfg, [s], threads=8, records=False)
            vals.append(o.rd[0, :40].sum() / s.n_photons)
        fd = (vals[0] - vals[1]) / (2 * h)
        # roulette decisions that flip under the perturbation add zero-mean
        # noise; each layer is checked against the largest component
        assert abs(fd - gmua[0, l]) < 0.01 * abs(gmua[0, l]) + 0.01 * scale, (l, fd, gmua[0, l])


def test_pmc_matches_central_difference_for_mus_total():
    """mus changes the trajectories; with common random numbers the FD of the
    total Rd is still far less noisy than independent runs."""
    cfg, stacks = W.multilayer_5(n_photons=60_000)
    base = O.run_layered(cfg, stacks, threads=8)
    grad_rd = np.ones((1, cfg.nr + 1))
    _, gmus = O.layered_backward(base, grad_rd)
    l = 0
    h = 0.05 * stacks[0].layer[l].mus
    vals = []
    for sgn in (+1, -1):
        s = P.Stack.from_buffer_copy(stacks[0])
        s.layer[l].mus = stacks[0].layer[l].mus + sgn * h
        o = O.run_layered(cfg, [s], threads=8, records=False)
        vals.append(o.rd[0].sum() / s.n_photons)
    fd = (vals[0] - vals[1]) / (2 * h)
    assert fd == pytest.approx(gmus[0, l], rel=0.15, abs=2e-5), (fd, gmus[0, l])
