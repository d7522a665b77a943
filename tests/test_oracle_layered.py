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

import fd_util
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


THIN5 = dict(thick=[0.01, 0.015, 0.02, 0.03, 0.0], mua=[0.5, 1.0, 2.0, 1.0, 0.5],
             mus=[100.0, 150.0, 120.0, 200.0, 100.0], n=[1.4, 1.38, 1.42, 1.36, 1.4])


def thin_stack(n_photons):
    """Five thin two-term layers that detected light visits down to the
    semi-infinite fifth: every layer's gradient is non-zero."""
    t = THIN5
    return W.make_stack(t["thick"], t["mua"], t["mus"], t["n"], n_photons, phase=W.TWO_TERM_SKIN)


def _oracle_fd_check(cfg, stack, K, m, label):
    """pMC (oracle backward) vs central FD with common random numbers per
    batch of m photons (fd_util has the bar)."""
    nl = stack.n_layers
    pa, ps, fa, fs = [], [], [], []
    gr = fd_util.grad_rd_groups(cfg.nr)
    for k in range(K):
        base = P.Stack.from_buffer_copy(stack)
        base.n_photons = m
        o = O.run_layered(cfg, [base], first=k * m, threads=8)
        ga, gs = np.zeros((len(gr), nl)), np.zeros((len(gr), nl))
        for i in range(len(gr)):
            gmua, gmus = O.layered_backward(o, gr[i:i + 1])
            ga[i], gs[i] = gmua[0, :nl], gmus[0, :nl]
        pert = fd_util.perturbed(base, nl)
        op = O.run_layered(cfg, pert, first=k * m, threads=8, records=False)
        f1, f2 = fd_util.fd_groups(op.rd / m, base, nl)
        pa.append(ga), ps.append(gs), fa.append(f1), fs.append(f2)
    ra = fd_util.judge(np.array(pa), np.array(fa), label + " mua")
    rs = fd_util.judge(np.array(ps), np.array(fs), label + " mus")
    print(fd_util.table(label + " dRd/dmua", ra))
    print(fd_util.table(label + " dRd/dmus", rs))
    return ra, rs


def test_pmc_matches_central_difference_per_layer_thin_stack():
    """All five layers, both parameters, five Rd(r) groups, N = 8 x 25000:
    dRd/dmua meets the 1 % bar on every group the noise floor resolves
    (continuous absorption leaves the paths unchanged under mua, so the CRN
    difference is nearly noise-free); dRd/dmus is held to its measured noise
    floor (the score estimator's own variance is 7-40 % per component here:
    the 1 % bar for mus is checked at 4e8 photons on the GPU,
    tests/test_gpu_layered.py)."""
    cfg = P.LayeredConfig.default(nr=200, dr=0.005)
    ra, rs = _oracle_fd_check(cfg, thin_stack(1), 8, 25_000, "thin")
    assert ra["counts"]["1%"] >= 20 and ra["counts"]["zero"] == 0


def test_pmc_matches_central_difference_per_layer_c4_stack():
    """The C4 stack itself (multilayer_5): its 0.19 cm first layer (mua 3.8)
    keeps detected light out of layers 3-5, whose gradients are exactly 0 on
    both sides; layer 1's dRd/dmua meets the 1 % bar on the resolved groups."""
    cfg, stacks = W.multilayer_5(n_photons=1)
    ra, rs = _oracle_fd_check(cfg, stacks[0], 8, 12_500, "C4")
    assert ra["counts"]["zero"] >= 15 and ra["counts"]["1%"] >= 3


# ------------------------------------------- absorbed-weight sensitivities ---
def test_absorbed_sensitivity_to_mua_non_scattering_slab():
    """A non-scattering slab with matched indices: every photon crosses it
    straight, absorbing A = 1 - e^{-mua d}, so dA/dmua = d e^{-mua d} exactly,
    photon by photon."""
    d, mua, n = 0.05, 7.0, 20_000
    cfg = P.LayeredConfig.default(nr=10, dr=1.0, records=0)
    st = W.make_stack([d, 0.0], [mua, 1.0], [0.0, 100.0], [1.0, 1.0], n, g=0.9)
    st.n_layers = 1
    st.layer[0].thickness = d
    o = O.run_layered(cfg, [st], threads=4, records=False, sens=True)
    assert o.absorbed[0, 0] / n == pytest.approx(1 - math.exp(-mua * d), rel=1e-6)
    assert o.sens_mua[0, 0, 0] / n == pytest.approx(d * math.exp(-mua * d), rel=1e-6)


def test_absorbed_sensitivity_to_mus_vanishes_without_deflection(tmp_path_factory):
    """g = 1: scattering never changes a path, so the absorbed weight cannot
    depend on mus: the estimator's mean is 0 within its noise.  Without the
    score's growth inside each segment (planted mistake 5) it is not."""
    n, K = 800_000, 16
    cfg = P.LayeredConfig.default(nr=10, dr=1.0, records=0)

    def per_batch(L=None):
        out = []
        for k in range(K):
            st = W.make_stack([0.02, 0.03], [3.0, 5.0], [80.0, 150.0], [1.0, 1.0], n // K, g=1.0)
            o = O.run_layered(cfg, [st], first=k * (n // K), threads=8, records=False, sens=True, L=L)
            out.append(o.sens_mus[0, :2, :2].ravel() / (n // K))
        return np.array(out)

    ok = per_batch()
    # light reaches layer 1 after layer 0, so d(absorbed in 0)/dmus_1 is 0
    assert not ok[:, 1].any()
    use = [0, 2, 3]
    z = ok.mean(0)[use] / (ok.std(0, ddof=1)[use] / math.sqrt(K))
    assert np.abs(z).max() < 4.5, z
    bad = per_batch(O.mutant(5, str(tmp_path_factory.mktemp("m5"))))
    zb = bad.mean(0)[use] / (bad.std(0, ddof=1)[use] / math.sqrt(K))
    assert np.abs(zb).max() > 10, "the g = 1 pin misses the within-segment term"


def _oracle_absorbed_fd_check(stack, K, m, label):
    """Absorbed weight per layer: perturbation-MC sensitivities against
    central FD with common random numbers, per batch (fd_util's bar with the
    absorbing layer as the group)."""
    cfg = P.LayeredConfig.default(nr=10, dr=1.0, records=0)
    nl = stack.n_layers
    pa, ps, fa, fs = [], [], [], []
    for k in range(K):
        base = P.Stack.from_buffer_copy(stack)
        base.n_photons = m
        o = O.run_layered(cfg, [base], first=k * m, threads=8, records=False, sens=True)
        pa.append(o.sens_mua[0, :nl, :nl] / m)
        ps.append(o.sens_mus[0, :nl, :nl] / m)
        op = O.run_layered(cfg, fd_util.perturbed(base, nl), first=k * m, threads=8, records=False)
        f1, f2 = fd_util.fd_values(op.absorbed[:, :nl] / m, base, nl)
        fa.append(f1), fs.append(f2)
    ra = fd_util.judge(np.array(pa), np.array(fa), label + " dA/dmua")
    rs = fd_util.judge(np.array(ps), np.array(fs), label + " dA/dmus")
    print(fd_util.table(label + " dA_l/dmua_t (group = l)", ra))
    print(fd_util.table(label + " dA_l/dmus_t (group = l)", rs))
    return ra, rs


def test_absorbed_sensitivities_match_central_difference_thin_stack():
    """d(absorbed in layer l)/d(mua_t, mus_t) for all 25 layer pairs of the
    thin stack, N = 8 x 25000, against CRN central differences."""
    ra, rs = _oracle_absorbed_fd_check(thin_stack(1), 8, 25_000, "thin")
    assert ra["counts"]["1%"] >= 20
