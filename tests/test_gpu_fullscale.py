"""GPU at the benchmark sizes, through properties that do not depend on
running the oracle over the same photons: weight balance, the fluence
checksum, mirror symmetry, and Rd against an independent oracle sample
(disjoint photon ids) within its statistical error and within 0.5 %."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from nmc_amd.layered import LayeredMC
from nmc_amd.metal_mc import MetalMC, simulate
from oracle import oracle as O

from parity_util import energy_balance

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_rd(w, batches: int, n: int, first: int):
    """Rd per photon of `batches` independent oracle batches of n photons."""
    vals = []
    for b in range(batches):
        t = O.run(w, n, first=first + b * n, threads=16, fluence=False).totals
        vals.append(t["rd"] / n)
    return np.array(vals)


def test_skin_vessel_1e8_properties(gpu):
    w = W.skin_vessel()
    n = 100_000_000
    _say(f"[fullscale] mcxyz skin-vessel n={n}")
    got = simulate(w, n)
    t = got.totals
    assert t["photons"] == n
    assert abs(energy_balance(t) - 1.0) < 1e-6
    # every deposit inside the grid lands in exactly one voxel
    assert got.fluence.sum() == pytest.approx(t["absorbed"], rel=1e-9)
    assert got.ryx.sum() == pytest.approx(t["rd"], rel=1e-9)
    assert int(got.det2d.sum()) == t["detections"]
    # the geometry is mirror-symmetric in y (vessel along x, beam on axis)
    ryx = got.ryx.reshape(w.rp.Nx, w.rp.Ny)
    lo, hi = ryx[:, : w.rp.Ny // 2].sum(), ryx[:, w.rp.Ny // 2:].sum()
    assert abs(lo - hi) < 0.01 * (lo + hi)
    rd = t["rd"] / n
    _say("[fullscale] oracle: 32 independent batches of 100000 photons")
    ref = _oracle_rd(w, 32, 100_000, first=1 << 40)
    sem = ref.std(ddof=1) / np.sqrt(len(ref))
    assert abs(rd - ref.mean()) < 4.0 * sem + 1e-12, (rd, ref.mean(), sem)
    assert abs(rd - ref.mean()) < 0.005 * ref.mean()


def test_skin_vessel_1e8_fp32_tile_properties(gpu):
    """The benchmark's configuration: FP32 LDS tile (sums within a workgroup
    in FP32, FP64 above), 1e8 photons."""
    from nmc_amd import params as P
    from nmc_amd.metal_mc import MetalMC

    w = W.skin_vessel()
    n = 100_000_000
    _say(f"[fullscale] mcxyz skin-vessel, FP32 tile, n={n}")
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(P.Tuning.make(mode=1))
        assert mc.get_tuning().mode == 1
        mc.set_num_runs_seed(1, 0)
        mc.send_compute_command(n)
        got = mc.get_compute_results()
    t = got.totals
    assert t["photons"] == n
    assert abs(energy_balance(t) - 1.0) < 1e-6
    # FP32 partial sums per workgroup: the grid total matches the FP64
    # absorbed tally to FP32 accuracy
    assert got.fluence.sum() == pytest.approx(t["absorbed"], rel=1e-5)
    assert got.ryx.sum() == pytest.approx(t["rd"], rel=1e-9)
    ryx = got.ryx.reshape(w.rp.Nx, w.rp.Ny)
    lo, hi = ryx[:, : w.rp.Ny // 2].sum(), ryx[:, w.rp.Ny // 2:].sum()
    assert abs(lo - hi) < 0.01 * (lo + hi)
    F = got.fluence.reshape(w.rp.Nz, w.rp.Nx, w.rp.Ny)
    a, b = F[:, :, : w.rp.Ny // 2].sum(), F[:, :, w.rp.Ny // 2:][:, :, ::-1].sum()
    assert abs(a - b) < 0.01 * (a + b)
    rd = t["rd"] / n
    _say("[fullscale] oracle: 32 independent batches of 100000 photons")
    ref = _oracle_rd(w, 32, 100_000, first=1 << 40)
    sem = ref.std(ddof=1) / np.sqrt(len(ref))
    assert abs(rd - ref.mean()) < 4.0 * sem + 1e-12, (rd, ref.mean(), sem)
    assert abs(rd - ref.mean()) < 0.005 * ref.mean()


def test_mcsub_1e8_properties(gpu):
    w = W.mcsub_semi_infinite()
    n = 100_000_000
    _say(f"[fullscale] mcsub n={n}")
    got = simulate(w, n)
    t = got.totals
    assert abs(energy_balance(t) - 1.0) < 1e-6
    assert got.rd_ra.sum() == pytest.approx(t["rd"], rel=1e-9)
    assert got.a_rz.sum() == pytest.approx(t["absorbed"], rel=1e-9)
    assert t["specular"] / n == pytest.approx(((1.33 - 1.0) / 2.33) ** 2, rel=1e-6)
    _say("[fullscale] oracle: 32 independent batches of 100000 photons")
    ref = _oracle_rd(w, 32, 100_000, first=1 << 40)
    sem = ref.std(ddof=1) / np.sqrt(len(ref))
    rd = t["rd"] / n
    assert abs(rd - ref.mean()) < 4.0 * sem + 1e-12, (rd, ref.mean(), sem)
    assert abs(rd - ref.mean()) < 0.005 * ref.mean()


def test_multilayer_1e7_properties(gpu):
    cfg, stacks = W.multilayer_5()
    _say(f"[fullscale] layered 5-layer n={stacks[0].n_photons}")
    with LayeredMC(cfg, stacks) as mc:
        mc.run()
        rd, ab, tots, nrec = mc.results()
    t = tots[0]
    assert t["photons"] == stacks[0].n_photons
    assert abs(energy_balance(t) - 1.0) < 1e-6
    assert rd.sum() == pytest.approx(t["rd"], rel=1e-9)
    assert ab.sum() == pytest.approx(t["absorbed"], rel=1e-9)
    assert nrec == t["detections"]


def test_fp32_tile_long_run_is_split(gpu):
    """A run above 1e8 photons with the FP32 tile goes out as launches of at
    most 1e8 photons: the same photons as one FP64-tile launch (integer
    tallies equal), fluence within the FP32 tile's bound."""
    w = W.skin_vessel()
    n = 100_000_000 + 3_000_000
    _say(f"[fullscale] fp32 tile, one run of {n} photons")
    out = []
    for mode in (1, 0):
        with MetalMC(w.sim) as mc:
            mc.prepare_data(w.rp, w.tp, w.v, w.ext)
            mc.set_tuning(P.Tuning.make(mode=mode))
            mc.set_num_runs_seed(1, 0)
            mc.send_compute_command(n)
            out.append(mc.get_compute_results())
    f32, f64 = out
    for k in P.TOTAL_INT_FIELDS:
        assert f32.totals[k] == f64.totals[k], k
    assert f32.totals["photons"] == n
    assert np.array_equal(f32.det2d, f64.det2d)
    # FP32 sums of one workgroup's share of a <= 1e8-photon launch
    np.testing.assert_allclose(f32.fluence, f64.fluence, rtol=1e-4, atol=1e-12 * f64.fluence.max())
    assert abs(f32.fluence.sum() - f64.fluence.sum()) <= 1e-6 * f64.fluence.sum()


def _batch_sums(run_batch, K):
    """Sum and sum of squares over K batches of (fluence, ryx) and the count of
    batches in which each bin is non-zero."""
    acc = None
    for k in range(K):
        t = run_batch(k)
        arrs = (t.fluence, t.ryx)
        if acc is None:
            acc = [[np.zeros_like(a), np.zeros_like(a), np.zeros(a.shape, np.int32)] for a in arrs]
        for (s1, s2, nz), a in zip(acc, arrs):
            s1 += a
            s2 += a * a
            nz += a != 0
    return acc


def test_skin_vessel_1e7_benchmark_config_every_bin(gpu):
    """The benchmark's configuration (skin-vessel 200^3, FP32 LDS tile with
    the visits outside it as binned records, tuning.mode 2) at 1e7 photons,
    against the oracle with sigma per bin from K = 32 batch means (batches of
    312500 consecutive photon ids).
    (a) Same photons [0, 1e7): integer tallies bit-exact; totals Rd, Tt, A
        within 0.5 % (in fact to FP64 summation order); every fluence voxel
        and R(x,y) pixel within 3 sigma (and within the FP32 tile's 2e-5).
    (b) Disjoint photons, GPU [1e7, 2e7) against the oracle's [0, 1e7): the
        totals Rd, Tt, A within 4.5 sigma of each other (Tt, ~8e-5 of the
        light, carries ~1 % noise at this N); the per-bin z of the two totals,
        over bins non-zero in >= 30 of 32 batches on both sides, exceeds 3
        in no more than the t-tail share plus a binomial allowance."""
    from scipy import stats

    w = W.skin_vessel()
    K, m = 32, 312_500
    n = K * m
    _say(f"[fullscale] oracle, skin-vessel, {K} batches of {m}")
    ora_b = []

    def ora_batch(k):
        t = O.run(w, m, first=k * m, threads=16)
        ora_b.append(t.totals)
        return t
    ora = _batch_sums(ora_batch, K)

    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(P.Tuning.make(mode=2))
        _say(f"[fullscale] gpu fp32 tile + records, photons [0, {n})")
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command(n)
        same = mc.get_compute_results()
        gpu_b = []

        def gpu_batch(k):
            mc.reset_tallies()
            mc.set_num_runs_seed(w.ext.seed, n + k * m)
            mc.send_compute_command(m)
            t = mc.get_compute_results()
            gpu_b.append(t.totals)
            return t
        _say(f"[fullscale] gpu fp32 tile + records, {K} batches of disjoint photons")
        gpu = _batch_sums(gpu_batch, K)

    # (a) same photons
    ora_tot = {q: sum(b[q] for b in ora_b) for q in ora_b[0]}
    for q in P.TOTAL_INT_FIELDS:
        assert same.totals[q] == ora_tot[q], q
    for q in ("rd", "tt", "absorbed"):
        assert abs(same.totals[q] - ora_tot[q]) <= 0.005 * ora_tot[q], q
        assert same.totals[q] == pytest.approx(ora_tot[q], rel=1e-9), q
    for name, (s1, s2, _), a in (("fluence", ora[0], same.fluence), ("ryx", ora[1], same.ryx)):
        sig = np.sqrt(np.maximum(s2 - s1 * s1 / K, 0.0) * K / (K - 1))   # sd of the total
        err = np.abs(a - s1)
        assert (err <= 3 * sig + 1e-12 * s1.max()).all(), name
        np.testing.assert_allclose(a, s1, rtol=2e-5, atol=1e-9)

    # (b) independent photons
    for q in ("rd", "tt", "absorbed"):
        a = np.array([b[q] for b in gpu_b])
        b_ = np.array([b[q] for b in ora_b])
        sig = np.sqrt(K * (a.var(ddof=1) + b_.var(ddof=1)))
        _say(f"[fullscale] {q}: gpu {a.sum() / n:.6e} oracle {b_.sum() / n:.6e} "
             f"(rel {a.sum() / b_.sum() - 1:+.2e}, sigma {sig / b_.sum():.2e})")
        assert abs(a.sum() - b_.sum()) <= 4.5 * sig, q
    p_tail = 2 * stats.t.sf(3.0, df=2 * (K - 1))
    for name, o, g in (("fluence", ora[0], gpu[0]), ("ryx", ora[1], gpu[1])):
        use = (o[2] >= 30) & (g[2] >= 30)
        var = lambda s: np.maximum(s[1] - s[0] * s[0] / K, 0.0) * K / (K - 1)
        z = (g[0] - o[0])[use] / np.sqrt(var(g)[use] + var(o)[use])
        frac = float((np.abs(z) > 3).mean())
        allowed = p_tail + 5 * np.sqrt(p_tail / z.size)
        _say(f"[fullscale] {name}: {z.size} bins, |z| > 3 in {frac:.4%} (allowed {allowed:.4%}), "
             f"mean z {z.mean():+.3f}")
        assert frac <= allowed, (name, frac, allowed)
        assert abs(z.mean()) < 5 / np.sqrt(z.size) + 0.05, name
