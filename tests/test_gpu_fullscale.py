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
