"""GPU: per-tissue fluence sensitivities of the voxel walk (nmc_mcxyz_sens).

- the sensitivity kernel variant against the oracle's restatement on the same
  photons: the unscattered two-tissue column (whose mua sensitivities the
  oracle tests pin to the closed form), a scattering three-tissue grid, its
  two-term phase-function version, and the skin-vessel grid; integer
  tallies bit-exact, fluence and sensitivities to the FP64 summation-order
  limit;
- the fluence of a sensitivity run equals that of a plain run;
- at the benchmark's skin-vessel configuration (2e6 photons), the weight
  absorbed in each tissue: perturbation-MC sums against central differences
  in every tissue's mua with common random numbers (mua moves no path, so
  the two agree to 1 %);
- selection, reset and argument handling through the C-ABI;
- graph replays of a sensitivity run."""
from __future__ import annotations

import copy
import sys

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from nmc_amd._lib import NmcError
from nmc_amd.metal_mc import MetalMC
from oracle import oracle as O

from parity_util import assert_same_tallies
from test_oracle_voxel_sens import column, three_tissue

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_sens(w, ids, n, first=0):
    with MetalMC(w.sim, 0) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_sensitivity(ids)
        mc.set_num_runs_seed(w.ext.seed, first)
        mc.send_compute_command(n)
        t = mc.get_compute_results()
        da, ds = mc.sensitivity_results()
    return t, da, ds


def _close(a, b, label):
    scale = max(1e-300, float(np.abs(b).max()))
    err = np.abs(a - b)
    ok = err <= RTOL * np.abs(b) + 1e-13 * scale
    assert ok.all(), f"{label}: differs at {np.flatnonzero(~ok.ravel())[:10]}, max {err.max():.3e}"


def two_term_three_tissue(n):
    w = three_tissue(n)
    w.rp.mckernelflag = P.SIM_MCXYZ_TT
    w.sim = P.SIM_MCXYZ_TT
    for t in (1, 2, 3):
        w.tp.gf[t], w.tp.af[t], w.tp.gb[t], w.tp.ab[t], w.tp.CC[t] = 0.9, 0.5, 0.3, 0.5, 0.8
    return w


CASES = {"column": (column, [1, 2], 2000),
         "three_tissue": (three_tissue, [1, 2, 3], 20000),
         "two_term": (two_term_three_tissue, [3, 1], 20000),
         "skin_vessel": (lambda n: W.skin_vessel(n), [1, 2, 3, 4], 20000)}


@pytest.mark.parametrize("case", list(CASES))
def test_sensitivities_equal_oracle(gpu, case):
    make, ids, n = CASES[case]
    w = make(n)
    _say(f"[sens] {case} ids={ids} n={n}")
    ref, ra, rs = O.run_sens(w, ids, n, threads=8)
    got, ga, gs = gpu_sens(w, ids, n)
    assert_same_tallies(got, ref, case)
    _close(ga, ra, f"{case} d/dmua")
    _close(gs, rs, f"{case} d/dmus")
    assert np.abs(ra).max() > 0 and np.abs(rs).max() > 0


def test_sensitivity_run_keeps_the_fluence(gpu):
    w = three_tissue(20000)
    _say("[sens] fluence with and without sensitivities")
    with MetalMC(w.sim, 0) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(P.Tuning.make(mode=0))
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command(20000)
        plain = mc.get_compute_results()
        mc.reset_tallies()
        mc.set_sensitivity([2])
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command(20000)
        sens = mc.get_compute_results()
        da, ds = mc.sensitivity_results()
        assert np.abs(da).max() > 0
        # reset zeroes them, an empty selection turns the variant off
        mc.reset_tallies()
        assert np.all(mc.sensitivity_results()[0] == 0.0)
        mc.set_sensitivity([])
        with pytest.raises(NmcError):
            mc.sensitivity_results()
    assert_same_tallies(sens, plain, "sens vs plain")


def test_selection_checks(gpu):
    w = three_tissue(10)
    with MetalMC(w.sim, 0) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        for bad in ([1, 1], list(range(1, 10)), [19], [-1]):
            with pytest.raises(NmcError):
                mc.set_sensitivity(bad)
    ext = P.ExtConfig.default()
    ext.flags |= P.FLAG_NO_FLUENCE
    with MetalMC(w.sim, 0) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, ext)
        with pytest.raises(NmcError):
            mc.set_sensitivity([1])


def test_skin_vessel_mua_sensitivities_match_crn_central_difference(gpu):
    n, h = 2_000_000, 0.01
    w = W.skin_vessel(n)
    v = np.asarray(w.v).reshape(w.rp.Nz, w.rp.Nx, w.rp.Ny)
    tissues = (1, 2, 3, 4)

    def by_tissue(F):
        F = F.reshape(v.shape)
        return np.array([F[v == k].sum() for k in tissues])

    _say(f"[sens] skin-vessel {n} photons, CRN central differences in mua")
    t, da, ds = gpu_sens(w, list(tissues), n)
    pm = np.array([by_tissue(da[j]) for j in range(4)]) / n
    fd = []
    for t_id in tissues:
        row = []
        for f in (1 + h, 1 - h):
            w2 = copy.deepcopy(w)
            w2.tp.muav[t_id] = w.tp.muav[t_id] * f
            with MetalMC(w2.sim, 0) as mc:
                mc.prepare_data(w2.rp, w2.tp, w2.v, w2.ext)
                mc.set_tuning(P.Tuning.make(mode=0))
                mc.set_num_runs_seed(w2.ext.seed, 0)
                mc.send_compute_command(n)
                row.append(by_tissue(mc.get_compute_results().fluence))
        fd.append((row[0] - row[1]) / (2 * h * w.tp.muav[t_id]) / n)
    fd = np.array(fd)
    _say(f"[sens] pMC\n{pm}\nFD\n{fd}")
    big = np.abs(fd) > 1e-4 * np.abs(fd).max()
    assert np.all(np.abs(pm - fd)[big] <= 0.01 * np.abs(fd)[big]), (pm, fd)


def test_sensitivities_in_graph_replays(gpu):
    """Two graph replays of a sensitivity run equal the oracle on both photon
    ranges together (the replays advance the photon ids)."""
    w = three_tissue(10000)
    _say("[sens] graph replays 2 x 10000")
    ref, ra, rs = O.run_sens(w, [1, 3], 20000, threads=8)
    with MetalMC(w.sim, 0) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_sensitivity([1, 3])
        mc.set_graph_mode(True)
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command(10000)
        mc.send_compute_command(10000)
        got = mc.get_compute_results()
        ga, gs = mc.sensitivity_results()
    assert_same_tallies(got, ref, "graph sens")
    _close(ga, ra, "graph d/dmua")
    _close(gs, rs, "graph d/dmus")
