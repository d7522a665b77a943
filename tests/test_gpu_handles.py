"""Handle state on the GPU: a rejected tuning request changes nothing, and an
FP32-tile run is split into the launches tuning.split asks for."""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from nmc_amd._lib import NmcError
from nmc_amd.metal_mc import MetalMC
from oracle import oracle as O

from parity_util import assert_same_tallies

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(mc, w, n, first=0):
    mc.reset_tallies()
    mc.set_num_runs_seed(w.ext.seed, first)
    mc.send_compute_command(n)
    return mc.get_compute_results()


@pytest.mark.parametrize("bad", [dict(mode=2, block=128), dict(mode=7), dict(refill_min=0, block=96)])
def test_rejected_mcsub_tuning_keeps_the_handle(gpu, bad):
    w = W.mcsub_semi_infinite()
    n = 5000
    _say(f"[handles] mcsub rejected tuning {bad}")
    ref = O.run(w, n, threads=4)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, None, w.ext)
        mc.set_tuning(P.Tuning.make(mode=1, block=512))
        before = mc.get_tuning().as_dict()
        with pytest.raises(NmcError):
            mc.set_tuning(P.Tuning.make(**bad))
        assert mc.get_tuning().as_dict() == before
        got = _run(mc, w, n)
    assert_same_tallies(got, ref, "mcsub after a rejected tuning")


def test_rejected_mcsub_tuning_after_global_mode(gpu):
    """From global-replica tallies, a rejected switch keeps the replicas."""
    w = W.mcsub_semi_infinite()
    n = 5000
    _say("[handles] mcsub global mode, rejected tuning")
    ref = O.run(w, n, threads=4)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, None, w.ext)
        mc.set_tuning(P.Tuning.make(mode=2))
        before = mc.get_tuning().as_dict()
        assert before["mode"] == 2
        with pytest.raises(NmcError):
            mc.set_tuning(P.Tuning.make(mode=1, block=300))
        assert mc.get_tuning().as_dict() == before
        got = _run(mc, w, n)
    assert_same_tallies(got, ref, "mcsub global mode after a rejected tuning")


def test_rejected_mcxyz_tuning_keeps_the_handle(gpu):
    w = W.tiny_voxel_case("mismatch")
    n = 3000
    _say("[handles] mcxyz rejected tuning")
    ref = O.run(w, n, threads=4)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(P.Tuning.make(block=512, tile=(2, 2, 1, 8, 8, 6)))
        before = mc.get_tuning().as_dict()
        for bad in (dict(tile=(0, 0, 0, 400, 400, 400)), dict(mode=3), dict(block=2048),
                    dict(split=-1, mode=1)):
            with pytest.raises(NmcError):
                mc.set_tuning(P.Tuning.make(**bad))
            assert mc.get_tuning().as_dict() == before
        got = _run(mc, w, n)
    assert_same_tallies(got, ref, "mcxyz after a rejected tuning")


def test_fp32_tile_split_launches(gpu):
    """tuning.split photons per workgroup: a run of n photons is
    ceil(n / (split * grid)) launches of the same photons (integer tallies
    equal the oracle's), each flushing its FP32 tile in FP64."""
    w = W.skin_vessel()
    n = 40_000
    _say(f"[handles] fp32 tile split, n={n}")
    ref = O.run(w, n, threads=8)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(P.Tuning.make(mode=1))
        t = mc.get_tuning()
        assert t.split == 390625
        grid = t.blocks_per_cu * _num_cu()
        got1 = _run(mc, w, n)
        assert mc.last_launch_count() == 1
        mc.set_tuning(P.Tuning.make(mode=1, split=7))
        assert mc.get_tuning().split == 7
        got = _run(mc, w, n)
        assert mc.last_launch_count() == math.ceil(n / (7 * grid))
        # graph replays report the launches the graph holds
        mc.set_graph_mode(True)
        _run(mc, w, n)
        g2 = _run(mc, w, n)
        assert mc.last_launch_count() == math.ceil(n / (7 * grid))
        mc.set_graph_mode(False)
        # the FP64 tile is never split
        mc.set_tuning(P.Tuning.make(mode=0, split=7))
        assert mc.get_tuning().split == 0
        _run(mc, w, n)
        assert mc.last_launch_count() == 1
    for g in (got1, got, g2):
        for k in P.TOTAL_INT_FIELDS:
            assert g.totals[k] == ref.totals[k], k
        assert np.array_equal(g.det2d, ref.det2d)
        np.testing.assert_allclose(g.fluence, ref.fluence, rtol=2e-5, atol=1e-9)


def _num_cu() -> int:
    import torch
    return torch.cuda.get_device_properties(0).multi_processor_count


def test_add_results_accumulates_like_get_compute_results(gpu):
    """nmc_add_results adds into the caller's arrays (getComputeResults'
    semantics); nmc_get_results overwrites them."""
    w = W.tiny_voxel_case("mismatch", n_photons=2000)
    _say("[handles] add_results")
    with MetalMC(w.sim, 0) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command(2000)
        got = mc.get_compute_results()
        acc = mc.get_compute_results()
        acc.fluence[:] = 1.5
        acc.ryx[:] = 0.25
        acc.det2d[:] = 3
        before = dict(acc.totals)
        mc.add_compute_results(acc)
        mc.add_compute_results(acc)
    np.testing.assert_array_equal(acc.fluence, 1.5 + 2 * got.fluence)
    np.testing.assert_array_equal(acc.ryx, 0.25 + 2 * got.ryx)
    np.testing.assert_array_equal(acc.det2d, 3 + 2 * got.det2d)
    for k, v in got.totals.items():
        assert acc.totals[k] == before[k] + 2 * v, k
