"""GPU: batched mcsub launches, graph replays and launch tuning.

A batch item must reproduce, photon for photon, the plain run of its medium
(integer tallies bit-exact against the oracle with the item's stream); graph
replays must equal consecutive plain runs; and no launch-tuning choice (block
size, refill batch, queue chunk, LDS tile box) may change any result.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from nmc_amd.metal_mc import MetalMC
from oracle import oracle as O

from parity_util import RTOL, assert_same_tallies

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_item(w, it, n):
    """Oracle run of one work item as a plain mcsub configuration."""
    w.rp.n1, w.rp.n2 = it["n1"], it["n2"]
    w.tp.muav[1], w.tp.musv[1], w.tp.gv[1] = it["mua"], it["mus"], it["g"]
    w.ext.stream_id = it["stream_id"]
    return O.run(w, n, threads=4)


def test_batch_items_match_oracle(gpu):
    w, items, _ = W.spectral_sweep(n_wavelengths=24, budget_scale=0.003)
    items[3]["g"] = 0.0
    items[5]["n1"] = 1.0
    w.ext.flags = 0                      # keep A(r,z) too
    w.ext.na = 4
    _say(f"[batch] 24 items, {sum(i['n_photons'] for i in items)} photons")
    with MetalMC(P.SIM_MCSUB) as mc:
        mc.prepare_batch(w.rp, items, w.ext)
        mc.run_batch()
        tally, tots = mc.get_batch_results()
    n_ra = (w.ext.nr + 1) * w.ext.na
    for j in (0, 3, 5, 11, 23):
        it = items[j]
        ww = W.mcsub_semi_infinite()
        ww.ext.na = 4
        ref = _oracle_item(ww, it, it["n_photons"])
        for k in P.TOTAL_INT_FIELDS:
            assert tots[j][k] == ref.totals[k], (j, k, tots[j][k], ref.totals[k])
        for k in P.TOTAL_FLOAT_FIELDS:
            assert np.isclose(tots[j][k], ref.totals[k], rtol=RTOL, atol=1e-9), (j, k)
        np.testing.assert_allclose(tally[j, :n_ra], ref.rd_ra, rtol=RTOL, atol=1e-12)
        np.testing.assert_allclose(tally[j, n_ra:], ref.a_rz, rtol=RTOL, atol=1e-12)


def test_batch_uneven_budgets_all_consumed(gpu):
    w, items, _ = W.spectral_sweep(n_wavelengths=128, budget_scale=0.002)
    _say(f"[batch] 128 items, {sum(i['n_photons'] for i in items)} photons")
    with MetalMC(P.SIM_MCSUB) as mc:
        mc.prepare_batch(w.rp, items, w.ext)
        mc.run_batch()
        _, tots = mc.get_batch_results()
    for j, it in enumerate(items):
        assert tots[j]["photons"] == it["n_photons"], j
        bal = (tots[j]["specular"] + tots[j]["rd"] + tots[j]["absorbed"] + tots[j]["roulette_loss"]
               + tots[j]["capped"] - tots[j]["roulette_gain"]) / it["n_photons"]
        assert abs(bal - 1.0) < 1e-5, j


@pytest.mark.parametrize("placement", [0, 2])
def test_plain_mcsub_graph_replay(gpu, placement):
    """Replays equal consecutive photons, with the tallies in LDS (0 picks it
    here) and in the global replicas (2), whose reduction must leave them
    cleared for the next replay."""
    w = W.mcsub_semi_infinite()
    _say(f"[graph] mcsub 3 replays x 20000, placement {placement}")
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, None, w.ext)
        mc.set_tuning(P.Tuning.make(mode=placement))
        assert mc.get_tuning().mode == (1 if placement == 0 else 2)
        mc.set_graph_mode(True)
        for _ in range(3):
            mc.send_compute_command(20000)
        g = mc.get_compute_results()
    ref = O.run(w, 60000, threads=8)
    assert_same_tallies(g, ref, "graph")


def test_mcsub_async_budget_change(gpu):
    """Back-to-back asynchronous runs of different sizes: the second run's
    budget must not reach the first, still in flight."""
    w = W.mcsub_semi_infinite()
    _say("[async] mcsub 30000 + 10000 photons, no wait between")
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, None, w.ext)
        mc.send_compute_command_async(30000)
        mc.send_compute_command_async(10000)
        mc.synchronize()
        g = mc.get_compute_results()
    ref = O.run(w, 40000, threads=8)
    assert_same_tallies(g, ref, "async-budget")


def test_mcxyz_graph_replay(gpu):
    w = W.skin_vessel()
    _say("[graph] mcxyz 2 replays x 10000")
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_graph_mode(True)
        mc.send_compute_command(10000)
        mc.send_compute_command(10000)
        g = mc.get_compute_results()
    ref = O.run(w, 20000, threads=8)
    assert_same_tallies(g, ref, "graph-xyz")


def test_mcxyz_graph_seed_change(gpu):
    """A new seed between replays must reach the captured launch.  The two
    seeds differ in both key words yet agree in key0 ^ (key1 * 0x9E3779B9),
    so a graph cache keyed on a 32-bit mix of the key would replay the first
    seed's stream."""
    w = W.skin_vessel()
    seed_a = 5
    seed_b = (1 << 32) | (5 ^ 0x9E3779B9)
    _say("[graph] mcxyz seed change 2 x 8000")
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_graph_mode(True)
        mc.set_num_runs_seed(seed_a, 0)
        mc.send_compute_command(8000)
        mc.reset_tallies()
        mc.set_num_runs_seed(seed_b, 0)
        mc.send_compute_command(8000)
        g = mc.get_compute_results()
    w.ext.seed = seed_b
    ref = O.run(w, 8000, threads=8)
    assert_same_tallies(g, ref, "graph-seed")


TUNINGS = [
    P.Tuning.make(block=256, refill_min=1, tile=(0, 0, 0, -1, 0, 0)),
    P.Tuning.make(block=512, refill_min=8, chunk=64),
    P.Tuning.make(block=1024, refill_min=32, chunk=1000, tile=(90, 95, 1, 20, 10, 30)),
    P.Tuning.make(block=1024, refill_min=64, blocks_per_cu=1),
]


@pytest.mark.parametrize("ti", range(len(TUNINGS)))
def test_mcxyz_tuning_invariance(gpu, ti):
    w = W.skin_vessel()
    _say(f"[tuning] {TUNINGS[ti].as_dict()}")
    ref = O.run(w, 15000, threads=8)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(TUNINGS[ti])
        mc.set_num_runs_seed(1, 0)
        mc.send_compute_command(15000)
        got = mc.get_compute_results()
        eff = mc.get_tuning()
    assert eff.block == TUNINGS[ti].block
    assert_same_tallies(got, ref, f"tuning {ti}")


SUB_TUNINGS = [
    P.Tuning.make(block=256, refill_min=1),
    P.Tuning.make(block=512, refill_min=8, chunk=64, mode=2),
    P.Tuning.make(block=1024, refill_min=32, chunk=1000, mode=1),
    P.Tuning.make(block=256, refill_min=4, blocks_per_cu=2, mode=2),
]


@pytest.mark.parametrize("ti", range(len(SUB_TUNINGS)))
def test_mcsub_tuning_invariance(gpu, ti):
    """Block size (256/512/1024 lanes), refill batch, queue chunk and tally
    placement (LDS or global replicas) never change an mcsub result."""
    w = W.mcsub_semi_infinite()
    _say(f"[tuning] mcsub {SUB_TUNINGS[ti].as_dict()}")
    ref = O.run(w, 20000, threads=8)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, None, w.ext)
        mc.set_tuning(SUB_TUNINGS[ti])
        mc.set_num_runs_seed(1, 0)
        mc.send_compute_command(20000)
        got = mc.get_compute_results()
        eff = mc.get_tuning()
    assert eff.block == SUB_TUNINGS[ti].block
    assert_same_tallies(got, ref, f"mcsub tuning {ti}")


def test_default_tile_is_used(gpu):
    w = W.skin_vessel(n_photons=1000)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        t = mc.get_tuning()
    _say(f"[tuning] default {t.as_dict()}")
    assert t.tile[3] * t.tile[4] * t.tile[5] > 1000


def test_mcsub_batch_autotuner(gpu, tmp_path, monkeypatch):
    """The batched-mcsub tuner returns a launch shape the library accepts,
    caches it, and the tuned launch gives the default launch's tallies."""
    from nmc_amd import autotune as A

    monkeypatch.setattr(A, "CACHE", str(tmp_path / "tuning.json"))
    w, items, _ = W.spectral_sweep(n_wavelengths=16, budget_scale=0.001)
    _say(f"[autotune] mcsub batch, 16 items, {sum(i['n_photons'] for i in items)} photons")
    t = A.autotune_mcsub_batch(w.rp, items, w.ext)
    assert t.block in (256, 512, 1024) and t.mode in (1, 2), t.as_dict()
    assert A.autotune_mcsub_batch(w.rp, items, w.ext).as_dict() == t.as_dict()   # cached
    out = []
    for tuning in (None, t):
        with MetalMC(P.SIM_MCSUB) as mc:
            mc.prepare_batch(w.rp, items, w.ext)
            if tuning is not None:
                mc.set_tuning(tuning)
            mc.run_batch()
            out.append(mc.get_batch_results())
    (ta, sa), (tb, sb) = out
    for j in range(len(items)):
        for k in P.TOTAL_INT_FIELDS:
            assert sa[j][k] == sb[j][k], (j, k)
    np.testing.assert_allclose(ta, tb, rtol=RTOL, atol=1e-12)
