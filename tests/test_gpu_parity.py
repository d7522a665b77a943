"""GPU parity: the HIP library against the CPU oracle on identical Philox streams.

Integer tallies (event counts, detections, escapes, per-pixel detection counts)
must be bit-exact; float tallies agree to the FP64 summation-order limit.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from nmc_amd.metal_mc import MetalMC, simulate
from oracle import oracle as O

from parity_util import assert_same_tallies, energy_balance

pytestmark = pytest.mark.gpu

TINY_KINDS = ["mismatch", "matched", "gaussian", "iso", "rect", "point", "manual", "infinite",
              "surface_only", "transmission", "two_term"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", TINY_KINDS)
def test_tiny_voxel_cases_bit_exact(gpu, kind):
    w = W.tiny_voxel_case(kind, n_photons=3000)
    _say(f"[parity] tiny {kind} n=3000")
    ref = O.run(w, threads=4)
    got = simulate(w)
    assert_same_tallies(got, ref, kind)
    assert abs(energy_balance(got.totals) - 1.0) < 1e-5


def test_skin_vessel_bit_exact(gpu):
    w = W.skin_vessel()
    _say("[parity] skin-vessel 200^3 n=20000")
    ref = O.run(w, 20000, threads=8)
    got = simulate(w, 20000)
    assert_same_tallies(got, ref, "skinvessel")


def test_skin_vessel_offset_stream(gpu):
    """Photons [first, first+n) on the GPU equal the oracle's same id range."""
    w = W.skin_vessel()
    _say("[parity] skin-vessel offset ids")
    ref = O.run(w, 5000, first=123_456_789_000, threads=8)
    got = simulate(w, 5000, first_photon=123_456_789_000)
    assert_same_tallies(got, ref, "offset")


def test_mcsub_bit_exact(gpu):
    w = W.mcsub_semi_infinite()
    _say("[parity] mcsub n=20000")
    ref = O.run(w, 20000, threads=8)
    got = simulate(w, 20000)
    assert_same_tallies(got, ref, "mcsub")


@pytest.mark.parametrize("mcflag", [1, 2])
def test_mcsub_launch_modes(gpu, mcflag):
    w = W.mcsub_semi_infinite()
    w.rp.mcflag = mcflag
    w.rp.radius, w.rp.waist, w.rp.zfocus = 0.02, 0.01, 0.5
    # beams enter at the surface z = 0; the point source sits inside at zs
    w.rp.zs = 0.05 if mcflag == 2 else 0.0
    _say(f"[parity] mcsub mcflag={mcflag}")
    ref = O.run(w, 5000, threads=4)
    got = simulate(w, 5000)
    assert_same_tallies(got, ref, f"mcsub{mcflag}")


def test_visit_counts_bit_exact(gpu):
    """The autotuner's pilot mode: the fluence array counts voxel visits."""
    w = W.skin_vessel()
    w.ext.flags = P.FLAG_COUNT_VISITS
    _say("[parity] visit counts n=5000")
    ref = O.run(w, 5000, threads=8)
    got = simulate(w, 5000)
    assert np.array_equal(got.fluence, ref.fluence)
    # one count per face crossing (the voxel left) and one per photon that ends
    # inside the grid; detector-plane arrivals do not change voxel
    t = ref.totals
    ends_inside = t["photons"] - t["escapes_bottom"] - t["escapes_side"] - t["escapes_top"]
    assert got.fluence.sum() == t["substeps_cross"] - t["surface_hits"] + ends_inside


@pytest.mark.parametrize("two_term", [False, True])
def test_no_fluence_variant(gpu, two_term):
    """NMC_FLAG_NO_FLUENCE runs the kernel variant that flushes no voxel
    visits: every other tally (totals, R(x,y), detections) is the oracle's."""
    w = W.tiny_voxel_case("two_term", n_photons=8000) if two_term else W.skin_vessel()
    _say(f"[parity] no-fluence variant two_term={two_term} n=8000")
    ref = O.run(w, 8000, threads=8, fluence=False)
    got = simulate(w, 8000, fluence=False)
    assert got.fluence is None
    assert_same_tallies(got, ref, "no-fluence")


def test_fp32_tile(gpu):
    """FP32 LDS tile (tuning.mode 1): integer tallies bit-exact; fluence within
    FP32 summation error of the oracle (sums of one workgroup's visits in FP32,
    across workgroups in FP64)."""
    w = W.skin_vessel()
    n = 200_000
    _say("[parity] skin-vessel fp32 tile n=200000")
    ref = O.run(w, n, threads=8)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(P.Tuning.make(mode=1))
        t = mc.get_tuning()
        assert t.mode == 1 and t.tile[3] * t.tile[4] * t.tile[5] > 20000   # twice the FP64 box
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command(n)
        got = mc.get_compute_results()
    for k in P.TOTAL_INT_FIELDS:
        assert got.totals[k] == ref.totals[k], k
    assert np.array_equal(got.det2d, ref.det2d)
    np.testing.assert_allclose(got.fluence, ref.fluence, rtol=2e-5, atol=1e-9)
    assert abs(got.fluence.sum() - ref.fluence.sum()) <= 1e-6 * ref.fluence.sum()


def _run_tuned(w, n, tuning, graph_runs=0, async_runs=0):
    """Runs n photons under `tuning`: one blocking run, `graph_runs` graph
    replays of it, or `async_runs` asynchronous runs of n / async_runs
    photons each (consecutive ids, no wait between them)."""
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(tuning)
        eff = mc.get_tuning()
        out = []
        if graph_runs:
            mc.set_graph_mode(True)
        for _ in range(max(1, graph_runs)):
            mc.reset_tallies()
            mc.set_num_runs_seed(w.ext.seed, 0)
            if async_runs:
                for _ in range(async_runs):
                    mc.send_compute_command_async(n // async_runs)
            else:
                mc.send_compute_command(n)
            out.append(mc.get_compute_results())
    return eff, out


def _assert_fp32_class(got, ref, label):
    """Integer tallies bit-exact; fluence within the FP32 tile's bound."""
    for k in P.TOTAL_INT_FIELDS:
        assert got.totals[k] == ref.totals[k], (label, k)
    assert np.array_equal(got.det2d, ref.det2d), label
    np.testing.assert_allclose(got.fluence, ref.fluence, rtol=2e-5, atol=1e-9, err_msg=label)
    assert abs(got.fluence.sum() - ref.fluence.sum()) <= 1e-6 * ref.fluence.sum(), label


@pytest.mark.parametrize("case", ["skin", "overflow", "split", "two_term", "graph", "async",
                                  "unfused", "unfused_split", "small_tile_fused"])
def test_binned_deposit_records(gpu, case):
    """tuning.mode 2 (FP32 tile + binned deposit records): the same photons as
    the oracle, fluence within the FP32 tile's bound; with 1 % of the pilot's
    slice capacity most visits take the atomic fallback.  Fused reduction:
    split launches (each reducing the previous one's records in its drained
    tail), asynchronous runs (each reducing the previous run's), a tile too
    small for the automatic choice with the fusion forced on; unfused: every
    launch followed by its own reduction; graph replays (unfused by design)
    equal the oracle every time."""
    w = W.tiny_voxel_case("two_term") if case == "two_term" else W.skin_vessel()
    n = 20_000 if case == "two_term" else 200_000
    _say(f"[parity] binned deposit records: {case}, n={n}")
    ref = O.run(w, n, threads=8)
    small = (90, 90, 0, 20, 20, 10) if case == "small_tile_fused" else None
    t = P.Tuning.make(mode=2, rec_cap_pct=1 if case == "overflow" else 0,
                      split=50 if case in ("split", "unfused_split") else 0,
                      rec_fuse=-1 if case.startswith("unfused") else (1 if small else 0),
                      tile=small)
    eff, outs = _run_tuned(w, n, t, graph_runs=3 if case == "graph" else 0,
                           async_runs=4 if case == "async" else 0)
    assert eff.mode == 2 and eff.tile[3] > 0
    # the tiny grid's whole tile is smaller than a bucket's sums: unfused
    fused = not case.startswith("unfused") and case != "two_term"
    assert eff.rec_fuse == (1 if fused else -1), eff.rec_fuse
    for i, got in enumerate(outs):
        _assert_fp32_class(got, ref, f"{case} #{i}")


def test_record_fusion_choice_and_pending_records(gpu):
    """The automatic choice leaves a tile smaller than a bucket's FP64 sums
    unfused; records pending from an asynchronous run are dropped by
    reset_tallies (they belong to the cleared tallies) and reduced by
    set_tuning, so a later run's fluence holds exactly its own photons."""
    w = W.skin_vessel()
    n = 100_000
    _say("[parity] record fusion: choice, reset and set_tuning with records pending")
    ref = O.run(w, n, threads=8)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_tuning(P.Tuning.make(mode=2, tile=(90, 90, 0, 20, 20, 10)))
        assert mc.get_tuning().rec_fuse == -1
        mc.set_tuning(P.Tuning.make(mode=2))
        assert mc.get_tuning().rec_fuse == 1
        # a run whose records are still pending, then a reset: not counted
        mc.set_num_runs_seed(w.ext.seed, 5_000_000)
        mc.send_compute_command_async(n)
        mc.reset_tallies()
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command_async(n)
        got = mc.get_compute_results()
        _assert_fp32_class(got, ref, "after reset")
        # pending records reduced before a new tuning frees their buffer
        mc.reset_tallies()
        mc.set_num_runs_seed(w.ext.seed, 0)
        mc.send_compute_command_async(n)
        mc.set_tuning(P.Tuning.make(mode=2, rec_fuse=-1))
        got = mc.get_compute_results()
        _assert_fp32_class(got, ref, "after set_tuning")


def test_autotuner_runs(gpu):
    from nmc_amd.autotune import autotune_mcxyz

    w = W.skin_vessel()
    _say("[autotune] skin-vessel")
    t = autotune_mcxyz(w, n_trial=2_000_000, use_cache=False)
    assert t.block in (512, 1024) and t.tile[3] * t.tile[4] * t.tile[5] > 0


def test_batches_compose(gpu):
    """Two runs over consecutive id ranges accumulate to the single-run tallies."""
    w = W.skin_vessel()
    _say("[parity] batch composition 2 x 50000")
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_num_runs_seed(1, 0)
        mc.send_compute_command(50000)
        mc.send_compute_command(50000)
        two = mc.get_compute_results()
    one = simulate(w, 100000)
    for k in P.TOTAL_INT_FIELDS:
        assert two.totals[k] == one.totals[k], k
    assert np.array_equal(two.det2d, one.det2d)
    np.testing.assert_allclose(two.fluence, one.fluence, rtol=1e-9, atol=1e-12)


def test_reset_and_determinism(gpu):
    w = W.skin_vessel()
    _say("[parity] determinism")
    a = simulate(w, 50000)
    b = simulate(w, 50000)
    for k in P.TOTAL_INT_FIELDS:
        assert a.totals[k] == b.totals[k], k
    np.testing.assert_allclose(a.fluence, b.fluence, rtol=1e-12, atol=0)


def _ragged(nx, ny, nz, n_photons=4000):
    """A tiny case on an nx x ny x nz grid with an absorber on the +x side, so
    that the voxel-index convention (iz*Ny*Nx + ix*Ny + iy) shows in the
    tallies; isotropic point source at the grid centre."""
    w = W.tiny_voxel_case("iso", n_photons=n_photons)
    d = 0.02
    w.rp.Nx, w.rp.Ny, w.rp.Nz = nx, ny, nz
    w.rp.xs, w.rp.ys, w.rp.zs = 0.0, 0.0, 0.5 * nz * d
    w.rp.zsurf = 0.0

    def lab(X, Y, Z):
        t = np.full(X.shape, 2, np.uint8)
        t[X > 0.02] = 3
        return t

    w.v = W.layered_volume(nx, ny, nz, d, d, lab)
    return w


@pytest.mark.parametrize("shape", [(11, 7, 5), (7, 11, 5), (1, 1, 9), (3, 1, 2), (2, 5, 1)])
def test_ragged_grids_bit_exact(gpu, shape):
    w = _ragged(*shape)
    _say(f"[parity] ragged grid {shape}")
    ref = O.run(w, threads=4)
    got = simulate(w)
    assert_same_tallies(got, ref, f"ragged {shape}")
    assert abs(energy_balance(got.totals) - 1.0) < 1e-5
    nx, ny, nz = shape
    F = np.asarray(got.fluence).reshape(nz, nx, ny)      # axes (iz, ix, iy)
    if nx >= 7:
        # the absorber (mua 8 against 2) sits at x > 0.02: more weight is
        # deposited in the +x columns than in their mirror images
        assert F[:, nx - 3:, :].sum() > F[:, :3, :].sum()


def test_empty_run(gpu):
    w = W.tiny_voxel_case("mismatch", n_photons=1000)
    with MetalMC(w.sim) as mc:
        mc.prepare_data(w.rp, w.tp, w.v, w.ext)
        mc.set_num_runs_seed(1, 0)
        mc.send_compute_command(0)
        got = mc.get_compute_results()
    assert got.totals["photons"] == 0
    assert all(got.totals[k] == 0 for k in P.TOTAL_INT_FIELDS)
    assert not np.asarray(got.fluence).any()
