"""GPU: the reference notebooks' own runs, reproduced.

example0.ipynb (homogeneous slab) and example3.ipynb (skin-vessel Rd over
350-750 nm) printed the Rd their Metal build obtained; those runs used a
different random generator and carry the reference's documented defects
(notably the hard kill after a fixed number of scattering events), so the
comparison is a loose sanity bound, stated per test. Photon-for-photon parity
with the oracle on the same configurations is exact.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

from nmc_amd import diffusion as D
from nmc_amd import spectral
from nmc_amd import workloads as W
from nmc_amd.metal_mc import simulate
from oracle import oracle as O

from parity_util import assert_same_tallies

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_example0_parity_and_reference_rd(gpu):
    w = W.example0_slab()
    _say("[notebooks] example0 parity n=3000")
    ref = O.run(w, 3000, threads=8)
    assert_same_tallies(simulate(w, 3000), ref, "example0")
    n = 2_000_000
    _say(f"[notebooks] example0 n={n}")
    got = simulate(w, n)
    rd = got.totals["rd"] / n
    # the same medium through the independent mcsub estimator (all-or-none
    # escape at the surface, albedo weighting, no voxels): same Rd once its
    # launch specular reflection is taken out
    sub = W.mcsub_semi_infinite()
    sub.rp.n1, sub.rp.n2 = 1.4, 1.001
    sub.tp.muav[1], sub.tp.musv[1], sub.tp.gv[1] = 1.0, 100.0, 0.9
    _say(f"[notebooks] example0 medium through mcsub n={n}")
    s = simulate(sub, n, fluence=False).totals
    rd_sub = s["rd"] / (n - s["specular"])
    assert abs(rd - rd_sub) < 0.005 * rd, (rd, rd_sub)
    # reference printout "Total Rd = 0.26601276" (1e5 photons, its own RNG and
    # detector-plane handling): ours is 2.4 % lower, consistently with mcsub
    assert abs(rd - 0.26601276) < 0.04 * 0.26601276, rd
    # diffusion theory (Farrell) is a ~10 % approximation here
    assert abs(rd - D.rd_farrell(1.0, 10.0, 1.4 / 1.001)) < 0.15 * rd
    # on-axis fluence decays close to mu_eff (the notebook's fit gave 6.53 /mm)
    phi = D.axial_profile(got.fluence, 200, 200, 200)
    z = (np.arange(200) + 0.5) * 0.02
    _, b = D.fit_exponential(z, phi, 0.3, 1.3)
    assert abs(b - D.mu_eff(1.0, 10.0)) < 0.25 * D.mu_eff(1.0, 10.0), b


def test_example3_skin_vessel_parity(gpu):
    lib = spectral.SpectralLibrary.from_json(os.path.join(GOLD, "spectral_lib.json"))
    w = spectral.skin_vessel(530, lib)
    _say("[notebooks] example3 530 nm parity n=2000")
    assert_same_tallies(simulate(w, 2000), O.run(w, 2000, threads=8), "skinvessel530")


def test_example3_rd_spectrum(gpu):
    """Rd(nm) of the skin-vessel model against example3's printout. Up to
    590 nm (photons of < 3e4 events) the two agree within 4 % (ours runs
    1-2 % lower, as for example0); beyond, the
    reference's hard kill of photons past MAX_SCATT = 10000 scatterings
    (kernels.metal:747, definitions.h:19; no counterpart here: roulette only)
    lowers its Rd, so ours is the larger, by less than 8 %."""
    lib = spectral.SpectralLibrary.from_json(os.path.join(GOLD, "spectral_lib.json"))
    with open(os.path.join(GOLD, "notebook_tissues.json")) as f:
        printed = {t["nm"]: t["rd_reference_mc"] for t in json.load(f)["example3"]}
    for nm in range(350, 591, 20):
        w = spectral.skin_vessel(nm, lib)
        _say(f"[notebooks] example3 {nm} nm n=100000")
        t = simulate(w, 100_000, fluence=False).totals
        rd = t["rd"] / t["photons"]
        assert abs(rd - printed[nm]) < 0.04 * printed[nm], (nm, rd, printed[nm])
    nm = 650
    _say(f"[notebooks] example3 {nm} nm n=30000")
    t = simulate(spectral.skin_vessel(nm, lib), 30_000, fluence=False).totals
    rd = t["rd"] / t["photons"]
    assert printed[nm] * 0.99 < rd < printed[nm] * 1.08, (nm, rd, printed[nm])
