"""Diffusion-theory helpers (nmc_amd/diffusion.py) against the values
NMC-samples/example0.ipynb printed for its run (mua 1, mus' 10, n 1.4/1.001):
"R_diffusion.theory (Farrell): 0.2978437731984608" and
"mueff (diffusion) = 5.744563 1/mm"; the fit against exact exponentials."""
from __future__ import annotations

import math

import numpy as np
import pytest

from nmc_amd import diffusion as D
from nmc_amd import workloads as W


def test_farrell_matches_example0_printout():
    assert D.rd_farrell(1.0, 100.0 * (1 - 0.9), 1.4 / 1.001) == pytest.approx(0.2978437731984608,
                                                                               rel=1e-13)


def test_mu_eff_matches_example0_printout():
    assert round(D.mu_eff(1.0, 10.0), 6) == 5.744563
    assert D.mu_eff(0.3, 20.0) == pytest.approx(math.sqrt(3 * 0.3 * 20.3))


def test_farrell_limits():
    # matched boundary: A = (1 + ri)/(1 - ri) with ri(1) = 0.6681+0.0636+0.7099-1.4399
    assert D.internal_reflection(1.0) == pytest.approx(0.0017)
    # more absorption, less reflectance; a mismatched boundary reflects light
    # back in, where more of it is absorbed: less reflectance
    assert D.rd_farrell(2.0, 10.0, 1.4) < D.rd_farrell(1.0, 10.0, 1.4)
    assert D.rd_farrell(0.1, 10.0, 1.4) < D.rd_farrell(0.1, 10.0, 1.0)
    assert 0.0 < D.rd_farrell(1e-4, 10.0, 1.33) < 1.0


def test_fit_exponential_recovers_exact_curve():
    z = np.linspace(0.01, 3.0, 150)
    phi = 2.5e-3 * np.exp(-4.2 * z)
    phi[::17] = 0.0                              # zeros are skipped, as log(0) would be
    a, b = D.fit_exponential(z, phi, 0.2, 1.3)
    assert a == pytest.approx(2.5e-3, rel=1e-10) and b == pytest.approx(4.2, rel=1e-12)
    with pytest.raises(ValueError):
        D.fit_exponential(z, phi, 5.0, 6.0)


def test_axial_profile_layout():
    nx, ny, nz = 4, 6, 5
    f = np.arange(nx * ny * nz, dtype=float)
    prof = D.axial_profile(f, nx, ny, nz)
    assert list(prof) == [W.voxel_index(nx // 2, ny // 2, iz, nx, ny) for iz in range(nz)]


def test_example0_workload_matches_notebook():
    w = W.example0_slab()
    rp = w.rp
    assert (rp.Nx, rp.dx, rp.boundaryflag, rp.zsurf, rp.radius) == (200, pytest.approx(0.02), 2,
                                                                     0.0, pytest.approx(1.2))
    assert (w.tp.muav[1], w.tp.musv[1]) == (1.0, 100.0) and w.v.min() == w.v.max() == 1
