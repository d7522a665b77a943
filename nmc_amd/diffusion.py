"""Diffusion-theory cross-checks of the notebooks (NMC-samples/example0.ipynb):
Farrell's diffuse reflectance of a semi-infinite medium, the effective
attenuation coefficient, and the single-exponential fit of the on-axis
fluence phi(z) ~ a exp(-b z) that the notebook compares with mu_eff.

    rd_farrell(mua, musp, n_ratio)      Farrell, Patterson & Wilson, Med. Phys. 19:879 (1992),
                                        internal reflection by the polynomial fit in n used by
                                        example0.ipynb (get_rd_farrell)
    mu_eff(mua, musp)                   sqrt(mua / D), D = 1 / (3 (mua + musp))
    fit_exponential(z, phi, z0, z1)     least squares on log phi over z0 < z <= z1
    axial_profile(fluence, rp)          phi at the grid centre column, z-ordered
"""
from __future__ import annotations

import math

import numpy as np


def internal_reflection(n_ratio: float) -> float:
    """Diffuse internal reflection of the boundary for index ratio n = n_in / n_out."""
    n = float(n_ratio)
    return 0.6681 + 0.0636 * n + 0.7099 / n - 1.4399 / (n * n)


def rd_farrell(mua: float, musp: float, n_ratio: float) -> float:
    """Total diffuse reflectance of a semi-infinite medium lit by a broad
    collimated beam (dipole source at depth z0 = 1/(mua + musp))."""
    ri = internal_reflection(n_ratio)
    A = (1.0 + ri) / (1.0 - ri)
    z0 = 1.0 / (mua + musp)
    ap = musp / (mua + musp)
    me = mu_eff(mua, musp)
    return float(ap * math.exp(-me * z0) / 2.0
                 * (1.0 + math.exp(-(4.0 / 3.0) * A * math.sqrt(3.0 * (1.0 - ap)))))


def mu_eff(mua: float, musp: float) -> float:
    """Effective attenuation coefficient of diffusion theory."""
    D = 1.0 / (3.0 * (mua + musp))
    return math.sqrt(mua / D)


def fit_exponential(z: np.ndarray, phi: np.ndarray, z_min: float, z_max: float):
    """(a, b) of phi ~ a exp(-b z) fitted to log phi over z_min < z <= z_max
    (points with phi <= 0 dropped)."""
    z = np.asarray(z, np.float64)
    phi = np.asarray(phi, np.float64)
    sel = (z > z_min) & (z <= z_max) & (phi > 0)
    if sel.sum() < 2:
        raise ValueError("fewer than two positive points in the fit range")
    slope, icpt = np.polyfit(z[sel], np.log(phi[sel]), 1)
    return float(np.exp(icpt)), float(-slope)


def axial_profile(fluence: np.ndarray, nx: int, ny: int, nz: int):
    """phi(z) along the centre column (ix = Nx/2, iy = Ny/2) of a grid stored in
    the file order iz*Ny*Nx + ix*Ny + iy."""
    f = np.asarray(fluence).reshape(nz, nx, ny)
    return f[:, nx // 2, ny // 2]
