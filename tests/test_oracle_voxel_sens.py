"""Oracle: per-voxel sensitivities of the absorbed weight to each tissue's mua
and mus (nmco_mcxyz_sens), pinned by

- the closed form of unscattered light (g = 1) through two tissues stacked
  in depth: every photon takes the same straight path, so the
  perturbation-MC derivative of each voxel's deposit equals the analytic
  derivative of exp(-sum mua_t L_t) (1 - exp(-mua_k dz)), to float rounding;
  and with the path independent of mus, the mus sensitivity has mean zero;
- central finite differences with common random numbers (the same photon
  ids at mua_t (1 +- h), mus_t (1 +- h)) on a scattering three-tissue grid,
  for the weight absorbed in each tissue;
- planted mistakes (oracle builds with NMCO_MUTANT 6 and 7) that each pin
  catches."""
from __future__ import annotations

import copy

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from oracle import oracle as O

N_COL, D_COL, Z_SPLIT = 10, 0.05, 4
MUA_COL = {1: 3.0, 2: 1.5}


def column(n_photons: int = 500) -> W.Workload:
    """Pencil beam down a 10^3 grid: tissue 1 above voxel layer 4, tissue 2
    below, g = 1 (no deflection)."""
    n, d = N_COL, D_COL
    rp = W._rp("col2", Nx=n, Ny=n, Nz=n, dx=d, dy=d, dz=d, Nt=2, mcflag=4, boundaryflag=1,
               zs=0.0, zsurf=-1.0, mckernelflag=P.SIM_MCXYZ)
    tp = W._tp({1: (MUA_COL[1], 40.0, 1.0), 2: (MUA_COL[2], 60.0, 1.0)})
    v = np.ones((n, n, n), np.uint8)
    v[Z_SPLIT:] = 2
    return W.Workload("col2", P.SIM_MCXYZ, rp, tp, v.ravel(), P.ExtConfig.default(), n_photons)


def column_closed_form():
    """(D_k, dD_k/dmua_1, dD_k/dmua_2) per voxel layer k of the column, per photon."""
    n, d = N_COL, D_COL
    z = np.arange(n) * d
    zc = Z_SPLIT * d
    m1, m2 = MUA_COL[1], MUA_COL[2]
    top = z < zc
    L1 = np.where(top, z, zc)
    L2 = np.where(top, 0.0, z - zc)
    mk = np.where(top, m1, m2)
    a = np.exp(-m1 * L1 - m2 * L2)
    D = a * (1 - np.exp(-mk * d))
    own = a * d * np.exp(-mk * d)
    dD1 = -L1 * D + np.where(top, own, 0.0)
    dD2 = -L2 * D + np.where(top, 0.0, own)
    return D, dD1, dD2


def _column_check(t, da, n):
    D, dD1, dD2 = column_closed_form()
    c = N_COL // 2
    col = t.fluence.reshape(N_COL, N_COL, N_COL)[:, c, c] / n
    np.testing.assert_allclose(col, D, rtol=2e-5)
    np.testing.assert_allclose(da[0][:, c, c] / n, dD1, rtol=2e-5, atol=1e-9)
    np.testing.assert_allclose(da[1][:, c, c] / n, dD2, rtol=2e-5, atol=1e-9)
    off = np.ones((N_COL, N_COL), bool)
    off[c, c] = False
    assert np.all(da[:, :, off] == 0.0)


def test_column_mua_sensitivity_equals_closed_form():
    n = 500
    t, da, ds = O.run_sens(column(n), [1, 2], threads=2)
    assert t.totals["escapes_bottom"] == n
    _column_check(t, da, n)


def test_column_mua_pin_fails_on_planted_mistake(tmp_path):
    n = 500
    t, da, ds = O.run_sens(column(n), [1, 2], threads=2, L=O.mutant(6, str(tmp_path)))
    with pytest.raises(AssertionError):
        _column_check(t, da, n)


def _column_mus_z(L=None, batches=16, per=2000):
    """z-scores of the batch means of the column's mus sensitivities, summed
    over each tissue's voxels, for both tissues' mus (expected mean: 0)."""
    w = column(per)
    vals = []
    for b in range(batches):
        t, da, ds = O.run_sens(w, [1, 2], first=b * per, threads=4, L=L)
        c = N_COL // 2
        s = ds[:, :, c, c] / per
        vals.append([s[j][:Z_SPLIT].sum() for j in range(2)] + [s[j][Z_SPLIT:].sum() for j in range(2)])
    vals = np.array(vals)
    m, se = vals.mean(0), vals.std(0, ddof=1) / np.sqrt(batches)
    return np.where(se > 0, m / np.where(se > 0, se, 1.0), 0.0), m


def test_column_mus_sensitivity_has_mean_zero():
    z, m = _column_mus_z()
    assert np.all(np.abs(z) < 4.5), (z, m)


def test_column_mus_pin_fails_on_planted_mistake(tmp_path):
    z, m = _column_mus_z(L=O.mutant(7, str(tmp_path)))
    assert np.any(np.abs(z) > 4.5), (z, m)


# ---------------------------------------------------------------------------
# CRN central differences on a scattering heterogeneous grid
# ---------------------------------------------------------------------------
def three_tissue(n_photons: int) -> W.Workload:
    """14^3 grid of 0.02 cm voxels: a 0.06 cm top slab (tissue 1), a sphere
    of radius 0.08 cm (tissue 3) in tissue 2; g 0.8 / 0.9 / 0.85, matched n,
    flat beam of radius 0.05 cm, every face escaping."""
    n, d = 14, 0.02
    rp = W._rp("t3", Nx=n, Ny=n, Nz=n, dx=d, dy=d, dz=d, Nt=3, mcflag=0, boundaryflag=1,
               zs=0.0001, radius=0.05, waist=0.05, zsurf=0.0, mckernelflag=P.SIM_MCXYZ,
               det_state=P.DET_REFLECTANCE)
    tp = W._tp({1: (4.0, 60.0, 0.8), 2: (0.8, 90.0, 0.9), 3: (12.0, 40.0, 0.85)})
    c = (np.arange(n) + 0.5) * d
    Z, X, Y = np.meshgrid(c, c - n * d / 2, c - n * d / 2, indexing="ij")
    v = np.full((n, n, n), 2, np.uint8)
    v[Z < 0.06] = 1
    v[(X ** 2 + Y ** 2 + (Z - 0.16) ** 2) < 0.08 ** 2] = 3
    return W.Workload("t3", P.SIM_MCXYZ, rp, tp, v.ravel(), P.ExtConfig.default(), n_photons)


def _absorbed_by_tissue(w, F):
    v = np.asarray(w.v).reshape(F.shape)
    return np.array([F[v == k].sum() for k in (1, 2, 3)])


def _perturbed(w, which, t, f):
    w2 = copy.deepcopy(w)
    if which == "mua":
        w2.tp.muav[t] = w.tp.muav[t] * f
    else:
        w2.tp.musv[t] = w.tp.musv[t] * f
    return w2


@pytest.mark.parametrize("which,h,tol,per", [("mua", 0.01, 0.01, 6000), ("mus", 0.05, None, 40000)])
def test_three_tissue_sensitivity_matches_crn_central_difference(which, h, tol, per):
    """d(weight absorbed in tissue k)/d(param of tissue t) for all k, t: the
    perturbation-MC sums of the per-voxel sensitivities against CRN central
    differences.  mua moves no path, so the two agree to the 1 % bar (they
    agree to ~1e-4); mus moves paths and is judged by the spread of 8
    batches of 40 000 photons (4.5 sigma of the difference, or 1 %; the
    standard error is 1-5 % of the larger entries)."""
    batches = 8
    w = three_tissue(per)
    shape = (w.rp.Nz, w.rp.Nx, w.rp.Ny)
    pm, fd = [], []
    for b in range(batches):
        first = b * per
        t0, da, ds = O.run_sens(w, [1, 2, 3], first=first, threads=8)
        S = da if which == "mua" else ds
        pm.append(np.array([_absorbed_by_tissue(w, S[j].reshape(shape)) for j in range(3)]))
        row = []
        for j, tis in enumerate((1, 2, 3)):
            base = w.tp.muav[tis] if which == "mua" else w.tp.musv[tis]
            up = O.run(_perturbed(w, which, tis, 1 + h), first=first, threads=8)
            dn = O.run(_perturbed(w, which, tis, 1 - h), first=first, threads=8)
            row.append((_absorbed_by_tissue(w, up.fluence.reshape(shape)) -
                        _absorbed_by_tissue(w, dn.fluence.reshape(shape))) / (2 * h * base))
        fd.append(np.array(row))
    pm, fd = np.array(pm) / per, np.array(fd) / per       # [batch, t, k]
    diff = pm - fd
    mp, mf = pm.mean(0), fd.mean(0)
    se = diff.std(0, ddof=1) / np.sqrt(batches)
    bar = 0.01 * np.abs(mf) if tol else np.maximum(0.01 * np.abs(mf), 4.5 * se)
    assert np.all(np.abs(mp - mf) <= bar + 1e-12), (mp, mf, se)


def test_sens_argument_checks():
    w = column(10)
    with pytest.raises(RuntimeError):
        O.run_sens(w, [1, 1])
    with pytest.raises(RuntimeError):
        O.run_sens(w, list(range(1, 10)))
    with pytest.raises(RuntimeError):
        O.run_sens(w, [19])
