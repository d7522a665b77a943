"""Spectral tissue model of the reference notebooks: optical properties of the
nine tissue types at a wavelength, from the spectral library of absorbers,
and the notebooks' skin-vessel geometry, feeding the transport kernels (one
run per wavelength, or one batched mcsub launch over many wavelengths).

Follows NMC-samples/example2.ipynb (tissue_parameters, the 9-entry tissue list
and the T(y,x,z) skin-vessel build) and example3.ipynb (the 350-750 nm sweep):

    mus'(nm) = mus'500 * (f_ray * (nm/500)^-4 + (1 - f_ray) * (nm/500)^-b_mie)
    mua(nm)  = MU(nm) . [B*S, B*(1-S), W, M]       MU = oxy-Hb, deoxy-Hb, water, melanin
    mus      = mus' / (1 - g)

with MU interpolated linearly in the library's 1-nm table.
"""
from __future__ import annotations

import dataclasses
import json
import math
import os

import numpy as np

from . import params as P
from .workloads import Workload

# (name, B, S, W, M, mus'500, f_ray, b_mie, g) per tissue type, or
# (name, mua, mus, g) for the fixed types (example2.ipynb, tissue_parameters)
TISSUES = {
    1: ("air", 1e-4, 1.0, 1.0),
    2: ("water", None, 10.0, 1.0),          # mua = water absorption
    3: ("blood", 1.00, 0.75, 0.95, 0.0, 10.0, 0.0, 1.0, 0.90),
    4: ("dermis", 0.002, 0.67, 0.65, 0.0, 43.6, 0.41, 0.562, 0.90),
    5: ("epidermis", 0.0, 0.75, 0.75, 0.03, 66.7, 0.29, 0.689, 0.90),
    6: ("skull", 0.0005, 0.75, 0.35, 0.0, 30.0, 0.0, 1.0, 0.90),
    7: ("gray matter", 0.01, 0.75, 0.75, 0.0, 20.0, 0.2, 1.0, 0.90),
    8: ("white matter", 0.01, 0.75, 0.75, 0.0, 20.0, 0.2, 1.0, 0.90),
    9: ("standard tissue", 1.0, 100.0, 0.90),
}


@dataclasses.dataclass
class SpectralLibrary:
    """Absorption spectra on a wavelength table (nm)."""

    nm: np.ndarray
    oxy: np.ndarray
    deoxy: np.ndarray
    water: np.ndarray
    mel: np.ndarray

    @classmethod
    def from_mat(cls, path: str) -> "SpectralLibrary":
        """The reference's spectralLIB.mat (MAT v5; scipy's reader executes nothing)."""
        from scipy.io import loadmat
        d = loadmat(path)
        return cls(*(np.asarray(d[k], np.float64).ravel()
                     for k in ("nmLIB", "muaoxy", "muadeoxy", "muawater", "muamel")))

    @classmethod
    def from_json(cls, path: str) -> "SpectralLibrary":
        with open(path) as f:
            d = json.load(f)
        return cls(*(np.asarray(d[k], np.float64)
                     for k in ("nmLIB", "muaoxy", "muadeoxy", "muawater", "muamel")))

    @classmethod
    def load(cls, path: str) -> "SpectralLibrary":
        return cls.from_mat(path) if path.endswith(".mat") else cls.from_json(path)

    def absorbers(self, nm: float) -> np.ndarray:
        """MU(nm) = [oxy-Hb, deoxy-Hb, water, melanin]."""
        return np.array([np.interp(nm, self.nm, a) for a in (self.oxy, self.deoxy, self.water,
                                                             self.mel)])


def tissue_list(nm: float, lib: SpectralLibrary) -> dict[int, dict]:
    """Properties of tissue types 1..9 at `nm` (keys name, mua, mus, g, musp)."""
    MU = lib.absorbers(nm)
    x = nm / 500.0
    out = {}
    for k, t in TISSUES.items():
        if len(t) == 4:
            name, mua, mus, g = t
            if mua is None:
                mua = float(MU[2])
            out[k] = {"name": name, "mua": float(mua), "mus": float(mus), "g": float(g),
                      "musp": float(mus) * (1.0 - float(g))}
        else:
            name, B, S, W, M, musp500, fray, bmie, g = t
            musp = musp500 * (fray * x ** -4 + (1.0 - fray) * x ** -bmie)
            mua = float(MU @ np.array([B * S, B * (1.0 - S), W, M]))
            out[k] = {"name": name, "mua": mua, "mus": musp / (1.0 - g), "g": g, "musp": musp}
    return out


def tissue_params(nm: float, lib: SpectralLibrary) -> P.TissueParams:
    """TissueParams with types 1..9 filled at `nm` (Nt = 9)."""
    tp = P.TissueParams()
    for k, t in tissue_list(nm, lib).items():
        tp.muav[k], tp.musv[k], tp.gv[k] = t["mua"], t["mus"], t["g"]
    return tp


def skin_vessel_volume(n: int = 200, binsize: float = 0.0005, zsurf: float = 0.01,
                       epidermis: float = 0.006, vessel_radius: float = 0.01) -> np.ndarray:
    """The notebooks' label volume (example2.ipynb): dermis (4), air (1) above
    zsurf, epidermis (5) below it, a blood (3) cylinder along y at x = 0,
    depth n/2 * binsize; returned in the file order iz*Ny*Nx + ix*Ny + iy."""
    nx = ny = nz = n
    d = binsize
    T = np.full((ny, nx, nz), 4, np.uint8)
    air = int(round(zsurf / d))
    if air >= 1:
        T[:, :, :air] = 1
    e1 = max(int(round(zsurf / d)) + 1, 1)
    e2 = min(int(round((zsurf + epidermis) / d)), nz)
    if e2 >= e1:
        T[:, :, e1 - 1:e2] = 5
    x = (np.arange(1, nx + 1) - nx / 2.0) * d
    z = np.arange(1, nz + 1) * d
    X, Z = np.meshgrid(x, z, indexing="xy")          # (nz, nx)
    inside = np.sqrt(X ** 2 + (Z - nz / 2.0 * d) ** 2) <= vessel_radius
    for iz in range(nz):
        cols = inside[iz]
        if cols.any():
            T[:, cols, iz] = 3
    return np.ascontiguousarray(T.ravel(order="F"))


def skin_vessel_run_params(n: int = 200, binsize: float = 0.0005) -> P.RunParams:
    """The run settings of example2.ipynb / example3.ipynb."""
    rp = P.RunParams()
    rp.myname = b"skinvessel"
    rp.time_min = 10.0
    rp.Nx = rp.Ny = rp.Nz = n
    rp.dx = rp.dy = rp.dz = binsize
    rp.mcflag, rp.launchflag, rp.boundaryflag = 0, 0, 2
    rp.xs, rp.ys, rp.zs = 0.0, 0.0, 0.0101
    rp.xfocus, rp.yfocus, rp.zfocus = 0.0, 0.0, math.inf
    rp.ux0, rp.uy0 = 0.7, 0.4
    rp.uz0 = math.sqrt(max(0.0, 1.0 - 0.7 ** 2 - 0.4 ** 2))
    rp.radius, rp.waist, rp.zsurf = 0.03, 0.03, 0.01
    rp.det_radius, rp.xd = 15.0, 0.0
    rp.n1, rp.n2 = 1.4, 1.001
    rp.na, rp.lambda_, rp.lc = 90.0, 532.0e-6, 0.001
    rp.speckleflag, rp.semiflag, rp.det_state = 0, 1, P.DET_REFLECTANCE
    rp.mckernelflag = P.SIM_MCXYZ
    rp.Nt = 9
    return rp


def skin_vessel(nm: float, lib: SpectralLibrary, n_photons: int = 1_000_000, n: int = 200,
                binsize: float = 0.0005, seed: int = 1) -> Workload:
    """The notebooks' 9-tissue skin-vessel run at wavelength `nm`."""
    rp = skin_vessel_run_params(n, binsize)
    rp.lambda_ = nm * 1e-6
    return Workload(f"skinvessel_{int(nm)}nm", P.SIM_MCXYZ, rp, tissue_params(nm, lib),
                    skin_vessel_volume(n, binsize), P.ExtConfig.default(seed=seed), n_photons)


def sweep_items(wavelengths, tissue: int, lib: SpectralLibrary, budgets, n1: float = 1.4,
                n2: float = 1.0, first_stream: int = 0) -> list[dict]:
    """Batched-mcsub work items (one per wavelength) for a homogeneous medium of
    tissue type `tissue`: the media of nmc_prepare_batch."""
    items = []
    for j, (nm, b) in enumerate(zip(wavelengths, budgets)):
        t = tissue_list(float(nm), lib)[tissue]
        items.append({"mua": t["mua"], "mus": t["mus"], "g": t["g"], "n1": n1, "n2": n2,
                      "stream_id": first_stream + j, "n_photons": int(b)})
    return items


def default_library() -> SpectralLibrary:
    """The library from $NMC_SPECTRAL_LIB (a .mat or .json file)."""
    path = os.environ.get("NMC_SPECTRAL_LIB")
    if not path:
        raise FileNotFoundError("set NMC_SPECTRAL_LIB to the spectral library (.mat or .json)")
    return SpectralLibrary.load(path)
