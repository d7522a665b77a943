"""Synthetic workloads: the benchmark configurations and small parity cases.

Each builder returns a `Workload` holding the reference-format structs
(`RunParams`, `TissueParams`), the uint8 voxel map (mcxyz) and an `ExtConfig`.
Units are cm throughout.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import params as P


@dataclasses.dataclass
class Workload:
    name: str
    sim: int
    rp: P.RunParams
    tp: P.TissueParams
    v: np.ndarray | None          # uint8, index iz*Ny*Nx + ix*Ny + iy
    ext: P.ExtConfig
    n_photons: int


def _rp(name: str, **kw) -> P.RunParams:
    rp = P.RunParams()
    rp.myname = name.encode()[: P.STRLEN - 1]
    rp.n1 = rp.n2 = 1.0
    rp.na = 90.0
    rp.det_radius = 1e6
    rp.zfocus = 1e12
    for k, v in kw.items():
        setattr(rp, k, v)
    return rp


def _tp(props: dict[int, tuple]) -> P.TissueParams:
    tp = P.TissueParams()
    for t, p in props.items():
        tp.muav[t], tp.musv[t], tp.gv[t] = p[:3]
        if len(p) > 3:
            tp.gf[t], tp.af[t], tp.gb[t], tp.ab[t], tp.CC[t] = p[3:8]
    return tp


def voxel_index(ix, iy, iz, nx, ny):
    """Linear voxel index of kernels.metal:482 (iz*Ny*Nx + ix*Ny + iy)."""
    return iz * ny * nx + ix * ny + iy


def layered_volume(nx, ny, nz, dx, dz, labeler) -> np.ndarray:
    """uint8 map labelled by voxel centre, stored in the reference's order.

    labeler(xc, yc, zc) takes broadcast arrays of voxel centres (x, y centred on
    the grid axis, z from 0 at the top) and returns the tissue id per voxel.
    """
    xc = (np.arange(nx) + 0.5 - nx / 2) * dx
    yc = (np.arange(ny) + 0.5 - ny / 2) * dx
    zc = (np.arange(nz) + 0.5) * dz
    Z, X, Y = np.meshgrid(zc, xc, yc, indexing="ij")     # axes (iz, ix, iy)
    t = labeler(X, Y, Z).astype(np.uint8)
    return np.ascontiguousarray(t).reshape(-1)


def skin_vessel(n_photons: int = 100_000_000, n: int = 200, seed: int = 1) -> Workload:
    """200^3 skin-vessel mcxyz case: air / epidermis / dermis / blood cylinder.

    0.005 cm voxels, air above z = 0.01, 60 um epidermis, dermis below, a blood
    cylinder along x of radius 0.05 cm centred 0.06 cm deep; n = 1 everywhere;
    flat collimated beam of radius 0.06 cm launched at z = 0.0101 with the
    detector plane at zsurf = 0.01 and escape at every grid face.
    """
    d = 1.0 / n
    rp = _rp("skinvessel", Nx=n, Ny=n, Nz=n, dx=d, dy=d, dz=d, Nt=4,
             mcflag=0, launchflag=0, boundaryflag=1, xs=0.0, ys=0.0, zs=0.0101,
             radius=0.06, waist=0.06, zsurf=0.01, mckernelflag=P.SIM_MCXYZ,
             det_state=P.DET_REFLECTANCE)
    tp = _tp({1: (1e-4, 1.0, 1.0), 2: (15.0, 300.0, 0.9), 3: (0.5, 200.0, 0.9),
              4: (230.0, 100.0, 0.98)})

    def lab(X, Y, Z):
        t = np.full(X.shape, 3, np.uint8)
        t[(Z >= 0.01) & (Z < 0.016)] = 2
        t[(Y ** 2 + (Z - 0.06) ** 2) <= 0.05 ** 2] = 4
        t[Z < 0.01] = 1
        return t

    v = layered_volume(n, n, n, d, d, lab)
    return Workload("mcxyz_skinvessel_%d" % n, P.SIM_MCXYZ, rp, tp, v,
                    P.ExtConfig.default(seed=seed), n_photons)


def mcsub_semi_infinite(n_photons: int = 1_000_000, seed: int = 1) -> Workload:
    """Semi-infinite homogeneous mcsub case: mua 1, mus 100, g 0.9, n 1.33/1.0,
    pencil beam, Rd(r, alpha) 100 x 30 with dr 0.01, A(r, z) 100 x 100."""
    rp = _rp("mcsub", mcflag=0, boundaryflag=1, radius=0.0, n1=1.33, n2=1.0,
             mckernelflag=P.SIM_MCSUB)
    tp = _tp({1: (1.0, 100.0, 0.9)})
    ext = P.ExtConfig.default(seed=seed, nr=100, nz=100, na=30, dr=0.01, dz=0.01)
    return Workload("mcsub_semi_infinite", P.SIM_MCSUB, rp, tp, None, ext, n_photons)


def spectral_sweep(n_wavelengths: int = 128, seed: int = 3, budget_scale: float = 1.0):
    """128-wavelength sweep on the semi-infinite mcsub geometry.

    lambda_i = 400 + 500 i/127 nm.  From numpy default_rng(seed): mus(lambda) =
    a (lambda/500)^-b with a ~ U(50, 300) /cm and b ~ U(0.5, 2); mua(lambda) is
    a sum of three Gaussian peaks (centres U(400, 900) nm, widths U(10, 60) nm,
    heights U(50, 300) /cm) scaled so its maximum is at most 300 /cm; budgets
    floor(10^U(5, 7)) photons.  g = 0.9, n = 1.33 against air, pencil beam.
    Returns (Workload of the geometry, list of WorkItem dicts, wavelengths).
    """
    rng = np.random.default_rng(seed)
    lam = 400.0 + 500.0 * np.arange(n_wavelengths) / max(1, n_wavelengths - 1)
    a = rng.uniform(50, 300)
    b = rng.uniform(0.5, 2.0)
    mus = a * (lam / 500.0) ** (-b)
    c = rng.uniform(400, 900, 3)
    wd = rng.uniform(10, 60, 3)
    amp = rng.uniform(50, 300, 3)
    mua = (amp[None, :] * np.exp(-(lam[:, None] - c[None, :]) ** 2 / (2 * wd[None, :] ** 2))).sum(1)
    if mua.max() > 300.0:
        mua *= 300.0 / mua.max()
    budgets = np.floor(10 ** rng.uniform(5, 7, n_wavelengths) * budget_scale).astype(np.int64)
    w = mcsub_semi_infinite()
    w.ext.na = 1
    w.ext.flags = P.FLAG_NO_FLUENCE
    items = [dict(mua=float(np.float32(mua[i])), mus=float(np.float32(mus[i])), g=0.9, n1=1.33,
                  n2=1.0, stream_id=i, n_photons=int(max(1, budgets[i])))
             for i in range(n_wavelengths)]
    return w, items, lam


TWO_TERM_SKIN = dict(gf=0.9, af=0.5, gb=0.3, ab=0.5, cc=0.8)   # HG(0.9)x0.8 + HG(-0.3)x0.2


def make_stack(thick, mua, mus, n, n_photons, stream_id=0, phase=None, g=0.9,
               n_above=1.0, n_below=1.0) -> P.Stack:
    """A Stack struct; the last layer is semi-infinite if its thickness <= 0."""
    s = P.Stack()
    s.n_layers = len(mua)
    s.n_above, s.n_below = n_above, n_below
    s.stream_id = stream_id
    s.n_photons = int(n_photons)
    for l in range(len(mua)):
        L = s.layer[l]
        L.thickness, L.mua, L.mus, L.n, L.g = thick[l], mua[l], mus[l], n[l], g
        if phase:
            L.phase = 1
            L.gf, L.af, L.gb, L.ab, L.cc = (phase[k] for k in ("gf", "af", "gb", "ab", "cc"))
    return s


def multilayer_5(n_photons: int = 10_000_000, seed: int = 4):
    """Differentiable 5-layer case: from default_rng(seed), thickness
    U(0.01, 0.2) cm (the 5th layer semi-infinite), mua U(0.1, 10) /cm, mus
    U(50, 300) /cm, n U(1.33, 1.45); two-term HG (0.9, -0.3, 0.8); pencil beam;
    Rd(r) in 200 bins of 0.005 cm.  Returns (LayeredConfig, [Stack])."""
    rng = np.random.default_rng(seed)
    thick = rng.uniform(0.01, 0.2, 5)
    thick[-1] = 0.0
    mua = rng.uniform(0.1, 10.0, 5)
    mus = rng.uniform(50.0, 300.0, 5)
    n = rng.uniform(1.33, 1.45, 5)
    st = make_stack(thick, mua, mus, n, n_photons, phase=TWO_TERM_SKIN)
    return P.LayeredConfig.default(nr=200, dr=0.005), [st]


def inverse_fit_instances(n_inst: int = 256, seed: int = 5, budget_scale: float = 1.0):
    """Batched inverse-fit case: n_inst independent 3-layer stacks (ranges as
    multilayer_5, the 3rd layer semi-infinite), budgets floor(10^U(5, log10 5e6)),
    Rd(r) in 100 bins of 0.01 cm.  Returns (LayeredConfig, [Stack] with the
    hidden properties)."""
    rng = np.random.default_rng(seed)
    stacks = []
    for i in range(n_inst):
        thick = rng.uniform(0.01, 0.2, 3)
        thick[-1] = 0.0
        mua = rng.uniform(0.1, 10.0, 3)
        mus = rng.uniform(50.0, 300.0, 3)
        n = rng.uniform(1.33, 1.45, 3)
        budget = int(np.floor(10 ** rng.uniform(5, np.log10(5e6)) * budget_scale))
        stacks.append(make_stack(thick, mua, mus, n, max(1, budget), stream_id=i,
                                 phase=TWO_TERM_SKIN))
    return P.LayeredConfig.default(nr=100, dr=0.01), stacks


def tiny_voxel_case(kind: str = "mismatch", n_photons: int = 2000, seed: int = 7,
                    n: int = 16) -> Workload:
    """Small mcxyz cases for parity: every launch mode / boundary mode can be set
    through `kind`, keeping the oracle run under a second."""
    d = 0.02
    base = dict(Nx=n, Ny=n, Nz=n, dx=d, dy=d, dz=d, Nt=3, mcflag=0, launchflag=0,
                boundaryflag=1, xs=0.0, ys=0.0, zs=0.0201, radius=0.05, waist=0.05,
                zsurf=0.02, mckernelflag=P.SIM_MCXYZ, det_state=P.DET_REFLECTANCE,
                det_radius=0.1, na=60.0)
    tp = _tp({1: (0.01, 1.0, 1.0), 2: (2.0, 60.0, 0.8), 3: (8.0, 30.0, 0.5),
              4: (1.0, 40.0, 0.9, 0.9, 0.5, 0.3, 0.5, 0.8)})
    if kind == "mismatch":
        base.update(n1=1.4, n2=1.0)
    elif kind == "matched":
        base.update(n1=1.0, n2=1.0, det_radius=1e6, na=90.0)
    elif kind == "gaussian":
        base.update(mcflag=1, n1=1.37, n2=1.0, zs=0.0, radius=0.03, waist=0.02, zfocus=0.1)
    elif kind == "iso":
        base.update(mcflag=2, zs=0.1, xs=0.01, ys=-0.02, n1=1.33, n2=1.0)
    elif kind == "rect":
        base.update(mcflag=3, radius=0.08, n1=1.33, n2=1.0)
    elif kind == "point":
        base.update(mcflag=4, n1=1.33, n2=1.0)
    elif kind == "manual":
        base.update(launchflag=1, xs=0.01, ys=0.0, zs=0.05, ux0=0.6, uy0=0.0, uz0=0.8,
                    n1=1.33, n2=1.0)
    elif kind == "infinite":
        base.update(boundaryflag=0, n1=1.33, n2=1.0)
    elif kind == "surface_only":
        base.update(boundaryflag=2, n1=1.33, n2=1.0)
    elif kind == "transmission":
        base.update(det_state=P.DET_TRANSMITTANCE, zsurf=0.2, n1=1.4, n2=1.0, zs=0.0201,
                    det_radius=0.2, xd=0.05)
    elif kind == "two_term":
        base.update(mckernelflag=P.SIM_MCXYZ_TT, n1=1.4, n2=1.0)
    else:
        raise ValueError(kind)
    rp = _rp("tiny_" + kind, **base)

    def lab(X, Y, Z):
        t = np.full(X.shape, 2, np.uint8)
        t[(X ** 2 + (Z - 0.12) ** 2) <= 0.06 ** 2] = 3
        t[Z < 0.02] = 1
        if kind == "two_term":
            t[t == 2] = 4
        return t

    v = layered_volume(n, n, n, d, d, lab)
    if kind == "two_term":
        tp.gf[1], tp.af[1], tp.gb[1], tp.ab[1], tp.CC[1] = 0.0, 0.5, 0.0, 0.5, 1.0
        tp.gf[3], tp.af[3], tp.gb[3], tp.ab[3], tp.CC[3] = 0.7, 0.7, 0.2, 1.0, 0.6
    return Workload("tiny_" + kind, rp.mckernelflag, rp, tp, v,
                    P.ExtConfig.default(seed=seed), n_photons)


def example0_slab(n_photons: int = 100_000, seed: int = 1) -> Workload:
    """The homogeneous run of NMC-samples/example0.ipynb: 200^3 voxels of
    0.02, mua 1, mus 100, g 0.9, n 1.4 against 1.001, a flat beam of radius
    1.2 entering at z = zsurf = 0, escape at the surface only."""
    n = 200
    rp = _rp("test", Nx=n, Ny=n, Nz=n, dx=0.02, dy=0.02, dz=0.02, Nt=1, mcflag=0, launchflag=0,
             boundaryflag=2, xs=0.0, ys=0.0, zs=0.0, zfocus=1e12, radius=1.2, waist=1.2,
             zsurf=0.0, det_radius=5.0, xd=0.0, n1=1.4, n2=1.001, na=90.0, lambda_=640.0e-6,
             lc=0.001, semiflag=1, time_min=1.0, mckernelflag=P.SIM_MCXYZ,
             det_state=P.DET_REFLECTANCE, ux0=0.7, uy0=0.4, uz0=(1 - 0.7 ** 2 - 0.4 ** 2) ** 0.5)
    tp = _tp({1: (1.0, 100.0, 0.9)})
    v = np.ones(n * n * n, np.uint8)
    return Workload("example0_slab", P.SIM_MCXYZ, rp, tp, v, P.ExtConfig.default(seed=seed),
                    n_photons)
