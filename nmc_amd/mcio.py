"""The reference's file contract: `<name>_H.mci` and `<name>_T.bin` in;
`_F.bin`, `_Rd.dat`, `_Ryx.dat`, `_DetPhot2D.dat`, `_props.m` and
`mcOUT0.dat` out, as the notebooks read them.

Readers follow NMC/NMCCode/mcio.h:16-189: one value per line, each read with
fgets(buf, 32) and parsed with sscanf "%f" / "%d" (so a line longer than 31
characters spills into the next read, exactly as there).  Writers follow the
printf formats of mcio.h:191-389 and the output code of main.m:209-566.
Where the reference reads garbage or overflows (no fopen/fread checks,
mcio.h:30-31,184-186; Nt unchecked against the 19-entry tissue arrays) these
raise `NmcError` instead.  Where it indexes a non-square grid wrongly
(main.m:320-333,529-563 use j + i*Nx) these follow the data layout
ix*Ny + iy, which is the same thing when Nx == Ny.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

from . import params as P
from ._lib import NmcError

PI_REF = 3.1415926    # definitions.h:60, used by the mcsub normalisation

_FLOAT = re.compile(r"\s*([+-]?(?:inf(?:inity)?|nan|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?))",
                    re.IGNORECASE)
_INT = re.compile(r"\s*([+-]?\d+)")

SIM_POLARIZATION_CBS = 3    # definitions.h:156, read like mcxyz (mcio.h:120-137)
SIM_RAMAN = 5               # definitions.h:158


def _scan_float(buf: str) -> float | None:
    """sscanf(buf, "%f"): the float32 value of the leading number, or None."""
    m = _FLOAT.match(buf)
    if not m:
        return None
    with np.errstate(over="ignore"):
        return float(np.float32(float(m.group(1))))


def _scan_int(buf: str) -> int | None:
    m = _INT.match(buf)
    if not m:
        return None
    v = int(m.group(1))
    return v if -(1 << 31) <= v < (1 << 31) else None


class _Reader:
    """fgets(buf, 32, fid): pieces of at most 31 characters, ending after a
    newline; a longer line continues in the next piece."""

    def __init__(self, text: str, path: str):
        self.text, self.pos, self.path = text, 0, path

    def _next(self, what: str) -> str:
        if self.pos >= len(self.text):
            raise NmcError(P.ERR_ARG, f"{self.path}: file ends before {what}")
        end = self.text.find("\n", self.pos)
        end = len(self.text) if end < 0 else end + 1
        end = min(end, self.pos + 31)
        piece = self.text[self.pos:end]
        self.pos = end
        return piece

    def f(self, what: str) -> float:
        v = _scan_float(self._next(what))
        if v is None:
            raise NmcError(P.ERR_ARG, f"{self.path}: {what} is not a number")
        return v

    def d(self, what: str) -> int:
        v = _scan_int(self._next(what))
        if v is None:
            raise NmcError(P.ERR_ARG, f"{self.path}: {what} is not an integer")
        return v


# (RunParams field, conversion) in file order, mcio.h:30-116
RUN_FIELDS = (
    ("time_min", "f"), ("Nx", "d"), ("Ny", "d"), ("Nz", "d"), ("dx", "f"), ("dy", "f"),
    ("dz", "f"), ("mcflag", "d"), ("launchflag", "d"), ("boundaryflag", "d"), ("xs", "f"),
    ("ys", "f"), ("zs", "f"), ("xfocus", "f"), ("yfocus", "f"), ("zfocus", "f"), ("ux0", "f"),
    ("uy0", "f"), ("uz0", "f"), ("radius", "f"), ("waist", "f"), ("zsurf", "f"),
    ("det_radius", "f"), ("xd", "f"), ("n1", "f"), ("n2", "f"), ("na", "f"), ("lambda_", "f"),
    ("lc", "f"), ("speckleflag", "d"), ("semiflag", "d"), ("det_state", "d"),
    ("mckernelflag", "d"),
)
TT_FIELDS = ("gf", "af", "gb", "ab", "CC")
READABLE_KERNELS = (P.SIM_MCSUB, P.SIM_MCXYZ, P.SIM_MCXYZ_TT, SIM_POLARIZATION_CBS, SIM_RAMAN)


def read_run_params(name: str) -> tuple[P.RunParams, P.TissueParams]:
    """ReadRunParams (mcio.h:16-171): parses `<name>_H.mci`."""
    path = name + "_H.mci"
    try:
        with open(path, "r", newline="") as f:
            text = f.read()
    except OSError as e:
        raise NmcError(P.ERR_ARG, f"cannot open {path}: {e.strerror}") from None
    rp, tp = P.RunParams(), P.TissueParams()
    rp.myname = name.encode()[: P.STRLEN - 1]
    rd = _Reader(text, path)
    for field, kind in RUN_FIELDS:
        setattr(rp, field, rd.f(field) if kind == "f" else rd.d(field))
    k = rp.mckernelflag
    if k not in READABLE_KERNELS:
        raise NmcError(P.ERR_UNSUPPORTED, f"{path}: unknown or unsupported kernel {k}")
    rp.Nt = rd.d("Nt")
    if not 0 <= rp.Nt < P.NTISS:
        raise NmcError(P.ERR_ARG, f"{path}: Nt = {rp.Nt} is not in 0..{P.NTISS - 1}")
    for i in range(1, rp.Nt + 1):
        tp.muav[i] = rd.f(f"muav[{i}]")
        tp.musv[i] = rd.f(f"musv[{i}]")
        if k == P.SIM_MCXYZ_TT:
            for nm in TT_FIELDS:
                getattr(tp, nm)[i] = rd.f(f"{nm}[{i}]")
        else:
            tp.gv[i] = rd.f(f"gv[{i}]")
    return rp, tp


def import_tissue(name: str, rp: P.RunParams) -> np.ndarray:
    """ImportBinaryTissueFile (mcio.h:173-189): the Nx*Ny*Nz bytes of
    `<name>_T.bin`, linear index iz*Ny*Nx + ix*Ny + iy."""
    path = name + "_T.bin"
    n = int(rp.Nx) * int(rp.Ny) * int(rp.Nz)
    if n <= 0:
        raise NmcError(P.ERR_ARG, f"{path}: grid {rp.Nx} x {rp.Ny} x {rp.Nz} is empty")
    try:
        v = np.fromfile(path, dtype=np.uint8, count=n)
    except OSError as e:
        raise NmcError(P.ERR_ARG, f"cannot open {path}: {e.strerror}") from None
    if v.size != n:
        raise NmcError(P.ERR_ARG, f"{path}: {v.size} bytes, expected Nx*Ny*Nz = {n}")
    return v


def _num(v: float) -> str:
    if math.isnan(v):
        return "nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return repr(float(v))


def write_mci(name: str, rp: P.RunParams, tp: P.TissueParams) -> str:
    """Writes `<name>_H.mci` in the reader's order, one value per line, each
    float in its shortest exact form (so reading it back is lossless)."""
    if not 0 <= rp.Nt < P.NTISS:
        raise NmcError(P.ERR_ARG, f"Nt = {rp.Nt} is not in 0..{P.NTISS - 1}")
    lines =["%d" % getattr(rp, f) if kind == "d" else _num(getattr(rp, f))
             for f, kind in RUN_FIELDS]
    lines.append("%d" % rp.Nt)
    for i in range(1, rp.Nt + 1):
        vals = [tp.muav[i], tp.musv[i]]
        if rp.mckernelflag == P.SIM_MCXYZ_TT:
            vals += [getattr(tp, nm)[i] for nm in TT_FIELDS]
        else:
            vals.append(tp.gv[i])
        lines += [_num(x) for x in vals]
    path = name + "_H.mci"
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def write_tissue(name: str, v: np.ndarray) -> str:
    path = name + "_T.bin"
    np.ascontiguousarray(v, np.uint8).tofile(path)
    return path


_KERNEL_TEXT = {
    P.SIM_MCSUB: "performing mcsub",
    P.SIM_MCXYZ: "performing mcxyz",
    P.SIM_MCXYZ_TT: "performing mcxyztt",
    SIM_POLARIZATION_CBS: "performing polarization simulations using electric field tracking method",
    SIM_RAMAN: "performing mcraman",
}
_BOUNDARY_TEXT = {0: "boundaryflag = 0, so no boundaries.",
                  1: "boundaryflag = 1, so escape at all boundaries.",
                  2: "boundaryflag = 2, so escape at surface only."}


def run_parameters_text(rp: P.RunParams, tp: P.TissueParams, v: np.ndarray) -> str:
    """PrintRunParameters (mcio.h:191-296) as a string.  The reference exits
    on a bad boundaryflag or kernel; this raises NmcError."""
    if rp.boundaryflag not in _BOUNDARY_TEXT:
        raise NmcError(P.ERR_ARG, "improper boundaryflag")
    if rp.mckernelflag not in _KERNEL_TEXT:
        raise NmcError(P.ERR_UNSUPPORTED, "unknown or unsupported kernel")
    o = ["time_min = %0.2f min" % rp.time_min,
         "Nx = %d, dx = %0.4f [mm]" % (rp.Nx, rp.dx),
         "Ny = %d, dy = %0.4f [mm]" % (rp.Ny, rp.dy),
         "Nz = %d, dz = %0.4f [mm]" % (rp.Nz, rp.dz),
         "xs = %0.4f [mm]" % rp.xs,
         "ys = %0.4f [mm]" % rp.ys,
         "zs = %0.4f [mm]" % rp.zs,
         "mcflag = %d [mm]" % rp.mcflag]
    o += {0: ["launching uniform flat-field beam"], 1: ["launching Gaissian beam"],
          2: ["launching isotropic point source"]}.get(rp.mcflag, [])
    o += ["xfocus = %0.4f [mm]" % rp.xfocus,
          "yfocus = %0.4f [mm]" % rp.yfocus,
          "zfocus = %0.2e [mm]" % rp.zfocus]
    if rp.launchflag == 1:
        o += ["Launchflag ON, so launch the following:",
              "ux0 = %0.4f [mm]" % rp.ux0, "uy0 = %0.4f [mm]" % rp.uy0,
              "uz0 = %0.4f [mm]" % rp.uz0]
    else:
        o += ["Launchflag OFF, so program calculates launch angles.",
              "radius = %0.4f [mm]" % rp.radius, "waist  = %0.4f [mm]" % rp.waist]
    o += ["zsurf  = %0.4f [mm]" % rp.zsurf, _BOUNDARY_TEXT[rp.boundaryflag],
          _KERNEL_TEXT[rp.mckernelflag], "# of tissues available, Nt = %d" % rp.Nt]
    for i in range(1, rp.Nt + 1):
        o.append("muav[%d] = %0.4f [mm^-1]" % (i, tp.muav[i]))
        o.append("musv[%d] = %0.4f [mm^-1]" % (i, tp.musv[i]))
        if rp.mckernelflag == P.SIM_MCXYZ_TT:
            o += ["%s(%d) = %0.4f;" % (nm, i, getattr(tp, nm)[i]) for nm in TT_FIELDS]
        else:
            o.append("  gv[%d] = %0.4f [--]\n" % (i, tp.gv[i]))
    o += ["n1 = %0.4f [mm]" % rp.n1, "n2 = %0.4f [mm]" % rp.n2,
          "central axial profile of tissue types:"]
    ix, iy = rp.Nx // 2, rp.Ny // 2
    o.append("".join("%d" % np.int8(v[iz * rp.Ny * rp.Nx + ix * rp.Ny + iy])
                     for iz in range(rp.Nz)))
    return "\n".join(o) + "\n\n"


def save_optical_properties(rp: P.RunParams, tp: P.TissueParams, name: str) -> str:
    """SaveOpticalProperties (mcio.h:298-322): `<name>_props.m`."""
    path = name + "_props.m"
    with open(path, "w") as f:
        for i in range(1, rp.Nt + 1):
            f.write("muav(%d) = %0.4f;\n" % (i, tp.muav[i]))
            f.write("musv(%d) = %0.4f;\n" % (i, tp.musv[i]))
            if rp.mckernelflag == P.SIM_MCXYZ_TT:
                for nm in TT_FIELDS:
                    f.write("%s(%d) = %0.4f;\n\n" % (nm, i, getattr(tp, nm)[i]))
            else:
                f.write("gv(%d) = %0.4f;\n\n" % (i, tp.gv[i]))
    return path


def save_mcsub_file(path: str, J: np.ndarray, F: np.ndarray, S: float, A: float, E: float,
                    mua: float, mus: float, g: float, n1: float, n2: float, mcflag: int,
                    radius: float, waist: float, xs: float, ys: float, zs: float, dr: float,
                    dz: float, nphotons: float) -> str:
    """SaveFile (mcio.h:327-389): the mcsub table `mcOUT0.dat`.

    J has NR+1 entries and F has shape (NR+1, NZ+1), both indexed from 1 as in
    the reference ([0] unused); the reference keeps F flat at ir*NR + iz, which
    is the same table when NR == NZ."""
    NR, NZ = F.shape[0] - 1, F.shape[1] - 1
    J = np.asarray(J, np.float32)
    F = np.asarray(F, np.float32)
    if J.shape[0] != NR + 1:
        raise NmcError(P.ERR_ARG, "J and F disagree on the number of radial bins")
    out = ["%0.3e\tmua, absorption coefficient [1/mm]\n" % mua,
           "%0.4f\tmus, scattering coefficient [1/mm]\n" % mus,
           "%0.4f\tg, anisotropy [-]\n" % g,
           "%0.4f\tn1, refractive index of tissue\n" % n1,
           "%0.4f\tn2, refractive index of outside medium\n" % n2,
           "%d\tmcflag\n" % mcflag,
           "%0.4f\tradius, radius of flat beam or 1/e radius of Gaussian beam [mm]\n" % radius,
           "%0.4f\twaist, 1/e waist of focus [mm]\n" % waist,
           "%0.4f\txs, x position of isotropic source [mm]\n" % xs,
           "%0.4f\tys, y\n" % ys,
           "%0.4f\tzs, z\n" % zs,
           "%d\tNR\n" % NR,
           "%d\tNZ\n" % NZ,
           "%0.5f\tdr\n" % dr,
           "%0.5f\tdz\n" % dz,
           "%0.1e\tNphotons\n" % nphotons,
           "%1.6e\tSpecular reflectance\n" % S,
           "%1.6e\tAbsorbed fraction\n" % A,
           "%1.6e\tEscaping fraction\n" % E]
    row = ["%0.1f" % 0.0]
    for ir in range(1, NR + 1):
        r2, r1 = dr * ir, dr * (ir - 1)
        row.append("\t%1.5f" % (2.0 / 3 * (r2 * r2 + r2 * r1 + r1 * r1) / (r1 + r2)))
    out.append("".join(row) + "\n")
    out.append("%0.1f" % 0.0 + "".join("\t%1.12e" % J[ir] for ir in range(1, NR + 1)) + "\n")
    for iz in range(1, NZ + 1):
        out.append("%1.5f" % ((iz - 0.5) * dz)
                   + "".join("\t %1.6e" % F[ir, iz] for ir in range(1, NR + 1)) + "\n")
    with open(path, "w") as t:
        t.write("".join(out))
    return path


def fluence_rate(deposit: np.ndarray, rp: P.RunParams, tp: P.TissueParams, v: np.ndarray,
                 n_photons: int) -> np.ndarray:
    """main.m:481-490: deposit / (dx dy dz N mua[v]) in float32, NaN/Inf -> 0."""
    temp = np.float32(float(np.float32(np.float32(rp.dx) * np.float32(rp.dy))
                            * np.float32(rp.dz)) * float(n_photons))
    mua = np.frombuffer(bytes(tp), np.float32, P.NTISS)[np.asarray(v, np.uint8)]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        out = np.asarray(deposit).astype(np.float32) / (temp * mua)
    out[~np.isfinite(out)] = 0.0
    return out.astype(np.float32)


def save_rd(path: str, rd: float) -> str:
    """main.m:250-255 / 513-519: "%0.4f\\n"."""
    with open(path, "w") as f:
        f.write("%0.4f\n" % rd)
    return path


def save_grid_text(path: str, grid: np.ndarray, nx: int, ny: int, integer: bool) -> str:
    """main.m:522-563: Nx lines of Ny values, "\\t %1.6e" (NaN/Inf -> 0) or
    "\\t %d"; `grid` in the data layout ix*Ny + iy."""
    g = np.asarray(grid).reshape(nx, ny)
    if not integer:
        g = g.astype(np.float32)
    with open(path, "w") as f:
        for i in range(nx):
            if integer:
                f.write("".join("\t %d" % int(x) for x in g[i]) + "\n")
            else:
                f.write("".join("\t %1.6e" % (x if np.isfinite(x) else 0.0) for x in g[i])
                        + "\n")
    return path


def write_mcxyz_outputs(name: str, rp: P.RunParams, tp: P.TissueParams, v: np.ndarray,
                        n_photons: int, fluence: np.ndarray, rd_total: float, ryx: np.ndarray,
                        det2d: np.ndarray) -> list[str]:
    """main.m:474-566: `_F.bin`, `_Rd.dat`, `_Ryx.dat`, `_DetPhot2D.dat`."""
    F = fluence_rate(fluence, rp, tp, v, n_photons)
    paths = [name + "_F.bin"]
    F.tofile(paths[0])
    paths.append(save_rd(name + "_Rd.dat", rd_total / n_photons))
    paths.append(save_grid_text(name + "_Ryx.dat", ryx, rp.Nx, rp.Ny, integer=False))
    paths.append(save_grid_text(name + "_DetPhot2D.dat", det2d, rp.Nx, rp.Ny, integer=True))
    return paths


def mcsub_tables(rd_ra: np.ndarray, a_rz: np.ndarray | None, nr: int, nz: int, na: int,
                 dr: float, dz: float, mua: float, n_photons: int):
    """main.m:221-236: J[ir] = escaping weight / (2 pi r dr N) and
    F[ir][iz] = deposit / (2 pi r dr dz N mua), r = (ir - 1/2) dr, 1-based with
    the overflow bin last (NR = nr + 1 rows)."""
    NR, NZ = nr + 1, nz + 1
    J = np.zeros(NR + 1, np.float32)
    F = np.zeros((NR + 1, NZ + 1), np.float32)
    ej = np.asarray(rd_ra, np.float64).reshape(NR, na).sum(axis=1)
    dep = None if a_rz is None else np.asarray(a_rz, np.float64).reshape(NR, NZ)
    for ir in range(1, NR + 1):
        r = float(np.float32((ir - 0.5) * dr))
        ring = 2.0 * PI_REF * r * dr
        J[ir] = np.float32(np.float32(ej[ir - 1]) / (ring * n_photons))
        if dep is not None:
            F[ir, 1:] = (dep[ir - 1].astype(np.float32).astype(np.float64)
                         / (ring * dz * n_photons * mua)).astype(np.float32)
    return J, F


def write_mcsub_outputs(name: str, rp: P.RunParams, tp: P.TissueParams, ext: P.ExtConfig,
                        n_photons: int, totals: dict, rd_ra: np.ndarray,
                        a_rz: np.ndarray | None) -> list[str]:
    """main.m:203-263: `_Rd.dat` (escaping + specular fraction) and `mcOUT0.dat`
    next to it (the reference writes it to the working directory)."""
    n = float(n_photons)
    S, A, E = totals["specular"] / n, totals["absorbed"] / n, totals["rd"] / n
    paths = [save_rd(name + "_Rd.dat", float(np.float32(E) + np.float32(S)))]
    J, F = mcsub_tables(rd_ra, a_rz, ext.nr, ext.nz, ext.na, ext.dr, ext.dz, tp.muav[1], n_photons)
    out = os.path.join(os.path.dirname(name), "mcOUT0.dat")
    paths.append(save_mcsub_file(out, J, F, S, A, E, tp.muav[1], tp.musv[1], tp.gv[1], rp.n1,
                                 rp.n2, rp.mcflag, rp.radius, rp.waist, rp.xs, rp.ys, rp.zs,
                                 ext.dr, ext.dz, n))
    return paths
