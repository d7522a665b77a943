"""`nmc <name>`: the reference's command-line driver (NMC/NMCCode/main.m) on the
HIP library.

    python -m nmc_amd.cli <name> [--photons N] [--seed S] [--device D]

reads `<name>_H.mci` and `<name>_T.bin`, echoes the run parameters
(PrintRunParameters), writes `<name>_props.m`, transports the photons on the
GPU and writes the files the notebooks read:
  mcxyz / mcxyztt: `<name>_F.bin`, `<name>_Rd.dat`, `<name>_Ryx.dat`,
                   `<name>_DetPhot2D.dat`                       (main.m:474-566)
  mcsub:           `<name>_Rd.dat`, `mcOUT0.dat` beside it      (main.m:203-263)

As in the reference, the photon count defaults to 1e5 (definitions.h:16) and
mcsub takes its beam from fixed settings (a pencil beam at the origin onto an
air/tissue surface, definitions.h:37-49) and its optical properties from
tissue 1 and n1/n2 of the .mci file; `--file-beam` takes the beam from the
file instead.  The mcsub bins are NR = NZ = 101 (the last one overflow) of
0.002 (definitions.h:31,52-53) unless overridden.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import mcio
from . import params as P
from ._lib import NmcError
from .metal_mc import MetalMC

PHOTON_TOTAL = 100_000        # definitions.h:16
MCSUB_BINS = 101              # definitions.h:31 (overflow bin included)
MCSUB_DR = MCSUB_DZ = 0.002   # definitions.h:52-53
# definitions.h:37-49: the beam mcsub is compiled with
MCSUB_BEAM = {"mcflag": 0, "radius": 0.0, "waist": 0.10, "zfocus": 1.0, "xs": 0.0, "ys": 0.0,
              "zs": 0.0, "boundaryflag": 1}

_WELCOME = {P.SIM_MCSUB: "Welcome to mcsub ...", P.SIM_MCXYZ: "Welcome to mcxyz...",
            P.SIM_MCXYZ_TT: "Welcome to mcxyztt..."}


def run(name: str, n_photons: int = PHOTON_TOTAL, seed: int = 1, device: int = 0,
        file_beam: bool = False, nr: int | None = None, nz: int | None = None,
        dr: float | None = None, dz: float | None = None, out=None) -> dict:
    """Runs one `<name>` job; returns the written paths and the totals."""
    out = out if out is not None else sys.stdout

    def say(s: str) -> None:
        print(s, file=out, flush=True)

    if n_photons <= 0:
        raise NmcError(P.ERR_ARG, "the photon count must be positive")
    say(f"name = {name}")
    rp, tp = mcio.read_run_params(name)
    sim = rp.mckernelflag
    if sim not in _WELCOME:
        raise NmcError(P.ERR_UNSUPPORTED,
                       f"kernel {sim} (polarisation / Raman) is not built in this library")
    say(_WELCOME[sim])
    if sim == P.SIM_MCSUB and not os.path.exists(name + "_T.bin"):
        v = np.zeros(max(int(rp.Nx) * int(rp.Ny) * int(rp.Nz), 0), np.uint8)
    else:
        v = mcio.import_tissue(name, rp)
    if v.size:
        out.write(mcio.run_parameters_text(rp, tp, v))
    props = mcio.save_optical_properties(rp, tp, name)

    ext = P.ExtConfig.default(seed=seed)
    if sim == P.SIM_MCSUB:
        run_rp = P.RunParams.from_buffer_copy(rp)
        if not file_beam:
            for k, val in MCSUB_BEAM.items():
                setattr(run_rp, k, val)
        ext.nr = (nr if nr is not None else MCSUB_BINS - 1)
        ext.nz = (nz if nz is not None else MCSUB_BINS - 1)
        ext.dr = dr if dr is not None else MCSUB_DR
        ext.dz = dz if dz is not None else MCSUB_DZ
        vox = None
    else:
        run_rp, vox = rp, v
    t0 = time.perf_counter()
    with MetalMC(sim, device) as mc:
        mc.prepare_data(run_rp, tp, vox, ext)
        mc.set_num_runs_seed(seed, 0)
        mc.send_compute_command(n_photons)
        res = mc.get_compute_results()
        kernel_ms = mc.last_kernel_ms()
    wall = time.perf_counter() - t0
    say(f"Total execution time (GPU):{wall:f}\nSimulation rate {int(n_photons / wall)} "
        f"[photons/sec] (kernel {kernel_ms:.3f} ms)")
    tot = res.totals
    if sim == P.SIM_MCSUB:
        n = float(n_photons)
        E, A, S = tot["rd"] / n, tot["absorbed"] / n, tot["specular"] / n
        say(f"Escaping fraction is: {E:g}\nAbsorbed fraction is: {A:g}\nReflectance is: {S:g}\n"
            f"Total is: {E + A + S:g}\nRd is: {E + S:g}")
        paths = mcio.write_mcsub_outputs(name, run_rp, tp, ext, n_photons, tot, res.rd_ra,
                                         res.a_rz)
    else:
        say("Total Rd = %0.8f" % (tot["rd"] / n_photons))
        paths = mcio.write_mcxyz_outputs(name, rp, tp, v, n_photons, res.fluence, tot["rd"],
                                         res.ryx, res.det2d)
    for p in paths:
        say(f"saving {p}")
    return {"paths": [props] + paths, "totals": tot, "kernel_ms": kernel_ms, "wall_s": wall,
            "ext": ext, "run_params": run_rp}


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="nmc", description=__doc__.split("\n\n")[0])
    ap.add_argument("name", help="reads <name>_H.mci and <name>_T.bin")
    ap.add_argument("--photons", type=float, default=PHOTON_TOTAL)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--file-beam", action="store_true",
                    help="mcsub: take the beam from the .mci file instead of the fixed pencil beam")
    ap.add_argument("--nr", type=int)
    ap.add_argument("--nz", type=int)
    ap.add_argument("--dr", type=float)
    ap.add_argument("--dz", type=float)
    a = ap.parse_args(argv)
    try:
        run(a.name, int(a.photons), a.seed, a.device, a.file_beam, a.nr, a.nz, a.dr, a.dz)
    except (NmcError, OSError, ValueError) as e:   # library errors, unreadable or malformed files
        print(f"nmc: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
