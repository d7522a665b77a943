"""GPU: the `nmc <name>` driver end to end.

Input files are written in the reference's format, the driver runs on the
GPU, and every output file is compared with the same writers applied to the
oracle's tallies for the same photons (integer files identical, float files
to float32 rounding).
"""
from __future__ import annotations

import io
import os
import subprocess
import sys

import numpy as np
import pytest

from nmc_amd import cli, mcio
from nmc_amd import params as P
from nmc_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    print(msg, flush=True)


def _table(path):
    return [line.split("\t") for line in open(path).read().splitlines()]


def test_cli_mcxyz_files_match_oracle(gpu, tmp_path):
    w = W.tiny_voxel_case("mismatch")
    n = 20000
    name = str(tmp_path / "tiny")
    mcio.write_mci(name, w.rp, w.tp)
    mcio.write_tissue(name, w.v)
    _say("[cli] mcxyz tiny n=20000")
    buf = io.StringIO()
    got = cli.run(name, n, seed=w.ext.seed, out=buf)
    assert "performing mcxyz" in buf.getvalue()
    rp, tp = mcio.read_run_params(name)
    ref = O.run(W.Workload("tiny", P.SIM_MCXYZ, rp, tp, w.v, got["ext"], n), n, threads=4)
    for k in P.TOTAL_INT_FIELDS:
        assert got["totals"][k] == ref.totals[k], k
    oname = str(tmp_path / "oracle")
    mcio.write_mcxyz_outputs(oname, rp, tp, w.v, n, ref.fluence, ref.totals["rd"], ref.ryx,
                             ref.det2d)
    for suffix in ("_DetPhot2D.dat", "_Rd.dat"):
        assert open(name + suffix).read() == open(oname + suffix).read(), suffix
    np.testing.assert_allclose(np.fromfile(name + "_F.bin", np.float32),
                               np.fromfile(oname + "_F.bin", np.float32), rtol=1e-6, atol=0)
    a = np.array([[float(x) for x in r[1:]] for r in _table(name + "_Ryx.dat")])
    b = np.array([[float(x) for x in r[1:]] for r in _table(oname + "_Ryx.dat")])
    assert a.shape == (rp.Nx, rp.Ny)
    np.testing.assert_allclose(a, b, rtol=2e-6, atol=0)
    assert os.path.exists(name + "_props.m")


def test_cli_mcsub_files_match_oracle(gpu, tmp_path):
    w = W.mcsub_semi_infinite()
    rp = P.RunParams.from_buffer_copy(w.rp)
    rp.Nx = rp.Ny = rp.Nz = 1
    rp.Nt = 1
    name = str(tmp_path / "sub")
    mcio.write_mci(name, rp, w.tp)
    mcio.write_tissue(name, np.ones(1, np.uint8))
    n = 20000
    _say("[cli] mcsub n=20000")
    got = cli.run(name, n, seed=1, out=io.StringIO())
    rp2, tp2 = mcio.read_run_params(name)
    ref = O.run(W.Workload("sub", P.SIM_MCSUB, got["run_params"], tp2, None, got["ext"], n), n,
                threads=4)
    for k in P.TOTAL_INT_FIELDS:
        assert got["totals"][k] == ref.totals[k], k
    odir = tmp_path / "o"
    odir.mkdir()
    mcio.write_mcsub_outputs(str(odir / "sub"), got["run_params"], tp2, got["ext"], n,
                             ref.totals, ref.rd_ra, ref.a_rz)
    assert open(name + "_Rd.dat").read() == open(str(odir / "sub_Rd.dat")).read()
    mine, theirs = _table(str(tmp_path / "mcOUT0.dat")), _table(str(odir / "mcOUT0.dat"))
    assert len(mine) == 19 + 2 + 101
    assert mine[:19] == theirs[:19]          # parameter header and S, A, E
    A =np.array([[float(x) for x in r] for r in mine[19:]])
    B = np.array([[float(x) for x in r] for r in theirs[19:]])
    np.testing.assert_allclose(A, B, rtol=2e-6, atol=0)


def test_cli_command_line(gpu, tmp_path):
    w = W.tiny_voxel_case("two_term")
    name = str(tmp_path / "tt")
    mcio.write_mci(name, w.rp, w.tp)
    mcio.write_tissue(name, w.v)
    _say("[cli] python -m nmc_amd.cli tt --photons 5000")
    r = subprocess.run([sys.executable, "-m", "nmc_amd.cli", name, "--photons", "5000"],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "performing mcxyztt" in r.stdout and "saving" in r.stdout
    F = np.fromfile(name + "_F.bin", np.float32)
    assert F.size == w.rp.Nx * w.rp.Ny * w.rp.Nz and np.isfinite(F).all() and F.max() > 0
    r = subprocess.run([sys.executable, "-m", "nmc_amd.cli", str(tmp_path / "missing")],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "cannot open" in r.stderr
