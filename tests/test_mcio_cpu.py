"""File contract (nmc_amd/mcio.py) on the CPU.

The reference's file routines (NMC/NMCCode/mcio.h) cannot be compiled here:
definitions.h needs Objective-C's YES/NO, which only the absent Foundation
headers define.  The checks therefore pin the formats against the printf
strings of mcio.h / main.m and the normalisations against closed forms.
"""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from nmc_amd import mcio
from nmc_amd import params as P
from nmc_amd import workloads as W
from nmc_amd._lib import NmcError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def notebook_mci(path_name: str, rp: P.RunParams, tp: P.TissueParams, props=None) -> None:
    """The notebooks' writer (NMC-samples/example2.ipynb:533-562): fixed
    decimals, `inf` for an infinite focus, tissues 1..Nt.  `props` gives the
    tissues' (mua, mus, g) in double precision, as the notebooks hold them."""
    f4 = lambda x: f"{x:.4f}"  # noqa: E731
    lines = [f"{rp.time_min:.2f}", str(rp.Nx), str(rp.Ny), str(rp.Nz), f4(rp.dx), f4(rp.dy),
             f4(rp.dz), str(rp.mcflag), str(rp.launchflag), str(rp.boundaryflag), f4(rp.xs),
             f4(rp.ys), f4(rp.zs), f4(rp.xfocus), f4(rp.yfocus), f4(rp.zfocus), f4(rp.ux0),
             f4(rp.uy0), f4(rp.uz0), f4(rp.radius), f4(rp.waist), f4(rp.zsurf),
             f4(rp.det_radius), f4(rp.xd), f4(rp.n1), f4(rp.n2), f4(rp.na), f4(rp.lambda_),
             f4(rp.lc), str(rp.speckleflag), str(rp.semiflag), str(rp.det_state),
             str(rp.mckernelflag), str(rp.Nt)]
    for i in range(1, rp.Nt + 1):
        mua, mus, g = props[i - 1] if props else (tp.muav[i], tp.musv[i], tp.gv[i])
        lines += [f4(mua), f4(mus)]
        if rp.mckernelflag < 2:
            lines.append(f4(g))
        else:
            lines += [f4(getattr(tp, nm)[i]) for nm in mcio.TT_FIELDS]
    with open(path_name + "_H.mci", "w") as f:
        f.write("\n".join(lines) + "\n")


def small_case(kernel=P.SIM_MCXYZ):
    w = W.tiny_voxel_case("two_term" if kernel == P.SIM_MCXYZ_TT else "mismatch")
    rp = P.RunParams.from_buffer_copy(w.rp)
    rp.mckernelflag = kernel
    rp.time_min = 10.0
    rp.lambda_, rp.lc = 0.532, 0.01
    rp.Nt = 4
    tp = P.TissueParams.from_buffer_copy(w.tp)
    # the file holds g for one-term kernels and (gf, af, gb, ab, CC) for the
    # two-term kernel only (mcio.h:118-167)
    for nm in (("gv",) if kernel == P.SIM_MCXYZ_TT else mcio.TT_FIELDS):
        for i in range(P.NTISS):
            getattr(tp, nm)[i] = 0.0
    return rp, tp, w.v


def struct_bytes(rp, tp) -> bytes:
    """Both structs, without the name field (the caller's path)."""
    return bytes(rp)[P.STRLEN:] + bytes(tp)


def test_mci_round_trip_is_lossless(tmp_path):
    for kernel in (P.SIM_MCXYZ, P.SIM_MCXYZ_TT, P.SIM_MCSUB):
        rp, tp, _ = small_case(kernel)
        rp.zfocus = math.inf
        name = str(tmp_path / f"k{kernel}")
        mcio.write_mci(name, rp, tp)
        rp2, tp2 = mcio.read_run_params(name)
        assert struct_bytes(rp2, tp2) == struct_bytes(rp, tp)
        assert rp2.myname == name.encode()[:31]


def test_notebook_style_file_parses_to_float32_of_the_text(tmp_path):
    rp, tp, _ = small_case()
    rp.zfocus = math.inf
    rp.dx = rp.dy = rp.dz = 0.0005
    name = str(tmp_path / "nb")
    notebook_mci(name, rp, tp)
    rp2, tp2 = mcio.read_run_params(name)
    assert math.isinf(rp2.zfocus)
    assert rp2.dx == np.float32(0.0005) and rp2.Nx == rp.Nx
    assert tp2.musv[2] == np.float32(float(f"{tp.musv[2]:.4f}"))


def test_tissue_file_layout_matches_the_notebook_transpose(tmp_path):
    """T[Ny, Nx, Nz].ravel(order='F') (example2.ipynb:562) is index
    iz*Ny*Nx + ix*Ny + iy (kernels.metal:482)."""
    ny, nx, nz = 3, 4, 5
    T = np.arange(ny * nx * nz, dtype=np.uint8).reshape(ny, nx, nz)
    rp = P.RunParams(Nx=nx, Ny=ny, Nz=nz)
    name = str(tmp_path / "t")
    with open(name + "_T.bin", "wb") as f:
        f.write(T.ravel(order="F").tobytes())
    v = mcio.import_tissue(name, rp)
    for iy, ix, iz in [(0, 0, 0), (2, 3, 4), (1, 2, 3), (2, 0, 1)]:
        assert v[W.voxel_index(ix, iy, iz, nx, ny)] == T[iy, ix, iz]


def test_reader_errors(tmp_path):
    name = str(tmp_path / "bad")
    with pytest.raises(NmcError):
        mcio.read_run_params(name)                       # no file
    rp, tp, v = small_case()
    mcio.write_mci(name, rp, tp)
    text = open(name + "_H.mci").read().split("\n")
    open(name + "_H.mci", "w").write("\n".join(text[:20]))
    with pytest.raises(NmcError, match="ends before"):
        mcio.read_run_params(name)                       # truncated
    text[1] = "many"
    open(name + "_H.mci", "w").write("\n".join(text))
    with pytest.raises(NmcError, match="Nx"):
        mcio.read_run_params(name)                       # not a number
    rp.mckernelflag = 7
    mcio.write_mci(name, rp, tp)
    with pytest.raises(NmcError, match="kernel"):
        mcio.read_run_params(name)
    rp.mckernelflag = 1
    mcio.write_mci(name, rp, tp)
    text = open(name + "_H.mci").read().split("\n")
    text[len(mcio.RUN_FIELDS)] = "19"
    open(name + "_H.mci", "w").write("\n".join(text) + "0\n" * 60)
    with pytest.raises(NmcError, match="Nt"):
        mcio.read_run_params(name)                       # would overflow the 19-entry arrays
    rp.Nt = 19
    with pytest.raises(NmcError, match="Nt"):
        mcio.write_mci(name, rp, tp)
    rp.Nt = 4
    mcio.write_tissue(name, v[:-1])
    with pytest.raises(NmcError, match="bytes"):
        mcio.import_tissue(name, rp)


def test_long_line_spills_into_the_next_field(tmp_path):
    """fgets(buf, 32): a 40-character first line is read as two values."""
    rp, tp, _ = small_case()
    name = str(tmp_path / "long")
    mcio.write_mci(name, rp, tp)
    lines = open(name + "_H.mci").read().split("\n")
    lines[0] = "1" + "0" * 29 + "7" + "0" * 8 + "5"   # 31 chars, then "000000005"
    open(name + "_H.mci", "w").write("\n".join(lines))
    rp2, _ = mcio.read_run_params(name)
    assert rp2.time_min == np.float32(1e30 + 7)        # first 31 characters
    assert rp2.Nx == 5                                 # the rest of the line
    assert rp2.Ny == rp.Nx                             # everything after shifts by one


def test_output_formats(tmp_path):
    p = str(tmp_path / "g.dat")
    mcio.save_grid_text(p, np.array([1.0, 0.25, np.nan, 3e-9]), 2, 2, integer=False)
    assert open(p).read() == "\t 1.000000e+00\t 2.500000e-01\n\t 0.000000e+00\t 3.000000e-09\n"
    mcio.save_grid_text(p, np.array([1, 0, 7, 12], np.uint32), 2, 2, integer=True)
    assert open(p).read() == "\t 1\t 0\n\t 7\t 12\n"
    mcio.save_rd(p, 0.071736)
    assert open(p).read() == "0.0717\n"


def test_fluence_rate_normalisation(tmp_path):
    rp, tp, v = small_case()
    v = v.copy()
    v[0] = 0                    # tissue 0 has mua 0: Inf/NaN -> 0 (main.m:484-489)
    dep = np.random.default_rng(1).random(v.size) * 0.1
    dep[1] = 0.0
    n = 12345
    F = mcio.fluence_rate(dep, rp, tp, v, n)
    assert F.dtype == np.float32 and F[0] == 0.0
    mua = np.array([tp.muav[t] for t in v], np.float64)
    ok = mua > 0
    want = dep[ok] / (rp.dx * rp.dy * rp.dz * n * mua[ok])
    np.testing.assert_allclose(F[ok], want, rtol=4e-7)


def test_mcsub_tables_closed_form():
    """J summed back over its rings, 2 pi r dr N J, returns the escaping weight
    per radial bin; F likewise with dz mua."""
    nr, nz, na, dr, dz, mua, n = 5, 4, 3, 0.002, 0.003, 1.5, 1000
    rng = np.random.default_rng(2)
    rd_ra = rng.random((nr + 1) * na)
    a_rz = rng.random((nr + 1) * (nz + 1))
    J, F = mcio.mcsub_tables(rd_ra, a_rz, nr, nz, na, dr, dz, mua, n)
    assert J.shape == (nr + 2,) and F.shape == (nr + 2, nz + 2) and J[0] == 0
    r = (np.arange(1, nr + 2) - 0.5) * dr
    np.testing.assert_allclose(J[1:] * 2 * math.pi * r * dr * n,
                               rd_ra.reshape(nr + 1, na).sum(1), rtol=1e-6)
    np.testing.assert_allclose(F[1:, 1:] * (2 * math.pi * r * dr * dz * n * mua)[:, None],
                               a_rz.reshape(nr + 1, nz + 1), rtol=1e-6)


def test_cli_rejects_bad_inputs_before_any_device_work(tmp_path, capsys):
    """The nmc driver fails with a message and status 1 (the reference exits
    or reads garbage) for a missing file, an unsupported kernel and a bad
    photon count; none of these reaches the GPU."""
    from nmc_amd import cli

    assert cli.main([str(tmp_path / "missing")]) == 1
    assert "cannot open" in capsys.readouterr().err
    rp, tp, v = small_case()
    rp.mckernelflag = 3                      # polarisation: readable, not built here
    name = str(tmp_path / "pol")
    mcio.write_mci(name, rp, tp)
    mcio.write_tissue(name, v)
    assert cli.main([name]) == 1
    assert "not built" in capsys.readouterr().err
    rp.mckernelflag = 1
    mcio.write_mci(name, rp, tp)
    assert cli.main([name, "--photons", "0"]) == 1
    assert "positive" in capsys.readouterr().err
