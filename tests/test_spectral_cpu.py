"""Spectral tissue model (nmc_amd/spectral.py) against what the reference
notebooks printed: the tissue tables of example2.ipynb (532 nm) and
example3.ipynb (350-750 nm, 21 wavelengths), and example2's whole
PrintRunParameters echo, which also fixes the skin-vessel geometry through
its central axial profile of tissue types."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from nmc_amd import mcio, spectral
from nmc_amd import params as P

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def lib():
    return spectral.SpectralLibrary.from_json(os.path.join(GOLD, "spectral_lib.json"))


@pytest.fixture(scope="module")
def printed():
    with open(os.path.join(GOLD, "notebook_tissues.json")) as f:
        return json.load(f)


def _check_table(model: dict, table: dict, nm):
    for k, (mua, mus, g) in table.items():
        t = model[int(k)]
        # printed with 4 decimals (g with 3 in example3)
        assert abs(t["mua"] - mua) <= 0.5e-4 + 1e-9, (nm, k, t["mua"], mua)
        assert abs(t["mus"] - mus) <= 0.5e-4 + 1e-9 * mus, (nm, k, t["mus"], mus)
        assert abs(t["g"] - g) <= 0.5e-3, (nm, k)


def test_example2_table(lib, printed):
    ex = printed["example2"]
    _check_table(spectral.tissue_list(ex["nm"], lib), ex["tissues"], ex["nm"])


def test_example3_tables(lib, printed):
    tables = printed["example3"]
    assert [t["nm"] for t in tables] == list(range(350, 751, 20))
    for t in tables:
        _check_table(spectral.tissue_list(t["nm"], lib), t["tissues"], t["nm"])


def test_example2_echo_and_geometry(lib, printed, tmp_path):
    """The notebook's pipeline with our parts: spectral model and run settings
    -> the notebook's 4-decimal .mci writer -> our reader and label volume ->
    our echo, equal to the text the reference printed for its 532 nm run."""
    from test_mcio_cpu import notebook_mci

    rp = spectral.skin_vessel_run_params()
    tp = spectral.tissue_params(532, lib)
    t = spectral.tissue_list(532, lib)
    name = str(tmp_path / "skinvessel")
    notebook_mci(name, rp, tp, props=[(t[k]["mua"], t[k]["mus"], t[k]["g"]) for k in range(1, 10)])
    mcio.write_tissue(name, spectral.skin_vessel_volume())
    rp2, tp2 = mcio.read_run_params(name)
    text = mcio.run_parameters_text(rp2, tp2, mcio.import_tissue(name, rp2))
    assert text.rstrip("\n") == printed["example2"]["echo"].rstrip("\n")


def test_geometry_layout(lib):
    v = spectral.skin_vessel_volume(n=40, binsize=0.0025)
    T = v.reshape(40, 40, 40, order="F")           # T[iy, ix, iz]
    assert (T[:, :, :4] == 1).all()                # air above zsurf = 0.01
    assert (T[:, :, 4:6] == 5).all()               # epidermis to 0.016
    assert (T[:, 19, 19] == 3).all()               # vessel along y at x = 0, z = 0.05
    assert set(np.unique(v)) == {1, 3, 4, 5}


def test_sweep_items_and_mat_loader(lib):
    items = spectral.sweep_items([400, 500, 600], 4, lib, [10, 20, 30])
    assert [i["n_photons"] for i in items] == [10, 20, 30]
    assert [i["stream_id"] for i in items] == [0, 1, 2]
    d500 = spectral.tissue_list(500, lib)[4]
    assert items[1]["mua"] == d500["mua"] and items[1]["mus"] == pytest.approx(43.6 / 0.1)
    mat = "/root/reference/NMC-samples/spectralLIB.mat"
    if os.path.exists(mat):
        m = spectral.SpectralLibrary.from_mat(mat)
        for nm in (350.5, 532, 999):
            np.testing.assert_allclose(m.absorbers(nm), lib.absorbers(nm), rtol=1e-15)


def test_tissue_params_struct(lib):
    tp = spectral.tissue_params(600, lib)
    assert isinstance(tp, P.TissueParams)
    assert tp.muav[0] == 0.0 and tp.gv[3] == pytest.approx(0.9)
    assert tp.musv[9] == 100.0
