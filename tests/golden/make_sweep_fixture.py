"""Writes tests/golden/spectral_sweep_items.json: the 128 work items (mua, mus,
budget per wavelength) of the seeded spectral-sweep workload, so that a change
to the generator in nmc_amd/workloads.py is caught.
Run: python tests/golden/make_sweep_fixture.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from nmc_amd.workloads import spectral_sweep  # noqa: E402

w, items, lam = spectral_sweep()
with open(os.path.join(HERE, "spectral_sweep_items.json"), "w") as f:
    json.dump({"seed": 3, "wavelength_nm": [float(x) for x in lam], "items": items}, f, indent=0)
print("total photons", sum(i["n_photons"] for i in items))
