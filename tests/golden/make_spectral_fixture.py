"""Writes the spectral-model fixtures from the reference's data files.

  spectral_lib.json      the absorption spectra of NMC-samples/spectralLIB.mat
                         (MAT v5, read with scipy.io.loadmat, which executes
                         nothing): nmLIB and the oxy-/deoxy-haemoglobin, water
                         and melanin columns, 300-1000 nm
  notebook_tissues.json  the optical-property tables the reference notebooks
                         printed for their runs (text of the stored outputs):
                         example2.ipynb at 532 nm (the whole PrintRunParameters
                         echo of its skin-vessel run, tissues 1..9) and
                         example3.ipynb at 350..750 nm in steps of 20 (its own
                         table, tissues 0..8, with the Rd it reported)

Run in a container that has the reference checked out:
    python tests/golden/make_spectral_fixture.py /root/reference
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def spectra(ref: str) -> dict:
    from scipy.io import loadmat
    d = loadmat(os.path.join(ref, "NMC-samples", "spectralLIB.mat"))
    keys = ("nmLIB", "muaoxy", "muadeoxy", "muawater", "muamel")
    return {k: [float(x) for x in d[k].ravel()] for k in keys}


def _outputs(path: str) -> list[str]:
    with open(path) as f:
        nb = json.load(f)
    out = []
    for c in nb["cells"]:
        for o in c.get("outputs", []):
            if "text" in o:
                out.append("".join(o["text"]))
    return out


def example2_table(ref: str) -> dict:
    text = "\n".join(_outputs(os.path.join(ref, "NMC-samples", "example2.ipynb")))
    tissues = {}
    for m in re.finditer(r"muav\[(\d+)\] = ([0-9.]+) \[mm\^-1\]\nmusv\[\1\] = ([0-9.]+) "
                         r"\[mm\^-1\]\n  gv\[\1\] = ([0-9.]+)", text):
        tissues[int(m.group(1))] = [float(m.group(2)), float(m.group(3)), float(m.group(4))]
    i = text.index("time_min = ")
    j = text.index("central axial profile of tissue types:\n", i)
    j = text.index("\n", j + len("central axial profile of tissue types:\n")) + 1
    return {"nm": 532, "tissues": {str(k): v for k, v in sorted(tissues.items())},
            "echo": text[i:j]}


def example3_tables(ref: str) -> list[dict]:
    text = "\n".join(_outputs(os.path.join(ref, "NMC-samples", "example3.ipynb")))
    tables = []
    for block in re.split(r"=== Optical properties at ", text)[1:]:
        nm = int(re.match(r"(\d+) nm ===", block).group(1))
        rows = {}
        for m in re.finditer(r"^\s*(\d+) \|\s*([0-9.]+) \|\s*([0-9.]+) \| ([0-9.]+) \|", block, re.M):
            rows[str(int(m.group(1)) + 1)] = [float(m.group(2)), float(m.group(3)),
                                              float(m.group(4))]
        rd = re.search(r"Rd \(MC\) @ %d nm: ([0-9.]+)" % nm, block)
        tables.append({"nm": nm, "tissues": rows, "rd_reference_mc": float(rd.group(1)) if rd else None})
    return tables


def main(ref: str) -> None:
    with open(os.path.join(HERE, "spectral_lib.json"), "w") as f:
        json.dump(spectra(ref), f, separators=(",", ":"))
    with open(os.path.join(HERE, "notebook_tissues.json"), "w") as f:
        json.dump({"source": "NMC-samples/example2.ipynb, example3.ipynb stored outputs",
                   "units": "as printed by the notebooks (cm^-1 values, labelled mm^-1)",
                   "example2": example2_table(ref), "example3": example3_tables(ref)}, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
