"""Writes yb_cross_sections.json: the Yb3+ absorption and emission
cross-section tables the reference's YDFA model interpolates
(src/gain_simulation/Yb_absorption.npz and Yb_emission.npz, read by
get_ydfa_cross_sections, gain.jl:86-112: wavelength in nm, intensity in
units of 1e-27 m^2).  Data only; read with numpy's loader, no pickles.
Run from the repository root:  python tests/golden/make_yb_cross_sections.py"""
from __future__ import annotations

import json
import pathlib

import numpy as np

SRC = pathlib.Path("/root/reference/src/gain_simulation")
OUT = pathlib.Path(__file__).resolve().parent / "yb_cross_sections.json"


def main():
    tables = {}
    for name in ("absorption", "emission"):
        with np.load(SRC / f"Yb_{name}.npz", allow_pickle=False) as d:
            tables[name] = {"wavelength_nm": [float(x) for x in d["wavelength"]],
                            "intensity": [float(x) for x in d["intensity"]]}
    tables["scale_m2"] = 1e-27
    tables["source"] = "src/gain_simulation/Yb_absorption.npz, Yb_emission.npz (gain.jl:86-112)"
    OUT.write_text(json.dumps(tables, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
