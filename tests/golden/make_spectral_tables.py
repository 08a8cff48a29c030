"""Writes tests/golden/spectral_tables.json from the reference's own data tables.

The 36-bin Smits (2000) basis spectra and the 36-bin spectrum->RGB matrix are
src/spectral.cxx:13-45, the bin centres src/spectral.hxx:287-289 and the
normalised solar spectrum src/light.hxx:260.  They are data that both the GPU
library and the oracle must use unchanged, so they are lifted once as a fixture.
Run from the repository root: python tests/golden/make_spectral_tables.py
"""
import json
import os
import re

REF = "/root/reference/src"


def _rows(block):
    return [[float(x) for x in r.split(",") if x.strip()] for r in re.findall(r"\{([^{}]*)\}", block)]


def main():
    spectral = open(os.path.join(REF, "spectral.cxx")).read()
    smits = _rows(re.search(r"spectra\[7\]\[NBINS\]\s*=\s*\{(.*?)\};", spectral, re.S).group(1))
    to_rgb = _rows(re.search(r"mat_spectrum_to_rgb\[3\]\[NBINS\]\s*=\s*\{(.*?)\};", spectral, re.S).group(1))
    header = open(os.path.join(REF, "spectral.hxx")).read()
    centers = [float(x) for x in re.search(r"wavelengths\[NBINS\]\s*=\s*\{(.*?)\};", header, re.S).group(1).split(",")]
    light = open(os.path.join(REF, "light.hxx")).read()
    sun = [float(x) for x in re.search(r"emission_spectrum <<(.*?);", light, re.S).group(1).split(",")]
    assert len(smits) == 7 and all(len(r) == 36 for r in smits)
    assert len(to_rgb) == 3 and all(len(r) == 36 for r in to_rgb)
    assert len(centers) == 36 and len(sun) == 36
    out = {"bin_centers": centers, "smits": smits, "spec_to_rgb": to_rgb, "sun": sun,
           "smits_order": ["white", "red", "green", "blue", "yellow", "cyan", "magenta"]}
    path = os.path.join(os.path.dirname(__file__), "spectral_tables.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
