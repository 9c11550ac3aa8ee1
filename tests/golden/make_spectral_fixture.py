"""Writes spectral_npol4.json: the reference's npol = 4 GLL/GLJ node,
weight and differentiation tables and its FFT lucky-number table, read as numbers from
/root/reference/SOLVER/src/preloop/se_model/quad/spectrals.hpp (lines 73-102)
so the tests can check the closed-form tables of axisem3d_amd.spectral
against them without the reference being present.

Run from the repository root: python tests/golden/make_spectral_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/SOLVER/src/preloop/se_model/quad/spectrals.hpp"
SRC_NR = "/root/reference/SOLVER/src/preloop/nr_field/NrField.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "spectral_npol4.json")


def main():
    text = open(SRC).read()
    block = text[text.index("#elif _NPOL == 4"):text.index("#elif _NPOL == 5")]
    data = {"source": "SOLVER/src/preloop/se_model/quad/spectrals.hpp:73-102"}
    for name in ("pGLL", "pGLJ", "wGLL", "wGLJ", "gGLL", "gGLJ"):
        s = block[block.index(name):]
        s = s[s.index("{") + 1:s.index("}")]
        data[name] = [float(v) for v in re.findall(r"[+-]?\d+\.\d+", s)]
    nr = open(SRC_NR).read()
    nr = nr[nr.index("sLuckyNumbers = {"):]
    nr = nr[:nr.index("};")]
    data["lucky_source"] = "SOLVER/src/preloop/nr_field/NrField.cpp:107-133"
    data["lucky"] = [int(v) for v in re.findall(r"\d+", nr)]
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
