"""Writes tests/golden/thermo_subset.inp: the He, CO2, N2, O2 and Ar records
of the NASA CEA thermodynamic database (data/thermo.inp of the reference,
NASA TP-2002-211556; Apache License 2.0, Copyright (c) 2025 United States
Government as represented by the Administrator of NASA), in their original
layout.  Run in the build container, where the reference is present."""
import os

SRC = "/root/reference/data/thermo.inp"
NAMES = ["He", "CO2", "N2", "O2", "Ar"]
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "thermo_subset.inp")


def main():
    lines = open(SRC).read().splitlines()
    recs = {}
    for i, ln in enumerate(lines):
        if not ln or ln.startswith("!") or i + 1 >= len(lines):
            continue
        name = ln[0:24].split()[0] if ln[0:24].strip() else ""
        if name in NAMES and name not in recs:
            try:
                n = int(lines[i + 1][0:2])
            except ValueError:
                continue
            recs[name] = lines[i:i + 2 + 3 * n]
    with open(OUT, "w") as fh:
        fh.write("! Subset of the NASA CEA thermodynamic database (thermo.inp, NASA TP-2002-211556).\n")
        fh.write("! Apache License 2.0; Copyright (c) 2025 United States Government as represented by the\n")
        fh.write("! Administrator of the National Aeronautics and Space Administration.\n")
        fh.write("thermo\n")
        for nm in NAMES:
            fh.write("\n".join(recs[nm]) + "\n")
        fh.write("END PRODUCTS\n")


if __name__ == "__main__":
    main()
