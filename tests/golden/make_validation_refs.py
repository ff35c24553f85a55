"""Copies the reference's stored shock-capturing profiles, test/refs/
shu_osher.csv (Shu-Osher at t = 1.8) and test/refs/woodward_colella.csv
(Woodward-Colella at t = 0.038), into tests/golden/ unchanged: 800 rows of
x, rho, u, p each, written by the reference at 4x resolution and interpolated
to the test grid (test/validation.jl:81-93).  They are data files of the
reference's own test suite.  Run in the build container, where the reference
is present."""
import os
import shutil

SRC = "/root/reference/test/refs"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    for name in ("shu_osher.csv", "woodward_colella.csv"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(HERE, name))


if __name__ == "__main__":
    main()
