"""Write the golden fixtures under tests/golden/.

* boys.json: F_m(T) = Gamma(m+1/2) P(m+1/2, T) / (2 T^(m+1/2)) from scipy's regularised
  incomplete gamma function (F_m(0) = 1/(2m+1)), m = 0..16, T on a grid covering the
  Taylor and asymptotic regimes.
* energies_uhf.json: the UHF goldens test0007-0012 (Li doublet, O2 triplet).
* energies.json: the reference's golden RHF energies, read as text from
  oldTests.tgz:tests/chronus-ref.val (test0001-0003; tolerance 5e-10 from tests/rununit.py:39),
  with the geometry of tests/test0001.inp.
Each record also carries the golden multipoles of the same line: dipole (3 values, Debye),
quadrupole (xx, xy, xz, yy, yz, zz; Debye Angstrom) and octupole (the 10 i <= j <= k
components; Debye Angstrom^2), compared at 5e-5, 5e-5 and 5e-3 (tests/rununit.py:39,311-333).

Run once in the container that has the reference mounted:  python tests/golden/make_golden.py
"""
import json
import os
import tarfile

import numpy as np
from scipy.special import gamma, gammainc

HERE = os.path.dirname(os.path.abspath(__file__))


def boys_scipy(m, T):
    if T == 0.0:
        return 1.0 / (2 * m + 1)
    a = m + 0.5
    return float(gamma(a) * gammainc(a, T) / (2.0 * T ** a))


def multipoles(v):
    """name/E/dipole(3)/quadrupole(6)/octupole(10)/SCF (oldTests.tgz:tests/refval.py:20-57)."""
    return {"dipole": [float(x) for x in v[2:5]], "quadrupole": [float(x) for x in v[5:11]],
            "octupole": [float(x) for x in v[11:21]]}


def main():
    Ts = sorted(set([0.0, 1e-12, 1e-8, 1e-4, 0.01, 0.049, 0.05, 0.051, 0.1, 0.2499, 0.5, 1.0, 1.7, 2.5, 3.3,
                     5.0, 7.77, 10.0, 12.5, 15.0, 19.9, 25.0, 29.99, 30.0, 33.0, 35.95, 35.99, 36.0, 36.01,
                     40.0, 45.5, 50.0, 75.0, 100.0, 250.0, 1000.0, 1e4]))
    rows = []
    for T in Ts:
        rows.append({"T": T, "F": [boys_scipy(m, T) for m in range(17)]})
    with open(os.path.join(HERE, "boys.json"), "w") as f:
        json.dump({"source": "scipy.special.gammainc", "mmax": 16, "rows": rows}, f, indent=1)

    ref_tgz = "/root/reference/oldTests.tgz"
    energies = {}
    with tarfile.open(ref_tgz) as t:
        txt = t.extractfile("tests/chronus-ref.val").read().decode()
    basis = {"test0001": "sto-3g", "test0002": "6-31g", "test0003": "cc-pvdz"}
    for line in txt.splitlines():
        v = line.split("/")
        if v[0] in basis and v[-1].strip() == "SCF":
            energies[v[0]] = {"basis": basis[v[0]], "E": float(v[1]), **multipoles(v)}
    uhf = {"test0007": ("sto-3g", "Li"), "test0008": ("6-31g", "Li"), "test0009": ("cc-pvdz", "Li"),
           "test0010": ("sto-3g", "O2"), "test0011": ("6-31g", "O2"), "test0012": ("cc-pvdz", "O2")}
    uhf_rec = {}
    for line in txt.splitlines():
        v = line.split("/")
        if v[0] in uhf and v[-1].strip() == "SCF":
            uhf_rec[v[0]] = {"basis": uhf[v[0]][0], "system": uhf[v[0]][1], "E": float(v[1]), **multipoles(v)}
    out_uhf = {"source": "oldTests.tgz:tests/chronus-ref.val, decks oldTests.tgz:tests/test0007-0012_serial_direct.inp",
               "tolerance": 5e-10,
               "multipole_tolerance": {"dipole": 5e-5, "quadrupole": 5e-5, "octupole": 5e-3},
               "systems": {"Li": {"symbols": ["Li"], "angstrom": [[0.0, 0.0, 0.0]], "mult": 2},
                           "O2": {"symbols": ["O", "O"], "angstrom": [[0.0, 0.0, 0.608586], [0.0, 0.0, -0.608586]],
                                  "mult": 3}},
               "tests": uhf_rec}
    with open(os.path.join(HERE, "energies_uhf.json"), "w") as f:
        json.dump(out_uhf, f, indent=1)
    out = {"source": "oldTests.tgz:tests/chronus-ref.val", "tolerance": 5e-10,
           "multipole_tolerance": {"dipole": 5e-5, "quadrupole": 5e-5, "octupole": 5e-3},
           "geometry_angstrom": {"O": [0.0, -0.07579184359, 0.0], "H1": [0.866811829, 0.6014357793, 0.0],
                                 "H2": [-0.866811829, 0.6014357793, 0.0]},
           "tests": energies}
    with open(os.path.join(HERE, "energies.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
