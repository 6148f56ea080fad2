"""Input decks and .xyz (CPU): the reference's deck rules (parseInput.py / knownKeywords.py,
parseMolecule.py:120-175, parseQM.py:84-140) and round trips."""
import os

import numpy as np
import pytest

from chronusq_public_amd import molecule as M
from chronusq_public_amd.inputdeck import DeckError, parse_deck, parse_deck_text, read_xyz, write_xyz

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")

BASE = """[Molecule]
charge = 0
mult = 1
geom:
 O 0 -0.07579184359 0
 H 0.866811829 0.6014357793 0
 H -0.866811829 0.6014357793 0
[QM]
reference = Real RHF
job = SCF
[BASIS]
basis = STO-3G
"""


@pytest.mark.parametrize("name,basis", [("water_sto3g", "sto-3g"), ("water_631g", "6-31g"),
                                        ("water_ccpvdz", "cc-pvdz")])
def test_reference_decks_parse(name, basis):
    d = parse_deck(os.path.join(DECKS, name + ".inp"))
    assert d.basis == basis and d.reference == "RHF" and d.job == "SCF" and d.guess == "SAD"
    ref = M.water_test0001()
    assert list(d.molecule.Z) == list(ref.Z)
    assert np.abs(d.molecule.coords_bohr - ref.coords_bohr).max() < 1e-12
    assert d.molecule.nelectron == 10
    assert (d.ene_tol, d.den_tol, d.max_iter) == (1e-10, 1e-8, 256)


def test_keywords_are_case_insensitive_and_typed():
    d = parse_deck_text(BASE.replace("job = SCF", "JOB = scf") + "[SCF]\nmaxiter = 50\nenetol = 1e-9\nguess = core\n")
    assert d.max_iter == 50 and d.ene_tol == 1e-9 and d.guess == "CORE"


def test_atomic_numbers_and_isotopes_in_geometry():
    d = parse_deck_text(BASE.replace(" O 0 -0.07579184359 0", " 8 16 0 -0.07579184359 0"))
    assert list(d.molecule.Z) == [8, 1, 1]


@pytest.mark.parametrize("edit,msg", [
    (lambda s: s + "[SCF]\nbogus = 1\n", "unknown keyword"),
    (lambda s: s.replace("mult = 1\n", ""), "MULT"),
    (lambda s: s.replace("[BASIS]\nbasis = STO-3G\n", ""), "BASIS"),
    (lambda s: s + "[RT]\n", "real-time"),
    (lambda s: s.replace("Real RHF", "B3LYP"), "not supported"),
    (lambda s: s.replace("Real RHF", "Complex RHF"), "complex"),
    (lambda s: s.replace("job = SCF", "job = RESP"), "SCF only"),
    (lambda s: s.replace(" H 0.866811829 0.6014357793 0", " Xx 0.8 0.6 0"), "Atomic Symbol"),
    (lambda s: s.replace(" H 0.866811829 0.6014357793 0", " H 0.8 0.6"), "Geometry"),
])
def test_invalid_decks_raise(edit, msg):
    with pytest.raises(DeckError, match=msg):
        parse_deck_text(edit(BASE))


def test_single_atom_defaults_to_core_guess():
    d = parse_deck_text(BASE.replace(""" O 0 -0.07579184359 0
 H 0.866811829 0.6014357793 0
 H -0.866811829 0.6014357793 0""", " Ne 0 0 0"))
    assert d.guess == "CORE" and d.molecule.natom == 1


def test_xyz_round_trip(tmp_path):
    mol = M.water_cluster(3, seed=4, box_ang=M.scaled_box(3))
    p = tmp_path / "w3.xyz"
    write_xyz(mol, str(p), comment="(H2O)3 seed 4")
    back = read_xyz(str(p))
    assert list(back.Z) == list(mol.Z)
    assert np.abs(back.coords_bohr - mol.coords_bohr).max() < 1e-9


def test_restart_file_round_trip_in_the_reference_layout(tmp_path):
    """/SCF/DensityScalar and /SCF/DensityMz, N x N doubles holding the column-major buffer
    (the dataset's row-major [i][j] is M(j, i)); shape and missing-data errors."""
    import numpy as np
    from chronusq_public_amd import restart as R
    rng = np.random.default_rng(0)
    A, Bm = rng.normal(size=(6, 6)), rng.normal(size=(6, 6))
    path = str(tmp_path / "x.bin")
    R.write_densities(path, A, Bm)
    P, Mz = R.read_densities(path, 6, open_shell=True)
    assert np.array_equal(P, A) and np.array_equal(Mz, Bm)
    R.write_densities(path, 2 * A)                 # overwrite in place
    assert np.array_equal(R.read_densities(path, 6)[0], 2 * A)
    # raw dataset order: read the file's buffer without the transpose
    L = R._lib()
    f = L.H5Fopen(path.encode(), R.H5F_ACC_RDONLY, 0)
    d = L.H5Dopen2(f, b"/SCF/DensityScalar", 0)
    raw = np.empty((6, 6))
    assert L.H5Dread(d, L.native_double, 0, 0, 0, raw.ctypes.data) >= 0
    L.H5Dclose(d)
    L.H5Fclose(f)
    assert np.array_equal(raw, 2 * A.T)
    with pytest.raises(R.RestartError):
        R.read_densities(path, 7)
    with pytest.raises(R.RestartError):
        R.read_densities(str(tmp_path / "missing.bin"), 6)
    R.write_densities(str(tmp_path / "closed.bin"), A)
    with pytest.raises(R.RestartError):
        R.read_densities(str(tmp_path / "closed.bin"), 6, open_shell=True)
