"""GPU multipole integrals (cq_multipole: the libint2 emultipole1..3 blocks of
AOIntegrals::OneEDriver, src/aointegrals/aointegrals_builders.cpp:38-55) against the oracle's
Hermite-moment restatement, about the origin and about a shifted centre, in one- and
several-molecule contexts; and the device SCF's dipole, quadrupole and octupole
(SingleSlater::computeMultipole, include/singleslater/singleslater_properties.h:115-172)
against the reference's golden values of test0001-0003 (RHF water) and test0007-0012 (UHF Li,
O2) at the tolerances of tests/rununit.py:39."""
import json
import os

import numpy as np
import pytest

import oracle
from chronusq_public_amd import basis as B
from chronusq_public_amd import molecule as M
from chronusq_public_amd.aointegrals import AOIntegrals

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name,basis", [("h2o", "6-31g*"), ("h2o4", "cc-pvdz")])
@pytest.mark.parametrize("origin", [None, (0.4, -1.3, 0.9)])
def test_multipole_integrals_match_oracle(name, basis, origin):
    mol = M.water_test0001() if name == "h2o" else M.water_cluster(4, seed=3, box_ang=M.scaled_box(4))
    bs = B.build_basis(basis, mol.Z, mol.coords_bohr)
    print(f"[multipole] {name}/{basis} origin={origin}", flush=True)
    ao = AOIntegrals(bs)
    got = np.stack([m.cpu().numpy() for m in ao.computeMultipole(3, origin)])
    ref = oracle.OracleBasis(bs).multipole(3, (0.0, 0.0, 0.0) if origin is None else origin)
    for k in range(19):
        assert np.abs(got[k] - ref[k]).max() < 1e-12 * max(1.0, np.abs(ref[k]).max()), k
    # orders 1 and 2 are the leading blocks of order 3 (up to the order of the atomic sums
    # over primitive slices)
    d2 = np.stack([m.cpu().numpy() for m in ao.computeMultipole(2, origin)])
    assert np.abs(d2 - got[:9]).max() < 1e-14 * max(1.0, np.abs(got[:9]).max())


def test_multipole_batch_context_blocks():
    """A context of two molecules: each molecule's diagonal block is its own integrals, the
    off-diagonal blocks are zero (pairs never span molecules)."""
    from chronusq_public_amd.aointegrals import AOIntegrals as AO
    m1 = M.water_test0001()
    m2 = M.alkane(2)
    b1 = B.build_basis("6-31g*", m1.Z, m1.coords_bohr)
    b2 = B.build_basis("6-31g*", m2.Z, m2.coords_bohr)
    ao = AO.batch([b1, b2])
    got = np.stack([m.cpu().numpy() for m in ao.computeMultipole(3)])
    n1 = b1.nbasis
    r1 = oracle.OracleBasis(b1).multipole(3)
    r2 = oracle.OracleBasis(b2).multipole(3)
    for k in range(19):
        assert np.abs(got[k][:n1, :n1] - r1[k]).max() < 1e-12 * max(1.0, np.abs(r1[k]).max())
        assert np.abs(got[k][n1:, n1:] - r2[k]).max() < 1e-12 * max(1.0, np.abs(r2[k]).max())
        assert np.abs(got[k][:n1, n1:]).max() == 0.0


def _check(mp, rec, tol, label):
    dip = mp["dipole"]
    quad = np.array([mp["quadrupole"][i, j] for i, j in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]])
    octu = np.array([mp["octupole"][i, j, k] for i in range(3) for j in range(i, 3) for k in range(j, 3)])
    errs = (np.abs(dip - rec["dipole"]).max(), np.abs(quad - rec["quadrupole"]).max(),
            np.abs(octu - rec["octupole"]).max())
    print(f"[multipole] {label}: max|d dipole| {errs[0]:.2e} D, |d quad| {errs[1]:.2e}, |d oct| {errs[2]:.2e}", flush=True)
    assert errs[0] < tol["dipole"] and errs[1] < tol["quadrupole"] and errs[2] < tol["octupole"], label


def test_rhf_golden_multipoles():
    from chronusq_public_amd.singleslater import SingleSlater
    d = json.load(open(os.path.join(GOLD, "energies.json")))
    mol = M.water_test0001()
    for test, rec in d["tests"].items():
        ss = SingleSlater(mol, rec["basis"], tau=1e-14, guess="SAD")
        r = ss.SCF3()
        assert abs(r.energy - rec["E"]) < d["tolerance"]
        _check(ss.computeMultipole(r.P), rec, d["multipole_tolerance"], test)


def test_uhf_golden_multipoles():
    from chronusq_public_amd.singleslater import SingleSlater
    d = json.load(open(os.path.join(GOLD, "energies_uhf.json")))
    for test, rec in d["tests"].items():
        sysd = d["systems"][rec["system"]]
        mol = M.Molecule.from_angstrom(sysd["symbols"], sysd["angstrom"], 0, sysd["mult"])
        ss = SingleSlater(mol, rec["basis"], ref="UHF", tau=1e-14, guess="SAD")
        r = ss.SCF3()
        assert abs(r.energy - rec["E"]) < d["tolerance"]
        _check(ss.computeMultipole(r.P), rec, d["multipole_tolerance"], test)
