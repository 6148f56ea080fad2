"""CPU tests of the host logic: basis construction, seeded geometries, the C ABI surface."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from chronusq_public_amd import basis as B
from chronusq_public_amd import molecule as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_water_cluster_is_seeded_and_respects_min_distance():
    a = M.water_cluster(16, 0)
    b = M.water_cluster(16, 0)
    c = M.water_cluster(16, 1)
    assert np.array_equal(a.coords_bohr, b.coords_bohr)
    assert not np.array_equal(a.coords_bohr, c.coords_bohr)
    O = a.coords_bohr[0::3] * B.BOHR_PER_ANGSTROM_INV
    d = np.linalg.norm(O[:, None] - O[None], axis=-1) + np.eye(16) * 100
    assert d.min() >= 2.7
    assert (O >= 0).all() and (O <= 10).all()
    # rigid monomers: O-H 0.9584 A, H-O-H 104.48 deg from (0, +-0.757, 0.587)
    for w in range(16):
        x = a.coords_bohr[3 * w:3 * w + 3] * B.BOHR_PER_ANGSTROM_INV
        r1, r2 = np.linalg.norm(x[1] - x[0]), np.linalg.norm(x[2] - x[0])
        assert abs(r1 - math.hypot(0.757, 0.587)) < 1e-12 and abs(r2 - r1) < 1e-12


def test_headline_basis_dimensions():
    mol = M.water_cluster(16, 0)
    bs = B.build_basis("cc-pvdz", mol.Z, mol.coords_bohr)
    assert bs.nshell == 192 and bs.nbasis == 384
    ls = [s.l for s in bs.shells]
    assert ls.count(0) == 112 and ls.count(1) == 64 and ls.count(2) == 16
    assert all(s.pure for s in bs.shells if s.l == 2)
    # config-1 count under the reference's rules (SP split, pure d): 18, not 25
    w = M.water_monomer()
    assert B.build_basis("6-31g*", w.Z, w.coords_bohr).nbasis == 18
    assert B.build_basis("6-31g*", w.Z, w.coords_bohr, force_cart=True).nbasis == 19


def test_sp_split_shares_exponents():
    ref = B.parse_gbs(B.basis_path("6-31g"))
    O = ref[8]
    assert [s.l for s in O] == [0, 0, 1, 0, 1]
    assert O[1].alpha == O[2].alpha and O[3].alpha == O[4].alpha


def test_contraction_normalisation_unit_self_overlap():
    for name in ("sto-3g", "6-31g*", "cc-pvdz"):
        for z, shells in B.parse_gbs(B.basis_path(name)).items():
            if z not in (1, 6, 8):
                continue
            for rs in shells:
                c = B.normalize_coefficients(rs.l, rs.alpha, rs.coef)
                a = np.asarray(rs.alpha)
                l = rs.l
                S = 0.0
                for p in range(len(a)):
                    for q in range(len(a)):
                        S += c[p] * c[q] * B._dfact(2 * l - 1) * math.pi ** 1.5 / (2 ** l * (a[p] + a[q]) ** (l + 1.5))
                assert abs(S - 1.0) < 1e-13


def test_nuclear_repulsion_water():
    mol = M.water_test0001()
    r = mol.coords_bohr
    e = 8 * 1 / np.linalg.norm(r[0] - r[1]) + 8 / np.linalg.norm(r[0] - r[2]) + 1 / np.linalg.norm(r[1] - r[2])
    assert abs(mol.nuclear_repulsion() - e) < 1e-12


def test_alkane_composition():
    for n in (1, 2, 5, 12):
        mol = M.alkane(n)
        assert (mol.Z == 6).sum() == n and (mol.Z == 1).sum() == 2 * n + 2


def _header_functions():
    names = set()
    for h in ("cqgpu.h", "cqgpu_guess.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(cq_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from chronusq_public_amd import _lib
    lib_path = _lib.LIB_PATH
    assert os.path.exists(lib_path), "libcqgpu.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (cq_\w+)", out))
    declared = _header_functions()
    assert declared and set(declared) <= exported, set(declared) - exported
    assert set(_lib.EXPORTED) == set(declared)
    L = _lib.load()   # loads without a GPU; no compute call is made
    for name in declared:
        assert hasattr(L, name)
    assert L.cq_num_class_pairs() == 21


def test_cpp_layer_compiles_with_reference_signature(tmp_path):
    """The header-only layer accepts ChronusQ's argument types (vector<reference_wrapper<Op>>,
    vector<ERI_CONTRACTION_TYPE>, vector<double>) and links against libcqgpu.so."""
    src = tmp_path / "shim.cpp"
    src.write_text(r'''
#include "cqgpu_aointegrals.hpp"
#include <vector>
struct Mat {   // column-major stand-in for Eigen::Map<RealMatrix>
  typedef double Scalar;
  std::vector<double> v; int n;
  explicit Mat(int n_) : v((size_t)n_ * n_, 0.0), n(n_) {}
  double* data() { return v.data(); }
  const double* data() const { return v.data(); }
  int rows() const { return n; } int cols() const { return n; } int outerStride() const { return n; }
};
enum ERI_CONTRACTION_TYPE { COULOMB, EXCHANGE, PAIR };
int run(cq_ctx_t* ctx) {
  Mat P(7), J(7), K(7);
  std::vector<std::reference_wrapper<Mat>> X{P, P}, AX{J, K};
  std::vector<ERI_CONTRACTION_TYPE> types{COULOMB, EXCHANGE};
  std::vector<double> scale{1.0, -0.5};
  cqgpu::newTwoEContractDirect(ctx, X, AX, types, scale);
  return 0;
}
int main() { return 0; }
''')
    exe = tmp_path / "shim"
    r = subprocess.run(["g++", "-std=c++14", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        os.path.join(ROOT, "chronusq_public_amd", "libcqgpu.so")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
