"""Gaussian-94 basis-set parsing and shell construction (host side).

Mirrors ChronusQ's ``BasisSet`` for the pieces the G[D] path consumes:

* ``parse_gbs`` follows ``BasisSet::parseGlobal`` (src/basisset/basisset_read.cpp:81-270):
  one shell per G94 record, ``SP`` split into an S and a P shell that share exponents,
  ``pure = (L > 1)`` unless ``force_cart``; reference shells stable-sorted by atomic number.
* ``build_shells`` follows ``BasisSet::constructLocal`` (src/basisset/basisset_gen.cpp:30-74):
  shells in atom order, then file order, centred on the atom (bohr).
* Normalisation follows the libint2 ``Shell`` constructor (primitive norm of the x^l
  component, then unit contracted self-overlap) followed by ``BasisSet::renormShells``
  (src/basisset/basisset_misc.cpp:95-113), which divides every coefficient by the square
  root of the first function's self-overlap.
* ``shell_offsets`` follows ``BasisSet::makeMapSh2Bf`` (src/basisset/basisset_map.cpp:29-40).

Function order inside a shell: Cartesian ``x^l ... z^l`` lexicographic (include/basisset.h:103-109)
with every component carrying the x^l normalisation; spherical ``m = -l..l``
(src/basisset/basisset_read.cpp:960-990).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Sequence

import numpy as np

BOHR_PER_ANGSTROM_INV = 0.5291772083000001  # phys.bohr, include/global.h:224

ELEMENTS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
            "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"]
SYMBOL_TO_Z = {s.upper(): z for z, s in enumerate(ELEMENTS) if z > 0}

_L_SYMBOLS = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6}

BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
_BASIS_FILES = {
    "sto-3g": "sto3g.gbs",
    "6-31g": "6-31g.gbs",
    "6-31g*": "6-31g*.gbs",
    "cc-pvdz": "cc-pvdz.gbs",
}


def basis_path(name: str) -> str:
    key = name.strip().lower()
    if key not in _BASIS_FILES:
        raise ValueError(f"unknown basis set '{name}' (known: {sorted(_BASIS_FILES)})")
    return os.path.join(BASIS_DIR, _BASIS_FILES[key])


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def cart_components(l: int):
    """Cartesian exponent triples in x^l ... z^l lexicographic order."""
    out = []
    for i in range(l + 1):
        lx = l - i
        for j in range(i + 1):
            ly = i - j
            out.append((lx, ly, l - lx - ly))
    return out


def _dfact(n: int) -> float:
    """Double factorial with (-1)!! = 0!! = 1."""
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


@dataclass
class RefShell:
    l: int
    pure: bool
    alpha: List[float]
    coef: List[float]          # as read from the file (unnormalised)


@dataclass
class Shell:
    l: int
    pure: bool
    alpha: np.ndarray
    coef: np.ndarray           # normalised contraction coefficients
    origin: np.ndarray         # bohr
    atom: int

    @property
    def size(self) -> int:
        return 2 * self.l + 1 if self.pure else ncart(self.l)

    @property
    def nprim(self) -> int:
        return len(self.alpha)


def _to_float(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def parse_gbs(path: str, force_cart: bool = False) -> Dict[int, List[RefShell]]:
    """Parse a Gaussian-94 basis file into reference shells keyed by atomic number."""
    recs: List[List[str]] = [[]]
    with open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            if len(line.strip()) == 0 or line.lstrip().startswith("!"):
                continue
            if line.strip() == "****":
                recs.append([])
                continue
            recs[-1].append(line)
    out: Dict[int, List[RefShell]] = {}
    for rec in recs:
        if not rec:
            continue
        sym = rec[0].split()[0].upper()
        if sym not in SYMBOL_TO_Z:
            continue
        z = SYMBOL_TO_Z[sym]
        shells: List[RefShell] = []
        i = 1
        while i < len(rec):
            toks = rec[i].split()
            sym_l = toks[0].upper()
            depth = int(toks[1])
            exps, c1, c2 = [], [], []
            for k in range(depth):
                t2 = rec[i + 1 + k].split()
                exps.append(_to_float(t2[0]))
                c1.append(_to_float(t2[1]))
                if sym_l == "SP":
                    c2.append(_to_float(t2[2]))
            i += 1 + depth
            if sym_l == "SP":
                shells.append(RefShell(0, False, exps, c1))
                shells.append(RefShell(1, False, list(exps), c2))
            else:
                L = _L_SYMBOLS[sym_l]
                shells.append(RefShell(L, (L > 1) and not force_cart, exps, c1))
        out[z] = shells
    return out


def normalize_coefficients(l: int, alpha: Sequence[float], coef: Sequence[float]) -> np.ndarray:
    """Libint-convention normalisation followed by the renormShells division.

    Each primitive is scaled by the norm of its x^l component,
    sqrt(2^l (2a)^(l+3/2) / (pi^(3/2) (2l-1)!!)); the contraction is then scaled to unit
    self-overlap.  renormShells (basisset_misc.cpp:95-113) divides by sqrt(<phi0|phi0>) once
    more; that overlap is 1 up to rounding, and the same division is applied here.
    """
    a = np.asarray(alpha, dtype=np.float64)
    c = np.asarray(coef, dtype=np.float64).copy()
    sqrt_pi3 = math.pi ** 1.5
    df = _dfact(2 * l - 1)
    for p in range(len(a)):
        two_a = 2.0 * a[p]
        c[p] *= math.sqrt(2.0 ** l * two_a ** (l + 1.5) / (sqrt_pi3 * df))

    def self_overlap(cc):
        s = 0.0
        for p in range(len(a)):
            for q in range(p + 1):
                g = a[p] + a[q]
                s += (1.0 if p == q else 2.0) * df * sqrt_pi3 * cc[p] * cc[q] / (2.0 ** l * g ** (l + 1.5))
        return s

    c *= 1.0 / math.sqrt(self_overlap(c))
    c /= math.sqrt(self_overlap(c))
    return c


@dataclass
class BasisSet:
    name: str
    shells: List[Shell] = field(default_factory=list)

    @property
    def nshell(self) -> int:
        return len(self.shells)

    @property
    def nbasis(self) -> int:
        return sum(s.size for s in self.shells)

    @property
    def max_l(self) -> int:
        return max(s.l for s in self.shells)

    @property
    def max_prim(self) -> int:
        return max(s.nprim for s in self.shells)

    def shell_offsets(self) -> np.ndarray:
        """mapSh2Bf: first basis function of every shell."""
        off = np.zeros(self.nshell, dtype=np.int64)
        n = 0
        for i, s in enumerate(self.shells):
            off[i] = n
            n += s.size
        return off

    def flat(self):
        """Flat arrays for the C ABI: l, pure, nprim, alpha, coef, origin (3*nshell, bohr)."""
        l = np.array([s.l for s in self.shells], dtype=np.int32)
        pure = np.array([1 if s.pure else 0 for s in self.shells], dtype=np.int32)
        nprim = np.array([s.nprim for s in self.shells], dtype=np.int32)
        alpha = np.concatenate([s.alpha for s in self.shells]).astype(np.float64)
        coef = np.concatenate([s.coef for s in self.shells]).astype(np.float64)
        origin = np.concatenate([s.origin for s in self.shells]).astype(np.float64)
        return l, pure, nprim, alpha, coef, origin


_REFERENCE_SHELLS: Dict[tuple, Dict[int, list]] = {}


def _reference_shells(name: str, force_cart: bool) -> Dict[int, list]:
    """Per element: (l, pure, alpha, normalised coefficients) of its reference shells, parsed
    and normalised once per process (read-only arrays, shared by every atom built from them)."""
    key = (basis_path(name), force_cart)
    if key not in _REFERENCE_SHELLS:
        ref = parse_gbs(key[0], force_cart=force_cart)
        out = {}
        for z, shells in ref.items():
            lst = []
            for rs in shells:
                alpha = np.asarray(rs.alpha, dtype=np.float64)
                coef = normalize_coefficients(rs.l, rs.alpha, rs.coef)
                alpha.setflags(write=False)
                coef.setflags(write=False)
                lst.append((rs.l, rs.pure, alpha, coef))
            out[z] = lst
        _REFERENCE_SHELLS[key] = out
    return _REFERENCE_SHELLS[key]


def build_basis(name: str, atomic_numbers: Sequence[int], coords_bohr: np.ndarray,
                force_cart: bool = False) -> BasisSet:
    """constructLocal: one copy of the element's reference shells per atom, in atom order."""
    ref = _reference_shells(name, force_cart)
    coords_bohr = np.asarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
    bs = BasisSet(name)
    for iat, z in enumerate(atomic_numbers):
        if z not in ref:
            raise ValueError(f"Atomic Number {z} not found in current Basis Set")
        for l, pure, alpha, coef in ref[z]:
            bs.shells.append(Shell(l=l, pure=pure, alpha=alpha, coef=coef, origin=coords_bohr[iat].copy(), atom=iat))
    return bs
