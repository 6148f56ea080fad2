"""Molecules and the seeded geometries the G[D] workloads are defined on.

* Units: input in Angstrom, converted with ``cart /= phys.bohr`` (include/molecule.h:97,
  include/global.h:224).
* Nuclear repulsion: pairwise Z_i Z_j / r_ij (src/molecule/molecule.cpp:159-170).
* Water geometry of tests/test0001.inp:6-12 (the golden-energy decks) and the monomer
  used for the clusters: O at the origin, H at (0, +-0.757, 0.587) Angstrom.
* ``water_cluster`` builds (H2O)_m: rigid monomers, uniformly random orientations
  (Shoemake unit quaternion from three uniforms), oxygen centres uniform in a cube,
  redrawn when any O-O distance is below 2.7 Angstrom.  The generator is ``numpy``'s
  ``MT19937`` seeded with the integer seed, so the same seed gives the same cluster on every
  machine.
* ``alkane`` builds all-trans C_nH_{2n+2} with C-C 1.54, C-H 1.09 Angstrom, tetrahedral angles.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

from .basis import BOHR_PER_ANGSTROM_INV, SYMBOL_TO_Z, ELEMENTS


@dataclass
class Molecule:
    Z: np.ndarray               # atomic numbers
    coords_bohr: np.ndarray     # (natom, 3)
    charge: int = 0
    mult: int = 1

    @classmethod
    def from_angstrom(cls, symbols: Sequence[str], coords_ang, charge=0, mult=1):
        Z = np.array([SYMBOL_TO_Z[s.upper()] for s in symbols], dtype=np.int64)
        xyz = np.asarray(coords_ang, dtype=np.float64).reshape(-1, 3) / BOHR_PER_ANGSTROM_INV
        return cls(Z, xyz, charge, mult)

    @property
    def natom(self) -> int:
        return len(self.Z)

    @property
    def nelectron(self) -> int:
        return int(self.Z.sum()) - self.charge

    def nuclear_repulsion(self) -> float:
        """sum_{i<j} Z_i Z_j / |R_i - R_j| (bohr, hartree)."""
        R = np.asarray(self.coords_bohr, dtype=np.float64)
        Z = np.asarray(self.Z, dtype=np.float64)
        i, j = np.triu_indices(self.natom, k=1)
        d = np.sqrt(((R[i] - R[j]) ** 2).sum(axis=1))
        return float(np.sum(Z[i] * Z[j] / d))

    def symbols(self) -> List[str]:
        return [ELEMENTS[z] for z in self.Z]

    def to_xyz(self, comment: str = "") -> str:
        ang = self.coords_bohr * BOHR_PER_ANGSTROM_INV
        lines = [str(self.natom), comment]
        for s, (x, y, z) in zip(self.symbols(), ang):
            lines.append(f"{s:2s} {x:20.12f} {y:20.12f} {z:20.12f}")
        return "\n".join(lines) + "\n"


def water_test0001() -> Molecule:
    """Water of tests/test0001.inp:6-12 (golden-energy geometry)."""
    return Molecule.from_angstrom(
        ["O", "H", "H"],
        [[0.0, -0.07579184359, 0.0],
         [0.866811829, 0.6014357793, 0.0],
         [-0.866811829, 0.6014357793, 0.0]])


WATER_MONOMER_ANG = np.array([[0.0, 0.0, 0.0],
                              [0.0, 0.757, 0.587],
                              [0.0, -0.757, 0.587]])


def water_monomer() -> Molecule:
    return Molecule.from_angstrom(["O", "H", "H"], WATER_MONOMER_ANG)


def _quat_to_rot(u1: float, u2: float, u3: float) -> np.ndarray:
    """Shoemake's uniform random rotation from three uniforms in [0,1)."""
    a = math.sqrt(1.0 - u1)
    b = math.sqrt(u1)
    q = (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
         b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(m: int, seed: int, box_ang: float = 10.0, min_oo_ang: float = 2.7,
                  max_attempts: int = 100000) -> Molecule:
    """Seeded (H2O)_m: random orientations, O centres uniform in a box_ang cube."""
    rng = np.random.Generator(np.random.MT19937(seed))
    centres: List[np.ndarray] = []
    atoms: List[np.ndarray] = []
    attempts = 0
    while len(centres) < m:
        attempts += 1
        if attempts > max_attempts:
            raise RuntimeError("water_cluster: too many rejected placements")
        u = rng.random(3)
        R = _quat_to_rot(*u)
        c = rng.random(3) * box_ang
        if any(np.linalg.norm(c - o) < min_oo_ang for o in centres):
            continue
        centres.append(c)
        atoms.append(WATER_MONOMER_ANG @ R.T + c)
    xyz = np.concatenate(atoms, axis=0)
    return Molecule.from_angstrom(["O", "H", "H"] * m, xyz)


def scaled_box(m: int, box16: float = 10.0) -> float:
    """Cube edge for (H2O)_m at the density of 16 waters in a box16 cube."""
    return box16 * (m / 16.0) ** (1.0 / 3.0)


def alkane(n: int) -> Molecule:
    """All-trans C_nH_{2n+2}: C-C 1.54, C-H 1.09 Angstrom, tetrahedral angles."""
    cc, ch = 1.54, 1.09
    tet = math.acos(-1.0 / 3.0)
    half = tet / 2.0
    # zig-zag backbone in the xy plane along x
    dx = cc * math.sin(half)
    dy = cc * math.cos(half)
    carbons = [np.array([i * dx, (dy if i % 2 else 0.0), 0.0]) for i in range(n)]
    syms: List[str] = ["C"] * n
    xyz: List[np.ndarray] = list(carbons)
    for i, c in enumerate(carbons):
        # in-plane outward direction (away from the backbone zig-zag side)
        side = -1.0 if i % 2 == 0 else 1.0
        nbrs = []
        if i > 0:
            nbrs.append(carbons[i - 1])
        if i < n - 1:
            nbrs.append(carbons[i + 1])
        if n == 1:
            # methane: tetrahedron
            for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]):
                xyz.append(c + ch * np.array(v, dtype=float) / math.sqrt(3.0))
                syms.append("H")
            continue
        bis = np.array([0.0, side, 0.0])
        hdirs = []
        # two out-of-plane hydrogens, bisected by the in-plane outward direction
        ang = half
        for sz in (1.0, -1.0):
            hdirs.append(np.array([0.0, side * math.cos(ang), sz * math.sin(ang)]))
        if len(nbrs) == 1:
            # terminal carbon: third H continues the zig-zag
            away = c - nbrs[0]
            away /= np.linalg.norm(away)
            # reflect the backbone bond across the local bisector for an in-plane H
            hx = np.array([math.copysign(dx, away[0]), -side * 0.0 + (dy if i % 2 == 0 else -dy), 0.0])
            hx /= np.linalg.norm(hx)
            hdirs.append(hx)
        for d in hdirs:
            d = d / np.linalg.norm(d)
            xyz.append(c + ch * d)
            syms.append("H")
        del bis
    return Molecule.from_angstrom(syms, np.array(xyz))
