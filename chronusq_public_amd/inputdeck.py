"""ChronusQ input decks and .xyz geometries.

The deck format is the reference's INI file (src/python/src/parse/parseInput.py:159-...,
keyword table src/python/src/parse/meta/knownKeywords.py): sections [Molecule], [QM],
[BASIS], [INTS], [SCF], [MISC] (and [RT], which is rejected: real-time propagation is not
on this path).  Section and keyword names are case-insensitive; unknown keywords and missing
required ones are errors, as in the reference.  The geometry (parseMolecule.py:120-175) is
one atom per line, `<symbol or Z> [isotope] x y z`, in Angstrom.

Only the jobs this path serves run: JOB = SCF with a real RHF/UHF reference (the G[D] path);
DFT functionals, complex/two-component references, RT, response and correlated methods raise
DeckError with the reason.
"""
from __future__ import annotations

import configparser
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from .basis import SYMBOL_TO_Z
from .molecule import Molecule


class DeckError(ValueError):
    """Invalid or unsupported input (the reference's CErrMsg / 'Input Error')."""


# section -> keyword -> (type, required); types: I int, D float, S string, B bool, D3 3 floats
KNOWN: Dict[str, Dict[str, tuple]] = {
    "MOLECULE": {"CHARGE": ("I", True), "MULT": ("I", True), "GEOM": ("S", True), "PRINT": ("I", False)},
    "QM": {"REFERENCE": ("S", True), "JOB": ("S", True), "EXCHANGE": ("S", False), "CORR": ("S", False),
           "DFT_NRAD": ("I", False), "DFT_NANG": ("I", False), "DFT_SCREEN": ("B", False),
           "DFT_SCRTOL": ("D", False), "PRINT": ("I", False)},
    "INTS": {"ALG": ("S", False)},
    "BASIS": {"BASIS": ("S", True), "FORCECART": ("B", False)},
    "SCF": {"DENTOL": ("D", False), "ENETOL": ("D", False), "MAXITER": ("I", False), "FIELD": ("D3", False),
            "GUESS": ("S", False), "DIIS": ("B", False), "ITP": ("B", False), "DT": ("D", False),
            "PRINT": ("I", False)},
    "MISC": {"NSMP": ("I", False), "PRINT": ("I", False)},
    "RT": {},
}
REQUIRED_SECTIONS = ("MOLECULE", "QM", "BASIS")


@dataclass
class Deck:
    molecule: Molecule
    basis: str
    reference: str            # "RHF" or "UHF"
    job: str = "SCF"
    guess: str = "SAD"
    ene_tol: float = 1e-10
    den_tol: float = 1e-8
    max_iter: int = 256
    diis: bool = True
    integral_algorithm: str = "DIRECT"
    raw: Dict[str, Dict[str, str]] = field(default_factory=dict)


def _convert(sec: str, key: str, typ: str, value: str):
    try:
        if typ == "I":
            return int(value)
        if typ == "D":
            return float(value)
        if typ == "B":
            v = value.strip().upper()
            if v in ("TRUE", "ON", "YES", "1"):
                return True
            if v in ("FALSE", "OFF", "NO", "0"):
                return False
            raise ValueError(value)
        if typ == "D3":
            v = [float(x) for x in value.split()]
            if len(v) != 3:
                raise ValueError(value)
            return v
        return value
    except ValueError as e:
        raise DeckError(f"Input Error: invalid value for {sec}.{key}: {value!r}") from e


def parse_geometry(text: str, charge: int, mult: int) -> Molecule:
    """`<ATM> X Y Z` or `<ATM> <ISO> X Y Z` (Angstrom); ATM a symbol or an atomic number."""
    symbols, xyz = [], []
    for line in text.strip().splitlines():
        f = line.split()
        if not f:
            continue
        if len(f) not in (4, 5):
            raise DeckError("Input Error: Invalid Geometry Specification")
        sym = f[0]
        if any(ch.isdigit() for ch in sym):
            z = int(sym)
            inv = {v: k for k, v in SYMBOL_TO_Z.items()}
            if z not in inv:
                raise DeckError("Input Error: Invalid Atomic Symbol or Mass Number")
            sym = inv[z]
        elif sym.upper() not in SYMBOL_TO_Z:
            raise DeckError("Input Error: Invalid Atomic Symbol or Mass Number")
        symbols.append(sym.upper())
        xyz.append([float(v) for v in f[-3:]])
    if not symbols:
        raise DeckError("Input Error: empty geometry")
    return Molecule.from_angstrom(symbols, np.array(xyz), charge=charge, mult=mult)


def parse_deck_text(text: str) -> Deck:
    cp = configparser.ConfigParser(interpolation=None, strict=False)
    cp.optionxform = str
    try:
        cp.read_string(text)
    except configparser.Error as e:
        raise DeckError(f"Input Error: {e}") from e
    raw: Dict[str, Dict[str, str]] = {}
    for sec in cp.sections():
        S = sec.upper()
        if S not in KNOWN:
            raise DeckError(f"Input Error: unknown section [{sec}]")
        raw[S] = {}
        for key, val in cp.items(sec):
            K = key.upper()
            if K not in KNOWN[S]:
                raise DeckError(f"Input Error: unknown keyword {K} in section [{sec}]")
            raw[S][K] = val
    for S in REQUIRED_SECTIONS:
        if S not in raw:
            raise DeckError(f"Input Error: required section [{S}] missing")
    for S, keys in KNOWN.items():
        for K, (typ, req) in keys.items():
            if req and S in raw and K not in raw[S]:
                raise DeckError(f"Input Error: required keyword {K} missing in [{S}]")
    if "RT" in raw:
        raise DeckError("real-time propagation is not on the direct-SCF G[D] path")
    v = {S: {K: _convert(S, K, KNOWN[S][K][0], x) for K, x in d.items()} for S, d in raw.items()}
    mol = parse_geometry(v["MOLECULE"]["GEOM"], v["MOLECULE"]["CHARGE"], v["MOLECULE"]["MULT"])
    # reference: [REAL|COMPLEX] <RHF|UHF|...> (parseQM.py:113-140)
    ref_words = v["QM"]["REFERENCE"].upper().split()
    if "COMPLEX" in ref_words or "X2C" in ref_words:
        raise DeckError("complex / two-component references are not supported by this G[D] path")
    ref = ref_words[-1]
    if ref not in ("RHF", "UHF"):
        raise DeckError(f"reference {ref} is not supported (Hartree-Fock RHF/UHF only; no DFT)")
    job = v["QM"]["JOB"].upper()
    if job != "SCF":
        raise DeckError(f"job {job} is not supported (SCF only)")
    if "CORR" in v["QM"]:
        raise DeckError("correlated methods are not supported")
    scf = v.get("SCF", {})
    if "FIELD" in scf and any(abs(x) > 0 for x in scf["FIELD"]):
        raise DeckError("static fields are not supported by this SCF driver")
    guess = scf.get("GUESS", "SAD").upper()
    # LEARNED (the network guess of learned_guess.py) has no reference counterpart; READ
    # takes the densities of the restart file <input basename>.bin
    if guess not in ("SAD", "CORE", "READ", "LEARNED"):
        raise DeckError(f"guess {guess} is not supported (SAD, CORE, READ or LEARNED)")
    if mol.natom == 1:
        guess = "CORE"           # the reference defaults to CORE for one atom (parseQM.py:84-87)
    alg = v.get("INTS", {}).get("ALG", "DIRECT").upper()
    if alg not in ("DIRECT", "INCORE"):
        raise DeckError(f"integral algorithm {alg} unknown")
    if v["BASIS"].get("FORCECART", False):
        raise DeckError("FORCECART is not supported (pure d shells as in the reference default)")
    return Deck(molecule=mol, basis=v["BASIS"]["BASIS"].lower(), reference=ref, job=job, guess=guess,
                ene_tol=scf.get("ENETOL", 1e-10), den_tol=scf.get("DENTOL", 1e-8),
                max_iter=scf.get("MAXITER", 256), diis=scf.get("DIIS", True), integral_algorithm=alg,
                raw=raw)


def parse_deck(path: str) -> Deck:
    with open(path) as f:
        return parse_deck_text(f.read())


def read_xyz(path_or_text: str, charge: int = 0, mult: int = 1) -> Molecule:
    """Standard .xyz: atom count, comment line, `symbol x y z` lines (Angstrom)."""
    text = path_or_text
    if "\n" not in path_or_text:
        with open(path_or_text) as f:
            text = f.read()
    lines = text.strip().splitlines()
    try:
        n = int(lines[0].split()[0])
    except (ValueError, IndexError) as e:
        raise DeckError("xyz: first line must hold the atom count") from e
    body = lines[2:2 + n]
    if len(body) != n:
        raise DeckError("xyz: fewer atom lines than the count")
    return parse_geometry("\n".join(body), charge, mult)


def write_xyz(mol: Molecule, path: Optional[str] = None, comment: str = "") -> str:
    text = mol.to_xyz(comment)
    if path:
        with open(path, "w") as f:
            f.write(text)
    return text
