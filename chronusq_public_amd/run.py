"""Run a ChronusQ input deck on the GPU path:  python -m chronusq_public_amd.run deck.inp

Follows the reference's job sequence for SCF jobs (runCQJob,
src/python/src/standardJobs.py:55-99: molecule and basis set-up, guess, SCF3, properties),
with the device SCF driver of this package.  INTS ALG = INCORE runs the direct contraction:
the reference's in-core and direct paths give the same G[D] for symmetric densities.  The
converged densities go to the restart file <input basename>.bin (src/fileio/fileio.cpp:58-69),
which GUESS = READ reads back.
"""
from __future__ import annotations

import os
import sys
import time

from .inputdeck import DeckError, parse_deck


def run_deck(path: str, out=sys.stdout, restart_file: str = None) -> dict:
    from .singleslater import SingleSlater
    deck = parse_deck(path)
    mol = deck.molecule
    restart_file = restart_file or os.path.splitext(path)[0] + ".bin"
    t0 = time.perf_counter()
    # the reference's SCF always damps the orthonormal Fock by 0.1 while |dE| > 1e-6
    # (include/singleslater.h:472-473)
    ss = SingleSlater(mol, deck.basis, ref=deck.reference, guess=deck.guess, ene_tol=deck.ene_tol,
                      den_tol=deck.den_tol, max_iter=deck.max_iter, ndiis=10 if deck.diis else 0,
                      restart_file=restart_file, damp=0.1)
    res = ss.SCF3()
    wall = time.perf_counter() - t0
    print(f"  Molecule: {mol.natom} atoms, charge {mol.charge}, multiplicity {mol.mult}", file=out)
    print(f"  Basis: {deck.basis} ({ss.N} functions)   Reference: {deck.reference}   Guess: {deck.guess}",
          file=out)
    if deck.integral_algorithm == "INCORE":
        print("  INTS ALG = INCORE: computed with the direct contraction", file=out)
    print(f"  SCF converged in {res.iterations} iterations ({wall:.3f} s)", file=out)
    print(f"  Total Energy = {res.energy:.10f} Eh", file=out)
    print(f"  Restart file: {restart_file}", file=out)
    return {"energy": res.energy, "iterations": res.iterations, "converged": res.converged, "wall_s": wall,
            "restart_file": restart_file}


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m chronusq_public_amd.run <deck.inp>", file=sys.stderr)
        return 2
    try:
        run_deck(argv[0])
    except DeckError as e:
        print(f"Input Error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
