"""Oracle FP64 RHF energies of the workload molecules -> tests/golden/scf_oracle.json.

The energies come from the oracle's numpy restatement of SCF3 (oracle/scf.py) with its C++
restatement of newTwoEContractDirect (include/aointegrals/aointegrals_contract.h:24-194) at
tau = 1e-12, converged to |dE| < 1e-10 and ||dP||_F < 1e-8 (the reference's rule,
include/singleslater/singleslater_scfutils.h:262-355).  The oracle itself is pinned by the
reference's golden water energies (tests/golden/energies.json).  Molecules:

* (H2O)16 / cc-pVDZ, cluster seed 0 (the headline G[D] workload and the mixed-precision one);
* CH4, C12H26 (standard all-trans geometry) and the first (H2O)8 of the 256-molecule batch
  (chronusq_public_amd.batch.c4_molecules(0)), 6-31G*;
* the first held-out C12H26 of the learned-guess evaluation set (train_guess.evaluation_molecules), 6-31G*.

Run once (CPU only, a few minutes on 8 cores):  python tests/golden/make_scf_fixtures.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from chronusq_public_amd import basis as B          # noqa: E402
from chronusq_public_amd import molecule as M       # noqa: E402
from chronusq_public_amd.batch import c4_molecules  # noqa: E402
from chronusq_public_amd.train_guess import evaluation_molecules  # noqa: E402
from oracle import scf as oscf                       # noqa: E402


def workload():
    """(key, molecule, basis, description) of every fixture."""
    c4 = c4_molecules(0)
    k8 = next(k for k, m in enumerate(c4) if m.natom == 24)
    held = evaluation_molecules(0, 1)[0][1]
    return [
        ("h2o16_ccpvdz_seed0", M.water_cluster(16, seed=0), "cc-pvdz", "(H2O)16 cluster seed 0"),
        ("ch4_631gs", M.alkane(1), "6-31g*", "CH4, standard geometry"),
        ("c12h26_631gs", M.alkane(12), "6-31g*", "C12H26, standard all-trans geometry"),
        ("h2o8_631gs_c4", c4[k8], "6-31g*", f"c4_molecules(0)[{k8}], the batch's first (H2O)8"),
        ("c12h26_631gs_heldout0", held, "6-31g*", "evaluation_molecules(0, 1)[0]: held-out displaced C12H26"),
    ]


def main():
    nthreads = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 8))
    out = {"source": "oracle/scf.py rhf (numpy SCF3 restatement) with the oracle direct contraction",
           "tau": 1e-12, "ene_tol": 1e-10, "den_tol": 1e-8, "molecules": {}}
    for key, mol, basis, desc in workload():
        bs = B.build_basis(basis, mol.Z, mol.coords_bohr)
        t0 = time.time()
        r = oscf.rhf(mol, bs, tau=1e-12, max_iter=128, nthreads=nthreads)
        assert r["converged"], key
        out["molecules"][key] = {"description": desc, "basis": basis, "nbf": bs.nbasis, "natom": mol.natom,
                                 "E": r["E"], "iterations": r["iterations"],
                                 "coords_checksum": float(np.sum(np.abs(mol.coords_bohr)))}
        print(f"{key}: E = {r['E']:.10f} in {r['iterations']} iterations, {time.time() - t0:.1f} s", flush=True)
    with open(os.path.join(HERE, "scf_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
