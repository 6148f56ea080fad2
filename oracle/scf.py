"""numpy restatement of the RHF SCF driver -- TEST INFRASTRUCTURE ONLY.

Follows SingleSlater<T>::SCF3 (include/singleslater/singleslater_scf.h:31-269) for the
closed-shell real case:

* X = S^{-1/2} (Loewdin, src/aointegrals/aointegrals_builders.cpp:736-769).
* formPT (singleslater_fock.h:30-93): PT = 2 J[P] - K[P] via the direct contraction with
  X = {P, P}, types = {COULOMB, EXCHANGE}, scales = {1, -1/2}.
* formFock (singleslater_fock.h:99-274): F_S = 2 H + PT; the orthonormal Fock that is
  diagonalised is 1/2 X F_S X (singleslater_scfutils.h:567-569).
* formDensity (singleslater_misc.h:27-56): P = 2 C_occ C_occ^T.
* computeEnergy (singleslater_properties.h:25-112): E = Tr(P H) + 1/4 Tr(P PT) + E_nuc.
* CDIIS (singleslater_diis.h:34-156, include/extrapolate.h:60-142) on the orthonormal
  commutator FP - PF, depth 10.
* Convergence (singleslater_scfutils.h:262-355): |dE| < eneTol and ||dP||_F < denTol.
"""
from __future__ import annotations

import numpy as np

from . import COULOMB, EXCHANGE, OracleBasis


def lowdin(S: np.ndarray) -> np.ndarray:
    w, U = np.linalg.eigh(S)
    return (U / np.sqrt(w)) @ U.T


def rhf(mol, basis, tau: float = 1e-12, max_iter: int = 128, ene_tol: float = 1e-10,
        den_tol: float = 1e-8, ndiis: int = 10, nthreads: int = 8, P0=None, fock_hook=None):
    """Closed-shell RHF.  Returns dict with E, P, iterations, and per-iteration PT history.

    fock_hook(P) -> PT replaces the oracle's direct contraction when given (used to drive
    the SCF with another G[D] implementation while keeping every other step here)."""
    ob = OracleBasis(basis)
    S, T, V = ob.one_electron(mol)
    H = T + V
    X = lowdin(S)
    Q = ob.schwarz(nthreads)
    nocc = mol.nelectron // 2
    enuc = mol.nuclear_repulsion()

    def form_pt(P):
        if fock_hook is not None:
            return fock_hook(P)
        J = np.zeros_like(P, order="F")
        K = np.zeros_like(P, order="F")
        ob.contract_direct(Q, tau, [P, P], [J, K], [COULOMB, EXCHANGE], [1.0, -0.5], nthreads)
        return J + K

    def density(F):
        Fo = X @ F @ X
        e, C = np.linalg.eigh(Fo)
        Cao = X @ C[:, :nocc]
        return 2.0 * Cao @ Cao.T

    if P0 is None:
        P = density(H)      # core guess
    else:
        P = P0.copy()
    E_old = 0.0
    hist_F, hist_e = [], []
    it = 0
    E = None
    converged = False
    for it in range(max_iter):
        PT = form_pt(np.asfortranarray(P))
        F = H + 0.5 * PT
        E = float(np.sum(P * H) + 0.25 * np.sum(P * PT) + enuc)
        Fo = X @ F @ X
        Si = np.linalg.inv(X)
        Po = Si @ P @ Si.T
        err = Fo @ Po - Po @ Fo
        hist_F.append(F.copy())
        hist_e.append(err)
        if len(hist_F) > ndiis:
            hist_F.pop(0)
            hist_e.pop(0)
        if len(hist_F) > 1:
            n = len(hist_F)
            Bm = -np.ones((n + 1, n + 1))
            Bm[n, n] = 0.0
            for i in range(n):
                for j in range(n):
                    Bm[i, j] = np.sum(hist_e[i] * hist_e[j])
            rhs = np.zeros(n + 1)
            rhs[n] = -1.0
            try:
                c = np.linalg.solve(Bm, rhs)[:n]
                F = sum(ci * Fi for ci, Fi in zip(c, hist_F))
            except np.linalg.LinAlgError:
                pass
        Pn = density(F)
        dP = np.linalg.norm(Pn - P)
        dE = E - E_old
        E_old = E
        P = Pn
        if it > 0 and abs(dE) < ene_tol and dP < den_tol:
            converged = True
            break
    # energy at the final density
    PT = form_pt(np.asfortranarray(P))
    E = float(np.sum(P * H) + 0.25 * np.sum(P * PT) + enuc)
    return {"E": E, "P": P, "iterations": it + 1, "converged": converged, "H": H, "S": S, "X": X,
            "Q": Q, "enuc": enuc}
