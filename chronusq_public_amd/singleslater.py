"""Device-resident Hartree-Fock driver around the GPU G[D] (mirror of ChronusQ's
``SingleSlater<double>`` for real RHF / UHF).

Everything stays in FP64 on the GPU: one-electron integrals and G[D] come from libcqgpu.so,
the dense linear algebra (orthonormalisation, CDIIS, eigensolve) runs through torch on the
same device.  Steps and conventions follow the reference:

* ``formPT`` (include/singleslater/singleslater_fock.h:30-93): RHF X = {P, P},
  types = {COULOMB, EXCHANGE}, scales = {1, -1/2}, so PT = 2 J[P] - K[P]; UHF adds
  {Mz, EXCHANGE, -1/2}, PT_z = -K[Mz].
* ``formFock`` (singleslater_fock.h:99-274): F_S = 2 H + PT (the orthonormal Fock that is
  diagonalised is 1/2 X F_S X, singleslater_scfutils.h:567-569); here F = H + PT/2 directly.
* ``computeEnergy`` (singleslater_properties.h:25-112): E = Tr(P H) + 1/4 Tr(P PT) + E_nuc
  (UHF adds 1/4 Tr(Mz PT_z)).
* ``formDensity`` (singleslater_misc.h:27-56): P = 2 C_occ C_occ^T (RHF), P_a + P_b / P_a - P_b
  (UHF).
* CDIIS (singleslater_diis.h:34-156, include/extrapolate.h:60-142): depth 10, orthonormal
  commutator FP - PF; a singular B system falls back to the plain step.
* damping (singleslater_scfutils.h:1065-1111): F_o <- (1-d) F_o + d F_o(prev) while
  |dE| > 1e-6 (off by default here: it only lengthens the path to the same energy).
* convergence (singleslater_scfutils.h:262-355): |dE| < eneTol and ||dP||_F < denTol
  (1e-10, 1e-8, include/singleslater.h:448-449); non-convergence raises.
* density step: the reference diagonalises the orthonormal Fock every iteration
  (singleslater_scfutils.h:567-581).  Here RHF does so while |dE| > 0.1 Eh and then rotates
  the previous orbitals to first order in F_ov (pseudo-diagonalisation, GEMMs only: the
  eigensolve is launch-bound on the GPU); the fixed point F_ov = 0 is the same, and so is the
  converged energy.
* SAD guess (singleslater_guess.h:99-222): one UHF per unique element at its default
  multiplicity, atomic scalar densities on the diagonal blocks, spin-averaged, one Fock build
  and diagonalisation.  The atomic UHFs use CDIIS (the reference switches DIIS off for them).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import basis as B
from .aointegrals import AOIntegrals, COULOMB, EXCHANGE
from .molecule import Molecule

# ground-state multiplicities used for the atomic guesses
DEFAULT_MULT = {1: 2, 2: 1, 3: 2, 4: 1, 5: 2, 6: 3, 7: 4, 8: 3, 9: 2, 10: 1}

# Converged atomic UHF densities of the SAD guess, per (Z, multiplicity, basis name, device):
# the same atoms recur across molecules and SCF runs of one process.
_SAD_ATOMS: Dict[tuple, object] = {}


def clear_sad_cache():
    """Forget the cached atomic densities (the next SAD guess recomputes them)."""
    _SAD_ATOMS.clear()


def atomic_density(z: int, basis_name: str, dev, ndiis: int = 10, use_cache: bool = True):
    """Spin-summed density of the free atom (UHF at its default multiplicity, CORE guess, from
    the per-(Z, multiplicity, basis, device) cache) -- the SAD guess's atomic block
    (singleslater_guess.h:99-222)."""
    key = (z, DEFAULT_MULT.get(z, 1), basis_name, str(dev))
    if use_cache and key in _SAD_ATOMS:
        return _SAD_ATOMS[key]
    atom = Molecule(np.array([z]), np.zeros((1, 3)), 0, DEFAULT_MULT.get(z, 1))
    P = SingleSlater(atom, basis_name, ref="UHF", tau=1e-14, guess="CORE", ndiis=ndiis, max_iter=512,
                     device=dev).SCF3().P
    if use_cache:
        _SAD_ATOMS[key] = P
    return P


class SCFNotConverged(RuntimeError):
    """SCF failed to converge within maxSCFIter (singleslater_scf.h:253-255)."""


@dataclass
class SCFResult:
    energy: float
    iterations: int
    converged: bool
    energies: List[float] = field(default_factory=list)
    times: List[float] = field(default_factory=list)       # seconds since SCF start, per iteration
    gd_times: List[float] = field(default_factory=list)    # seconds per G[D]
    P: object = None
    Mz: object = None
    setup_s: float = 0.0
    setup_parts: Dict[str, float] = field(default_factory=dict)
    guess_s: float = 0.0
    guess_info: Dict[str, object] = field(default_factory=dict)
    phase_times: List[Dict[str, float]] = field(default_factory=list)   # with profile: per iteration


class SingleSlater:
    def __init__(self, mol: Molecule, basis, ref: str = "RHF", tau: float = 1e-12, guess: str = "SAD",
                 ndiis: int = 10, damp: float = 0.0, ene_tol: float = 1e-10, den_tol: float = 1e-8,
                 max_iter: int = 256, device=None, aoints: Optional[AOIntegrals] = None,
                 incremental: bool = False, rebuild_every: int = 8, sad_cache: bool = True,
                 precision: str = "fp64", mixed_threshold: float = 1e-5,
                 pseudo_diag: bool = True, pdiag_switch: float = 1.0, pdiag_first: bool = True,
                 guess_model: Optional[str] = None, judge_guess="auto", judge_idem_tol: float = 0.15,
                 restart_file: Optional[str] = None):
        import torch
        self.torch = torch
        self.mol = mol
        self.basis = basis if isinstance(basis, B.BasisSet) else B.build_basis(basis, mol.Z, mol.coords_bohr)
        self.ref = ref.upper()
        self.guess = guess.upper()
        self.ndiis, self.damp = ndiis, damp
        # incremental Fock build: PT(k) = PT(k-1) + G[P(k) - P(k-1)], the difference density
        # screened by its own block norms; a full build every `rebuild_every` iterations
        # (the reference carries disabled scaffolding for this, singleslater_scf.h:62-63,
        # singleslater_scfutils.h:423-564)
        self.incremental, self.rebuild_every = incremental, rebuild_every
        self.sad_cache = sad_cache
        # learned guess: network file (default: the one trained for this basis) and whether
        # the SAD density competes with it on energy
        # judge_guess: True always makes the SAD start too and keeps the lower-energy one; "auto"
        # does so only when the network's prediction is not nearly idempotent (relative error
        # ||P S P / 2 - P||_F / ||P||_F above judge_idem_tol, or a wrong electron count): a
        # prediction outside what the network has learned; False never
        if judge_guess not in (True, False, "auto"):
            raise ValueError("judge_guess must be True, False or 'auto'")
        self.guess_model, self.judge_guess, self.judge_idem_tol = guess_model, judge_guess, judge_idem_tol
        self.guess_info: Dict[str, object] = {}
        # restart file in the reference's layout (restart.py): the READ guess reads its
        # densities; SCF3 writes the converged ones into it
        self.restart_file = restart_file
        self.ene_tol, self.den_tol, self.max_iter = ene_tol, den_tol, max_iter
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        t0 = time.perf_counter()
        self.ao = aoints or AOIntegrals(self.basis)
        self.ao.thresholdSchwartz = tau
        # mixed: FP32 integrals for quartets whose bound (inside the incremental build:
        # Schwarz x Schwarz x max|dD|) is below mixed_threshold; full FP64 rebuilds stay FP64
        self.precision, self.mixed_threshold = precision, mixed_threshold
        # RHF density step: first-order orbital rotations of the previous (or the guess's)
        # orbitals once |dE| < pdiag_switch (pseudo-diagonalisation, GEMMs only), the eigensolve
        # before that and whenever an occupied-virtual gap closes
        self.pseudo_diag, self.pdiag_switch = pseudo_diag, pdiag_switch
        # pdiag_first: the first two iterations (whose energy change is not known yet) already
        # rotate the guess's orbitals
        self.pdiag_first = pdiag_first
        t1 = time.perf_counter()
        S, T, V = self.ao.computeAOOneE(mol, self.dev)
        self.S, self.H = S, T + V
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        # S^{-1/2} (Loewdin, aointegrals_builders.cpp:736-769) and S^{1/2}
        self.X, self.Xi = self._lowdin(S)
        torch.cuda.synchronize()
        self.setup_parts = {"context_s": t1 - t0, "one_e_s": t2 - t1, "lowdin_s": time.perf_counter() - t2}
        self.enuc = mol.nuclear_repulsion()
        nel = mol.nelectron
        if self.ref == "RHF":
            if nel % 2:
                raise ValueError("RHF needs an even electron count")
            self.nocc = (nel // 2, nel // 2)
        else:
            na = (nel + mol.mult - 1) // 2
            self.nocc = (na, nel - na)
        self.N = self.basis.nbasis
        self.setup_s = time.perf_counter() - t0

    # ------------------------------------------------------------------ pieces
    def _lowdin(self, S):
        """S^{-1/2}, S^{1/2} by the coupled Newton-Schulz iteration on A = S / c (c >= the
        largest eigenvalue, Gershgorin): Y <- Y T, Z <- T Z, T = (3 I - Z Y) / 2 converges
        quadratically to A^{1/2}, A^{-1/2} (GEMMs only; the eigensolve is launch-bound on the
        GPU).  Falls back to the eigendecomposition if the residual |Z Y - I| stays above
        1e-12 (a nearly singular S)."""
        torch = self.torch
        n = S.shape[0]
        eye = torch.eye(n, dtype=S.dtype, device=S.device)
        c = float(S.abs().sum(dim=1).max())
        Y, Z = S / c, eye.clone()
        for it in range(60):
            T = 1.5 * eye - 0.5 * (Z @ Y)
            Y, Z = Y @ T, T @ Z
            if it % 6 == 5 and float((Z @ Y - eye).abs().max()) < 1e-13:
                break
        if float((Z @ Y - eye).abs().max()) < 1e-12:
            Z = 0.5 * (Z + Z.T)
            Y = 0.5 * (Y + Y.T)
            return Z / c ** 0.5, Y * c ** 0.5
        w, U = torch.linalg.eigh(S)
        return (U * w.rsqrt()) @ U.T, (U * w.sqrt()) @ U.T

    def formPT(self, P, Mz=None):
        torch = self.torch
        N = self.N
        J = torch.zeros((N, N), dtype=torch.float64, device=self.dev)
        K = torch.zeros_like(J)
        if Mz is None:
            self.ao.newTwoEContractDirect([P, P], [J, K], [COULOMB, EXCHANGE], [1.0, -0.5])
            return J + K, None
        Kz = torch.zeros_like(J)
        self.ao.newTwoEContractDirect([P, P, Mz], [J, K, Kz], [COULOMB, EXCHANGE, EXCHANGE], [1.0, -0.5, -0.5])
        return J + K, Kz

    @staticmethod
    def _occupied_density(Cao):
        """C_occ C_occ^T made symmetric to the last bit.  A GEMM need not round elements (i,j)
        and (j,i) alike (at 1536 functions 1.7 M pairs differ by ~5e-16), and a density that is
        not exactly symmetric takes the contraction's non-symmetric path, which digests the
        transposed density as well: G[P] of (H2O)64/cc-pVDZ 633 ms instead of 150 ms."""
        P = Cao @ Cao.T
        return 0.5 * (P + P.T)

    def _diag_density(self, F, nocc, with_vectors=False):
        torch = self.torch
        Fo = self.X @ F @ self.X
        e, C = torch.linalg.eigh(Fo)
        Cao = self.X @ C[:, :nocc]
        if with_vectors:
            return self._occupied_density(Cao), C
        return self._occupied_density(Cao), e

    def _pseudo_diag_density(self, F, nocc, C):
        """Density from a first-order occupied-virtual rotation of the previous orbitals
        (orthonormal basis): kappa_ai = F_ai / (e_i - e_a) with e the diagonal of F in the
        previous orbitals, C_occ += C_vir kappa, C_vir -= C_occ kappa^T, then Newton-Schulz
        re-orthonormalisation C <- C (3 I - C^T C) / 2.  GEMMs only, in place of the
        eigensolve once the SCF is near its fixed point (F_ov = 0 there, so the converged
        density is the aufbau density the eigensolve gives).  Returns None when an
        occupied-virtual gap is below 1e-3 Eh (near-degenerate frontier orbitals)."""
        torch = self.torch
        Fo = self.X @ F @ self.X
        FC = Fo @ C
        eps = (C * FC).sum(dim=0)
        eo, ev = eps[:nocc], eps[nocc:]
        den = eo[None, :] - ev[:, None]
        if float(den.abs().min()) < 1e-3:
            return None
        kap = (C[:, nocc:].T @ FC[:, :nocc]) / den
        Cn = torch.cat([C[:, :nocc] + C[:, nocc:] @ kap, C[:, nocc:] - C[:, :nocc] @ kap.T], dim=1)
        for _ in range(3):
            Cn = 1.5 * Cn - 0.5 * Cn @ (Cn.T @ Cn)
        Cao = self.X @ Cn[:, :nocc]
        return self._occupied_density(Cao), Cn

    def _energy(self, P, PT, Mz=None, PTz=None):
        E = float((P * self.H).sum() + 0.25 * (P * PT).sum()) + self.enuc
        if Mz is not None:
            E += 0.25 * float((Mz * PTz).sum())
        return E

    # ------------------------------------------------------------------ properties
    # units of the reference's printed and golden multipoles (tests/rununit.py:312-313,
    # src/python/src/.chronusq_old.py.in:222): e bohr per Debye, Angstrom per bohr
    AU_PER_DEBYE = 0.3934303070
    ANGSTROM_PER_BOHR = 0.5291772083

    def computeMultipole(self, P, order: int = 3) -> Dict[str, np.ndarray]:
        """Dipole (Debye), quadrupole (Debye Angstrom, 3 x 3) and octupole (Debye Angstrom^2,
        3 x 3 x 3) of the total density P about the origin: SingleSlater::computeMultipole
        (include/singleslater/singleslater_properties.h:115-172) with the electronic parts of
        include/quantum/quantum_stdproperties.h:25-66 -- minus Tr(P r_i ...) over the GPU
        multipole integrals, plus sum_A Z_A R_Ai ... -- in the units of the reference's golden
        values.  The quadrupole is the plain second moment (not traceless), as compared there."""
        torch = self.torch
        M = self.ao.computeMultipole(order, None, self.dev)
        el = torch.stack([(P * m).sum() for m in M]).cpu().numpy()
        Z = self.mol.Z.astype(np.float64)
        R = self.mol.coords_bohr
        out: Dict[str, np.ndarray] = {}
        dip = -el[0:3] + Z @ R
        out["dipole"] = dip / self.AU_PER_DEBYE
        if order >= 2:
            q = np.zeros((3, 3))
            for k, (i, j) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
                q[i, j] = q[j, i] = -el[3 + k] + float(np.sum(Z * R[:, i] * R[:, j]))
            out["quadrupole"] = q * self.ANGSTROM_PER_BOHR / self.AU_PER_DEBYE
        if order >= 3:
            o = np.zeros((3, 3, 3))
            idx = [(i, j, k) for i in range(3) for j in range(i, 3) for k in range(j, 3)]
            for n, (i, j, k) in enumerate(idx):
                v = -el[9 + n] + float(np.sum(Z * R[:, i] * R[:, j] * R[:, k]))
                for a, b, c in {(i, j, k), (i, k, j), (j, i, k), (j, k, i), (k, i, j), (k, j, i)}:
                    o[a, b, c] = v
            out["octupole"] = o * self.ANGSTROM_PER_BOHR ** 2 / self.AU_PER_DEBYE
        return out

    def core_guess(self):
        na, nb = self.nocc
        Pa, _ = self._diag_density(self.H, na)
        if self.ref == "RHF":
            return 2.0 * Pa, None
        Pb, _ = self._diag_density(self.H, nb)
        return Pa + Pb, Pa - Pb

    def _atomic_densities(self, spherical: bool = False):
        """Block-diagonal superposition of atomic (UHF) densities, spin-averaged, from the
        per-(Z, multiplicity, basis, device) cache (singleslater_guess.h:99-222).  With
        spherical, every atomic block is averaged over rotations (the learned guess's
        baseline, which must not depend on the molecule's orientation): a block between two
        shells of one l becomes tr/(2l+1) times the identity, blocks between different l vanish
        (Schur's lemma for the shells' irreducible rotations); electron counts are kept."""
        torch = self.torch
        atom_of_bf = np.concatenate([[s.atom] * s.size for s in self.basis.shells])
        Ps = torch.zeros((self.N, self.N), dtype=torch.float64, device=self.dev)
        cache: Dict[int, object] = {z: atomic_density(z, self.basis.name, self.dev, self.ndiis, self.sad_cache)
                                    for z in dict.fromkeys(int(v) for v in self.mol.Z)}
        if spherical:
            for z in list(cache):
                key = ("spherical", z, DEFAULT_MULT.get(z, 1), self.basis.name, str(self.dev))
                if key not in _SAD_ATOMS:
                    _SAD_ATOMS[key] = self._spherical_average(z, cache[z])
                cache[z] = _SAD_ATOMS[key]
        for iat, z in enumerate(self.mol.Z):
            idx = np.nonzero(atom_of_bf == iat)[0]
            lo, hi = int(idx[0]), int(idx[-1]) + 1
            Ps[lo:hi, lo:hi] = cache[int(z)]
        return Ps

    def _spherical_average(self, z: int, blk):
        iat = int(np.nonzero(self.mol.Z == z)[0][0])
        shells = [(s.l, s.pure, s.size) for s in self.basis.shells if s.atom == iat]
        if any(l > 1 and not pure for (l, pure, _) in shells):
            raise ValueError("spherical averaging needs pure shells above l = 1")
        out = self.torch.zeros_like(blk)
        offs = np.concatenate([[0], np.cumsum([n for (_, _, n) in shells])])
        for a, (la, _, na) in enumerate(shells):
            for b, (lb, _, nb) in enumerate(shells):
                if la != lb:
                    continue
                sub = blk[offs[a]:offs[a] + na, offs[b]:offs[b] + nb]
                out[offs[a]:offs[a] + na, offs[b]:offs[b] + nb] = (
                    self.torch.eye(na, dtype=blk.dtype, device=blk.device) * (self.torch.trace(sub) / na))
        return out

    def _guess_from_fock(self, PT):
        F = self.H + 0.5 * PT
        na, nb = self.nocc
        if self.ref == "RHF":
            Pa, C = self._diag_density(F, na, with_vectors=True)
            P = 2.0 * Pa
            self._guess_C = (P, C)     # orbitals of the starting density (pseudo-diagonalisation)
            return P, None
        Pa, _ = self._diag_density(F, na)
        Pb, _ = self._diag_density(F, nb)
        return Pa + Pb, Pa - Pb

    def sad_guess(self):
        """Superposition of atomic (UHF) densities, spin-averaged, then one Fock build and
        diagonalisation (singleslater_guess.h:24-65,99-222)."""
        self.guess_info = {"used": "SAD"}
        if self.mol.natom <= 1:
            self.guess_info["used"] = "CORE"
            return self.core_guess()
        PT, _ = self.formPT(self._atomic_densities())
        return self._guess_from_fock(PT)

    def learned_guess(self):
        """Atomic densities plus network-predicted pair blocks and block corrections
        (learned_guess.py), then one Fock build and diagonalisation as for SAD.  When judged
        (judge_guess True, or "auto" with a prediction far from idempotent) the SAD starting
        density is made as well, and of the two idempotent starting densities the one with the
        lower (variational) energy E[P0] = Tr(P0 H) + 1/4 Tr(P0 PT[P0]) + E_nuc is kept; its
        PT[P0] becomes the SCF's first Fock build, so judging costs an eigensolve and two extra
        G[D] builds.  Molecules or bases the network does not cover take the SAD guess.  No
        reference counterpart."""
        from . import learned_guess as LG
        if self.mol.natom <= 1:
            return self.sad_guess()
        try:
            net = LG.load_model(self.basis.name, path=self.guess_model, device=self.dev)
            PL = LG.predict_density(net, self.mol, self.basis, self._atomic_densities(spherical=True))
        except LG.GuessUnavailable as e:
            out = self.sad_guess()
            self.guess_info["fallback"] = str(e)
            return out
        uhf = self.ref != "RHF"
        # the prediction's distance from an RHF density of the right electron count
        torch = self.torch
        PS = PL @ self.S
        idem = float(torch.linalg.norm(PS @ PL * 0.5 - PL) / torch.linalg.norm(PL))
        nel = float(torch.trace(PS))
        confident = idem <= self.judge_idem_tol and abs(nel - self.mol.nelectron) <= 0.05 * self.mol.nelectron
        P0_L = self._guess_from_fock(self.formPT(PL)[0])
        self.guess_info = {"used": "LEARNED", "idempotency_error": idem, "electrons": nel,
                           "judged": bool(self.judge_guess is True or (self.judge_guess == "auto" and not confident))}
        if not self.guess_info["judged"]:
            return P0_L
        C_L = self._guess_C
        P0_S = self._guess_from_fock(self.formPT(self._atomic_densities())[0])
        C_S = self._guess_C
        cand = []
        for name, (P0, Mz0) in (("LEARNED", P0_L), ("SAD", P0_S)):
            PT0, PTz0 = self.formPT(P0, Mz0 if uhf else None)
            cand.append((self._energy(P0, PT0, Mz0, PTz0), name, P0, Mz0, PT0, PTz0))
        self.guess_info.update({"E0_learned": cand[0][0], "E0_sad": cand[1][0]})
        E0, name, P0, Mz0, PT0, PTz0 = min(cand, key=lambda c: c[0])
        self.guess_info["used"] = name
        self._guess_C = C_L if name == "LEARNED" else C_S
        self._first_PT = (P0, PT0, PTz0)
        return P0, Mz0

    def read_guess(self):
        """Densities of a restart file (/SCF/DensityScalar, and DensityMz for UHF); the SCF's
        first iteration forms their Fock matrix (singleslater_guess.h:320-329)."""
        from . import restart
        if not self.restart_file:
            raise ValueError("the READ guess needs a restart file")
        P, Mz = restart.read_densities(self.restart_file, self.N, open_shell=self.ref != "RHF")
        t = lambda M: self.torch.from_numpy(np.ascontiguousarray(M)).to(self.dev)
        self.guess_info = {"used": "READ", "file": self.restart_file}
        return t(P), (t(Mz) if Mz is not None else None)

    # ------------------------------------------------------------------ SCF3
    def SCF3(self, P0=None, Mz0=None, reference_energy: Optional[float] = None, profile: bool = False) -> SCFResult:
        torch = self.torch
        t_start = time.perf_counter()
        self._first_PT = None
        self._guess_C = None
        if P0 is None:
            if self.guess == "SAD":
                P, Mz = self.sad_guess()
            elif self.guess == "LEARNED":
                P, Mz = self.learned_guess()
            elif self.guess == "CORE":
                P, Mz = self.core_guess()
                self.guess_info = {"used": "CORE"}
            elif self.guess == "READ":
                P, Mz = self.read_guess()
            else:
                raise ValueError(f"unknown guess {self.guess}")
        else:
            self.guess_info = {"used": "INPUT"}
            P, Mz = P0, Mz0
        first_PT, self._first_PT = self._first_PT, None
        guess_C, self._guess_C = self._guess_C, None
        torch.cuda.synchronize()
        guess_s = time.perf_counter() - t_start
        uhf = self.ref != "RHF"
        na, nb = self.nocc
        hist_F: List[tuple] = []
        hist_e: List[torch.Tensor] = []
        res = SCFResult(energy=0.0, iterations=0, converged=False, setup_s=self.setup_s, guess_s=guess_s,
                        setup_parts=dict(self.setup_parts), guess_info=dict(self.guess_info))
        E_old = 0.0
        last_dE = float("inf")
        # orthonormal-basis orbitals of the last density step (RHF); the guess's own to start
        C_orb = guess_C[1] if (guess_C is not None and guess_C[0] is P) else None
        Fo_prev = None
        P_prev = PT_prev = None
        B_hist = None
        damp = self.damp
        for it in range(self.max_iter):
            torch.cuda.synchronize()
            tg = time.perf_counter()
            # the difference-density build (and FP32 in it) only while the energy still moves by
            # more than 1e-7 Eh per step: its screening and rounding errors (~1e-9 in PT) would
            # hold the 1e-10 / 1e-8 criteria back near convergence
            if it == 0 and first_PT is not None and first_PT[0] is P:
                PT, PTz = first_PT[1], first_PT[2]
                kind = "guess"
            elif (self.incremental and not uhf and it % self.rebuild_every != 0
                    and abs(last_dE) > 1e-7):
                if self.precision == "mixed":
                    self.ao.setPrecision("mixed", self.mixed_threshold)
                dPT, _ = self.formPT(P - P_prev)
                self.ao.setPrecision("fp64")
                PT = PT_prev + dPT
                PTz = None
                kind = "incremental"
            else:
                PT, PTz = self.formPT(P, Mz if uhf else None)
                kind = "full"
            P_prev, PT_prev = P, PT
            torch.cuda.synchronize()
            res.gd_times.append(time.perf_counter() - tg)
            E = self._energy(P, PT, Mz, PTz)
            if profile:
                torch.cuda.synchronize()
                t_ph = time.perf_counter()
                q64, q32 = self.ao.precision_counts() if kind != "guess" else (0, 0)
                ph = {"gd": t_ph - tg, "quartets_fp64": q64, "quartets_fp32": q32, "build": kind}
            if uhf:
                Fa = self.H + 0.5 * (PT + PTz)
                Fb = self.H + 0.5 * (PT - PTz)
                Pa, Pb = 0.5 * (P + Mz), 0.5 * (P - Mz)
                errs = [self.X @ (Fa @ Pa @ self.S - self.S @ Pa @ Fa) @ self.X,
                        self.X @ (Fb @ Pb @ self.S - self.S @ Pb @ Fb) @ self.X]
                Fs = (Fa, Fb)
            else:
                F = self.H + 0.5 * PT
                errs = [self.X @ (F @ P @ self.S - self.S @ P @ F) @ self.X]
                Fs = (F,)
            if self.ndiis > 0:
                e_new = torch.cat([e.reshape(-1) for e in errs])
                hist_F.append(Fs)
                hist_e.append(e_new)
                # one new row of B = <e_i, e_j> per iteration, as a reduction over the N^2
                # entries (a k x N^2 x k GEMM maps to a single workgroup in rocBLAS)
                row = (torch.stack(hist_e) * e_new).sum(dim=1)
                m = 0 if B_hist is None else B_hist.shape[0]
                B_new = torch.empty((m + 1, m + 1), dtype=torch.float64, device=self.dev)
                if m:
                    B_new[:m, :m] = B_hist
                B_new[m, :] = row
                B_new[:, m] = row
                B_hist = B_new
                if len(hist_F) > self.ndiis:
                    hist_F.pop(0)
                    hist_e.pop(0)
                    B_hist = B_hist[1:, 1:]
                n = len(hist_F)
                if n > 1:
                    Bm = -torch.ones((n + 1, n + 1), dtype=torch.float64, device=self.dev)
                    Bm[:n, :n] = B_hist
                    Bm[n, n] = 0.0
                    rhs = torch.zeros(n + 1, dtype=torch.float64, device=self.dev)
                    rhs[n] = -1.0
                    # a singular system (vanishing commutators) keeps the plain step
                    # (singleslater_diis.h:60-64)
                    c, info = torch.linalg.solve_ex(Bm, rhs)
                    c = c[:n]
                    if int(info) == 0 and bool(torch.isfinite(c).all()):
                        Fs = tuple(sum(c[i] * hist_F[i][k] for i in range(n)) for k in range(len(Fs)))
            if profile:
                torch.cuda.synchronize()
                ph["energy_diis"] = time.perf_counter() - t_ph
                t_ph = time.perf_counter()
            if damp > 0.0 and Fo_prev is not None:
                Fs = tuple((1 - damp) * f + damp * fp for f, fp in zip(Fs, Fo_prev))
            Fo_prev = Fs
            if uhf:
                Pa_n, _ = self._diag_density(Fs[0], na)
                Pb_n, _ = self._diag_density(Fs[1], nb)
                Pn, Mzn = Pa_n + Pb_n, Pa_n - Pb_n
            else:
                out = None
                if self.pseudo_diag and C_orb is not None and (abs(last_dE) < self.pdiag_switch or
                                                               (it <= 1 and self.pdiag_first)):
                    out = self._pseudo_diag_density(Fs[0], na, C_orb)
                if out is None:
                    out = self._diag_density(Fs[0], na, with_vectors=True)
                Pa_n, C_orb = out
                Pn, Mzn = 2.0 * Pa_n, None
            if profile:
                torch.cuda.synchronize()
                ph["density"] = time.perf_counter() - t_ph
                t_ph = time.perf_counter()
            dP = float(torch.linalg.norm(Pn - P))
            if uhf:
                dP = max(dP, float(torch.linalg.norm(Mzn - Mz)))
            dE = E - E_old
            last_dE = dE
            E_old = E
            # damping off once |dE| < 1e-6, back on if the energy moves by more again
            # (singleslater_scfutils.h:262-280)
            if abs(dE) < 1e-6:
                damp = 0.0
            elif self.damp > 0.0 and damp <= 0.0:
                damp = self.damp
            P, Mz = Pn, Mzn
            res.energies.append(E)
            res.times.append(time.perf_counter() - t_start)
            if profile:
                ph["rest"] = time.perf_counter() - t_ph
                res.phase_times.append(ph)
            res.iterations = it + 1
            if it > 0 and abs(dE) < self.ene_tol and dP < self.den_tol:
                res.converged = True
                break
        if not res.converged:
            raise SCFNotConverged("SCF Failed to converge within MAXITER iterations")
        res.energy = res.energies[-1]
        res.P, res.Mz = P, Mz
        if self.restart_file:
            from . import restart
            restart.write_densities(self.restart_file, P.cpu().numpy(),
                                    Mz.cpu().numpy() if Mz is not None else None)
        return res


def time_to_energy(res: SCFResult, E_ref: float, tol: float = 1e-6) -> float:
    """Seconds from SCF start until |E - E_ref| <= tol holds for this and every later iteration."""
    t = None
    for E, ts in zip(res.energies, res.times):
        if abs(E - E_ref) <= tol:
            if t is None:
                t = ts
        else:
            t = None
    return float("nan") if t is None else t
