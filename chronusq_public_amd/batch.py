"""Batched RHF SCF of many uneven molecules on one GPU.  No reference counterpart (the
reference runs one molecule per job, src/python/src/standardJobs.py:55-99).

G[D]: the molecules are split into groups of at most group_functions basis functions; each
group is one library context over all its molecules' shells (cq_ctx_create_batch: shell pairs
and quartets only within a molecule), and one contraction of the group's block-diagonal
density -- one screening and one ERI launch per class pair for every molecule of the group --
gives each molecule's G[D] in its diagonal block (merged=False: one context and one call per
molecule, spread over four streams).  Everything else of the SCF runs once for the whole
batch on matrices padded to the largest basis:

* overlap padded with the identity, core Hamiltonian and Fock with zeros, so the padding
  block decouples; the orthonormal Fock's padding diagonal is set to 1e3 Eh, above every
  real orbital, so padding orbitals stay virtual and the densities vanish there;
* Loewdin S^{-1/2} by batched coupled Newton-Schulz iterations (as SingleSlater._lowdin);
* SAD guess from the per-element atomic densities, one Fock build, batched eigensolve;
* CDIIS per molecule (depth ndiis) on batched error vectors and B matrices, batched solve;
* batched eigensolve while |dE| > pdiag_switch, then batched first-order orbital rotations
  with Newton-Schulz re-orthonormalisation (SingleSlater._pseudo_diag_density);
* convergence per molecule by the reference's rules (|dE| < 1e-10 and ||dP||_F < 1e-8,
  singleslater_scfutils.h:262-355); converged molecules drop out of the G[D] calls and keep
  their density.

The per-molecule result is the same fixed point as SingleSlater.SCF3 (energies agree to the
convergence threshold; tests/test_gpu_batch.py).
"""
from __future__ import annotations

import ctypes
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from . import basis as B
from .aointegrals import AOIntegrals, COULOMB, EXCHANGE
from .molecule import Molecule, alkane, scaled_box, water_cluster
from .singleslater import atomic_density

PAD_LEVEL = 1.0e3


@dataclass
class BatchResult:
    energies: np.ndarray                    # final energy per molecule
    iterations: np.ndarray                  # SCF iterations per molecule
    converged: np.ndarray
    seconds: np.ndarray                     # time from the batch's start to each molecule's convergence
    setup_s: float = 0.0
    wall_s: float = 0.0
    history: List[np.ndarray] = field(default_factory=list)   # energies per iteration (B,)
    times: List[float] = field(default_factory=list)          # seconds since the start, per iteration


class BatchSCF:
    def __init__(self, mols: Sequence[Molecule], basis_name: str, tau: float = 1e-12, ndiis: int = 10,
                 ene_tol: float = 1e-10, den_tol: float = 1e-8, max_iter: int = 128, device=None,
                 pdiag_switch: float = 1e-1, precision: str = "fp64", mixed_threshold: float = 1e-5,
                 incremental: bool = False, rebuild_every: int = 8, launch_streams: int = 4,
                 streams: int = 4, setup_threads: int = 8, profile: bool = False, merged: bool = True,
                 group_functions: int = 4096):
        import torch
        self.torch = torch
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.mols = list(mols)
        self.basis_name = basis_name
        self.tau, self.ndiis = tau, ndiis
        self.ene_tol, self.den_tol, self.max_iter = ene_tol, den_tol, max_iter
        self.pdiag_switch = pdiag_switch
        self.precision, self.mixed_threshold = precision, mixed_threshold
        self.incremental, self.rebuild_every = incremental, rebuild_every
        # profile: host-side phase times (enqueue of the G[D] calls, waiting for the GPU,
        # eigensolves, the rest), with synchronisations that cost a little throughput
        self.profile = profile
        # the molecules' G[D] calls are spread round-robin over `streams` streams (default
        # priority) so that different molecules' kernels overlap
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(max(1, streams))]
        self.phase = {"gd_enqueue": 0.0, "gd_wait": 0.0, "diag": 0.0, "other": 0.0}
        t0 = time.perf_counter()
        self.bases = [B.build_basis(basis_name, m.Z, m.coords_bohr) for m in self.mols]
        self.N = np.array([b.nbasis for b in self.bases])
        self.Nmax = int(self.N.max())
        nel = np.array([m.nelectron for m in self.mols])
        if np.any(nel % 2):
            raise ValueError("the batched SCF is closed-shell RHF: every molecule needs an even electron count")
        self.nocc = nel // 2
        self.enuc = torch.tensor([m.nuclear_repulsion() for m in self.mols], dtype=torch.float64, device=self.dev)
        Bn, Nm = len(self.mols), self.Nmax
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.S = torch.zeros((Bn, Nm, Nm), **f64)
        T = torch.zeros((Bn, Nm, Nm), **f64)
        V = torch.zeros((Bn, Nm, Nm), **f64)
        L = _lib.load()
        self.merged = merged
        self.groups = []
        self.ao = []
        if merged:
            # one batch context per group of molecules (G[D] and the one-electron integrals)
            self._make_groups(group_functions, setup_threads)
            for g in self.groups:
                Sg, Tg, Vg = g["ao"].computeAOOneE_batch([self.mols[k] for k in g["members"]], self.dev)
                for dst, src in ((self.S, Sg), (T, Tg), (V, Vg)):
                    dst.view(-1)[g["pad"]] = src.view(-1)[g["dense"]]
        else:
            def make(k):
                # contexts (pair lists, Schwarz bounds) and one-electron integrals per molecule; the
                # library calls release the interpreter lock, so threads overlap them
                torch.cuda.set_device(self.dev)
                m, bs = self.mols[k], self.bases[k]
                ao = AOIntegrals(bs)
                ao.thresholdSchwartz = tau
                _lib.check(L.cq_set_launch_streams(ao._h, launch_streams), ao._h)
                Z = np.ascontiguousarray(m.Z, dtype=np.float64)
                xyz = np.ascontiguousarray(m.coords_bohr, dtype=np.float64)
                _lib.check(L.cq_one_electron(ao._h, m.natom, Z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             ctypes.c_void_p(self.S[k].data_ptr()), ctypes.c_void_p(T[k].data_ptr()),
                                             ctypes.c_void_p(V[k].data_ptr()), Nm, None), ao._h)
                return ao

            torch.cuda.synchronize(self.dev)    # S, T, V zeroed before the contexts' streams write them
            with ThreadPoolExecutor(max_workers=setup_threads) as ex:
                self.ao = list(ex.map(make, range(Bn)))
        self.H = T + V
        self.real = torch.zeros((Bn, Nm), dtype=torch.bool, device=self.dev)
        for k, n in enumerate(self.N):
            self.real[k, :n] = True
        pad = (~self.real).double()
        self.S += torch.diag_embed(pad)
        self.pad_level = torch.diag_embed(pad * PAD_LEVEL)
        self.occ = (torch.arange(Nm, device=self.dev)[None, :] < torch.from_numpy(self.nocc).to(self.dev)[:, None])
        self.X = self._lowdin(self.S)
        torch.cuda.synchronize()
        self.setup_s = time.perf_counter() - t0

    def _make_groups(self, max_functions: int, threads: int = 8):
        """merged: one batch context (cq_ctx_create_batch) per group of molecules of at most
        max_functions functions, created on `threads` threads; each G[D] of a group is one
        contraction over its block-diagonal density, one launch per class pair for all its
        molecules."""
        torch = self.torch
        members, cur, tot = [], [], 0
        for k in range(len(self.mols)):
            if cur and tot + int(self.N[k]) > max_functions:
                members.append(cur)
                cur, tot = [], 0
            cur.append(k)
            tot += int(self.N[k])
        if cur:
            members.append(cur)
        Nm = self.Nmax

        def make(mem):
            torch.cuda.set_device(self.dev)
            ao = AOIntegrals.batch([self.bases[k] for k in mem])
            ao.thresholdSchwartz = self.tau
            return ao

        with ThreadPoolExecutor(max_workers=threads) as ex:
            aos = list(ex.map(make, members))
        for mem, ao in zip(members, aos):
            n = ao.nBasis
            dense, pad = [], []
            off = 0
            for k in mem:
                nk = int(self.N[k])
                r = np.arange(nk)
                dense.append(((off + r)[:, None] * n + (off + r)[None, :]).ravel())
                pad.append((k * Nm * Nm + r[:, None] * Nm + r[None, :]).ravel())
                off += nk
            f64 = dict(dtype=torch.float64, device=self.dev)
            self.groups.append(dict(
                ao=ao, members=mem, n=n,
                dense=torch.from_numpy(np.concatenate(dense)).to(self.dev),
                pad=torch.from_numpy(np.concatenate(pad)).to(self.dev),
                X=torch.zeros((n, n), **f64), J=torch.zeros((n, n), **f64), K=torch.zeros((n, n), **f64)))

    def _formPT_merged(self, X, active: np.ndarray, mixed: bool):
        """One contraction per group, the groups spread round-robin over the streams so that
        their kernels overlap; each group writes its molecules' diagonal blocks of the result."""
        torch = self.torch
        act = torch.from_numpy(active).to(self.dev).double()[:, None, None]
        Xm = (X * act).reshape(-1)          # converged molecules contribute nothing (screened out)
        out = torch.zeros_like(X)
        main = torch.cuda.current_stream(self.dev)
        ready = torch.cuda.Event()
        ready.record(main)
        used = set()
        for i, g in enumerate(self.groups):
            if not active[g["members"]].any():
                continue
            st = self.streams[i % len(self.streams)]
            if st not in used:
                st.wait_event(ready)
                used.add(st)
            ao = g["ao"]
            with torch.cuda.stream(st):
                g["X"].zero_()
                g["X"].view(-1)[g["dense"]] = Xm[g["pad"]]
                g["J"].zero_()
                g["K"].zero_()
                if mixed:
                    ao.setPrecision("mixed", self.mixed_threshold)
                try:
                    ao.newTwoEContractDirect([g["X"], g["X"]], [g["J"], g["K"]], [COULOMB, EXCHANGE], [1.0, -0.5],
                                             stream=st.cuda_stream)
                finally:
                    if mixed:
                        ao.setPrecision("fp64")
                out.view(-1)[g["pad"]] = (g["J"] + g["K"]).view(-1)[g["dense"]]
        for st in used:
            done = torch.cuda.Event()
            done.record(st)
            main.wait_event(done)
        return out

    # ---------------------------------------------------------------- pieces
    def _lowdin(self, S):
        torch = self.torch
        n = S.shape[-1]
        eye = torch.eye(n, dtype=S.dtype, device=S.device).expand_as(S)
        c = S.abs().sum(dim=-1).amax(dim=-1)[:, None, None]
        Y, Z = S / c, eye.clone()
        for it in range(60):
            Tm = 1.5 * eye - 0.5 * torch.bmm(Z, Y)
            Y, Z = torch.bmm(Y, Tm), torch.bmm(Tm, Z)
            if it % 6 == 5 and float((torch.bmm(Z, Y) - eye).abs().max()) < 1e-13:
                break
        if float((torch.bmm(Z, Y) - eye).abs().max()) > 1e-12:
            w, U = torch.linalg.eigh(S)
            return (U * w.rsqrt()[:, None, :]) @ U.transpose(1, 2)
        Z = 0.5 * (Z + Z.transpose(1, 2))
        return Z / c.sqrt()

    def formPT(self, P, active: np.ndarray, dP=None, mixed: bool = False):
        """PT = 2J[P] - K[P] for the active molecules (one library call each, no host waits);
        with dP the contraction is of dP (the incremental build)."""
        torch = self.torch
        L = _lib.load()
        X = P if dP is None else dP
        if self.merged:
            return self._formPT_merged(X, active, mixed)
        J = torch.zeros_like(P)
        K = torch.zeros_like(P)
        main = torch.cuda.current_stream(self.dev)
        ready = torch.cuda.Event()
        ready.record(main)
        for st in self.streams:
            st.wait_event(ready)
        types = (ctypes.c_int * 2)(_lib.CQ_COULOMB, _lib.CQ_EXCHANGE)
        scales = (ctypes.c_double * 2)(1.0, -0.5)
        for j, k in enumerate(np.nonzero(active)[0]):
            ao = self.ao[k]
            stream = ctypes.c_void_p(self.streams[j % len(self.streams)].cuda_stream)
            if mixed:
                _lib.check(L.cq_set_precision_mode(ao._h, _lib.CQ_PREC_MIXED, self.mixed_threshold), ao._h)
            xp = ctypes.c_void_p(X[k].data_ptr())
            dX = (ctypes.c_void_p * 2)(xp, xp)
            dAX = (ctypes.c_void_p * 2)(ctypes.c_void_p(J[k].data_ptr()), ctypes.c_void_p(K[k].data_ptr()))
            _lib.check(L.cq_contract_direct_device(ao._h, 2, dX, dAX, types, scales, int(self.N[k]), self.Nmax,
                                                    stream), ao._h)
            if mixed:
                _lib.check(L.cq_set_precision_mode(ao._h, _lib.CQ_PREC_FP64, self.mixed_threshold), ao._h)
        for st in self.streams:
            done = torch.cuda.Event()
            done.record(st)
            main.wait_event(done)
        return J + K

    def _tick(self):
        return time.perf_counter()

    def _tock(self, key, t, sync=False):
        if not self.profile:
            return t
        if sync:
            self.torch.cuda.synchronize()
        now = time.perf_counter()
        self.phase[key] += now - t
        return now

    def _energy(self, P, PT):
        return (P * self.H).sum(dim=(1, 2)) + 0.25 * (P * PT).sum(dim=(1, 2)) + self.enuc

    def _diag(self, F):
        torch = self.torch
        Fo = torch.bmm(torch.bmm(self.X, F), self.X) + self.pad_level
        e, C = torch.linalg.eigh(Fo)
        Cocc = torch.bmm(self.X, C * self.occ[:, None, :].double())
        P = torch.bmm(Cocc, Cocc.transpose(1, 2))
        return P + P.transpose(1, 2), C     # 2 C C^T, exactly symmetric

    def _pseudo_diag(self, F, C):
        """Batched first-order occupied-virtual rotation + Newton-Schulz (SingleSlater's
        _pseudo_diag_density), with the occupied/virtual split per molecule as masks."""
        torch = self.torch
        Fo = torch.bmm(torch.bmm(self.X, F), self.X) + self.pad_level
        FC = torch.bmm(Fo, C)
        eps = (C * FC).sum(dim=1)                                   # (B, N)
        occ = self.occ.double()
        vir = 1.0 - occ
        den = eps[:, None, :] - eps[:, :, None]                      # [b, a, i] = e_i - e_a
        ov = vir[:, :, None] * occ[:, None, :]                        # a virtual, i occupied
        if float((den.abs() + (1.0 - ov) * 1e30).min()) < 1e-3:
            return None
        kap = torch.bmm(C.transpose(1, 2), FC) * ov / torch.where(ov > 0, den, torch.ones_like(den))
        Cn = C + torch.bmm(C, kap) - torch.bmm(C, kap.transpose(1, 2))
        for _ in range(3):
            Cn = 1.5 * Cn - 0.5 * torch.bmm(Cn, torch.bmm(Cn.transpose(1, 2), Cn))
        Cocc = torch.bmm(self.X, Cn * occ[:, None, :])
        P = torch.bmm(Cocc, Cocc.transpose(1, 2))
        return P + P.transpose(1, 2), Cn

    def sad_density(self):
        torch = self.torch
        P = torch.zeros((len(self.mols), self.Nmax, self.Nmax), dtype=torch.float64, device=self.dev)
        for k, (m, bs) in enumerate(zip(self.mols, self.bases)):
            off = 0
            for iat, z in enumerate(m.Z):
                n = sum(s.size for s in bs.shells if s.atom == iat)
                P[k, off:off + n, off:off + n] = atomic_density(int(z), self.basis_name, self.dev, self.ndiis)
                off += n
        return P

    # ---------------------------------------------------------------- SCF
    def run(self) -> BatchResult:
        torch = self.torch
        Bn = len(self.mols)
        t_start = time.perf_counter()
        all_on = np.ones(Bn, dtype=bool)
        P, C = self._diag(self.H + 0.5 * self.formPT(self.sad_density(), all_on))
        active = all_on.copy()
        iters = np.zeros(Bn, dtype=np.int64)
        done_t = np.full(Bn, np.nan)
        E_old = torch.zeros(Bn, dtype=torch.float64, device=self.dev)
        last_dE = np.full(Bn, np.inf)
        hist_F: List = []
        hist_e: List = []
        Bh = None
        P_prev = PT_prev = None
        res = BatchResult(np.zeros(Bn), iters, np.zeros(Bn, dtype=bool), done_t)
        for it in range(self.max_iter):
            act_t = torch.from_numpy(active).to(self.dev)
            tp = self._tick()
            if (self.incremental and it % self.rebuild_every != 0 and P_prev is not None
                    and np.all(np.abs(last_dE[active]) > 1e-7)):
                PT = PT_prev + self.formPT(P, active, dP=P - P_prev, mixed=self.precision == "mixed")
            else:
                PT = self.formPT(P, active)
            tp = self._tock("gd_enqueue", tp)
            if self.profile:
                torch.cuda.synchronize()
                tp = self._tock("gd_wait", tp)
            # converged molecules keep their last Fock build
            if PT_prev is not None:
                PT = torch.where(act_t[:, None, None], PT, PT_prev)
            P_prev, PT_prev = P, PT
            E = self._energy(P, PT)
            F = self.H + 0.5 * PT
            err = torch.bmm(torch.bmm(self.X, torch.bmm(torch.bmm(F, P), self.S) - torch.bmm(torch.bmm(self.S, P), F)),
                            self.X)
            e_new = err.reshape(Bn, -1)
            hist_F.append(F)
            hist_e.append(e_new)
            row = torch.stack([(h * e_new).sum(dim=1) for h in hist_e], dim=1)     # (B, k)
            m = 0 if Bh is None else Bh.shape[1]
            Bnew = torch.empty((Bn, m + 1, m + 1), dtype=torch.float64, device=self.dev)
            if m:
                Bnew[:, :m, :m] = Bh
            Bnew[:, m, :] = row
            Bnew[:, :, m] = row
            Bh = Bnew
            if len(hist_F) > self.ndiis:
                hist_F.pop(0)
                hist_e.pop(0)
                Bh = Bh[:, 1:, 1:]
            n = len(hist_F)
            if n > 1:
                A = -torch.ones((Bn, n + 1, n + 1), dtype=torch.float64, device=self.dev)
                A[:, :n, :n] = Bh
                A[:, n, n] = 0.0
                rhs = torch.zeros((Bn, n + 1), dtype=torch.float64, device=self.dev)
                rhs[:, n] = -1.0
                c, info = torch.linalg.solve_ex(A, rhs)
                ok = (info == 0) & torch.isfinite(c).all(dim=1)
                c = torch.where(ok[:, None], c, torch.zeros_like(c))
                Fd = sum(c[:, i, None, None] * hist_F[i] for i in range(n))
                F = torch.where(ok[:, None, None], Fd, F)
            dE = (E - E_old).cpu().numpy()
            E_old = E
            out = None
            tp = self._tock("other", tp, sync=True)
            # the batched eigensolve costs little per molecule: the rotations start only once
            # every active molecule's energy moves by less than pdiag_switch
            if it > 0 and np.all(np.abs(dE[active]) < self.pdiag_switch):
                out = self._pseudo_diag(F, C)
            if out is None:
                out = self._diag(F)
            Pn, Cn = out
            tp = self._tock("diag", tp, sync=True)
            dP = torch.linalg.norm((Pn - P).reshape(Bn, -1), dim=1).cpu().numpy()
            keep = act_t[:, None, None]
            P = torch.where(keep, Pn, P)
            C = torch.where(keep, Cn, C)
            now = time.perf_counter() - t_start
            res.history.append(E.cpu().numpy())
            res.times.append(now)
            for k in np.nonzero(active)[0]:
                iters[k] = it + 1
                last_dE[k] = dE[k]
                if it > 0 and abs(dE[k]) < self.ene_tol and dP[k] < self.den_tol:
                    active[k] = False
                    done_t[k] = now
                    res.energies[k] = float(E[k])
                    res.converged[k] = True
            if not active.any():
                break
        for k in np.nonzero(active)[0]:
            res.energies[k] = float(E_old[k])
        res.setup_s = self.setup_s
        res.wall_s = time.perf_counter() - t_start
        return res

    def close(self):
        for ao in self.ao:
            ao.close()
        for g in self.groups:
            g["ao"].close()


def c4_molecules(seed: int = 0, n_alkane: int = 128, n_water: int = 128) -> List[Molecule]:
    """The batched workload: alkanes C_nH_{2n+2}, n uniform in [1, 12] (standard all-trans
    geometry), and (H2O)_m, m uniform in [1, 8], in a cube scaled to m (cluster seeds
    seed * 1000 + k); draws from MT19937(seed)."""
    rng = np.random.Generator(np.random.MT19937(seed))
    ns = rng.integers(1, 13, size=n_alkane)
    ms = rng.integers(1, 9, size=n_water)
    mols = [alkane(int(n)) for n in ns]
    mols += [water_cluster(int(m), seed * 1000 + k, box_ang=scaled_box(int(m))) for k, m in enumerate(ms)]
    return mols
