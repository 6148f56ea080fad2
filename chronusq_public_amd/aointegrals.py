"""Host-side mirror of ChronusQ's ``AOIntegrals`` for the direct G[D] path.

Same names and argument meaning as the reference operator interface
(include/aointegrals.h:468-502):

* ``COULOMB``, ``EXCHANGE`` (``ERI_CONTRACTION_TYPE``, include/aointegrals.h:468-472);
* ``newTwoEContract(X, AX, contractionList, scalingFactors)`` dispatches on
  ``integralAlgorithm`` (include/aointegrals.h:489-502) -- only ``DIRECT`` exists here; any
  other value raises, as the reference's ``CErr`` does;
* ``newTwoEContractDirect`` accumulates ``AX[i] += 2*scale*J[X]`` or ``2*scale*K[X]``
  (include/aointegrals/aointegrals_contract.h:24-194);
* ``computeSchwartz`` / ``schwartz`` (src/aointegrals/aointegrals_builders.cpp:469-518);
* ``thresholdSchwartz`` (include/aointegrals.h:334, default 1e-14).

Matrices may be numpy arrays (host; column-major is used as in Eigen, C-ordered arrays are
accepted as their transposes, which leaves J and K of any density consistent) or torch tensors
on the GPU (device-resident path, no host copies).  All arithmetic runs in libcqgpu.so on
the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import CQ_COULOMB as COULOMB, CQ_EXCHANGE as EXCHANGE, check

DIRECT = "DIRECT"
INCORE = "INCORE"


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class AOIntegrals:
    def __init__(self, basis, integral_algorithm: str = DIRECT):
        self.basis = basis
        self.integralAlgorithm = integral_algorithm
        L = _lib.load()
        l, pure, nprim, alpha, coef, origin = basis.flat()
        self._flat = (np.ascontiguousarray(l, np.int32), np.ascontiguousarray(pure, np.int32),
                      np.ascontiguousarray(nprim, np.int32), np.ascontiguousarray(alpha, np.float64),
                      np.ascontiguousarray(coef, np.float64), np.ascontiguousarray(origin, np.float64))
        f = self._flat
        h = ctypes.c_void_p()
        rc = L.cq_ctx_create(ctypes.byref(h), len(f[0]), _ip(f[0]), _ip(f[1]), _ip(f[2]), _dp(f[3]),
                             _dp(f[4]), _dp(f[5]))
        check(rc, None)
        self._h = h
        dims = [ctypes.c_int() for _ in range(5)]
        check(L.cq_get_dims(self._h, *[ctypes.byref(d) for d in dims]), self._h)
        self.nShell, self.nBasis, self.nCart, self.nShellPair, self.nPrimPair = (d.value for d in dims)

    @classmethod
    def batch(cls, bases: Sequence) -> "AOIntegrals":
        """One context for several molecules (cq_ctx_create_batch): their shells concatenated
        in order, shell pairs and quartets formed only within a molecule.  A contraction over
        the block-diagonal density of all molecules (functions in the same order) gives every
        molecule's G[D] in its diagonal block.  No reference counterpart."""
        self = cls.__new__(cls)
        self.basis = list(bases)
        self.integralAlgorithm = DIRECT
        parts = [b.flat() for b in self.basis]
        self._flat = tuple(np.ascontiguousarray(np.concatenate([p[i] for p in parts]),
                                                np.int32 if i < 3 else np.float64) for i in range(6))
        mol = np.ascontiguousarray(np.concatenate([np.full(len(p[0]), k, dtype=np.int32)
                                                   for k, p in enumerate(parts)]))
        self._mol = mol
        f = self._flat
        L = _lib.load()
        h = ctypes.c_void_p()
        check(L.cq_ctx_create_batch(ctypes.byref(h), len(f[0]), _ip(f[0]), _ip(f[1]), _ip(f[2]), _dp(f[3]),
                                    _dp(f[4]), _dp(f[5]), _ip(mol)), None)
        self._h = h
        dims = [ctypes.c_int() for _ in range(5)]
        check(L.cq_get_dims(self._h, *[ctypes.byref(d) for d in dims]), self._h)
        self.nShell, self.nBasis, self.nCart, self.nShellPair, self.nPrimPair = (d.value for d in dims)
        return self

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().cq_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ screening
    @property
    def thresholdSchwartz(self) -> float:
        v = ctypes.c_double()
        check(_lib.load().cq_get_schwarz_threshold(self._h, ctypes.byref(v)), self._h)
        return v.value

    @thresholdSchwartz.setter
    def thresholdSchwartz(self, tau: float):
        check(_lib.load().cq_set_schwarz_threshold(self._h, float(tau)), self._h)

    def setPrecision(self, mode: str = "fp64", threshold: float = 1e-5):
        """'fp64' (the reference's arithmetic) or 'mixed': quartets whose density-weighted
        Schwarz bound lies below `threshold` get FP32 integrals (FP64 Boys function and
        accumulation).  No reference counterpart."""
        code = {"fp64": _lib.CQ_PREC_FP64, "mixed": _lib.CQ_PREC_MIXED}.get(mode)
        if code is None:
            raise ValueError(f"unknown precision mode {mode!r}")
        check(_lib.load().cq_set_precision_mode(self._h, code, float(threshold)), self._h)

    DEFAULT_TILES = ((0, 0), (1, 256), (1, 512), (1, 1024), (2, 512), (4, 512), (4, 1024))

    def autotune(self, X, reps: int = 3, variants: Sequence[int] = (0, 1, 2),
                 tiles: Optional[Sequence[Tuple[int, int]]] = None, balance_streams: bool = False) -> dict:
        """Pick, per class pair, the fastest J/K digestion variant and tile for the RHF
        contraction of the device density X.  Variants: 0 K updates as FP64 atomics into HBM,
        1 ket-keyed groups with K columns in LDS, 2 also bra-keyed groups where the class pair
        allows them.  Tiles (max_group, union_bound) as in set_class_tile, (0, 0) being the
        built-in default.  Every (tile, variant) runs `reps` times with the class kernels
        serialised and timed by HIP events (cq_get_kernel_times + cq_get_screen_times: the
        screening is part of the cost); the fastest is set per class pair
        (cq_set_class_tile, cq_set_class_digestion).  With balance_streams the class pairs are
        then placed on the four streams longest first, each on the stream with the least
        kernel time so far (cq_set_class_stream); off by default, since the serialised times
        do not model the streams' overlap: on (H2O)16/cc-pVDZ the balanced placement ran
        G[D] in 4.17 ms against 4.04 ms for the built-in round robin.  Returns
        {class name: (variant, tile, {(tile, variant): mean ms})}.  No reference counterpart."""
        import torch
        L = _lib.load()
        nc = L.cq_num_class_pairs()
        tiles = list(tiles) if tiles is not None else list(self.DEFAULT_TILES)
        J = torch.zeros_like(X)
        K = torch.zeros_like(X)
        times = {}
        check(L.cq_set_launch_streams(self._h, 1), self._h)
        try:
            for tile in tiles:
                for k in range(nc):
                    check(L.cq_set_class_tile(self._h, k, 0, int(tile[0]), int(tile[1])), self._h)
                for v in variants:
                    for k in range(nc):
                        check(L.cq_set_class_digestion(self._h, k, int(v)), self._h)
                    self.newTwoEContractDirect([X, X], [J, K], [COULOMB, EXCHANGE], [1.0, -0.5])
                    check(L.cq_set_profiling(self._h, 1), self._h)
                    for _ in range(reps):
                        self.newTwoEContractDirect([X, X], [J, K], [COULOMB, EXCHANGE], [1.0, -0.5])
                    ms = (ctypes.c_double * nc)()
                    nl = (ctypes.c_longlong * nc)()
                    sms = (ctypes.c_double * nc)()
                    snl = (ctypes.c_longlong * nc)()
                    check(L.cq_get_kernel_times(self._h, ms, nl, nc), self._h)
                    check(L.cq_get_screen_times(self._h, sms, snl, nc), self._h)
                    check(L.cq_set_profiling(self._h, 0), self._h)
                    times[(tuple(tile), int(v))] = [(ms[k] / nl[k] if nl[k] else 0.0) +
                                                    (sms[k] / snl[k] if snl[k] else 0.0) for k in range(nc)]
        finally:
            check(L.cq_set_launch_streams(self._h, 4), self._h)
        names = list(self.quartet_counts().keys())
        out = {}
        chosen_ms = []
        for k in range(nc):
            (tile, v) = min(times, key=lambda tv: times[tv][k])
            check(L.cq_set_class_tile(self._h, k, 0, int(tile[0]), int(tile[1])), self._h)
            check(L.cq_set_class_digestion(self._h, k, v), self._h)
            out[names[k]] = (v, tile, {tv: times[tv][k] for tv in times})
            chosen_ms.append(times[(tile, v)][k])
        load = [0.0] * 4
        for k in sorted(range(nc), key=lambda k: -chosen_ms[k]):
            s = min(range(4), key=lambda i: load[i]) if balance_streams else -1
            check(L.cq_set_class_stream(self._h, k, s), self._h)
            if s >= 0:
                load[s] += chosen_ms[k]
        return out

    def set_class_tile(self, k: int, max_group: int = 0, union_bound: int = 0, lds_bytes: int = 0) -> None:
        """Tile of class pair k's digestion kernels (cq_set_class_tile): groups of up to
        max_group uniform pairs, task regions of up to union_bound swept pairs, a dynamic LDS
        budget; zeros select the defaults."""
        check(_lib.load().cq_set_class_tile(self._h, int(k), int(lds_bytes), int(max_group), int(union_bound)), self._h)

    def class_tile(self, k: int, variant: int = 1) -> Tuple[int, int, int, int]:
        """(max group, union bound, groups, kernel LDS bytes) planned for class pair k."""
        mg, ub, ng = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lds = ctypes.c_longlong()
        check(_lib.load().cq_get_class_tile(self._h, int(k), int(variant), ctypes.byref(mg), ctypes.byref(ub),
                                            ctypes.byref(ng), ctypes.byref(lds)), self._h)
        return mg.value, ub.value, ng.value, lds.value

    def class_digestion(self) -> List[int]:
        L = _lib.load()
        out = []
        for k in range(L.cq_num_class_pairs()):
            v = ctypes.c_int()
            check(L.cq_get_class_digestion(self._h, k, ctypes.byref(v)), self._h)
            out.append(v.value)
        return out

    def set_task_budget(self, entries: int) -> None:
        """Largest screened task list of one class pair, in 8-byte entries; class pairs with a
        larger worst case run in batches of uniform pairs (cq_set_task_budget)."""
        check(_lib.load().cq_set_task_budget(self._h, int(entries)), self._h)

    def task_plan(self) -> Tuple[int, int]:
        """(task entries allocated, largest number of batches of one class pair)."""
        e, b = ctypes.c_longlong(), ctypes.c_int()
        check(_lib.load().cq_get_task_plan(self._h, ctypes.byref(e), ctypes.byref(b)), self._h)
        return e.value, b.value

    def computeSchwartz(self) -> np.ndarray:
        Q = np.zeros((self.nShell, self.nShell))
        check(_lib.load().cq_get_schwarz(self._h, _dp(Q)), self._h)
        return Q

    @property
    def schwartz(self) -> np.ndarray:
        return self.computeSchwartz()

    # ------------------------------------------------------------------ contraction
    def newTwoEContract(self, X: Sequence, AX: Sequence, contractionList: Sequence[int],
                        scalingFactors: Sequence[float]) -> None:
        if self.integralAlgorithm == DIRECT:
            self.newTwoEContractDirect(X, AX, contractionList, scalingFactors)
        else:
            raise _lib.CQError(f"Integral algorithm {self.integralAlgorithm} is not available "
                               "(only DIRECT is implemented on the GPU)")

    def newTwoEContractDirect(self, X: Sequence, AX: Sequence, contractionList: Sequence[int],
                              scalingFactors: Sequence[float], stream=None) -> None:
        n = len(X)
        if not (len(AX) == n == len(contractionList) == len(scalingFactors)):
            raise ValueError("X, AX, contractionList and scalingFactors must have equal length")
        types = np.ascontiguousarray(contractionList, dtype=np.int32)
        scales = np.ascontiguousarray(scalingFactors, dtype=np.float64)
        L = _lib.load()
        if _is_torch_cuda(X[0]):
            import torch
            for t in list(X) + list(AX):
                if not (_is_torch_cuda(t) and t.dtype == torch.float64 and t.is_contiguous()):
                    raise TypeError("device matrices must be contiguous float64 CUDA tensors")
                if t.shape != (self.nBasis, self.nBasis):
                    raise ValueError("matrix shape does not match the basis")
            if stream is None:
                stream = torch.cuda.current_stream(X[0].device).cuda_stream
            xp = (ctypes.c_void_p * n)(*[t.data_ptr() for t in X])
            ap = (ctypes.c_void_p * n)(*[t.data_ptr() for t in AX])
            rc = L.cq_contract_direct_device(self._h, n, xp, ap, _ip(types), _dp(scales), self.nBasis,
                                             self.nBasis, ctypes.c_void_p(stream))
            check(rc, self._h)
            return
        xs: List[np.ndarray] = []
        for x in X:
            x = np.asarray(x)
            if x.dtype != np.float64 or x.shape != (self.nBasis, self.nBasis):
                raise ValueError("matrices must be float64 and match the basis dimension")
            xs.append(np.asfortranarray(x) if not x.flags.c_contiguous else x)
        # C-ordered arrays are passed as their transposes (see module docstring)
        for a in AX:
            if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not (a.flags.f_contiguous or a.flags.c_contiguous):
                raise ValueError("AX must be contiguous float64 numpy arrays")
        xp = (ctypes.POINTER(ctypes.c_double) * n)(*[_dp(x) for x in xs])
        ap = (ctypes.POINTER(ctypes.c_double) * n)(*[_dp(a) for a in AX])
        rc = L.cq_contract_direct(self._h, n, xp, ap, _ip(types), _dp(scales), self.nBasis, self.nBasis)
        check(rc, self._h)

    # ------------------------------------------------------------------ one-electron
    def computeAOOneE(self, mol, device=None):
        """S, T, V (torch float64 tensors on the GPU) -- AOIntegrals::computeAOOneE
        (src/aointegrals/aointegrals_builders.cpp:337-467).  Symmetric, so the column-major
        results are used as they are."""
        import torch
        dev = device or torch.device("cuda", torch.cuda.current_device())
        N = self.nBasis
        out = [torch.empty((N, N), dtype=torch.float64, device=dev) for _ in range(3)]
        Z = np.ascontiguousarray(mol.Z, dtype=np.float64)
        xyz = np.ascontiguousarray(mol.coords_bohr, dtype=np.float64)
        stream = torch.cuda.current_stream(dev).cuda_stream
        check(_lib.load().cq_one_electron(self._h, mol.natom, _dp(Z), _dp(xyz), ctypes.c_void_p(out[0].data_ptr()),
                                          ctypes.c_void_p(out[1].data_ptr()), ctypes.c_void_p(out[2].data_ptr()),
                                          N, ctypes.c_void_p(stream)), self._h)
        return tuple(out)

    def computeMultipole(self, order: int = 3, origin=None, device=None):
        """Electric multipole integrals <mu| x^i y^j z^k |nu> about `origin` (bohr; default the
        origin, as the reference's libint2 emultipole engines), 1 <= i + j + k <= order: a list
        of 3, 9 or 19 symmetric N x N float64 tensors on the GPU in the order x, y, z; xx, xy,
        xz, yy, yz, zz; xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz (cq_multipole).
        AOIntegrals::OneEDriver with emultipole1..3, src/aointegrals/aointegrals_builders.cpp:38-55."""
        import torch
        dev = device or torch.device("cuda", torch.cuda.current_device())
        N = self.nBasis
        nops = {1: 3, 2: 9, 3: 19}.get(order)
        if nops is None:
            raise ValueError("order must be 1, 2 or 3")
        out = torch.empty((nops, N, N), dtype=torch.float64, device=dev)
        C = None if origin is None else np.ascontiguousarray(origin, dtype=np.float64)
        stream = torch.cuda.current_stream(dev).cuda_stream
        check(_lib.load().cq_multipole(self._h, int(order), _dp(C) if C is not None else None,
                                       ctypes.c_void_p(out.data_ptr()), N, ctypes.c_void_p(stream)), self._h)
        return list(out.unbind(0))

    def computeAOOneE_batch(self, mols, device=None):
        """S, T, V of a batch context (AOIntegrals.batch) for the molecules in the same order:
        N x N tensors holding every molecule's own integrals in its diagonal block
        (cq_one_electron_batch).  No reference counterpart."""
        import torch
        dev = device or torch.device("cuda", torch.cuda.current_device())
        N = self.nBasis
        out = [torch.empty((N, N), dtype=torch.float64, device=dev) for _ in range(3)]
        Z = np.ascontiguousarray(np.concatenate([m.Z for m in mols]), dtype=np.float64)
        xyz = np.ascontiguousarray(np.concatenate([m.coords_bohr for m in mols]), dtype=np.float64)
        amol = np.ascontiguousarray(np.concatenate([np.full(m.natom, k, dtype=np.int32) for k, m in enumerate(mols)]))
        stream = torch.cuda.current_stream(dev).cuda_stream
        check(_lib.load().cq_one_electron_batch(self._h, len(Z), _dp(Z), _dp(xyz), _ip(amol),
                                                ctypes.c_void_p(out[0].data_ptr()), ctypes.c_void_p(out[1].data_ptr()),
                                                ctypes.c_void_p(out[2].data_ptr()), N, ctypes.c_void_p(stream)), self._h)
        return tuple(out)

    # ------------------------------------------------------------------ diagnostics
    def quartet_counts(self) -> dict:
        L = _lib.load()
        nc = L.cq_num_class_pairs()
        buf = (ctypes.c_longlong * nc)()
        nout = ctypes.c_int()
        check(L.cq_get_quartet_counts(self._h, buf, nc, ctypes.byref(nout)), self._h)
        out = {}
        for k in range(nc):
            ls = [ctypes.c_int() for _ in range(4)]
            L.cq_class_pair_info(k, *[ctypes.byref(x) for x in ls])
            name = "({}{}|{}{})".format(*["spdfg"[x.value] for x in ls])
            out[name] = int(buf[k])
        return out

    def precision_counts(self) -> Tuple[int, int]:
        """Surviving quartets of the last contraction with FP64 and with FP32 integrals."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        check(_lib.load().cq_get_precision_counts(self._h, ctypes.byref(a), ctypes.byref(b)), self._h)
        return a.value, b.value

    def quartets(self, shells4: np.ndarray) -> List[np.ndarray]:
        """Contracted quartets (s1 s2|s3 s4) in the shells' own basis (as libint returns them)."""
        s4 = np.ascontiguousarray(np.asarray(shells4, dtype=np.int32).reshape(-1, 4))
        sizes = [s.size for s in self.basis.shells]
        dims = [[sizes[i] for i in row] for row in s4]
        total = int(sum(np.prod(d) for d in dims))
        out = np.zeros(total)
        check(_lib.load().cq_debug_quartets(self._h, len(s4), _ip(s4), _dp(out)), self._h)
        res, off = [], 0
        for d in dims:
            n = int(np.prod(d))
            res.append(out[off:off + n].reshape(d))
            off += n
        return res


def _is_torch_cuda(x) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda
