"""Learned initial guess: atom-pair density blocks predicted by a small BF16 network.

No reference counterpart (the reference offers SAD, CORE, RANDOM and READ guesses,
include/singleslater/singleslater_guess.h:24-329).  The guess density is

    P_L = spherically averaged SAD atomic blocks + sum over neighbours j of a predicted
          correction to atom i's block (diagonal blocks) + predicted pair blocks P_ij
          (off-diagonal, atom pairs closer than R_CUT; farther pairs stay zero),

after which, as for SAD, one Fock build and diagonalisation give the starting orbitals.  The
SAD guess stays the fallback (SingleSlater.learned_guess: of the two idempotent starting
densities the one with the lower energy is kept, and SAD is used for elements or bases the
network was not trained on).

Every prediction lives in the pair's own frame: z along the pair axis (from atom i to atom
j, pairs ordered by (Z, index)), x towards the off-axis atom nearest the pair's midpoint (the
perpendicular component of its position), y = z x x.  The frame turns with the molecule, so the
guess is invariant under rotations of the molecule.  The network maps the rotation-invariant
features

    [64 Gaussian-smeared distance bins | element-pair one-hot |
     radial environment of atom i (8 bins x element) | the same for atom j]

to three 14 x 14 blocks (padded and masked): P_ij, the correction to atom i's block and the
correction to atom j's block, in that frame.  Blocks turn back to the molecule's axes with the
shells' rotation matrices: s 1, p the frame rotation itself, pure d a 5 x 5 matrix fitted once
from the library's solid harmonics (cq_solid_harmonics).

The network is a ReLU multilayer perceptron with BF16 weights and activations and FP32
accumulation on the matrix cores (include/cqgpu_guess.h), FP32 master weights and Adam
moments, trained by train_guess.py on converged densities of this package's device SCF.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

ELEMENTS = (1, 6, 8)            # H, C, O
PAIR_TYPES = [(a, b) for i, a in enumerate(ELEMENTS) for b in ELEMENTS[i:]]
NBINS = 64
R_MIN, R_MAX = 0.5, 10.0        # bohr: centres of the distance bins
SIGMA = (R_MAX - R_MIN) / (NBINS - 1)
R_CUT = 9.0                     # bohr: pairs beyond keep a zero block
ENV_BINS = 8
ENV_C = np.linspace(1.0, 9.0, ENV_BINS)
ENV_W = ENV_C[1] - ENV_C[0]
FEATURES = NBINS + len(PAIR_TYPES) + 2 * len(ELEMENTS) * ENV_BINS
MAXF = 14                       # largest atom block (C and O in 6-31G* and cc-pVDZ)
BLK = MAXF * MAXF
OUT = 3 * BLK                   # P_ij, correction of block ii, correction of block jj
HIDDEN = 256
PERP_MIN = 0.5                  # bohr: smallest off-axis distance of a frame-defining atom


class GuessUnavailable(RuntimeError):
    """The learned guess does not cover this molecule or basis (the caller falls back to SAD)."""


# ----------------------------------------------------------------------------- rotations
_SH_CACHE: Dict[int, np.ndarray] = {}


def solid_harmonics(l: int) -> np.ndarray:
    """(2l+1) x ncart(l) coefficients of the library's pure shells (m = -l..l)."""
    if l not in _SH_CACHE:
        nc = (l + 1) * (l + 2) // 2
        out = np.zeros((2 * l + 1) * nc)
        _lib.check(_lib.load().cq_solid_harmonics(l, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        _SH_CACHE[l] = out.reshape(2 * l + 1, nc)
    return _SH_CACHE[l]


def _monomials(l: int, pts: np.ndarray) -> np.ndarray:
    """x^a y^b z^c over the last axis, Cartesian order x^l ... z^l."""
    cols = []
    for i in range(l + 1):
        for j in range(i + 1):
            ax, ay, az = l - i, i - j, j
            cols.append(pts[..., 0] ** ax * pts[..., 1] ** ay * pts[..., 2] ** az)
    return np.stack(cols, axis=-1)


_PTS = np.random.default_rng(1234).normal(size=(16, 3))


def d_rotation(R: np.ndarray) -> np.ndarray:
    """U (..., 5, 5) with phi_local = U phi_lab for a pure d shell when r_local = R r.

    With Y_m(r) the library's d harmonics, Y(R r) = D^T Y(r) for a 5 x 5 matrix D; it is
    solved exactly from 16 points (least squares on a consistent system) and U = D^T."""
    T = solid_harmonics(2)
    Y = _monomials(2, _PTS) @ T.T                  # (16, 5)
    pinv = np.linalg.pinv(Y)                       # (5, 16)
    rot = np.einsum("kc,...dc->...kd", _PTS, R)    # R r_k
    Yr = _monomials(2, rot) @ T.T                  # (..., 16, 5)
    D = pinv @ Yr
    return np.swapaxes(D, -1, -2)


def atom_rotations(shells: Sequence[Tuple[int, bool]], R: np.ndarray) -> np.ndarray:
    """Block-diagonal U (n, MAXF, MAXF) of one atom's shells for n frames R (n, 3, 3)."""
    n = R.shape[0]
    U = np.zeros((n, MAXF, MAXF))
    Dd = None
    off = 0
    for (l, pure) in shells:
        size = 1 if l == 0 else (3 if l == 1 else 5)
        if off + size > MAXF:
            raise GuessUnavailable(f"atom block exceeds {MAXF} functions")
        if l == 0:
            U[:, off, off] = 1.0
        elif l == 1 and not pure:
            U[:, off:off + 3, off:off + 3] = R
        elif l == 2 and pure:
            if Dd is None:
                Dd = d_rotation(R)
            U[:, off:off + 5, off:off + 5] = Dd
        else:
            raise GuessUnavailable(f"no shell rotation for l={l} pure={pure}")
        off += size
    return U


# ----------------------------------------------------------------------------- geometry
@dataclass
class PairGeometry:
    """Everything the network and the assembly need for one molecule in one basis."""
    pairs: np.ndarray        # (np, 2) atom indices, (Z, index) ordered
    feats: np.ndarray        # (np, FEATURES) float32
    Ui: np.ndarray           # (np, MAXF, MAXF) rotations of atom i's shells into the pair frame
    Uj: np.ndarray
    first: np.ndarray        # (natom,) first basis function of each atom
    count: np.ndarray        # (natom,) functions of each atom
    N: int


def atom_layout(basis, natom: int):
    first = np.full(natom, -1, dtype=np.int64)
    count = np.zeros(natom, dtype=np.int64)
    shells: List[List[Tuple[int, bool]]] = [[] for _ in range(natom)]
    off = 0
    for s in basis.shells:
        if first[s.atom] < 0:
            first[s.atom] = off
        count[s.atom] += s.size
        shells[s.atom].append((int(s.l), bool(s.pure)))
        off += s.size
    return first, count, shells


def pair_list(Z: np.ndarray, X: np.ndarray) -> np.ndarray:
    na = len(Z)
    if na < 2:
        return np.zeros((0, 2), dtype=np.int64)
    i, j = np.triu_indices(na, 1)
    r = np.linalg.norm(X[j] - X[i], axis=1)
    keep = r < R_CUT
    i, j = i[keep], j[keep]
    swap = Z[i] > Z[j]
    a = np.where(swap, j, i)
    b = np.where(swap, i, j)
    return np.stack([a, b], axis=1).astype(np.int64)


def pair_frames(X: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    """Rows x, y, z of every pair frame, (np, 3, 3)."""
    i, j = pairs[:, 0], pairs[:, 1]
    d = X[j] - X[i]
    e = d / np.linalg.norm(d, axis=1, keepdims=True)
    mid = 0.5 * (X[i] + X[j])
    v = X[None, :, :] - X[i][:, None, :]                       # (np, na, 3)
    par = np.einsum("pnc,pc->pn", v, e)
    perp = v - par[..., None] * e[:, None, :]
    pn = np.linalg.norm(perp, axis=2)
    dm = np.linalg.norm(X[None, :, :] - mid[:, None, :], axis=2)
    dm = np.where(pn < PERP_MIN, np.inf, dm)
    rows = np.arange(len(pairs))
    dm[rows, i] = np.inf
    dm[rows, j] = np.inf
    k = np.argmin(dm, axis=1)
    ok = np.isfinite(dm[rows, k])
    x = perp[rows, k] / np.where(ok, pn[rows, k], 1.0)[:, None]
    # collinear environments: a fixed helper axis (the block is then cylindrical up to noise)
    helper = np.eye(3)[np.argmin(np.abs(e), axis=1)]
    hx = helper - np.einsum("pc,pc->p", helper, e)[:, None] * e
    hx /= np.linalg.norm(hx, axis=1, keepdims=True)
    x = np.where(ok[:, None], x, hx)
    y = np.cross(e, x)
    return np.stack([x, y, e], axis=1)


def environment(Z: np.ndarray, X: np.ndarray) -> np.ndarray:
    """Radial environment of every atom: per element, ENV_BINS Gaussian bins of the distances
    to the other atoms of that element, with a cosine cutoff at R_CUT; (natom, 3 * ENV_BINS)."""
    na = len(Z)
    r = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2)
    fc = np.where(r < R_CUT, 0.5 * (np.cos(np.pi * r / R_CUT) + 1.0), 0.0)
    np.fill_diagonal(fc, 0.0)
    g = np.exp(-0.5 * ((r[..., None] - ENV_C) / ENV_W) ** 2) * fc[..., None]   # (na, na, bins)
    out = np.zeros((na, len(ELEMENTS), ENV_BINS))
    for k, z in enumerate(ELEMENTS):
        out[:, k, :] = g[:, Z == z, :].sum(axis=1)
    return out.reshape(na, -1)


def pair_features(Z: np.ndarray, X: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    i, j = pairs[:, 0], pairs[:, 1]
    r = np.linalg.norm(X[j] - X[i], axis=1)
    centres = np.linspace(R_MIN, R_MAX, NBINS)
    F = np.zeros((len(pairs), FEATURES), dtype=np.float32)
    F[:, :NBINS] = np.exp(-0.5 * ((r[:, None] - centres) / SIGMA) ** 2)
    types = {t: k for k, t in enumerate(PAIR_TYPES)}
    for p, (a, b) in enumerate(zip(Z[i], Z[j])):
        t = types.get((int(a), int(b)))
        if t is None:
            raise GuessUnavailable(f"element pair ({a}, {b}) outside the learned guess {ELEMENTS}")
        F[p, NBINS + t] = 1.0
    env = environment(Z, X)
    o = NBINS + len(PAIR_TYPES)
    w = len(ELEMENTS) * ENV_BINS
    F[:, o:o + w] = env[i]
    F[:, o + w:o + 2 * w] = env[j]
    return F


def pair_geometry(mol, basis) -> PairGeometry:
    Z = np.asarray(mol.Z, dtype=np.int64)
    if any(int(z) not in ELEMENTS for z in Z):
        raise GuessUnavailable(f"elements {sorted(set(int(z) for z in Z))} outside {ELEMENTS}")
    X = np.asarray(mol.coords_bohr, dtype=np.float64)
    first, count, shells = atom_layout(basis, len(Z))
    if count.max() > MAXF:
        raise GuessUnavailable(f"atom blocks of {int(count.max())} functions exceed {MAXF}")
    pairs = pair_list(Z, X)
    npair = len(pairs)
    Ui = np.zeros((npair, MAXF, MAXF))
    Uj = np.zeros((npair, MAXF, MAXF))
    if npair:
        R = pair_frames(X, pairs)
        # atoms of one element share their shell layout: rotate per element
        for z in ELEMENTS:
            atoms = np.nonzero(Z == z)[0]
            if len(atoms) == 0:
                continue
            sh = shells[atoms[0]]
            for side, U in ((0, Ui), (1, Uj)):
                sel = Z[pairs[:, side]] == z
                if sel.any():
                    U[sel] = atom_rotations(sh, R[sel])
        feats = pair_features(Z, X, pairs)
    else:
        feats = np.zeros((0, FEATURES), dtype=np.float32)
    return PairGeometry(pairs, feats, Ui, Uj, first, count, int(count.sum()))


def block_index(g: PairGeometry):
    """Flat positions (row-major N x N) and the padded-block entries that fill them:
    (off, off_t, sel_off) for the P_ij blocks and their transposes, (di, sel_i), (dj, sel_j)
    for the corrections of the diagonal blocks of atoms i and j."""
    a = np.arange(MAXF)
    i, j = g.pairs[:, 0], g.pairs[:, 1]
    fi, fj, ci, cj = g.first[i], g.first[j], g.count[i], g.count[j]
    rows = fi[:, None, None] + a[None, :, None]
    cols = fj[:, None, None] + a[None, None, :]
    m = (a[None, :, None] < ci[:, None, None]) & (a[None, None, :] < cj[:, None, None])
    sel_off = m.reshape(-1)
    off = (rows * g.N + cols).reshape(-1)[sel_off]
    off_t = (cols * g.N + rows).reshape(-1)[sel_off]
    out = [off, off_t, sel_off]
    for f, c in ((fi, ci), (fj, cj)):
        r = f[:, None, None] + a[None, :, None]
        cc = f[:, None, None] + a[None, None, :]
        mm = ((a[None, :, None] < c[:, None, None]) & (a[None, None, :] < c[:, None, None])).reshape(-1)
        out += [(r * g.N + cc).reshape(-1)[mm], mm]
    return tuple(out)


# ----------------------------------------------------------------------------- the network
def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _check(rc: int):
    if rc != 0:
        raise RuntimeError("libcqgpu network kernel: " + _lib.load().cq_mlp_last_error().decode())


class BF16MLP:
    """dims[0] -> ... -> dims[-1] with ReLU between layers, on the GPU (BF16 MFMA kernels)."""

    def __init__(self, dims: Sequence[int] = (FEATURES, HIDDEN, HIDDEN, OUT), seed: int = 0, device=None):
        import torch
        self.torch = torch
        self.dims = [int(d) for d in dims]
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        g = torch.Generator().manual_seed(seed)
        self.W, self.b = [], []
        for din, dout in zip(self.dims[:-1], self.dims[1:]):
            w = torch.randn(din, dout, generator=g, dtype=torch.float32) * math.sqrt(2.0 / din)
            self.W.append(w.to(self.dev).contiguous())
            self.b.append(torch.zeros(dout, dtype=torch.float32, device=self.dev))
        self.Wbf = [torch.empty(w.shape, dtype=torch.int16, device=self.dev) for w in self.W]
        self.refresh_bf16()
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.step_count = 0
        self._cache = None

    @property
    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def refresh_bf16(self):
        L = _lib.load()
        for w, wb in zip(self.W, self.Wbf):
            _check(L.cq_mlp_to_bf16(_ptr(w), _ptr(wb), w.numel(), self.stream))

    def to_bf16(self, x):
        if x.dtype != self.torch.float32 or not x.is_contiguous() or x.device != self.dev:
            raise TypeError("network inputs are contiguous float32 tensors on the network's device")
        xb = self.torch.empty(x.shape, dtype=self.torch.int16, device=self.dev)
        _check(_lib.load().cq_mlp_to_bf16(_ptr(x), _ptr(xb), x.numel(), self.stream))
        return xb

    def forward(self, X):
        """X (n, dims[0]) float32 on the device -> output (n, dims[-1]) float32.  Keeps the bf16
        layer inputs and the pre-activations for backward."""
        torch = self.torch
        L = _lib.load()
        X = X.contiguous()
        n = X.shape[0]
        h_bf = self.to_bf16(X)
        self._cache = []
        nl = len(self.W)
        Z = None
        for li in range(nl):
            din, dout = self.dims[li], self.dims[li + 1]
            Z = torch.empty((n, dout), dtype=torch.float32, device=self.dev)
            _check(L.cq_mlp_gemm_bf16(_ptr(h_bf), din, 0, _ptr(self.Wbf[li]), dout, 0, _ptr(Z), dout, None,
                                      n, dout, din, _ptr(self.b[li]), 0, 0, self.stream))
            self._cache.append((h_bf, Z))
            if li < nl - 1:
                h_bf = self.to_bf16(torch.relu(Z))
        return Z

    def backward(self, dY):
        """Parameter gradients for dL/dY (float32, shape of the last forward output)."""
        torch = self.torch
        L = _lib.load()
        n = dY.shape[0]
        nl = len(self.W)
        gW: List = [None] * nl
        gb: List = [None] * nl
        dZ = dY.contiguous()
        for li in range(nl - 1, -1, -1):
            din, dout = self.dims[li], self.dims[li + 1]
            h_bf = self._cache[li][0]
            dZbf = self.to_bf16(dZ)
            gW[li] = torch.empty((din, dout), dtype=torch.float32, device=self.dev)
            # dW = H^T dZ, H stored n x din
            _check(L.cq_mlp_gemm_bf16(_ptr(h_bf), din, 1, _ptr(dZbf), dout, 0, _ptr(gW[li]), dout, None,
                                      din, dout, n, None, 0, 0, self.stream))
            gb[li] = torch.empty(dout, dtype=torch.float32, device=self.dev)
            _check(L.cq_mlp_column_sum(_ptr(dZ), n, dout, _ptr(gb[li]), self.stream))
            if li > 0:
                # dH = dZ W^T, W stored din x dout
                dH = torch.empty((n, din), dtype=torch.float32, device=self.dev)
                _check(L.cq_mlp_gemm_bf16(_ptr(dZbf), dout, 0, _ptr(self.Wbf[li]), dout, 1, _ptr(dH), din, None,
                                          n, din, dout, None, 0, 0, self.stream))
                dZn = torch.empty_like(dH)
                _check(L.cq_mlp_relu_backward(_ptr(dH), _ptr(self._cache[li - 1][1]), _ptr(dZn), None,
                                              dH.numel(), self.stream))
                dZ = dZn
        return gW, gb

    def mse(self, Y, T, M, w=None):
        """Masked mean squared error (sum m w (y - t)^2 / sum m) and its gradient."""
        torch = self.torch
        n, k = Y.shape
        dY = torch.empty_like(Y)
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        norm = float(max(1.0, float(M.sum())))
        _check(_lib.load().cq_mlp_mse(_ptr(Y), _ptr(T.contiguous()), _ptr(M.contiguous()),
                                      None if w is None else _ptr(w), n, k, norm, _ptr(dY), _ptr(loss),
                                      self.stream))
        return loss, dY

    def adam_step(self, gW, gb, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        L = _lib.load()
        self.step_count += 1
        for w, wb, g, m, v in zip(self.W, self.Wbf, gW, self.mW, self.vW):
            _check(L.cq_mlp_adam(_ptr(w), _ptr(g), _ptr(m), _ptr(v), _ptr(wb), w.numel(), lr, b1, b2, eps,
                                 self.step_count, wd, self.stream))
        for b, g, m, v in zip(self.b, gb, self.mb, self.vb):
            _check(L.cq_mlp_adam(_ptr(b), _ptr(g), _ptr(m), _ptr(v), None, b.numel(), lr, b1, b2, eps,
                                 self.step_count, 0.0, self.stream))

    def state(self) -> dict:
        out = {"dims": np.array(self.dims, dtype=np.int64)}
        for i, (w, b) in enumerate(zip(self.W, self.b)):
            out[f"W{i}"] = w.detach().cpu().numpy()
            out[f"b{i}"] = b.detach().cpu().numpy()
        return out

    @classmethod
    def from_state(cls, st: dict, device=None) -> "BF16MLP":
        net = cls([int(x) for x in st["dims"]], device=device)
        for i in range(len(net.W)):
            net.W[i].copy_(net.torch.from_numpy(np.asarray(st[f"W{i}"], dtype=np.float32)))
            net.b[i].copy_(net.torch.from_numpy(np.asarray(st[f"b{i}"], dtype=np.float32)))
        net.refresh_bf16()
        return net


# ----------------------------------------------------------------------------- models on disk
MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "guess_data")
CONSTANTS = {"NBINS": NBINS, "R_MIN": R_MIN, "R_MAX": R_MAX, "R_CUT": R_CUT, "ENV_BINS": ENV_BINS,
             "MAXF": MAXF, "FEATURES": FEATURES, "OUT": OUT}


def model_path(basis_name: str) -> str:
    key = basis_name.strip().lower().replace("*", "s").replace("-", "")
    return os.path.join(MODEL_DIR, f"pair_blocks_{key}.npz")


def save_model(net: BF16MLP, basis_name: str, path: Optional[str] = None, **meta) -> str:
    path = path or model_path(basis_name)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    st = net.state()
    st["basis"] = np.array(basis_name.lower())
    st["elements"] = np.array(ELEMENTS, dtype=np.int64)
    for k, v in CONSTANTS.items():
        st["const_" + k] = np.array(v)
    for k, v in meta.items():
        st["meta_" + k] = np.asarray(v)
    np.savez(path, **st)
    return path


_MODELS: Dict[Tuple[str, str], BF16MLP] = {}


def load_model(basis_name: str, path: Optional[str] = None, device=None) -> BF16MLP:
    """The trained network for a basis (cached per path and device); GuessUnavailable if none."""
    import torch
    path = path or model_path(basis_name)
    dev = device or torch.device("cuda", torch.cuda.current_device())
    key = (path, str(dev))
    if key in _MODELS:
        return _MODELS[key]
    if not os.path.exists(path):
        raise GuessUnavailable(f"no trained guess network for basis {basis_name} ({path})")
    with np.load(path, allow_pickle=False) as z:
        st = {k: z[k] for k in z.files}
    if str(st["basis"]) != basis_name.lower():
        raise GuessUnavailable(f"network in {path} was trained for {st['basis']}, not {basis_name}")
    for k, v in CONSTANTS.items():
        if not np.allclose(st["const_" + k], v):
            raise GuessUnavailable(f"network in {path} uses {k}={st['const_' + k]}, this module {v}")
    net = BF16MLP.from_state(st, device=dev)
    _MODELS[key] = net
    return net


# ----------------------------------------------------------------------------- assembly
def _sym(Y):
    return 0.5 * (Y + Y.transpose(1, 2))


def assemble_density(Y, g: PairGeometry, P_atoms, dev):
    """P_L (N x N float64 tensor) from the network output Y (np, OUT) and the block-diagonal
    SAD density P_atoms (N x N tensor)."""
    import torch
    P = P_atoms.clone()
    if len(g.pairs) == 0:
        return P
    off, off_t, sel_off, di, sel_i, dj, sel_j = (torch.from_numpy(x).to(dev) for x in block_index(g))
    Ui = torch.from_numpy(g.Ui).to(dev)
    Uj = torch.from_numpy(g.Uj).to(dev)
    Y = Y.double()
    n = Y.shape[0]
    Yo = Y[:, :BLK].reshape(n, MAXF, MAXF)
    Yi = _sym(Y[:, BLK:2 * BLK].reshape(n, MAXF, MAXF))
    Yj = _sym(Y[:, 2 * BLK:].reshape(n, MAXF, MAXF))
    UiT = Ui.transpose(1, 2).contiguous()
    UjT = Uj.transpose(1, 2).contiguous()
    lab = torch.bmm(torch.bmm(UiT, Yo), Uj).reshape(-1)[sel_off]
    flat = P.view(-1)
    flat[off] = lab
    flat[off_t] = lab
    flat.index_add_(0, di, torch.bmm(torch.bmm(UiT, Yi), Ui).reshape(-1)[sel_i])
    flat.index_add_(0, dj, torch.bmm(torch.bmm(UjT, Yj), Uj).reshape(-1)[sel_j])
    # U^T Y U rounds differently in its two triangles: make P exactly symmetric (the G[D]
    # contraction takes its slower non-symmetric path for any asymmetry)
    return 0.5 * (P + P.T)


def predict_density(net: BF16MLP, mol, basis, P_atoms):
    """The learned guess density P_L (before the Fock build) for one molecule."""
    import torch
    g = pair_geometry(mol, basis)
    if g.N != P_atoms.shape[0]:
        raise ValueError("atomic density does not match the basis")
    if len(g.pairs) == 0:
        return P_atoms.clone()
    X = torch.from_numpy(g.feats).to(net.dev)
    Y = net.forward(X)
    return assemble_density(Y, g, P_atoms, net.dev)
