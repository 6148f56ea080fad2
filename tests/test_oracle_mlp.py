"""Pins the network oracle (oracle/mlp.py) and the host side of the learned guess
(chronusq_public_amd/learned_guess.py: rotations, frames, features, block assembly).

The oracle's float64 form is checked by central finite differences, its bf16 rounding by known
answers; the rotation matrices by evaluating the library's own solid harmonics at points other
than those they were fitted on; the frames and features by rotating and translating molecules.
"""
import numpy as np
import pytest

from oracle import mlp as O
from chronusq_public_amd import basis as B
from chronusq_public_amd import learned_guess as LG
from chronusq_public_amd.molecule import Molecule, _quat_to_rot, alkane, water_cluster


def test_bf16_round_known_answers():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 0.0, -0.0,
                  np.inf, -np.inf, 3.4028235e38, 2 ** -130], dtype=np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0, -0.0, np.inf, -np.inf, np.inf,
                     2 ** -130], dtype=np.float32)
    got = O.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.isnan(O.bf16_round(np.array([np.nan], dtype=np.float32)))[0]
    # every bf16 value is a fixed point; neighbours differ by one unit in the 8th significant bit
    r = O.bf16_round(np.random.default_rng(0).normal(size=1000).astype(np.float32))
    assert np.array_equal(O.bf16_round(r), r)
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def _net(dims, seed=0):
    rng = np.random.default_rng(seed)
    W = [rng.normal(size=(a, b)) / np.sqrt(a) for a, b in zip(dims[:-1], dims[1:])]
    b = [rng.normal(size=b) * 0.1 for b in dims[1:]]
    return W, b


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    dims = [7, 11, 9, 5]
    W, b = _net(dims)
    X = rng.normal(size=(6, dims[0]))
    T = rng.normal(size=(6, dims[-1]))
    M = (rng.random(size=T.shape) > 0.3).astype(float)

    def loss():
        Y, _ = O.forward(X, W, b, bf16=False)
        return O.mse(Y, T, M)[0]

    Y, cache = O.forward(X, W, b, bf16=False)
    _, dY = O.mse(Y, T, M)
    gW, gb = O.backward(dY, W, cache, bf16=False)
    h = 1e-6
    for li in range(len(W)):
        for (i, j) in [(0, 0), (dims[li] - 1, dims[li + 1] - 1), (dims[li] // 2, 1)]:
            W[li][i, j] += h
            lp = loss()
            W[li][i, j] -= 2 * h
            lm = loss()
            W[li][i, j] += h
            assert abs((lp - lm) / (2 * h) - gW[li][i, j]) < 1e-7 * max(1.0, abs(gW[li][i, j]))
        for j in [0, dims[li + 1] - 1]:
            b[li][j] += h
            lp = loss()
            b[li][j] -= 2 * h
            lm = loss()
            b[li][j] += h
            assert abs((lp - lm) / (2 * h) - gb[li][j]) < 1e-7 * max(1.0, abs(gb[li][j]))


def test_mse_and_adam_closed_forms():
    Y = np.array([[1.0, 2.0], [3.0, 5.0]])
    T = np.array([[0.0, 2.0], [1.0, 1.0]])
    M = np.array([[1.0, 1.0], [1.0, 0.0]])
    loss, dY = O.mse(Y, T, M)
    assert loss == pytest.approx((1 + 0 + 4) / 3)
    assert np.allclose(dY, [[2 / 3, 0], [4 / 3, 0]])
    loss_w, _ = O.mse(Y, T, M, w=np.array([2.0, 0.5]))
    assert loss_w == pytest.approx((2 * 1 + 0.5 * 4) / 3)
    # the first Adam step moves every parameter by lr against the sign of its gradient
    p = np.array([1.0, -1.0, 0.5])
    g = np.array([0.3, -2.0, 1e-3])
    p1, m1, v1 = O.adam(p, g, np.zeros(3), np.zeros(3), step=1, lr=0.01, eps=0.0)
    assert np.allclose(p1, p - 0.01 * np.sign(g))


# ------------------------------------------------------------------------- host side of the guess
def _lib_or_skip():
    try:
        from chronusq_public_amd import _lib
        _lib.load()
    except OSError as e:
        pytest.fail(f"libcqgpu.so not built: {e}")


def test_d_rotation_rotates_the_library_harmonics():
    _lib_or_skip()
    T = LG.solid_harmonics(2)
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(50, 3))
    for _ in range(5):
        R = _quat_to_rot(*rng.random(3))
        U = LG.d_rotation(R)
        # phi_local_m(r) = Y_m(R r) = sum_n U[m, n] Y_n(r)
        lhs = LG._monomials(2, pts @ R.T) @ T.T
        rhs = (LG._monomials(2, pts) @ T.T) @ U.T
        assert np.abs(lhs - rhs).max() < 1e-12
        assert np.abs(U @ U.T - np.eye(5)).max() < 1e-12
    R1, R2 = _quat_to_rot(*rng.random(3)), _quat_to_rot(*rng.random(3))
    assert np.abs(LG.d_rotation(R1 @ R2) - LG.d_rotation(R1) @ LG.d_rotation(R2)).max() < 1e-12


@pytest.mark.parametrize("basis_name", ["6-31g*", "cc-pvdz"])
def test_frames_equivariant_features_invariant(basis_name):
    _lib_or_skip()
    m = water_cluster(5, 11, box_ang=6.0)
    rng = np.random.default_rng(5)
    Q = _quat_to_rot(*rng.random(3))
    m2 = Molecule(m.Z, m.coords_bohr @ Q.T + np.array([0.3, -1.0, 2.0]))
    g = LG.pair_geometry(m, B.build_basis(basis_name, m.Z, m.coords_bohr))
    g2 = LG.pair_geometry(m2, B.build_basis(basis_name, m2.Z, m2.coords_bohr))
    assert np.array_equal(g.pairs, g2.pairs)
    assert np.abs(g.feats - g2.feats).max() < 1e-6
    R = LG.pair_frames(m.coords_bohr, g.pairs)
    R2 = LG.pair_frames(m2.coords_bohr, g.pairs)
    assert np.abs(R2 - R @ Q.T).max() < 1e-12
    assert np.abs(np.einsum("pij,pkj->pik", R, R) - np.eye(3)).max() < 1e-12
    assert np.all(np.linalg.det(R) > 0)
    # pairs ordered by (Z, index), only within the cutoff
    Z = m.Z
    assert np.all((Z[g.pairs[:, 0]] < Z[g.pairs[:, 1]]) |
                  ((Z[g.pairs[:, 0]] == Z[g.pairs[:, 1]]) & (g.pairs[:, 0] < g.pairs[:, 1])))
    d = np.linalg.norm(m.coords_bohr[g.pairs[:, 0]] - m.coords_bohr[g.pairs[:, 1]], axis=1)
    assert d.max() < LG.R_CUT


def test_assembly_inverts_the_training_targets():
    """Local blocks made as the training targets are, turned back by the assembly, give the
    original density on every pair block and diagonal block (torch on the CPU)."""
    _lib_or_skip()
    import torch
    from chronusq_public_amd.train_guess import Dataset
    m = alkane(3)
    bs = B.build_basis("6-31g*", m.Z, m.coords_bohr)
    g = LG.pair_geometry(m, bs)
    rng = np.random.default_rng(7)
    A = rng.normal(size=(g.N, g.N))
    P = torch.from_numpy(A + A.T)
    Ps = torch.zeros_like(P)
    for f, c in zip(g.first, g.count):
        Ps[f:f + c, f:f + c] = P[f:f + c, f:f + c] * 0.5
    ds = Dataset(torch.device("cpu"))
    ds.add(g, P, Ps)
    part = ds.parts[0]
    # every pair of this molecule is within the cutoff: the pair blocks reproduce P off the diagonal
    assert len(g.pairs) == m.natom * (m.natom - 1) // 2
    Y = part["To"].double().clone()
    Pl = LG.assemble_density(Y, g, Ps, torch.device("cpu"))
    off = torch.ones_like(P, dtype=torch.bool)
    for f, c in zip(g.first, g.count):
        off[f:f + c, f:f + c] = False
    assert float((Pl - P)[off].abs().max()) < 1e-5
    # diagonal corrections: one pair's correction for atom i set to the local-frame target of that
    # atom's block, rotated back, restores it
    k = 0
    i = int(g.pairs[k, 0])
    f, c = int(g.first[i]), int(g.count[i])
    Ui = torch.from_numpy(g.Ui[k])
    T = torch.zeros(LG.MAXF, LG.MAXF, dtype=torch.float64)
    T[:c, :c] = (P - Ps)[f:f + c, f:f + c]
    Y2 = torch.zeros_like(Y)
    Y2[k, LG.BLK:2 * LG.BLK] = (Ui @ T @ Ui.T).reshape(-1)
    Pd = LG.assemble_density(Y2, g, Ps, torch.device("cpu"))
    assert float((Pd[f:f + c, f:f + c] - P[f:f + c, f:f + c]).abs().max()) < 1e-12
