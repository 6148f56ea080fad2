"""GPU parity: the HIP G[D] path (through the C ABI) against the CPU oracle.

Bar: 1e-10 absolute in every G[D] element (north_star), identical screening decisions
(surviving-quartet counts equal the oracle's for the same Schwarz matrix), Schwarz bounds
and individual contracted quartets to 1e-12.
"""
import numpy as np
import pytest

import oracle
from chronusq_public_amd import basis as B
from chronusq_public_amd import molecule as M
from chronusq_public_amd.aointegrals import AOIntegrals, COULOMB, EXCHANGE

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _setup(mol, name):
    bs = B.build_basis(name, mol.Z, mol.coords_bohr)
    return bs, AOIntegrals(bs), oracle.OracleBasis(bs)


@pytest.fixture(scope="module")
def water631():
    return _setup(M.water_test0001(), "6-31g*")


@pytest.fixture(scope="module")
def waterdz():
    return _setup(M.water_test0001(), "cc-pvdz")


def _sym_random(N, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (N, N))
    return np.asfortranarray(0.5 * (X + X.T))


def _gpu_jk(ao, Xs, types, scales, tau):
    ao.thresholdSchwartz = tau
    AX = [np.zeros((ao.nBasis, ao.nBasis), order="F") for _ in Xs]
    ao.newTwoEContract(Xs, AX, types, scales)
    return AX


def _oracle_jk(ob, Q, Xs, types, scales, tau):
    AX = [np.zeros((ob.nbf, ob.nbf), order="F") for _ in Xs]
    n = ob.contract_direct(Q, tau, Xs, AX, types, scales)
    return AX, n


@pytest.mark.parametrize("which", ["water631", "waterdz"])
def test_schwarz_matches_oracle(request, which):
    bs, ao, ob = request.getfixturevalue(which)
    Qg = ao.computeSchwartz()
    Qo = ob.schwarz()
    assert np.abs(Qg - Qg.T).max() == 0.0
    assert np.abs(Qg - Qo).max() <= 1e-12 * max(1.0, np.abs(Qo).max())


@pytest.mark.parametrize("which", ["water631", "waterdz"])
def test_every_quartet_matches_oracle(request, which):
    bs, ao, ob = request.getfixturevalue(which)
    ns = bs.nshell
    quads = [(s1, s2, s3, s4) for s1 in range(ns) for s2 in range(s1 + 1)
             for s3 in range(s1 + 1) for s4 in range((s2 if s1 == s3 else s3) + 1)]
    # and some non-canonical orders
    rng = np.random.default_rng(7)
    quads += [tuple(int(x) for x in rng.integers(0, ns, 4)) for _ in range(100)]
    got = ao.quartets(np.array(quads))
    worst = 0.0
    for q, g in zip(quads, got):
        worst = max(worst, np.abs(g - ob.quartet(*q)).max())
    assert worst < 1e-12, worst


@pytest.mark.parametrize("which", ["water631", "waterdz"])
@pytest.mark.parametrize("tau", [0.0, 1e-12, 1e-14])
def test_rhf_gd_random_density(request, which, tau):
    bs, ao, ob = request.getfixturevalue(which)
    X = _sym_random(bs.nbasis, 17)
    Qg = ao.computeSchwartz()
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], tau)
    (Jo, Ko), n = _oracle_jk(ob, Qg, [X, X], [0, 1], [1.0, -0.5], tau)
    assert np.abs(J - Jo).max() < TOL
    assert np.abs(K - Ko).max() < TOL
    assert sum(ao.quartet_counts().values()) == n


def test_nonsymmetric_density_matches_reference_scatter(water631):
    """A non-symmetric X reproduces the reference's direct scatter, not the in-core one."""
    bs, ao, ob = water631
    rng = np.random.default_rng(5)
    X = np.asfortranarray(rng.uniform(-1, 1, (bs.nbasis, bs.nbasis)))
    Qg = ao.computeSchwartz()
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [0.7, -0.3], 1e-12)
    (Jo, Ko), _ = _oracle_jk(ob, Qg, [X, X], [0, 1], [0.7, -0.3], 1e-12)
    assert np.abs(J - Jo).max() < TOL
    assert np.abs(K - Ko).max() < TOL
    assert np.abs(J - J.T).max() > 1e-6   # the reference's Coulomb scatter is order dependent


@pytest.mark.parametrize("eps", [2e-13, 1e-9])
def test_rounding_level_asymmetry_digests_as_symmetric(water631, eps):
    """X = S + eps*A (A antisymmetric, |X| ~ 1).  Below 1e-12 of max|X| the asymmetry is
    GEMM-rounding sized and X is digested as symmetric: J comes out symmetric to rounding,
    where the non-symmetric scatter would leave J - J^T ~ eps |J|.  Above it the reference's
    non-symmetric scatter runs.  Either way J and K match the oracle, which applies that
    scatter to X as given."""
    bs, ao, ob = water631
    rng = np.random.default_rng(9)
    S = rng.uniform(-1, 1, (bs.nbasis, bs.nbasis))
    A = rng.uniform(-1, 1, (bs.nbasis, bs.nbasis))
    X = np.asfortranarray(0.5 * (S + S.T) + eps * 0.5 * (A - A.T))
    Qg = ao.computeSchwartz()
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
    (Jo, Ko), _ = _oracle_jk(ob, Qg, [X, X], [0, 1], [1.0, -0.5], 1e-12)
    assert np.abs(J - Jo).max() < TOL
    assert np.abs(K - Ko).max() < TOL
    asym = np.abs(J - J.T).max() / np.abs(J).max()
    if eps < 1e-12:
        assert asym < 2e-15, asym
    else:
        assert asym > 1e-12, asym


def test_uhf_three_matrices_and_accumulate(waterdz):
    """SAD-style UHF call (singleslater_fock.h:49-55): {P,P,Mz}, {J,K,K}; AX accumulates."""
    bs, ao, ob = waterdz
    N = bs.nbasis
    P = _sym_random(N, 1)
    Mz = _sym_random(N, 2) * 0.1
    Qg = ao.computeSchwartz()
    ao.thresholdSchwartz = 1e-12
    AX = [np.asfortranarray(np.full((N, N), 0.25)) for _ in range(3)]
    ao.newTwoEContract([P, P, Mz], AX, [COULOMB, EXCHANGE, EXCHANGE], [1.0, -0.5, -0.5])
    AXo = [np.asfortranarray(np.full((N, N), 0.25)) for _ in range(3)]
    ob.contract_direct(Qg, 1e-12, [P, P, Mz], AXo, [0, 1, 1], [1.0, -0.5, -0.5])
    for a, b in zip(AX, AXo):
        assert np.abs(a - b).max() < TOL


def test_zero_scale_matrix_still_screens(water631):
    bs, ao, ob = water631
    N = bs.nbasis
    small = _sym_random(N, 3) * 1e-9
    big = _sym_random(N, 4)
    Qg = ao.computeSchwartz()
    AX = _gpu_jk(ao, [small, big], [COULOMB, EXCHANGE], [1.0, 0.0], 1e-10)
    AXo, n = _oracle_jk(ob, Qg, [small, big], [0, 1], [1.0, 0.0], 1e-10)
    assert np.abs(AX[0] - AXo[0]).max() < TOL
    assert np.abs(AX[1]).max() == 0.0
    assert sum(ao.quartet_counts().values()) == n


def test_device_tensor_path_matches_host(waterdz):
    import torch
    bs, ao, ob = waterdz
    N = bs.nbasis
    X = _sym_random(N, 9)
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
    dX = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    dJ = torch.zeros((N, N), dtype=torch.float64, device="cuda")
    dK = torch.zeros((N, N), dtype=torch.float64, device="cuda")
    ao.newTwoEContractDirect([dX, dX], [dJ, dK], [COULOMB, EXCHANGE], [1.0, -0.5])
    torch.cuda.synchronize()
    assert np.abs(dJ.cpu().numpy() - J).max() < 1e-12
    assert np.abs(dK.cpu().numpy() - K).max() < 1e-12


def test_errors_are_reported(water631):
    from chronusq_public_amd._lib import CQError
    bs, ao, ob = water631
    X = _sym_random(bs.nbasis, 1)
    with pytest.raises(CQError):
        ao.newTwoEContract([X], [np.zeros_like(X, order="F")], [2], [1.0])
    with pytest.raises(CQError):
        ao.thresholdSchwartz = -1.0


def test_water_tetramer_ccpvdz():
    """(H2O)4 / cc-pVDZ (96 functions, 48 shells): every class pair exercised."""
    mol = M.water_cluster(4, seed=3, box_ang=M.scaled_box(4))
    bs, ao, ob = _setup(mol, "cc-pvdz")
    X = _sym_random(bs.nbasis, 21)
    Qg = ao.computeSchwartz()
    Qo = ob.schwarz()
    assert np.abs(Qg - Qo).max() <= 1e-12 * max(1.0, np.abs(Qo).max())
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
    (Jo, Ko), n = _oracle_jk(ob, Qg, [X, X], [0, 1], [1.0, -0.5], 1e-12)
    assert np.abs(J - Jo).max() < TOL
    assert np.abs(K - Ko).max() < TOL
    assert sum(ao.quartet_counts().values()) == n


def test_full_size_water16_properties():
    """(H2O)16 / cc-pVDZ (384 functions): screening decisions bit-identical to the oracle,
    sampled quartets against the oracle, J/K symmetry and linearity at tau = 0."""
    mol = M.water_cluster(16, seed=0)
    bs, ao, ob = _setup(mol, "cc-pvdz")
    assert bs.nbasis == 384 and bs.nshell == 192
    N = bs.nbasis
    X1 = _sym_random(N, 1)
    X2 = _sym_random(N, 2)
    Qg = ao.computeSchwartz()
    J, K = _gpu_jk(ao, [X1, X1], [COULOMB, EXCHANGE], [1.0, 1.0], 1e-12)
    assert sum(ao.quartet_counts().values()) == ob.count_screened(Qg, 1e-12, [X1])
    assert np.abs(J - J.T).max() < 1e-10 and np.abs(K - K.T).max() < 1e-10
    # linearity with screening off
    Ja = _gpu_jk(ao, [X1], [COULOMB], [1.0], 0.0)[0]
    Jb = _gpu_jk(ao, [X2], [COULOMB], [1.0], 0.0)[0]
    Jab = _gpu_jk(ao, [np.asfortranarray(X1 + X2)], [COULOMB], [1.0], 0.0)[0]
    assert np.abs(Jab - Ja - Jb).max() < 1e-9 * max(1.0, np.abs(Jab).max())
    rng = np.random.default_rng(0)
    quads = [tuple(int(x) for x in rng.integers(0, bs.nshell, 4)) for _ in range(300)]
    got = ao.quartets(np.array(quads))
    for q, g in zip(quads, got):
        assert np.abs(g - ob.quartet(*q)).max() < 1e-12


@pytest.mark.parametrize("lds", [0, 1, 2])
def test_every_digestion_mode_matches_oracle(lds):
    """HBM-atomic, ket-keyed and bra-keyed LDS digestion against the oracle ((H2O)8 / cc-pVDZ:
    bra-keyed (dd|ss) segments hold more than one screening buffer of kets)."""
    from chronusq_public_amd import _lib
    mol = M.water_cluster(8, seed=11, box_ang=M.scaled_box(8))
    bs, ao, ob = _setup(mol, "cc-pvdz")
    _lib.check(_lib.load().cq_set_lds_digestion(ao._h, lds), ao._h)
    X = _sym_random(bs.nbasis, 5)
    Qg = ao.computeSchwartz()
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
    (Jo, Ko), n = _oracle_jk(ob, Qg, [X, X], [0, 1], [1.0, -0.5], 1e-12)
    assert np.abs(J - Jo).max() < TOL and np.abs(K - Ko).max() < TOL
    assert sum(ao.quartet_counts().values()) == n


def test_per_class_digestion_variants_match_oracle():
    """Class pairs on different digestion variants in one call (task tables 0 and 1 side by
    side) and the autotuner's choice: J and K against the oracle ((H2O)8 / cc-pVDZ)."""
    from chronusq_public_amd import _lib
    import torch
    L = _lib.load()
    mol = M.water_cluster(8, seed=11, box_ang=M.scaled_box(8))
    bs, ao, ob = _setup(mol, "cc-pvdz")
    X = _sym_random(bs.nbasis, 7)
    Qg = ao.computeSchwartz()
    (Jo, Ko), n = _oracle_jk(ob, Qg, [X, X], [0, 1], [1.0, -0.5], 1e-12)
    nc = L.cq_num_class_pairs()
    for k in range(nc):
        _lib.check(L.cq_set_class_digestion(ao._h, k, (2, 0, 1)[k % 3]), ao._h)
    assert ao.class_digestion() == [(2, 0, 1)[k % 3] for k in range(nc)]
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
    assert np.abs(J - Jo).max() < TOL and np.abs(K - Ko).max() < TOL
    assert sum(ao.quartet_counts().values()) == n
    choice = ao.autotune(torch.from_numpy(np.ascontiguousarray(X)).cuda(), reps=1)
    assert len(choice) == nc and all(v in (0, 1, 2) for v, _, _ in choice.values())
    assert ao.class_digestion() == [v for v, _, _ in choice.values()]
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
    assert np.abs(J - Jo).max() < TOL and np.abs(K - Ko).max() < TOL
    with pytest.raises(_lib.CQError):
        _lib.check(L.cq_set_class_digestion(ao._h, nc, 0), ao._h)


def test_mixed_precision_window_matches_fp64():
    """Mixed mode: quartets whose bound lies in [tau, 1e-5) get FP32 integrals.  On a
    difference-density-sized X ((H2O)4 / cc-pVDZ, |X| ~ 1e-3) most quartets fall in the FP32
    window; the screening decisions are unchanged (same surviving-quartet count) and G agrees
    with the FP64 result within the FP32 rounding of the window's contributions (each below
    1e-5, relative error ~1e-7): bound 1e-9 + 1e-6 max|G|."""
    mol = M.water_cluster(4, seed=3, box_ang=M.scaled_box(4))
    bs, ao, ob = _setup(mol, "cc-pvdz")
    X = np.asfortranarray(_sym_random(bs.nbasis, 31) * 1e-3)
    J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
    n64 = sum(ao.quartet_counts().values())
    ao.setPrecision("mixed", 1e-5)
    try:
        Jm, Km = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
        nmix = sum(ao.quartet_counts().values())
    finally:
        ao.setPrecision("fp64")
    assert nmix == n64
    for a, b in ((J, Jm), (K, Km)):
        err = np.abs(a - b).max()
        assert err < 1e-9 + 1e-6 * np.abs(a).max(), err
        assert err > 0.0   # the FP32 window was exercised
    # FP64 again after switching back: bitwise the first result
    J2, K2 = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
    assert np.abs(J2 - J).max() < 1e-13 and np.abs(K2 - K).max() < 1e-13


def test_precision_mode_errors():
    from chronusq_public_amd._lib import CQError
    bs, ao, ob = _setup(M.water_test0001(), "6-31g*")
    with pytest.raises(ValueError):
        ao.setPrecision("half")
    with pytest.raises(CQError):
        ao.setPrecision("mixed", 0.0)


def test_large_basis_lds_segments_match_hbm_digestion():
    """(H2O)32 / cc-pVDZ (768 functions, 800 Cartesian): LDS accumulators of up to 12 columns x
    800 rows (beyond 64 KiB for the d classes, within the device's per-workgroup limit) give the
    same J and K as the HBM-atomic digestion."""
    import torch
    from chronusq_public_amd import _lib
    mol = M.water_cluster(32, seed=0, box_ang=M.scaled_box(32))
    bs = B.build_basis("cc-pvdz", mol.Z, mol.coords_bohr)
    X = torch.from_numpy(np.ascontiguousarray(_sym_random(bs.nbasis, 9))).cuda()
    out = []
    for mode in (2, 0):
        ao = AOIntegrals(bs)
        ao.thresholdSchwartz = 1e-12
        _lib.check(_lib.load().cq_set_lds_digestion(ao._h, mode), ao._h)
        J = torch.zeros_like(X)
        K = torch.zeros_like(X)
        ao.newTwoEContractDirect([X, X], [J, K], [COULOMB, EXCHANGE], [1.0, -0.5])
        out.append((J, K, sum(ao.quartet_counts().values())))
    (J2, K2, n2), (J0, K0, n0) = out
    assert n2 == n0
    assert float((J2 - J0).abs().max()) < 1e-9 * max(1.0, float(J0.abs().max()))
    assert float((K2 - K0).abs().max()) < 1e-9 * max(1.0, float(K0.abs().max()))


@pytest.mark.parametrize("max_group,union_bound,lds", [(1, 64, 0), (4, 256, 0), (16, 1024, 160 * 1024), (8, 128, 24 * 1024)])
def test_group_tiles_match_oracle(max_group, union_bound, lds):
    """Groups of up to 16 uniform pairs sharing a shell, task regions of 64 .. 1024 swept pairs
    (J collected in LDS slots across a group), in every digestion variant: J and K against the
    oracle with identical surviving-quartet counts ((H2O)8 / cc-pVDZ)."""
    from chronusq_public_amd import _lib
    L = _lib.load()
    mol = M.water_cluster(8, seed=11, box_ang=M.scaled_box(8))
    bs, ao, ob = _setup(mol, "cc-pvdz")
    X = _sym_random(bs.nbasis, 13)
    Qg = ao.computeSchwartz()
    (Jo, Ko), n = _oracle_jk(ob, Qg, [X, X], [0, 1], [1.0, -0.5], 1e-12)
    nc = L.cq_num_class_pairs()
    for k in range(nc):
        ao.set_class_tile(k, max_group, union_bound, lds)
    assert max(ao.class_tile(k)[0] for k in range(nc)) > 1 or max_group == 1
    for lds_mode in (2, 1, 0):
        _lib.check(L.cq_set_lds_digestion(ao._h, lds_mode), ao._h)
        J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
        assert np.abs(J - Jo).max() < TOL and np.abs(K - Ko).max() < TOL, (lds_mode, max_group, union_bound)
        assert sum(ao.quartet_counts().values()) == n
    with pytest.raises(_lib.CQError):
        ao.set_class_tile(0, 17, 0, 0)
    with pytest.raises(_lib.CQError):
        ao.set_class_tile(0, 1, 32, 0)


def test_screen_fusion_matches_per_class_screening():
    """The class pairs screened in one launch per stream (the default), in one launch for all,
    or one launch per class pair on its stream, with a budget that leaves some class pairs in
    several batches (those screen per batch) and without: J and K against the oracle,
    identical surviving counts, for two density slots (UHF-style) and in the mixed-precision
    rounds ((H2O)8 / cc-pVDZ)."""
    from chronusq_public_amd import _lib
    L = _lib.load()
    mol = M.water_cluster(8, seed=11, box_ang=M.scaled_box(8))
    bs, ao, ob = _setup(mol, "cc-pvdz")
    X = _sym_random(bs.nbasis, 17)
    Y = np.asfortranarray(0.3 * _sym_random(bs.nbasis, 18))
    Qg = ao.computeSchwartz()
    (Jo, Ko, Ky), n = _oracle_jk(ob, Qg, [X, X, Y], [0, 1, 1], [1.0, -0.5, -0.5], 1e-12)
    for budget in (None, 60000):
        if budget:
            ao.set_task_budget(budget)
            assert ao.task_plan()[1] > 1
        for fusion in (1, 2, 0):
            print(f"[fusion] budget={budget} fusion={fusion}", flush=True)
            _lib.check(L.cq_set_screen_fusion(ao._h, fusion), ao._h)
            J, K, KY = _gpu_jk(ao, [X, X, Y], [COULOMB, EXCHANGE, EXCHANGE], [1.0, -0.5, -0.5], 1e-12)
            assert sum(ao.quartet_counts().values()) == n
            assert np.abs(J - Jo).max() < TOL and np.abs(K - Ko).max() < TOL and np.abs(KY - Ky).max() < TOL
    # mixed precision: the FP32 window's round screens again (fused or not): same counts
    ao.setPrecision("mixed", 1e-5)
    res = []
    try:
        for fusion in (1, 2, 0):
            _lib.check(L.cq_set_screen_fusion(ao._h, fusion), ao._h)
            J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
            res.append((J, K, ao.precision_counts()))
            for a, b in ((J, Jo), (K, Ko)):
                assert np.abs(a - b).max() < 1e-9 + 1e-6 * np.abs(b).max()
    finally:
        ao.setPrecision("fp64")
    assert res[0][2][1] > 0
    for r in res[1:]:
        assert r[2] == res[0][2]
        assert np.abs(r[0] - res[0][0]).max() < 1e-12 and np.abs(r[1] - res[0][1]).max() < 1e-12
    _lib.check(L.cq_set_screen_fusion(ao._h, 1), ao._h)
    assert L.cq_set_screen_fusion(ao._h, 3) != 0


@pytest.mark.parametrize("cpi", [1, 2, 3, 36])
def test_chunks_per_work_item_match_oracle(cpi):
    """The split class pairs ((dd|dd) in 36 chunks, (dp|dp) and (dd|dp) in 18, ...) with 1, 2, 3
    or all of a task region's chunks per work item, spread over the workgroup's wavefronts and
    flushed once, in every digestion variant: J and K against the oracle with identical
    surviving-quartet counts ((H2O)8 / cc-pVDZ)."""
    from chronusq_public_amd import _lib
    L = _lib.load()
    mol = M.water_cluster(8, seed=11, box_ang=M.scaled_box(8))
    bs, ao, ob = _setup(mol, "cc-pvdz")
    X = _sym_random(bs.nbasis, 23)
    Qg = ao.computeSchwartz()
    (Jo, Ko), n = _oracle_jk(ob, Qg, [X, X], [0, 1], [1.0, -0.5], 1e-12)
    for k in range(L.cq_num_class_pairs()):
        _lib.check(L.cq_set_class_chunk_group(ao._h, k, cpi), ao._h)
    for lds_mode in (2, 1, 0):
        print(f"[chunks] cpi={cpi} lds_mode={lds_mode}", flush=True)
        _lib.check(L.cq_set_lds_digestion(ao._h, lds_mode), ao._h)
        J, K = _gpu_jk(ao, [X, X], [COULOMB, EXCHANGE], [1.0, -0.5], 1e-12)
        assert np.abs(J - Jo).max() < TOL and np.abs(K - Ko).max() < TOL, (cpi, lds_mode)
        assert sum(ao.quartet_counts().values()) == n
    with pytest.raises(_lib.CQError):
        _lib.check(L.cq_set_class_chunk_group(ao._h, 0, -1), ao._h)
