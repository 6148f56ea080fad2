"""GPU: the batched SCF (chronusq_public_amd/batch.py) reaches every molecule's own SCF3 fixed
point -- same energy, same iteration count -- for an uneven batch (1 to 5 heavy atoms, padded
to the largest basis), in FP64 and with the mixed-precision incremental build."""
import numpy as np
import pytest

from chronusq_public_amd import molecule as M

pytestmark = pytest.mark.gpu


def _uneven():
    return [M.alkane(1), M.water_cluster(1, 5, box_ang=M.scaled_box(1)), M.alkane(3),
            M.water_cluster(3, 6, box_ang=M.scaled_box(3)), M.alkane(5), M.water_test0001()]


def test_batch_matches_single_molecule_scf():
    from chronusq_public_amd.batch import BatchSCF
    from chronusq_public_amd.singleslater import SingleSlater
    mols = _uneven()
    b = BatchSCF(mols, "6-31g*")
    r = b.run()
    b.close()
    assert r.converged.all()
    for k, m in enumerate(mols):
        # the batch's density-step schedule: eigensolve until |dE| < 0.1, then rotations
        s = SingleSlater(m, "6-31g*", pdiag_switch=0.1, pdiag_first=False).SCF3()
        assert abs(r.energies[k] - s.energy) < 1e-9, (k, r.energies[k], s.energy)
        assert abs(int(r.iterations[k]) - s.iterations) <= 1
        assert abs(r.energies[k] - SingleSlater(m, "6-31g*").SCF3().energy) < 1e-9


def test_batch_reaches_golden_water_energy():
    """Water of test0002 (RHF/6-31G, oldTests.tgz:tests/chronus-ref.val) padded beside a larger
    molecule: the reference's golden energy at its 5e-10 tolerance."""
    import json
    import os
    from chronusq_public_amd.batch import BatchSCF
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "energies.json")))
    r = BatchSCF([M.alkane(2), M.water_test0001()], "6-31g").run()
    assert abs(r.energies[1] - gold["tests"]["test0002"]["E"]) < gold["tolerance"]


def test_batch_mixed_incremental_reaches_fp64_energies():
    from chronusq_public_amd.batch import BatchSCF
    mols = _uneven()[:5]
    ref = BatchSCF(mols, "6-31g*").run()
    b = BatchSCF(mols, "6-31g*", precision="mixed", incremental=True)
    r = b.run()
    b.close()
    assert r.converged.all()
    assert np.abs(r.energies - ref.energies).max() < 1e-8


def test_batch_rejects_open_shells():
    from chronusq_public_amd.batch import BatchSCF
    oh = M.Molecule.from_angstrom(["O", "H"], [[0, 0, 0], [0, 0, 0.97]])
    with pytest.raises(ValueError):
        BatchSCF([oh], "6-31g*")


def test_batch_context_gives_every_molecules_own_gd():
    """cq_ctx_create_batch: one contraction over the block-diagonal density of three molecules
    gives, in each diagonal block, that molecule's own J and K (1e-10), zeros off the
    diagonal blocks, and the molecules' surviving-quartet counts summed; the one-electron
    integrals refuse a batch context."""
    import torch
    from chronusq_public_amd import basis as B
    from chronusq_public_amd._lib import CQError
    from chronusq_public_amd.aointegrals import AOIntegrals, COULOMB, EXCHANGE
    mols = [M.water_test0001(), M.alkane(2), M.water_cluster(2, 3, box_ang=M.scaled_box(2))]
    bases = [B.build_basis("6-31g*", m.Z, m.coords_bohr) for m in mols]
    big = AOIntegrals.batch(bases)
    big.thresholdSchwartz = 1e-12
    n = big.nBasis
    assert n == sum(b.nbasis for b in bases)
    rng = np.random.default_rng(3)
    X = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    blocks, off = [], 0
    for b in bases:
        A = rng.uniform(-1, 1, (b.nbasis, b.nbasis))
        A = 0.5 * (A + A.T)
        blocks.append((off, b.nbasis, A))
        X[off:off + b.nbasis, off:off + b.nbasis] = torch.from_numpy(A)
        off += b.nbasis
    J = torch.zeros_like(X)
    K = torch.zeros_like(X)
    big.newTwoEContractDirect([X, X], [J, K], [COULOMB, EXCHANGE], [1.0, -0.5])
    nq = sum(big.quartet_counts().values())
    mask = torch.zeros_like(X, dtype=torch.bool)
    total = 0
    for (o, nb, A), b in zip(blocks, bases):
        ao = AOIntegrals(b)
        ao.thresholdSchwartz = 1e-12
        x = torch.from_numpy(A).cuda()
        j = torch.zeros_like(x)
        k = torch.zeros_like(x)
        ao.newTwoEContractDirect([x, x], [j, k], [COULOMB, EXCHANGE], [1.0, -0.5])
        total += sum(ao.quartet_counts().values())
        assert float((J[o:o + nb, o:o + nb] - j).abs().max()) < 1e-10
        assert float((K[o:o + nb, o:o + nb] - k).abs().max()) < 1e-10
        mask[o:o + nb, o:o + nb] = True
    assert float(J[~mask].abs().max()) == 0.0 and float(K[~mask].abs().max()) == 0.0
    assert nq == total
    with pytest.raises(CQError):
        big.computeAOOneE(mols[0])
    # the batch's one-electron integrals: each molecule's own S, T, V in its diagonal block
    Sb, Tb, Vb = big.computeAOOneE_batch(mols)
    for (o, nb, _), b, m in zip(blocks, bases, mols):
        for got, ref in zip((Sb, Tb, Vb), AOIntegrals(b).computeAOOneE(m)):
            assert float((got[o:o + nb, o:o + nb] - ref).abs().max()) < 1e-12
    for M_ in (Sb, Tb, Vb):
        assert float(M_[~mask].abs().max()) == 0.0


def test_batch_per_molecule_calls_match_merged():
    from chronusq_public_amd.batch import BatchSCF
    mols = _uneven()[:4]
    a = BatchSCF(mols, "6-31g*", merged=True, group_functions=40)   # several groups
    ra = a.run()
    b = BatchSCF(mols, "6-31g*", merged=False)
    rb = b.run()
    assert len(a.groups) > 1
    assert ra.converged.all() and rb.converged.all()
    assert np.abs(ra.energies - rb.energies).max() < 1e-9


def test_batch_context_molecules_without_p_or_d_shells():
    """The LDS accumulators of a batch context hold, per class pair, one row per function of
    the swept pair class's shells in the largest molecule: molecules with no p or no d shells
    at all (H2/6-31G*: s only; NH3/6-31G: s and p) beside one with s, p and d (H2O/cc-pVDZ),
    the largest in the middle, give in each diagonal block the oracle's J and K (1e-10) with
    the oracle's surviving-quartet counts."""
    import torch
    import oracle
    from chronusq_public_amd import basis as B
    from chronusq_public_amd.aointegrals import AOIntegrals, COULOMB, EXCHANGE
    h2 = M.Molecule.from_angstrom(["H", "H"], [[0, 0, 0], [0, 0, 0.74]])
    nh3 = M.Molecule.from_angstrom(["N", "H", "H", "H"],
                                   [[0, 0, 0.12], [0.94, 0, -0.27], [-0.47, 0.81, -0.27], [-0.47, -0.81, -0.27]])
    parts = [(h2, "6-31g*"), (M.water_test0001(), "cc-pvdz"), (nh3, "6-31g"), (h2, "6-31g*")]
    bases = [B.build_basis(name, m.Z, m.coords_bohr) for m, name in parts]
    assert max(b.max_l for b in bases[0::3]) == 0 and bases[2].max_l == 1 and bases[1].max_l == 2
    big = AOIntegrals.batch(bases)
    big.thresholdSchwartz = 1e-12
    n = big.nBasis
    rng = np.random.default_rng(5)
    X = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    blocks, off = [], 0
    for b in bases:
        A = rng.uniform(-1, 1, (b.nbasis, b.nbasis))
        A = np.asfortranarray(0.5 * (A + A.T))
        blocks.append((off, b.nbasis, A))
        X[off:off + b.nbasis, off:off + b.nbasis] = torch.from_numpy(A)
        off += b.nbasis
    J = torch.zeros_like(X)
    K = torch.zeros_like(X)
    big.newTwoEContractDirect([X, X], [J, K], [COULOMB, EXCHANGE], [1.0, -0.5])
    nq = sum(big.quartet_counts().values())
    total = 0
    for (o, nb, A), b in zip(blocks, bases):
        ob = oracle.OracleBasis(b)
        Q = AOIntegrals(b).computeSchwartz()
        AX = [np.zeros((nb, nb), order="F") for _ in range(2)]
        total += ob.contract_direct(Q, 1e-12, [A, A], AX, [0, 1], [1.0, -0.5])
        assert float(np.abs(J[o:o + nb, o:o + nb].cpu().numpy() - AX[0]).max()) < 1e-10
        assert float(np.abs(K[o:o + nb, o:o + nb].cpu().numpy() - AX[1]).max()) < 1e-10
    assert nq == total
