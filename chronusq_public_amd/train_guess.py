"""Training and evaluation of the learned-guess network (learned_guess.py).  No reference
counterpart.

    python -m chronusq_public_amd.train_guess --basis 6-31g* --waters 2000 --alkanes 2000 \\
        --epochs 60 --out chronusq_public_amd/guess_data/pair_blocks_631gs.npz
    python -m chronusq_public_amd.train_guess --basis 6-31g* --evaluate --eval-waters 100 \\
        --eval-alkanes 100 --json result.json

Training set (seeded; evaluation seeds are disjoint):
  * (H2O)_m, m uniform in [2, 8], oxygen centres in a cube scaled to m at the density of
    sixteen waters in a 10 Angstrom cube, cluster seeds 10000 + k;
  * C_nH_{2n+2}, n uniform in [1, 10]: the standard all-trans geometry, every coordinate
    displaced by N(0, 0.05 Angstrom), then a uniformly random rotation, seeds 20000 + k.
Targets: the converged RHF density of the device SCF (SingleSlater.SCF3) in each pair's frame
(P_ij) and, per atom, the converged diagonal block minus the spherically averaged atomic (SAD)
block.
Evaluation: (H2O)_16 seeds 90000 + k (box 10 Angstrom) and displaced C12H26 seeds 95000 + k;
the SCF from the SAD guess against the SCF from the learned guess, iterations, wall time and
final energies.

Training runs entirely on the GPU: the network's forward and backward passes and Adam are the
BF16 MFMA kernels of libcqgpu (include/cqgpu_guess.h); the rotation of the per-pair diagonal
corrections into the molecule's axes and their sum per atom (linear maps) take their gradient
from torch.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from typing import List

import numpy as np

from . import learned_guess as LG
from .molecule import Molecule, _quat_to_rot, alkane, scaled_box, water_cluster

WATER_SEED, ALKANE_SEED = 10000, 20000
EVAL_WATER_SEED, EVAL_ALKANE_SEED = 90000, 95000


def displaced_alkane(n: int, seed: int, sigma_ang: float = 0.05) -> Molecule:
    rng = np.random.Generator(np.random.MT19937(seed))
    m = alkane(n)
    X = m.coords_bohr + rng.normal(scale=sigma_ang / 0.5291772083000001, size=m.coords_bohr.shape)
    R = _quat_to_rot(*rng.random(3))
    X = (X - X.mean(axis=0)) @ R.T
    return Molecule(m.Z.copy(), X, 0, 1)


def training_molecules(n_water: int, n_alkane: int) -> List[Molecule]:
    mols = []
    for k in range(n_water):
        rng = np.random.Generator(np.random.MT19937(WATER_SEED + k))
        m = int(rng.integers(2, 9))
        mols.append(water_cluster(m, WATER_SEED + k, box_ang=scaled_box(m)))
    for k in range(n_alkane):
        rng = np.random.Generator(np.random.MT19937(ALKANE_SEED + k))
        n = int(rng.integers(1, 11))
        mols.append(displaced_alkane(n, ALKANE_SEED + k))
    return mols


def evaluation_molecules(n_water: int, n_alkane: int):
    out = [("(H2O)16", water_cluster(16, EVAL_WATER_SEED + k)) for k in range(n_water)]
    out += [("C12H26", displaced_alkane(12, EVAL_ALKANE_SEED + k)) for k in range(n_alkane)]
    return out


class Dataset:
    """Features, frames and targets of many molecules, concatenated on the device."""

    def __init__(self, dev):
        self.dev = dev
        self.parts = []

    def add(self, g: LG.PairGeometry, P, Ps):
        import torch
        dev = self.dev
        npair = len(g.pairs)
        na = len(g.first)
        flatP = P.reshape(-1)
        off, off_t, sel_off, di, sel_i, dj, sel_j = (torch.from_numpy(x).to(dev) for x in LG.block_index(g))
        Ui = torch.from_numpy(g.Ui).to(dev)
        Uj = torch.from_numpy(g.Uj).to(dev)
        blk = torch.zeros(npair * LG.BLK, dtype=torch.float64, device=dev)
        blk[sel_off] = flatP[off]
        loc = torch.bmm(torch.bmm(Ui, blk.view(npair, LG.MAXF, LG.MAXF)), Uj.transpose(1, 2))
        To = torch.zeros((npair, LG.OUT), dtype=torch.float32, device=dev)
        Mo = torch.zeros_like(To)
        To[:, :LG.BLK] = loc.reshape(npair, -1).float()
        Mo[:, :LG.BLK] = sel_off.view(npair, -1).float()
        # per-atom diagonal targets: converged block minus the atomic block (padded, masked)
        a = np.arange(LG.MAXF)
        f, c = g.first, g.count
        rows = f[:, None, None] + a[None, :, None]
        cols = f[:, None, None] + a[None, None, :]
        m = ((a[None, :, None] < c[:, None, None]) & (a[None, None, :] < c[:, None, None])).reshape(-1)
        pos = torch.from_numpy((rows * g.N + cols).reshape(-1)[m]).to(dev)
        mt = torch.from_numpy(m).to(dev)
        Td = torch.zeros(na * LG.BLK, dtype=torch.float64, device=dev)
        Td[mt] = (P - Ps).reshape(-1)[pos]
        self.parts.append(dict(
            feats=torch.from_numpy(g.feats).to(dev), Ui=Ui.float(), Uj=Uj.float(), To=To, Mo=Mo,
            Td=Td.view(na, LG.BLK).float(), Md=mt.view(na, LG.BLK).float(),
            ai=torch.from_numpy(g.pairs[:, 0]).to(dev), aj=torch.from_numpy(g.pairs[:, 1]).to(dev)))

    def finalize(self):
        import torch
        keys = ["feats", "Ui", "Uj", "To", "Mo", "Td", "Md"]
        cat = {k: torch.cat([p[k] for p in self.parts]) for k in keys}
        npairs = np.array([len(p["feats"]) for p in self.parts])
        natoms = np.array([len(p["Td"]) for p in self.parts])
        self.p0 = np.concatenate([[0], np.cumsum(npairs)])
        self.a0 = np.concatenate([[0], np.cumsum(natoms)])
        ai = torch.cat([p["ai"] + int(self.a0[k]) for k, p in enumerate(self.parts)])
        aj = torch.cat([p["aj"] + int(self.a0[k]) for k, p in enumerate(self.parts)])
        self.__dict__.update(cat)
        self.ai, self.aj = ai, aj
        self.UiT = self.Ui.transpose(1, 2).contiguous()
        self.UjT = self.Uj.transpose(1, 2).contiguous()
        self.nmol = len(self.parts)
        self.parts = []

    def batch(self, mols):
        """Pair rows, atom rows and the pairs' atoms renumbered within the batch."""
        import torch
        pidx = np.concatenate([np.arange(self.p0[m], self.p0[m + 1]) for m in mols])
        aidx = np.concatenate([np.arange(self.a0[m], self.a0[m + 1]) for m in mols])
        remap = np.full(int(self.a0[-1]), -1, dtype=np.int64)
        remap[aidx] = np.arange(len(aidx))
        pidx_t = torch.from_numpy(pidx).to(self.dev)
        remap_t = torch.from_numpy(remap).to(self.dev)
        return (pidx_t, torch.from_numpy(aidx).to(self.dev), remap_t[self.ai[pidx_t]], remap_t[self.aj[pidx_t]])


def batch_loss(net, ds: Dataset, mols, w_diag: float = 1.0, train: bool = True):
    """Masked MSE of the pair blocks plus that of the summed diagonal corrections; with train,
    the gradient with respect to the network output."""
    import torch
    pidx, aidx, ai, aj = ds.batch(mols)
    X = ds.feats[pidx]
    Y = net.forward(X)
    lo, dYo = net.mse(Y, ds.To[pidx], ds.Mo[pidx])
    n = Y.shape[0]
    Yt = Y.detach().requires_grad_(train)
    Yi = LG._sym(Yt[:, LG.BLK:2 * LG.BLK].reshape(n, LG.MAXF, LG.MAXF))
    Yj = LG._sym(Yt[:, 2 * LG.BLK:].reshape(n, LG.MAXF, LG.MAXF))
    Li = torch.bmm(torch.bmm(ds.UiT[pidx], Yi), ds.Ui[pidx]).reshape(n, -1)
    Lj = torch.bmm(torch.bmm(ds.UjT[pidx], Yj), ds.Uj[pidx]).reshape(n, -1)
    D = torch.zeros((len(aidx), LG.BLK), dtype=torch.float32, device=ds.dev)
    D = D.index_add(0, ai, Li).index_add(0, aj, Lj)
    Md = ds.Md[aidx]
    ld = (((D - ds.Td[aidx]) * Md) ** 2).sum() / Md.sum().clamp(min=1.0)
    if not train:
        return float(lo), float(ld), None
    (w_diag * ld).backward()
    return float(lo), float(ld.detach()), dYo + Yt.grad


def build_dataset(mols, basis_name: str, dev, log_every: int = 200):
    import torch
    from .singleslater import SCFNotConverged, SingleSlater
    ds = Dataset(dev)
    t0 = time.perf_counter()
    skipped = 0
    for k, mol in enumerate(mols):
        try:
            ss = SingleSlater(mol, basis_name, ref="RHF", tau=1e-11, ene_tol=1e-9, den_tol=1e-6,
                              max_iter=128, device=dev)
            res = ss.SCF3()
        except SCFNotConverged:
            skipped += 1
            continue
        g = LG.pair_geometry(mol, ss.basis)
        Ps = ss._atomic_densities(spherical=True)
        if len(g.pairs):
            ds.add(g, res.P, Ps)
        ss.ao.close()
        if (k + 1) % log_every == 0:
            torch.cuda.synchronize()
            print(f"  densities {k + 1}/{len(mols)}  {time.perf_counter() - t0:.1f} s", flush=True)
    ds.finalize()
    return ds, skipped, time.perf_counter() - t0


def train(ds: Dataset, epochs: int, batch_mols: int = 32, lr: float = 2e-3, seed: int = 0, val_frac: float = 0.05,
          w_diag: float = 1.0, log=print):
    import torch
    rng = np.random.default_rng(seed)
    order = rng.permutation(ds.nmol)
    nval = max(1, int(val_frac * ds.nmol))
    val, tr = order[:nval], order[nval:]
    net = LG.BF16MLP(seed=seed, device=ds.dev)
    steps_per_epoch = math.ceil(len(tr) / batch_mols)
    total = epochs * steps_per_epoch
    step = 0
    t0 = time.perf_counter()
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(tr)
        lo_sum = ld_sum = 0.0
        for b in range(0, len(perm), batch_mols):
            lr_t = 1e-5 + 0.5 * (lr - 1e-5) * (1.0 + math.cos(math.pi * step / total))
            lo, ld, dY = batch_loss(net, ds, perm[b:b + batch_mols], w_diag=w_diag)
            gW, gb = net.backward(dY)
            net.adam_step(gW, gb, lr=lr_t)
            lo_sum += lo
            ld_sum += ld
            step += 1
        vo, vd = [], []
        for b in range(0, len(val), batch_mols):
            a, c, _ = batch_loss(net, ds, val[b:b + batch_mols], train=False)
            vo.append(a)
            vd.append(c)
        rec = dict(epoch=ep + 1, train_pair=lo_sum / steps_per_epoch, train_diag=ld_sum / steps_per_epoch,
                   val_pair=float(np.mean(vo)), val_diag=float(np.mean(vd)), seconds=time.perf_counter() - t0)
        hist.append(rec)
        if ep == 0 or (ep + 1) % 5 == 0 or ep + 1 == epochs:
            log(f"  epoch {ep + 1}/{epochs}  pair mse {rec['train_pair']:.3e} / val {rec['val_pair']:.3e}  "
                f"diag mse {rec['train_diag']:.3e} / val {rec['val_diag']:.3e}  {rec['seconds']:.1f} s")
    torch.cuda.synchronize()
    return net, hist


def evaluate(basis_name: str, n_water: int, n_alkane: int, dev, model_path=None, log=print, tau: float = 1e-12):
    import torch
    from .singleslater import SingleSlater
    rows = []
    for k, (label, mol) in enumerate(evaluation_molecules(n_water, n_alkane)):
        rec = {"molecule": label, "index": k}
        for variant, guess, judge in (("SAD", "SAD", False), ("LEARNED", "LEARNED", True),
                                      ("LEARNED_UNJUDGED", "LEARNED", False)):
            ss = SingleSlater(mol, basis_name, ref="RHF", tau=tau, guess=guess, device=dev, guess_model=model_path,
                              judge_guess=judge)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = ss.SCF3()
            torch.cuda.synchronize()
            info = {k: (float(v) if isinstance(v, (int, float)) else str(v)) for k, v in res.guess_info.items()}
            E_ref = rec["SAD"]["energy"] if variant != "SAD" else res.energy
            k6 = next(i for i, e in enumerate(res.energies) if abs(e - E_ref) < 1e-6)
            rec[variant] = dict(iterations=res.iterations, energy=res.energy, seconds=time.perf_counter() - t0,
                                guess_s=res.guess_s, first_energy=res.energies[0], iterations_to_1e6=k6 + 1,
                                seconds_to_1e6=res.times[k6], **info)
            ss.ao.close()
        rec["dE"] = max(abs(rec[v]["energy"] - rec["SAD"]["energy"]) for v in ("LEARNED", "LEARNED_UNJUDGED"))
        rows.append(rec)
        if (k + 1) % 20 == 0:
            log(f"  evaluated {k + 1}/{n_water + n_alkane}")
    summary = {}
    for label in sorted(set(r["molecule"] for r in rows)):
        rs = [r for r in rows if r["molecule"] == label]
        d = dict(n=len(rs), max_abs_dE=float(max(r["dE"] for r in rs)),
                 learned_used=int(sum(r["LEARNED"]["used"] == "LEARNED" for r in rs)))
        for v in ("SAD", "LEARNED", "LEARNED_UNJUDGED"):
            it = np.array([r[v]["iterations"] for r in rs])
            t = np.array([r[v]["seconds"] for r in rs])
            e1 = np.array([r[v]["first_energy"] - r[v]["energy"] for r in rs])
            k6 = np.array([r[v]["iterations_to_1e6"] for r in rs])
            t6 = np.array([r[v]["seconds_to_1e6"] for r in rs])
            d[v] = dict(iterations_mean=float(it.mean()), iterations_max=int(it.max()),
                        iterations_to_1e6_mean=float(k6.mean()), seconds_to_1e6_p50=float(np.median(t6)),
                        seconds_p50=float(np.median(t)), seconds_p99=float(np.percentile(t, 99)),
                        first_iteration_error_mean=float(e1.mean()))
        summary[label] = d
    return summary, rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--basis", default="6-31g*")
    ap.add_argument("--waters", type=int, default=2000)
    ap.add_argument("--alkanes", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--evaluate", action="store_true")
    ap.add_argument("--model", default=None)
    ap.add_argument("--eval-waters", type=int, default=100)
    ap.add_argument("--eval-alkanes", type=int, default=100)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        sys.exit("train_guess needs a GPU (the network kernels run on the device)")
    dev = torch.device("cuda", 0)
    out = {}
    if not a.evaluate:
        mols = training_molecules(a.waters, a.alkanes)
        print(f"training set: {len(mols)} molecules ({a.waters} water clusters, {a.alkanes} alkanes), "
              f"basis {a.basis}", flush=True)
        ds, skipped, t_data = build_dataset(mols, a.basis, dev)
        print(f"  {ds.nmol} densities ({skipped} not converged), {int(ds.p0[-1])} pairs, "
              f"{int(ds.a0[-1])} atoms, {t_data:.1f} s", flush=True)
        net, hist = train(ds, a.epochs, batch_mols=a.batch, lr=a.lr, log=lambda s: print(s, flush=True))
        path = LG.save_model(net, a.basis, a.out, n_molecules=ds.nmol, n_pairs=int(ds.p0[-1]), epochs=a.epochs,
                             val_pair=hist[-1]["val_pair"], val_diag=hist[-1]["val_diag"],
                             data_seconds=t_data, train_seconds=hist[-1]["seconds"])
        print(f"saved {path}", flush=True)
        out["training"] = dict(molecules=ds.nmol, skipped=skipped, pairs=int(ds.p0[-1]), data_seconds=t_data,
                               history=hist, model=path)
    else:
        summary, rows = evaluate(a.basis, a.eval_waters, a.eval_alkanes, dev, model_path=a.model,
                                 log=lambda s: print(s, flush=True))
        print(json.dumps(summary, indent=1), flush=True)
        out["evaluation"] = dict(summary=summary, rows=rows)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
