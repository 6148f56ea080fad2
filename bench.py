#!/usr/bin/env python3
"""Benchmark: direct-SCF G[D] on (H2O)16 / cc-pVDZ (384 functions) on one MI355X.

A step is one complete G[D] for RHF -- AOIntegrals::newTwoEContract with X = {P, P},
types = {COULOMB, EXCHANGE}, scales = {1, -1/2} (SingleSlater::formPT,
include/singleslater/singleslater_fock.h:38-47) -- with the density and outputs resident in
HBM.  value = surviving shell quartets per second (tau = 1e-12).

Extra fields:
  scf           the device SCF on the same system (SAD guess, CDIIS): energy, iterations, wall
                time, and time until |E - E_fp64| <= 1e-6 Eh, for the plain FP64 build and the
                incremental (difference-density) FP64 build with a full rebuild every 8 steps;
                "_cold_sad" runs recompute the SAD atomic densities, the others take them from
                the process cache filled by an earlier run; "mixed_incremental" adds FP32
                integrals for quartets with Schwarz x Schwarz x max|dD| < 1e-5 inside the
                incremental build (the full rebuilds stay FP64).
  roofline      FP64 algorithmic FLOPs (McMurchie-Davidson model, counted over the surviving
                quartets of the dominant class-pair kernel) / that kernel's mean launch time
                from HIP events recorded on its stream, against the FP64 vector peak.
  cpu_baseline  the CPU restatement of newTwoEContractDirect (oracle/, OpenMP over the
                host cores) on the same G[D], timed once.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--calibrate-copy]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the class-pair kernels come from the rocprofv3 counter summary of the
# same build (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors from a 256 MiB device copy in the
# same runs); the summary carries the sha256 of the libcqgpu.so it profiled, and the traffic is
# reported only when that matches the library this run loaded.
COUNTERS_MD = os.path.join(ROOT, "profiles", "gdirect_h2o16_ccpvdz_counters.md")
METRIC = "SCF time-to-energy (1e-6 Eh vs FP64) + G[D] quartets/s, (H2O)16 cc-pVDZ"


def H(n):
    return (n + 1) * (n + 2) * (n + 3) // 6


def H4(n):
    return (n + 1) * (n + 2) * (n + 3) * (n + 4) // 24


def c(l):
    return (l + 1) * (l + 2) // 2


def model_flops(la, lb, lc, ld, sums, nmat):
    """McMurchie-Davidson FLOP yardstick per class pair (kernel orientation, bra = (la lb)):
    K_ab K_cd [37 + 3L + 3 H4(L) + 2 H(Lab) H(Lcd) c(lc) c(ld)] + K_ab 2 H(Lab) c(la)c(lb)c(lc)c(ld)
    + S_sph + 24 n_sph nMat, summed over surviving quartets."""
    kk, k, ssph, nsph = sums
    Lab, Lcd = la + lb, lc + ld
    L = Lab + Lcd
    A = 37 + 3 * L + 3 * H4(L) + 2 * H(Lab) * H(Lcd) * c(lc) * c(ld)
    Bc = 2 * H(Lab) * c(la) * c(lb) * c(lc) * c(ld)
    return A * kk + Bc * k + ssph + 24.0 * nsph * nmat


def profiled_traffic(lib_path):
    """{class: HBM bytes per launch} from the counter summary, or None (no summary, or one
    taken from another build of the library)."""
    import hashlib
    import re
    try:
        text = open(COUNTERS_MD).read()
    except OSError:
        return None, "no counter summary"
    m = re.search(r"<!-- traffic-json: (\{.*?\}) -->", text, re.S)
    if not m:
        return None, "counter summary has no traffic record"
    rec = json.loads(m.group(1))
    with open(lib_path, "rb") as f:
        sha = hashlib.sha256(f.read()).hexdigest()
    if rec.get("lib_sha256") != sha:
        return None, f"counter summary profiles another build ({rec.get('lib_sha256', '?')[:12]} vs {sha[:12]})"
    return rec["traffic_bytes"], "rocprofv3 FETCH_SIZE/WRITE_SIZE of this build"


def synthetic_density(bs, seed=0):
    """Symmetric density with exponential decay in the centre distance, D = exp(-R/2) U,
    U uniform in [-1, 1] (seeded)."""
    cen = np.concatenate([[s.origin] * s.size for s in bs.shells])
    R = np.linalg.norm(cen[:, None, :] - cen[None, :, :], axis=-1)
    rng = np.random.default_rng(seed)
    U = rng.uniform(-1, 1, (bs.nbasis, bs.nbasis))
    U = 0.5 * (U + U.T)
    return np.exp(-0.5 * R) * U


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(bs, D, tau, nthreads, ncalls=3):
    """CPU restatement (not Libint) of newTwoEContractDirect on the same G[D]: one warm-up
    call, then the median of `ncalls` timed calls."""
    import oracle   # the CPU restatement, used only as the baseline leg
    ob = oracle.OracleBasis(bs)
    Q = ob.schwarz(nthreads)
    N = bs.nbasis
    Df = np.asfortranarray(D)
    times = []
    n = 0
    for i in range(ncalls + 1):
        J = np.zeros((N, N), order="F")
        K = np.zeros((N, N), order="F")
        t = time.perf_counter()
        n = ob.contract_direct(Q, tau, [Df, Df], [J, K], [0, 1], [1.0, -0.5], nthreads)
        if i:
            times.append(time.perf_counter() - t)
    dt = float(np.median(times))
    return {"value": n / dt, "unit": "quartets/s", "cores": nthreads, "kind": "port",
            "cpu": f"{cpu_model()} (nproc {os.cpu_count()})",
            "sample": f"CPU restatement (not Libint) of newTwoEContractDirect, Obara-Saika/HGP "
                      f"ERIs, OpenMP: the same G[D] ({n} surviving quartets), 1 warm-up + median "
                      f"of {ncalls} calls, {dt:.2f} s per call"}


def run_autotuned(bs, args, dD, dJ, dK, stream, nquart, e0, e1):
    """G[D] with per-class digestion variants and tiles chosen by AOIntegrals.autotune, timed
    over the same steps as the production configuration."""
    import torch
    from chronusq_public_amd.aointegrals import AOIntegrals, COULOMB, EXCHANGE
    ao_t = AOIntegrals(bs)
    ao_t.thresholdSchwartz = args.tau
    choice = ao_t.autotune(dD, reps=2)
    for _ in range(args.warmup):
        dJ.zero_()
        dK.zero_()
        ao_t.newTwoEContractDirect([dD, dD], [dJ, dK], [COULOMB, EXCHANGE], [1.0, -0.5], stream=stream.cuda_stream)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        dJ.zero_()
        dK.zero_()
        ao_t.newTwoEContractDirect([dD, dD], [dJ, dK], [COULOMB, EXCHANGE], [1.0, -0.5], stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_tuned = e0.elapsed_time(e1) / args.steps
    ao_t.close()
    return {"ms_per_step": ms_tuned, "quartets_per_s": nquart / (ms_tuned * 1e-3),
            "choice": {k: {"variant": v, "tile": list(t), "ms": tm[(t, v)]} for k, (v, t, tm) in choice.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tau", type=float, default=1e-12)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 256-molecule SCF")
    ap.add_argument("--no-autotune", action="store_true", help="skip the per-class digestion autotuner")
    ap.add_argument("--no-scf-variants", action="store_true",
                    help="run only the reference FP64 SCF (its converged density feeds the G[D] steps), "
                         "not the guess / incremental / mixed-precision variants")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run three 256 MiB device-to-device copies (a plain copy kernel of "
                         "known size, the scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path "
              "(G[D] would shard by quartet ranges with one all-reduce of G).", file=sys.stderr)
        sys.exit(2)

    import torch
    from chronusq_public_amd import basis as B, molecule as M
    from chronusq_public_amd.aointegrals import AOIntegrals, COULOMB, EXCHANGE
    from chronusq_public_amd import _lib
    import ctypes

    from chronusq_public_amd.singleslater import SingleSlater, time_to_energy

    dev = torch.device("cuda:0")
    mol = M.water_cluster(16, seed=args.seed)
    bs = B.build_basis("cc-pvdz", mol.Z, mol.coords_bohr)
    N = bs.nbasis

    # ---- SCF time-to-energy (device SCF: 1e integrals, SAD guess, CDIIS, eigensolve, G[D])
    from chronusq_public_amd.singleslater import clear_sad_cache
    w = M.water_monomer()   # warm-up: code objects, solver handles, the guess network's weights
    SingleSlater(w, "cc-pvdz", tau=args.tau).SCF3()
    SingleSlater(M.water_cluster(2, seed=1, box_ang=M.scaled_box(2)), "cc-pvdz", tau=args.tau,
                 guess="LEARNED").SCF3()

    def run_scf(inc, cold, precision="fp64", guess="SAD", judge="auto", profile=False):
        if cold:
            clear_sad_cache()   # the SAD guess recomputes its atomic UHF densities
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = SingleSlater(mol, bs, tau=args.tau, guess=guess, incremental=inc, rebuild_every=8,
                         precision=precision, mixed_threshold=1e-5, judge_guess=judge).SCF3(profile=profile)
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    r, wall = run_scf(False, True)
    E_ref, P_conv = r.energy, r.P
    # the oracle's FP64 energy of this cluster (tests/golden/scf_oracle.json, seed 0 only)
    E_oracle = None
    try:
        with open(os.path.join(ROOT, "tests", "golden", "scf_oracle.json")) as f:
            if args.seed == 0:
                E_oracle = json.load(f)["molecules"]["h2o16_ccpvdz_seed0"]["E"]
    except (OSError, KeyError):
        pass
    scf = {"E_fp64_device": E_ref, "E_fp64_oracle": E_oracle,
           "note": "dE_vs_oracle against the oracle's FP64 SCF energy; time_to_1e-6 against the device FP64 "
                   "energy (equal to the oracle's within 1e-9 Eh, tested). The SCF follows the reference's "
                   "rules (CDIIS depth 10, convergence |dE| < 1e-10 and ||dP||_F < 1e-8) without damping and "
                   "with pseudo-diagonalisation steps instead of an eigensolve in most iterations: iteration "
                   "counts are this build's, not the reference's schedule"}
    # SAD (cold: atomic UHF densities recomputed; warm: cached per process) and the learned
    # guess (network trained on small water clusters and alkanes; "judged": the SAD start is
    # made too and the lower-energy start kept)
    variants = [(name + ("_cold_sad" if cold else ""), inc, prec, "SAD", True, cold)
                for name, inc, prec in (("fp64", False, "fp64"), ("fp64_incremental", True, "fp64"),
                                        ("mixed_incremental", True, "mixed"))
                for cold in (True, False)]
    variants += [("fp64_learned", False, "fp64", "LEARNED", "auto", False),
                 ("fp64_learned_judged", False, "fp64", "LEARNED", True, False),
                 ("mixed_incremental_learned", True, "mixed", "LEARNED", "auto", False),
                 ("mixed_incremental_learned_judged", True, "mixed", "LEARNED", True, False)]
    if args.no_scf_variants:
        variants = []
    for key, inc, prec, guess, judge, cold in variants:
        prof = key in ("fp64_incremental", "mixed_incremental")
        r, wall = run_scf(inc, cold, prec, guess, judge, profile=prof)
        t6 = time_to_energy(r, E_ref, 1e-6)
        k6 = next(i for i, t in enumerate(r.times) if t >= t6)
        scf[key] = {"energy": r.energy, "dE_vs_fp64": r.energy - E_ref,
                    "dE_vs_oracle": (r.energy - E_oracle) if E_oracle is not None else None,
                    "iterations": r.iterations, "iterations_to_1e-6": k6 + 1,
                    "wall_s": wall, "setup_s": r.setup_s, "setup_parts": r.setup_parts,
                    "guess_s": r.guess_s, "guess_used": str(r.guess_info.get("used")),
                    "guess_judged": r.guess_info.get("judged"),
                    "iterations_s_to_1e-6": t6 - r.guess_s,
                    "first_iteration_error": r.energies[0] - E_ref,
                    "time_to_1e-6_s": r.setup_s + t6,
                    "gd_ms_mean": 1e3 * float(np.mean(r.gd_times))}
        if prof:
            scf[key]["per_iteration"] = [{"build": ph["build"], "gd_ms": 1e3 * ph["gd"],
                                          "quartets_fp64": ph["quartets_fp64"], "quartets_fp32": ph["quartets_fp32"]}
                                         for ph in r.phase_times]

    # ---- batched uneven molecules: 128 alkanes C_nH_{2n+2}, n in [1, 12], and 128 (H2O)_m,
    # m in [1, 8], 6-31G*, one batched SCF; the same molecules one by one for comparison
    batch = None
    if not args.no_batch:
        from chronusq_public_amd.batch import BatchSCF, c4_molecules
        mols4 = c4_molecules(args.seed)
        BatchSCF(mols4[:4] + mols4[-4:], "6-31g*").run()      # warm-up: code objects, atoms
        batch = {"molecules": len(mols4), "basis": "6-31g*"}
        for key, prec, inc in (("fp64", "fp64", False), ("mixed_incremental", "mixed", True)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bs4 = BatchSCF(mols4, "6-31g*", tau=args.tau, precision=prec, incremental=inc)
            rb = bs4.run()
            torch.cuda.synchronize()
            wall4 = time.perf_counter() - t0
            bs4.close()
            # completion_s: when each molecule converged, counted from the batch's start (its latency
            # in the batch); amortized: the batch wall time per molecule
            batch[key] = {"molecules_per_s": len(mols4) / wall4, "wall_s": wall4, "setup_s": rb.setup_s,
                          "amortized_s_per_molecule": wall4 / len(mols4),
                          "completion_s_p50": float(np.nanmedian(rb.seconds)),
                          "completion_s_p99": float(np.nanpercentile(rb.seconds, 99)),
                          "iterations_max": int(rb.iterations.max()), "converged": int(rb.converged.sum())}
            if key == "fp64":
                E4 = rb.energies
        nseq = 32
        seq_t = []
        dE4 = 0.0
        for k in range(nseq):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rs4 = SingleSlater(mols4[k * len(mols4) // nseq], "6-31g*", tau=args.tau).SCF3()
            torch.cuda.synchronize()
            seq_t.append(time.perf_counter() - t0)
            dE4 = max(dE4, abs(rs4.energy - E4[k * len(mols4) // nseq]))
        batch["one_by_one"] = {"molecules_per_s": nseq / sum(seq_t), "sample": f"{nseq} of the molecules, evenly spaced",
                               "scf_s_p50": float(np.median(seq_t)),
                               "scf_s_p99": float(np.percentile(seq_t, 99)),
                               "max_abs_dE_vs_batch": dE4}

    if args.calibrate_copy:
        src = torch.empty(32 << 20, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ao = AOIntegrals(bs)
    ao.thresholdSchwartz = args.tau
    D = P_conv.cpu().numpy() if P_conv is not None else synthetic_density(bs, args.seed)
    dD = torch.from_numpy(np.ascontiguousarray(D)).to(dev)
    dJ = torch.zeros((N, N), dtype=torch.float64, device=dev)
    dK = torch.zeros((N, N), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        dJ.zero_()
        dK.zero_()
        ao.newTwoEContractDirect([dD, dD], [dJ, dK], [COULOMB, EXCHANGE], [1.0, -0.5],
                                 stream=stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    L = _lib.load()
    ncp = L.cq_num_class_pairs()
    # timed region: the production configuration (class kernels over 4 streams)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = e0.elapsed_time(e1) / args.steps
    # per-kernel pass: kernels serialised on one stream, HIP events around each launch
    _lib.check(L.cq_set_launch_streams(ao._h, 1), ao._h)
    _lib.check(L.cq_set_profiling(ao._h, 1), ao._h)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (ctypes.c_double * ncp)()
    nl = (ctypes.c_longlong * ncp)()
    _lib.check(L.cq_get_kernel_times(ao._h, ms, nl, ncp), ao._h)
    sms = (ctypes.c_double * (ncp + 1))()      # entry ncp: the launches shared by several class pairs
    snl = (ctypes.c_longlong * (ncp + 1))()
    _lib.check(L.cq_get_screen_times(ao._h, sms, snl, ncp + 1), ao._h)
    fused_screen_ms = sms[ncp] / args.steps     # per G[D] (two launches per stream and call)
    _lib.check(L.cq_set_profiling(ao._h, 0), ao._h)
    _lib.check(L.cq_set_launch_streams(ao._h, 4), ao._h)
    counts = ao.quartet_counts()
    nquart = sum(counts.values())
    # per-class digestion variants picked by the autotuner (timed once, then the same K steps)
    autotuned = None
    if not args.no_autotune:
        autotuned = run_autotuned(bs, args, dD, dJ, dK, stream, nquart, e0, e1)
    ws = (ctypes.c_double * (4 * ncp))()
    _lib.check(L.cq_get_work_stats(ao._h, ws, ncp), ao._h)
    per_class = []
    total_flops = 0.0
    for k in range(ncp):
        ls = [ctypes.c_int() for _ in range(4)]
        L.cq_class_pair_info(k, *[ctypes.byref(x) for x in ls])
        la, lb, lc, ld = (x.value for x in ls)
        fl = model_flops(la, lb, lc, ld, [ws[4 * k + q] for q in range(4)], 2)
        total_flops += fl
        mean_ms = ms[k] / nl[k] if nl[k] else 0.0
        per_class.append({"class": "({}{}|{}{})".format(*["spdfg"[x] for x in (la, lb, lc, ld)]),
                          "quartets": int(list(counts.values())[k]), "prim_quartets": ws[4 * k],
                          "model_flop": fl, "kernel_ms": mean_ms,
                          "screen_ms": sms[k] / snl[k] if snl[k] else 0.0,
                          "tflops": fl / (mean_ms * 1e-3) / 1e12 if mean_ms > 0 else 0.0})
    prim_q = sum(ws[4 * k] for k in range(ncp))
    n_int = sum(ws[4 * k + 3] for k in range(ncp))
    dom = max(per_class, key=lambda r: r["kernel_ms"])
    achieved = dom["model_flop"] / (dom["kernel_ms"] * 1e-3) / 1e12
    traffic_tab, traffic_src = profiled_traffic(_lib.LIB_PATH)
    traffic = traffic_tab.get(dom["class"]) if traffic_tab else None
    # algorithmic bytes of one launch: the Cartesian density read once, J and K written once,
    # the primitive-pair table (96 B per primitive pair) read once
    ncart = ctypes.c_int()
    nprim = ctypes.c_int()
    _lib.check(L.cq_get_dims(ao._h, None, None, ctypes.byref(ncart), None, ctypes.byref(nprim)), ao._h)
    alg_bytes = 3 * 8 * ncart.value ** 2 + 96 * nprim.value
    value = nquart / (ms_per_step * 1e-3)
    out = {
        "metric": METRIC, "value": value, "unit": "quartets/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic input: seeded (H2O)16 geometry; density = its converged RHF density from the device SCF",
        "config": {"workload": "RHF G[D] = 2J[P] - K[P], (H2O)16 seed 0 / cc-pVDZ", "nbf": N,
                   "nshell": bs.nshell, "tau": args.tau, "surviving_quartets": nquart,
                   "model_gflop_per_step": total_flops / 1e9,
                   "prim_quartets_per_s": prim_q / (ms_per_step * 1e-3),
                   "integrals_per_s": n_int / (ms_per_step * 1e-3),
                   "serial_kernels_ms": sum(r["kernel_ms"] for r in per_class),
                   "serial_screen_ms": sum(r["screen_ms"] for r in per_class) + fused_screen_ms,
                   "fused_screen_ms": fused_screen_ms,
                   "all_kernels_tflops": total_flops / (sum(r["kernel_ms"] for r in per_class) * 1e-3) / 1e12,
                   "step_tflops": total_flops / (ms_per_step * 1e-3) / 1e12},
        "roofline": {"bound": "fp64-valu",
                     "pipe": "FP64 vector ALU (no ERI kernel issues MFMA; the FP64 vector peak is 78.6 TFLOP/s)",
                     "kernel": "eri_grp" + dom["class"], "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": alg_bytes,
                     "traffic_over_algorithmic": traffic / alg_bytes if traffic else None},
        "autotuned": autotuned,
        "scf": scf,
        "batch": batch,
        "classes": sorted(per_class, key=lambda r: -r["kernel_ms"]),
        "wall_s_timed": wall,
    }
    if not args.no_cpu:
        ncores = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(bs, D, args.tau, ncores)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
