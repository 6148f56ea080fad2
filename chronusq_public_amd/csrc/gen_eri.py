#!/usr/bin/env python3
"""Generate the per-class McMurchie-Davidson ERI kernels for gfx950.

One HIP device function per shell-quartet class (la lb | lc ld), fully unrolled: the
Hermite expansion coefficients E (bra and ket), the Hermite Coulomb integrals R_tuv, the
ket Hermite->Cartesian transform and the bra transform are straight-line code over
scalars, so every intermediate lives in a register.  The ket Cartesian components are
processed in chunks sized to keep the live set under a register budget; each chunk is
digested into J and K (forward updates) before the next one starts.

Two kernels are instantiated per class from the generated core:
  * eri_block_<cls>: writes the contracted Cartesian block (ab|cd) of each task (used for
    Schwarz bounds and for per-quartet parity tests);
  * eri_jk_<cls>: digests the block into the forward J/K accumulators.

The algorithm restated here replaces libint2::Engine::compute as called at
include/aointegrals/aointegrals_contract.h:138-143; the digestion reproduces the scatter at
include/aointegrals/aointegrals_contract.h:145-171 (see the J/K notes in gdirect.hip).

Usage: gen_eri.py OUTPUT_DIR
"""
from __future__ import annotations

import sys
from typing import Dict, List, Tuple

CLS_NAME = "spdfg"


def cart(l: int) -> List[Tuple[int, int, int]]:
    out = []
    for i in range(l + 1):
        lx = l - i
        for j in range(i + 1):
            ly = i - j
            out.append((lx, ly, l - lx - ly))
    return out


def ncart(l):
    return (l + 1) * (l + 2) // 2


def hermite_list(L):
    """(t,u,v) with t+u+v <= L."""
    out = []
    for n in range(L + 1):
        for t in range(n, -1, -1):
            for u in range(n - t, -1, -1):
                out.append((t, u, n - t - u))
    return out


# Arithmetic type of the generated quartet code: the template parameter R (double, or float
# for the mixed-precision kernels).  Literals are wrapped so that they do not promote R.
RT = "R"


def lit(v: float) -> str:
    return f"R({float(v)})"


class Emitter:
    def __init__(self):
        self.lines: List[str] = []
        self.ind = 1

    def __call__(self, s: str):
        self.lines.append("  " * self.ind + s)


# --------------------------------------------------------------------------- E coefficients
def gen_E(em: Emitter, pre: str, d: str, LA: int, LB: int, PA: str, PB: str, o2p: str, sign: bool) -> Dict:
    """Emit E^{ij}_t for one direction; returns dict (i,j,t) -> expression string
    (a variable name or a literal).  sign=True folds (-1)^t into the result (ket side)."""
    E: Dict[Tuple[int, int, int], str] = {(0, 0, 0): "1.0"}

    def term(coef: str, e: str):
        if e == "0" or coef == "0":
            return None
        if e == "1.0":
            return coef
        if coef == "1.0":
            return e
        return f"{coef}*{e}"

    for i in range(LA + 1):
        for j in range(LB + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                si, sj, X = i - 1, j, PA
            else:
                si, sj, X = i, j - 1, PB
            for t in range(i + j + 1):
                parts = []
                if t > 0 and (si, sj, t - 1) in E:
                    parts.append(term(o2p, E[(si, sj, t - 1)]))
                if (si, sj, t) in E:
                    parts.append(term(X, E[(si, sj, t)]))
                if (si, sj, t + 1) in E:
                    parts.append(term(lit(t + 1), E[(si, sj, t + 1)]))
                parts = [p for p in parts if p is not None]
                if not parts:
                    continue
                expr = " + ".join(parts)
                name = f"{pre}{d}_{i}_{j}_{t}"
                # plain aliases need no variable
                if len(parts) == 1 and ("*" not in parts[0]) and ("+" not in parts[0]):
                    E[(i, j, t)] = parts[0]
                else:
                    em(f"const {RT} {name} = {expr};")
                    E[(i, j, t)] = name
    if sign:
        out = {}
        for (i, j, t), e in E.items():
            out[(i, j, t)] = (e, -1 if t % 2 else 1)
        return out
    return {k: (v, 1) for k, v in E.items()}


def mul(*fs):
    """Product of (expr, sign) factors -> (expr or None for 1, sign)."""
    s = 1
    parts = []
    for e, sg in fs:
        s *= sg
        if e == "1.0":
            continue
        parts.append(e)
    return ("*".join(parts) if parts else None, s)


# --------------------------------------------------------------------------- R integrals
class RGen:
    def __init__(self, em: Emitter, L: int):
        self.em = em
        self.L = L
        self.memo: Dict[Tuple[int, int, int, int], str] = {}

    def get(self, n, t, u, v) -> str:
        key = (n, t, u, v)
        if key in self.memo:
            return self.memo[key]
        if t == 0 and u == 0 and v == 0:
            name = f"Rb{n}"
            self.memo[key] = name
            return name
        if t > 0:
            a = self.get(n + 1, t - 1, u, v)
            b = self.get(n + 1, t - 2, u, v) if t > 1 else None
            X, k = "PQx", t - 1
        elif u > 0:
            a = self.get(n + 1, t, u - 1, v)
            b = self.get(n + 1, t, u - 2, v) if u > 1 else None
            X, k = "PQy", u - 1
        else:
            a = self.get(n + 1, t, u, v - 1)
            b = self.get(n + 1, t, u, v - 2) if v > 1 else None
            X, k = "PQz", v - 1
        name = f"R{n}_{t}_{u}_{v}"
        if b is not None:
            expr = f"{X}*{a} + {b}" if k == 1 else f"{X}*{a} + {lit(k)}*{b}"
        else:
            expr = f"{X}*{a}"
        self.em(f"const {RT} {name} = {expr};")
        self.memo[key] = name
        return name


# --------------------------------------------------------------------------- class core
REG_SOFT = 125     # live doubles targeted by the chunk-by-c plan (about 250 VGPRs at most)
REG_HARD = 170     # above this even one c component per chunk spills: use (c,d) chunks


def plan_class(LA, LB, LC, LD):
    """Split the ket Cartesian components (c, d) into chunks processed one after the other.

    Returns (chunks, persist): chunks is a list of lists of (ic, idd); persist says whether
    J(1,2), K(a,d), K(b,d) are accumulated in registers across chunks (else every chunk
    issues its own atomics)."""
    NAB = ncart(LA) * ncart(LB)
    HAB = len(hermite_list(LA + LB))
    NC, ND = ncart(LC), ncart(LD)
    HL = len(hermite_list(LA + LB + LC + LD))
    persist_regs = NAB + ncart(LA) * ND + ncart(LB) * ND

    def regs(nk, persist):
        return (NAB + HAB) * nk + HL + (persist_regs if persist else 0) + 24

    if regs(ND, True) <= REG_HARD:
        cs = NC
        while cs > 1 and regs(cs * ND, True) > REG_SOFT:
            cs -= 1
        chunks = []
        for c0 in range(0, NC, cs):
            chunks.append([(ic, idd) for ic in range(c0, min(NC, c0 + cs)) for idd in range(ND)])
        return chunks, True
    pairs = [(ic, idd) for ic in range(NC) for idd in range(ND)]
    nk = max(1, (REG_HARD - 20 - HL - 24) // (NAB + HAB))
    return [pairs[i:i + nk] for i in range(0, len(pairs), nk)], False


def gen_class(LA, LB, LC, LD, kind: str) -> str:
    """kind: 'jk' (digest into J/K) or 'block' (write the contracted Cartesian block)."""
    name = f"{CLS_NAME[LA]}{CLS_NAME[LB]}{CLS_NAME[LC]}{CLS_NAME[LD]}"
    L = LA + LB + LC + LD
    ca, cb, cc, cd = cart(LA), cart(LB), cart(LC), cart(LD)
    NA, NB, NC, ND = len(ca), len(cb), len(cc), len(cd)
    herm_ab = hermite_list(LA + LB)
    chunks, persist = plan_class(LA, LB, LC, LD)
    em = Emitter()
    em.ind = 0
    em(f"// ---- class ({name[0]}{name[1]}|{name[2]}{name[3]}) kind={kind} chunks={len(chunks)} persist={persist}")
    if kind == "jk":
        # The wavefront's lanes share one uniform pair (the ket for DIG 0 and 1, the bra for DIG 2);
        # each lane holds one quartet with its own swept pair.
        # DIG 0: K updates are HBM atomics; DIG 1 (uniform ket): K columns c and d accumulate in LDS
        # (colC: the c columns, colD: the d columns, LN rows each); DIG 2 (uniform bra): K rows a
        # and b accumulate in LDS (colC: the a rows, colD: the b rows, LN columns each).
        # J of the uniform pair is reduced across the wavefront into LDS (jcdL); J of the lane's
        # own swept pair is added to its LDS slot jl (shared by the work item's quartets of that
        # pair, flushed once per item).
        # R: arithmetic type of the integrals and their digestion (LDS and HBM sums stay double)
        em("template <int NSLOT, int DIG, typename R = double>")
        em(f"__device__ __forceinline__ void quartet_jk_{name}(const QuartetArgs& q, const JKArgs& jk, const double* __restrict__ boys_lds, int chunk, double* colC, double* colD, double* jcdL, double* jl) {{")
    else:
        em(f"__device__ __forceinline__ void quartet_block_{name}(const QuartetArgs& q, double* __restrict__ out, const double* __restrict__ boys_lds) {{")
    em.ind = 1
    if kind != "jk":
        em("using R = double;")
    if kind == "jk" and persist:
        for ia in range(NA):
            for ib in range(NB):
                em(f"R jab_{ia}_{ib}[NSLOT] = {{}};")
        for ia in range(NA):
            for idd in range(ND):
                em(f"R kad_{ia}_{idd}[NSLOT] = {{}};")
        for ib in range(NB):
            for idd in range(ND):
                em(f"R kbd_{ib}_{idd}[NSLOT] = {{}};")
    for ci, ks in enumerate(chunks):
        if kind == "jk" and not persist:
            # chunks without shared accumulators are independent: one per wavefront
            em(f"if (chunk == {ci}) {{ // chunk {ci}: (c,d) components {ks}")
        else:
            em(f"{{ // chunk {ci}: (c,d) components {ks}")
        em.ind += 1
        for ia in range(NA):
            for ib in range(NB):
                for (ic, idd) in ks:
                    em(f"R o_{ia}_{ib}_{ic}_{idd} = R(0);")
        em("for (int ip = q.bra_p0; ip < q.bra_p1; ++ip) {")
        em.ind += 1
        em("const PrimPair bp = q.prim[ip];")
        em("const R bp_K = R(bp.K), bp_o2p = R(bp.o2p);")
        em("const R bp_PA0 = R(bp.PA[0]), bp_PA1 = R(bp.PA[1]), bp_PA2 = R(bp.PA[2]);")
        em("const R bp_PB0 = R(bp.PB[0]), bp_PB1 = R(bp.PB[1]), bp_PB2 = R(bp.PB[2]);")
        Ebra = {}
        for d, i in (("x", 0), ("y", 1), ("z", 2)):
            Ebra[d] = gen_E(em, "Eb", d, LA, LB, f"bp_PA{i}", f"bp_PB{i}", "bp_o2p", sign=False)
        for (t, u, v) in herm_ab:
            for (ic, idd) in ks:
                em(f"R W_{t}_{u}_{v}_{ic}_{idd} = R(0);")
        em("for (int iq = q.ket_p0; iq < q.ket_p1; ++iq) {")
        em.ind += 1
        em("const PrimPair kp = q.prim[iq];")
        # the Boys argument and PQ are formed in double, the rest in R
        em("const double pqd = bp.p + kp.p;")
        em("const double rsd = rsqrt_f64(pqd);")
        em("const double alphad = bp.p * kp.p * (rsd * rsd);")
        em("const double PQxd = bp.P[0] - kp.P[0], PQyd = bp.P[1] - kp.P[1], PQzd = bp.P[2] - kp.P[2];")
        em(f"R F[{L + 1}];")
        em(f"boys_eval<{L}>(alphad * (PQxd * PQxd + PQyd * PQyd + PQzd * PQzd), boys_lds, F);")
        em("const R PQx = R(PQxd), PQy = R(PQyd), PQz = R(PQzd);")
        em("const R alpha = R(alphad);")
        em("R pw = R(kp.K * rsd);   // K of both pairs carry sqrt(2 pi^(5/2)) each")
        for n in range(L + 1):
            em(f"const R Rb{n} = F[{n}] * pw;")
            if n < L:
                em("pw *= R(-2.0) * alpha;")
        em("const R kp_o2p = R(kp.o2p);")
        em("const R kp_PA0 = R(kp.PA[0]), kp_PA1 = R(kp.PA[1]), kp_PA2 = R(kp.PA[2]);")
        em("const R kp_PB0 = R(kp.PB[0]), kp_PB1 = R(kp.PB[1]), kp_PB2 = R(kp.PB[2]);")
        Eket = {}
        for d, i in (("x", 0), ("y", 1), ("z", 2)):
            Eket[d] = gen_E(em, "Ek", d, LC, LD, f"kp_PA{i}", f"kp_PB{i}", "kp_o2p", sign=True)
        rg = RGen(em, L)
        for (ic, idd) in ks:
            (kx, ky, kz), (lx, ly, lz) = cc[ic], cd[idd]
            terms = []
            for tau in range(kx + lx + 1):
                if (kx, lx, tau) not in Eket["x"]:
                    continue
                for nu in range(ky + ly + 1):
                    if (ky, ly, nu) not in Eket["y"]:
                        continue
                    for ph in range(kz + lz + 1):
                        if (kz, lz, ph) not in Eket["z"]:
                            continue
                        e, sg = mul(Eket["x"][(kx, lx, tau)], Eket["y"][(ky, ly, nu)], Eket["z"][(kz, lz, ph)])
                        cname = None
                        if e is not None and "*" in e:
                            cname = f"C_{ic}_{idd}_{tau}_{nu}_{ph}"
                            em(f"const R {cname} = {'-' if sg < 0 else ''}{e};")
                            sg = 1
                        elif e is not None:
                            cname = e
                        terms.append((tau, nu, ph, cname, sg))
            for (t, u, v) in herm_ab:
                # one FMA into the accumulator per term (an expression summed first and then
                # added would cost one more add per update)
                w = f"W_{t}_{u}_{v}_{ic}_{idd}"
                for (tau, nu, ph, cname, sg) in terms:
                    r = rg.get(0, t + tau, u + nu, v + ph)
                    if cname is None:
                        em(f"{w} {'+=' if sg > 0 else '-='} {r};")
                    else:
                        em(f"{w} = fmar({'' if sg > 0 else '-'}{cname}, {r}, {w});")
        em.ind -= 1
        em("}  // ket primitives")
        for ia, (ax, ay, az) in enumerate(ca):
            for ib, (bx, by, bz) in enumerate(cb):
                bterms = []
                for t in range(ax + bx + 1):
                    if (ax, bx, t) not in Ebra["x"]:
                        continue
                    for u in range(ay + by + 1):
                        if (ay, by, u) not in Ebra["y"]:
                            continue
                        for v in range(az + bz + 1):
                            if (az, bz, v) not in Ebra["z"]:
                                continue
                            e, _ = mul(Ebra["x"][(ax, bx, t)], Ebra["y"][(ay, by, u)], Ebra["z"][(az, bz, v)])
                            bn = f"B_{ia}_{ib}_{t}_{u}_{v}"
                            em(f"const R {bn} = bp_K{('*' + e) if e else ''};")
                            bterms.append((t, u, v, bn))
                for (ic, idd) in ks:
                    for (t, u, v, bn) in bterms:
                        em(f"o_{ia}_{ib}_{ic}_{idd} = fmar({bn}, W_{t}_{u}_{v}_{ic}_{idd}, o_{ia}_{ib}_{ic}_{idd});")
        em.ind -= 1
        em("}  // bra primitives")
        if kind == "block":
            for ia in range(NA):
                for ib in range(NB):
                    for (ic, idd) in ks:
                        idx = ((ia * NB + ib) * NC + ic) * ND + idd
                        em(f"atomicAdd(&out[{idx}], o_{ia}_{ib}_{ic}_{idd});")
        else:
            gen_digest_chunk(em, NA, NB, NC, ND, ks, persist)
        em.ind -= 1
        em("}")
    if kind == "jk" and persist:
        gen_digest_final(em, NA, NB, NC, ND)
    em.ind = 0
    em("}")
    if kind == "jk":
        em(f"constexpr int kChunks_{name} = {1 if persist else len(chunks)};")
        em(f"constexpr int kNC_{name} = {NC}, kND_{name} = {ND}, kNA_{name} = {NA}, kNB_{name} = {NB};")
    return "\n".join(em.lines) + "\n"


def _jk_addressing(em: Emitter):
    """Index forms for the Coulomb scatter in each pair's canonical order (s1 >= s2).

    The reference's updates G(1,2) += X(4,3) v, G(3,4) += X(2,1) v (and the transposed
    G(2,1) += X(3,4) v, G(4,3) += X(1,2) v; aointegrals_contract.h:153-157) refer to the
    canonical order inside each pair; the kernel may hold a pair reversed (bra_swap,
    ket_swap).  Address of element (i, j) of a block = base + i*stride_i + j*stride_j with
    i the kernel-order a (or c) component and j the b (or d) component:
      tab: (1,2) = (s1,s2) of the bra   xba: (2,1)
      tcd: (3,4) = (s3,s4) of the ket   xdc: (4,3)"""
    em("const int N = jk.ncart;")
    em("const int LN = jk.lds_n;   // LDS accumulator rows: the swept shells' functions")
    em("const int tab = q.bra_swap ? q.b0 * N + q.a0 : q.a0 * N + q.b0;")
    em("const int tab_a = q.bra_swap ? 1 : N, tab_b = q.bra_swap ? N : 1;")
    em("const int xba = q.bra_swap ? q.a0 * N + q.b0 : q.b0 * N + q.a0;")
    em("const int xba_a = q.bra_swap ? N : 1, xba_b = q.bra_swap ? 1 : N;")
    em("const int tcd = q.ket_swap ? q.d0 * N + q.c0 : q.c0 * N + q.d0;")
    em("const int tcd_c = q.ket_swap ? 1 : N, tcd_d = q.ket_swap ? N : 1;")
    em("const int xdc = q.ket_swap ? q.c0 * N + q.d0 : q.d0 * N + q.c0;")
    em("const int xdc_c = q.ket_swap ? N : 1, xdc_d = q.ket_swap ? 1 : N;")


def _ab(form, ia, ib):
    return f"{form} + {ia} * {form}_a + {ib} * {form}_b"


def _cd(form, ic, idd):
    return f"{form} + {ic} * {form}_c + {idd} * {form}_d"


def gen_digest_chunk(em: Emitter, NA, NB, NC, ND, ks, persist):
    """Digest one chunk of (c,d) components for every density slot.

    Coulomb (forward, canonical in-pair order): G(1,2) += X(4,3) v, G(3,4) += X(2,1) v.
    Exchange (forward, kernel order): G(a,c) += X(b,d) v/2, G(b,d) += X(a,c) v/2,
    G(a,d) += X(b,c) v/2, G(b,c) += X(a,d) v/2.  A non-symmetric density adds the same
    updates taken with X^T into GJ2 / GK2, which the finalisation adds transposed (the
    reference's remaining Coulomb and exchange updates).  With persist, G(1,2), G(a,d) and
    G(b,d) stay in registers across chunks; otherwise each chunk adds its partial sums."""
    cs = sorted(set(ic for ic, _ in ks))
    ds = sorted(set(idd for _, idd in ks))
    em("#pragma unroll")
    em("for (int s = 0; s < NSLOT; ++s) {")
    em.ind += 1
    em("const double* __restrict__ D = jk.D[s];")
    _jk_addressing(em)
    em("const R dg = R(q.deg);")
    em("const R hk = R(0.5 * q.deg);")
    em("const bool ns = jk.nonsym[s] != 0;")
    em("const bool live = q.deg != 0.0;   // lanes without a quartet in this pass skip their atomics")
    em("if (jk.do_j[s]) {")
    em.ind += 1
    if not persist:
        for ia in range(NA):
            for ib in range(NB):
                em(f"R jab_{ia}_{ib} = R(0);")
    for (ic, idd) in ks:
        em(f"{{ const R x43 = R(D[{_cd('xdc', ic, idd)}]);")
        em("  R jcd = R(0);")
        for ia in range(NA):
            for ib in range(NB):
                em(f"  jcd += o_{ia}_{ib}_{ic}_{idd} * R(D[{_ab('xba', ia, ib)}]);")
                tgt = f"jab_{ia}_{ib}[s]" if persist else f"jab_{ia}_{ib}"
                em(f"  {tgt} += o_{ia}_{ib}_{ic}_{idd} * x43;")
        em(f"  if constexpr (DIG == 2) {{ if (live) atomicAdd(&jl[{ic * ND + idd}], double(dg * jcd)); }}")
        em(f"  else wave_add_lds(&jcdL[{ic * ND + idd}], double(dg * jcd)); }}")
    if not persist:
        for ia in range(NA):
            for ib in range(NB):
                em(f"if constexpr (DIG == 2) wave_add_lds(&jcdL[{ia * NB + ib}], double(dg * jab_{ia}_{ib}));")
                em(f"else if (live) atomicAdd(&jl[{ia * NB + ib}], double(dg * jab_{ia}_{ib}));")
    em("if (ns) {")
    em.ind += 1
    for (ic, idd) in ks:
        em(f"{{ const R x34 = dg * R(D[{_cd('tcd', ic, idd)}]);")
        em("  R jcd = R(0);")
        for ia in range(NA):
            for ib in range(NB):
                em(f"  jcd += o_{ia}_{ib}_{ic}_{idd} * R(D[{_ab('tab', ia, ib)}]);")
        for ia in range(NA):
            for ib in range(NB):
                em(f"  if (live) atomicAdd(&jk.GJ2[s][{_ab('tab', ia, ib)}], double(o_{ia}_{ib}_{ic}_{idd} * x34));")
        em(f"  wave_add(&jk.GJ2[s][{_cd('tcd', ic, idd)}], double(dg * jcd), q.ket_uniform); }}")
    em.ind -= 1
    em("}")
    em.ind -= 1
    em("}")
    em("if (jk.do_k[s]) {")
    em.ind += 1

    def X(r0, ri, c0, ci, trans):
        if trans:
            return f"R(D[(q.{c0} + {ci}) * N + q.{r0} + {ri}])"
        return f"R(D[(q.{r0} + {ri}) * N + q.{c0} + {ci}])"

    NCOL = NC + ND

    def kadd(G, trans, row0, ri, colkind, ci, val):
        """One K update: row (row0 + ri), column c0 + ci (colkind 'c') or d0 + ci ('d')."""
        col0 = "q.c0" if colkind == "c" else "q.d0"
        glob = f"atomicAdd(&{G}[(q.{row0} + {ri}) * N + {col0} + {ci}], double({val}))"
        if trans:
            return f"if (live) {glob};"
        colp = "colC" if colkind == "c" else "colD"
        lds = f"atomicAdd(&{colp}[{ci} * LN + q.{row0[0]}r + {ri}], double({val}))"
        rowp = "colC" if row0 == "a0" else "colD"
        ldsb = f"atomicAdd(&{rowp}[{ri} * LN + {col0[:3]}r + {ci}], double({val}))"
        return (f"if (live) {{ if constexpr (DIG == 1) {lds}; else if constexpr (DIG == 2) {ldsb}; "
                f"else {glob}; }}")

    def kupdates(G, trans, with_ad_bd):
        # K(a,c) += X(b,d) ; K(b,c) += X(a,d)   (partial over the chunk's d)
        for ia in range(NA):
            for ic in cs:
                em("{ R acc = R(0);")
                for ib in range(NB):
                    for (jc, idd) in ks:
                        if jc == ic:
                            em(f"  acc += o_{ia}_{ib}_{ic}_{idd} * {X('b0', ib, 'd0', idd, trans)};")
                em(f"  {kadd(G, trans, 'a0', ia, 'c', ic, 'hk * acc')} }}")
        for ib in range(NB):
            for ic in cs:
                em("{ R acc = R(0);")
                for ia in range(NA):
                    for (jc, idd) in ks:
                        if jc == ic:
                            em(f"  acc += o_{ia}_{ib}_{ic}_{idd} * {X('a0', ia, 'd0', idd, trans)};")
                em(f"  {kadd(G, trans, 'b0', ib, 'c', ic, 'hk * acc')} }}")
        if not with_ad_bd:
            return
        # K(a,d) += X(b,c) ; K(b,d) += X(a,c)   (partial over the chunk's c)
        for ia in range(NA):
            for idd in ds:
                em("{ R acc = R(0);")
                for ib in range(NB):
                    for (ic, jd) in ks:
                        if jd == idd:
                            em(f"  acc += o_{ia}_{ib}_{ic}_{idd} * {X('b0', ib, 'c0', ic, trans)};")
                em(f"  {kadd(G, trans, 'a0', ia, 'd', idd, 'hk * acc')} }}")
        for ib in range(NB):
            for idd in ds:
                em("{ R acc = R(0);")
                for ia in range(NA):
                    for (ic, jd) in ks:
                        if jd == idd:
                            em(f"  acc += o_{ia}_{ib}_{ic}_{idd} * {X('a0', ia, 'c0', ic, trans)};")
                em(f"  {kadd(G, trans, 'b0', ib, 'd', idd, 'hk * acc')} }}")

    kupdates("jk.GK[s]", False, not persist)
    if persist:
        for ic in cs:
            for ib in range(NB):
                em(f"{{ const R xbc = {X('b0', ib, 'c0', ic, False)};")
                for ia in range(NA):
                    for (jc, idd) in ks:
                        if jc == ic:
                            em(f"  kad_{ia}_{idd}[s] += o_{ia}_{ib}_{ic}_{idd} * xbc;")
                em("}")
            for ia in range(NA):
                em(f"{{ const R xac = {X('a0', ia, 'c0', ic, False)};")
                for ib in range(NB):
                    for (jc, idd) in ks:
                        if jc == ic:
                            em(f"  kbd_{ib}_{idd}[s] += o_{ia}_{ib}_{ic}_{idd} * xac;")
                em("}")
    em("if (ns) {")
    em.ind += 1
    kupdates("jk.GK2[s]", True, True)
    em.ind -= 1
    em("}")
    em.ind -= 1
    em("}")
    em.ind -= 1
    em("}")


def gen_digest_final(em: Emitter, NA, NB, NC, ND):
    NCOL = NC + ND
    em("#pragma unroll")
    em("for (int s = 0; s < NSLOT; ++s) {")
    em.ind += 1
    em("const int N = jk.ncart;")
    em("const int LN = jk.lds_n;   // LDS accumulator rows: the swept shells' functions")
    em("const int tab = q.bra_swap ? q.b0 * N + q.a0 : q.a0 * N + q.b0;")
    em("const int tab_a = q.bra_swap ? 1 : N, tab_b = q.bra_swap ? N : 1;")
    em("const R dg = R(q.deg);")
    em("const R hk = R(0.5 * q.deg);")
    em("if (jk.do_j[s]) {")
    for ia in range(NA):
        for ib in range(NB):
            em(f"  if constexpr (DIG == 2) wave_add_lds(&jcdL[{ia * NB + ib}], double(dg * jab_{ia}_{ib}[s]));")
            em(f"  else if (q.deg != 0.0) atomicAdd(&jl[{ia * NB + ib}], double(dg * jab_{ia}_{ib}[s]));")
    em("}")
    em("if (q.deg != 0.0 && jk.do_k[s]) {")
    for ia in range(NA):
        for idd in range(ND):
            em(f"  if constexpr (DIG == 1) atomicAdd(&colD[{idd} * LN + q.ar + {ia}], double(hk * kad_{ia}_{idd}[s]));")
            em(f"  else if constexpr (DIG == 2) atomicAdd(&colC[{ia} * LN + q.dr + {idd}], double(hk * kad_{ia}_{idd}[s]));")
            em(f"  else atomicAdd(&jk.GK[s][(q.a0 + {ia}) * N + q.d0 + {idd}], double(hk * kad_{ia}_{idd}[s]));")
    for ib in range(NB):
        for idd in range(ND):
            em(f"  if constexpr (DIG == 1) atomicAdd(&colD[{idd} * LN + q.br + {ib}], double(hk * kbd_{ib}_{idd}[s]));")
            em(f"  else if constexpr (DIG == 2) atomicAdd(&colD[{ib} * LN + q.dr + {idd}], double(hk * kbd_{ib}_{idd}[s]));")
            em(f"  else atomicAdd(&jk.GK[s][(q.b0 + {ib}) * N + q.d0 + {idd}], double(hk * kbd_{ib}_{idd}[s]));")
    em("}")
    em.ind -= 1
    em("}")


def class_list(maxl=2):
    pairs = []
    for la in range(maxl + 1):
        for lb in range(la + 1):
            pairs.append((la, lb))
    return pairs


def orientation(X, Y):
    """Choose which pair class is the bra: the one giving the smaller W array
    (bra Hermite count x ket Cartesian count)."""
    (la, lb), (lc, ld) = X, Y
    w1 = len(hermite_list(la + lb)) * ncart(lc) * ncart(ld)
    w2 = len(hermite_list(lc + ld)) * ncart(la) * ncart(lb)
    return (X, Y) if w1 <= w2 else (Y, X)


def main(out_dir: str, maxl: int = 2):
    """Write one translation unit per class pair plus the registry include."""
    import os
    os.makedirs(out_dir, exist_ok=True)
    pcl = class_list(maxl)
    table = []
    for i, X in enumerate(pcl):
        for j, Y in enumerate(pcl):
            if j > i:
                continue
            bra, ket = orientation(X, Y)
            LA, LB = bra
            LC, LD = ket
            name = f"{CLS_NAME[LA]}{CLS_NAME[LB]}{CLS_NAME[LC]}{CLS_NAME[LD]}"
            L = LA + LB + LC + LD
            src = ["// GENERATED by gen_eri.py -- do not edit.",
                   '#include "../gdirect_kernels.h"', ""]
            src.append(gen_class(LA, LB, LC, LD, "jk"))
            src.append(gen_class(LA, LB, LC, LD, "block"))
            src.append(f"DEFINE_CLASS_KERNELS({name}, {L})")
            # bra-uniform groups pay off where the bra class has far fewer pairs than the ket
            # class (few bras survive per ket): different pair classes and a bra above (ss)
            segb = i != j and (LA, LB) != (0, 0)
            if segb:
                src.append(f"DEFINE_CLASS_SEGB({name}, {L})")
            # ket-keyed groups of an (ss) or (ps) bra class hold many bras per ket: 512-thread
            # workgroups share one Boys table and one set of K columns over 8 wavefronts
            seg_wg = 512 if (LA, LB) in ((0, 0), (1, 0)) else 256
            gb = (lambda R: f"eri_grpb_{name}<{R}>") if segb else (lambda R: "nullptr")
            src.append(f"void cq_register_{name}(ClassKernels& ck) {{")
            src.append(f"  ck = ClassKernels{{{i}, {j}, {pcl.index(bra)}, {pcl.index(ket)}, {L}, kChunks_{name}, \"{name}\", "
                       f"eri_block_{name}, "
                       f"{{{{eri_grp_{name}<{seg_wg}, 0, double>, eri_grp_{name}<{seg_wg}, 1, double>, {gb('double')}}}, "
                       f"{{eri_grp_{name}<{seg_wg}, 0, float>, eri_grp_{name}<{seg_wg}, 1, float>, {gb('float')}}}}}, "
                       f"kNC_{name}, kND_{name}, kNA_{name}, kNB_{name}, {seg_wg}}};")
            src.append("}")
            _write_if_changed(os.path.join(out_dir, f"eri_{name}.hip"), "\n".join(src) + "\n")
            table.append(name)
    reg = ["// GENERATED by gen_eri.py -- do not edit."]
    for name in table:
        reg.append(f"void cq_register_{name}(ClassKernels& ck);")
    reg.append(f"#define CQ_NUM_CLASS_PAIRS {len(table)}")
    reg.append("static inline void cq_register_all(ClassKernels* t) {")
    for k, name in enumerate(table):
        reg.append(f"  cq_register_{name}(t[{k}]);")
    reg.append("}")
    _write_if_changed(os.path.join(out_dir, "eri_registry.inc"), "\n".join(reg) + "\n")


def _write_if_changed(path: str, text: str):
    """Leave an unchanged file untouched so incremental builds skip it."""
    import os
    if os.path.exists(path) and open(path).read() == text:
        return
    with open(path, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 2)
