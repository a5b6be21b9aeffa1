#!/usr/bin/env python3
"""Generate straight-line FP64 ERI device code for every (s,p,d) shell-quartet class.

Method (written from the published recurrences, not from any library):

* Obara-Saika vertical recurrence (VRR) on primitive quartets builds
  [e0|f0]^(0) for |e| in [la, la+lb] and |f| in [lc, lc+ld] from Boys
  values F_m(T), m = 0..L:

    [e+1i,0|00]^(m) = PA_i [e0|00]^(m) + WP_i [e0|00]^(m+1)
                      + e_i/(2p) ([e-1i,0|00]^(m) - rho/p [e-1i,0|00]^(m+1))
    [e0|f+1i,0]^(m) = QC_i [e0|f0]^(m) + WQ_i [e0|f0]^(m+1)
                      + f_i/(2q) ([e0|f-1i,0]^(m) - rho/q [e0|f-1i,0]^(m+1))
                      + e_i/(2(p+q)) [e-1i,0|f0]^(m+1)

* The [e0|f0] are contracted over all primitive quartets (accumulators).
* Head-Gordon-Pople horizontal recurrence (HRR) on contracted integrals:

    (a, b+1i| = (a+1i, b| + AB_i (a, b|        |c, d+1i) = |c+1i, d) + CD_i |c, d)

Every class gets one device function

    eri_<la><lb><lc><ld>(bra prim pairs, ket prim pairs, AB, CD, boys table, out[])

writing the contracted cartesian block in CCA order with the A index slowest
(out[((a*nB + b)*nC + c)*nD + d]).  Classes are canonical: la >= lb, lc >= ld and
the bra pair class is not below the ket pair class (pair class order below).
Run as a script it writes ``eri_gen.inc`` next to itself and prints the FP64
operation count per class (used for the algorithmic FLOP model).
"""
from __future__ import annotations

import json
import os
import re
import sys
from functools import lru_cache
from typing import Dict, List, Tuple

LMAX = 2
PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]  # order used everywhere


def cart(l: int) -> List[Tuple[int, int, int]]:
    """CCA order (valence.F90:2364-2378): i=0..l, j=0..i -> (l-i, i-j, j)."""
    out = []
    for i in range(l + 1):
        for j in range(i + 1):
            out.append((l - i, i - j, j))
    return out


def quartet_classes() -> List[Tuple[int, int, int, int]]:
    out = []
    for bi, (la, lb) in enumerate(PAIR_CLASSES):
        for ki in range(bi + 1):
            lc, ld = PAIR_CLASSES[ki]
            out.append((la, lb, lc, ld))
    return out


def sub(v, i, k=1):
    w = list(v)
    w[i] -= k
    return tuple(w)


def add(v, i, k=1):
    w = list(v)
    w[i] += k
    return tuple(w)


def name3(v):
    return f"{v[0]}{v[1]}{v[2]}"


class Emitter:
    def __init__(self):
        self.lines: List[str] = []
        self.flops = 0

    def emit(self, s: str):
        self.lines.append(s)


def vrr_plan(la, lb, lc, ld):
    """Intermediates [e0|f0]^(m) needed for the class, in dependency order, with
    the recurrence used for each (greedy choice of the decrement direction)."""
    targets = [(e, f) for le in range(la, la + lb + 1) for e in cart(le)
               for lf in range(lc, lc + ld + 1) for f in cart(lf)]
    order: List[Tuple] = []
    recipe: Dict[Tuple, Tuple] = {}
    seen = set()

    def need(e, f, m):
        if min(e) < 0 or min(f) < 0:
            return None
        key = (e, f, m)
        if key in seen:
            return key
        # choose the decrement direction greedily: fewest new intermediates
        if sum(f) > 0:
            best = None
            for i in range(3):
                if f[i] == 0:
                    continue
                f1 = sub(f, i)
                deps = [(e, f1, m), (e, f1, m + 1)]
                if f1[i] > 0:
                    deps += [(e, sub(f1, i), m), (e, sub(f1, i), m + 1)]
                if e[i] > 0:
                    deps += [(sub(e, i), f1, m + 1)]
                new = sum(1 for d in deps if d not in seen)
                if best is None or new < best[0]:
                    best = (new, i)
            i = best[1]
            f1 = sub(f, i)
            parts = [("QC", i, need(e, f1, m)), ("WQ", i, need(e, f1, m + 1))]
            if f1[i] > 0:
                parts.append(("F2Q", i, f1[i], need(e, sub(f1, i), m), need(e, sub(f1, i), m + 1)))
            if e[i] > 0:
                parts.append(("E2S", i, e[i], need(sub(e, i), f1, m + 1)))
            recipe[key] = ("ket", parts)
        elif sum(e) > 0:
            best = None
            for i in range(3):
                if e[i] == 0:
                    continue
                e1 = sub(e, i)
                deps = [(e1, f, m), (e1, f, m + 1)]
                if e1[i] > 0:
                    deps += [(sub(e1, i), f, m), (sub(e1, i), f, m + 1)]
                new = sum(1 for d in deps if d not in seen)
                if best is None or new < best[0]:
                    best = (new, i)
            i = best[1]
            e1 = sub(e, i)
            parts = [("PA", i, need(e1, f, m)), ("WP", i, need(e1, f, m + 1))]
            if e1[i] > 0:
                parts.append(("E2P", i, e1[i], need(sub(e1, i), f, m), need(sub(e1, i), f, m + 1)))
            recipe[key] = ("bra", parts)
        else:
            recipe[key] = ("base", m)
        seen.add(key)
        order.append(key)
        return key

    for (e, f) in targets:
        need(e, f, 0)
    return order, recipe, targets


def gen_vrr(la, lb, lc, ld, em: Emitter):
    """Emit VRR for one primitive quartet; returns dict target->(var name)."""
    order, recipe, targets = vrr_plan(la, lb, lc, ld)

    def var(k):
        e, f, m = k
        return f"v{name3(e)}_{name3(f)}_{m}"

    xyz = "xyz"
    for k in order:
        r = recipe[k]
        if r[0] == "base":
            em.emit(f"    const double {var(k)} = F{r[1]};")
            continue
        terms = []
        fl = 0
        if r[0] == "bra":
            for part in r[1]:
                if part[0] == "PA":
                    terms.append(f"PA{xyz[part[1]]} * {var(part[2])}"); fl += 2
                elif part[0] == "WP":
                    terms.append(f"WP{xyz[part[1]]} * {var(part[2])}"); fl += 2
                elif part[0] == "E2P":
                    n = part[2]
                    terms.append(f"{float(n)!r} * oo2p * ({var(part[3])} - rop * {var(part[4])})"); fl += 5
        else:
            for part in r[1]:
                if part[0] == "QC":
                    terms.append(f"QC{xyz[part[1]]} * {var(part[2])}"); fl += 2
                elif part[0] == "WQ":
                    terms.append(f"WQ{xyz[part[1]]} * {var(part[2])}"); fl += 2
                elif part[0] == "F2Q":
                    n = part[2]
                    terms.append(f"{float(n)!r} * oo2q * ({var(part[3])} - roq * {var(part[4])})"); fl += 5
                elif part[0] == "E2S":
                    n = part[2]
                    terms.append(f"{float(n)!r} * oo2s * {var(part[3])}"); fl += 3
        em.flops += fl - 1
        em.emit(f"    const double {var(k)} = " + " + ".join(terms) + ";")
    return {(e, f): var((e, f, 0)) for (e, f) in targets}, len(order), targets


def gen_hrr(la, lb, lc, ld, accname, em: Emitter):
    """Emit bra then ket HRR on contracted accumulators; returns out index map."""
    fs = [f for lf in range(lc, lc + ld + 1) for f in cart(lf)]
    # bra: (a,b|f) for |a|=la, |b|=lb
    seen = {}
    order = []

    def bra(a, b):
        key = (a, b)
        if key in seen:
            return key
        if sum(b) == 0:
            seen[key] = ("base",)
        else:
            i = next(i for i in range(3) if b[i] > 0)
            b1 = sub(b, i)
            seen[key] = ("rec", i, bra(add(a, i), b1), bra(a, b1))
        order.append(key)
        return key

    finals_ab = [(a, b) for a in cart(la) for b in cart(lb)]
    for a, b in finals_ab:
        bra(a, b)

    def bvar(key, f):
        a, b = key
        if sum(b) == 0:
            return accname[(a, f)]
        return f"hb{name3(a)}_{name3(b)}_{name3(f)}"

    xyz = "xyz"
    for key in order:
        r = seen[key]
        if r[0] == "base":
            continue
        for f in fs:
            em.emit(f"    const double {bvar(key, f)} = {bvar(r[2], f)} + AB{xyz[r[1]]} * {bvar(r[3], f)};")
            em.flops += 2
    # ket: (ab|c,d) for |c|=lc, |d|=ld from (ab|f0)
    kseen = {}
    korder = []

    def ket(c, d):
        key = (c, d)
        if key in kseen:
            return key
        if sum(d) == 0:
            kseen[key] = ("base",)
        else:
            i = next(i for i in range(3) if d[i] > 0)
            d1 = sub(d, i)
            kseen[key] = ("rec", i, ket(add(c, i), d1), ket(c, d1))
        korder.append(key)
        return key

    finals_cd = [(c, d) for c in cart(lc) for d in cart(ld)]
    for c, d in finals_cd:
        ket(c, d)

    def kvar(ab, key):
        c, d = key
        if sum(d) == 0:
            return bvar(ab, c)
        return f"hk{name3(ab[0])}{name3(ab[1])}_{name3(c)}_{name3(d)}"

    out_map = {}
    nB, nC, nD = len(cart(lb)), len(cart(lc)), len(cart(ld))
    for ia, a in enumerate(cart(la)):
        for ib, b in enumerate(cart(lb)):
            ab = (a, b)
            for key in korder:
                r = kseen[key]
                if r[0] == "base":
                    continue
                em.emit(f"    const double {kvar(ab, key)} = {kvar(ab, r[2])} + CD{xyz[r[1]]} * {kvar(ab, r[3])};")
                em.flops += 2
            for ic, c in enumerate(cart(lc)):
                for idd, d in enumerate(cart(ld)):
                    idx = ((ia * nB + ib) * nC + ic) * nD + idd
                    em.emit(f"    out[{idx}] = {kvar(ab, (c, d))};")
    return out_map


# Primitive-loop option per class (off by default): KET_UNROLL[class] = 2
# emits two ket primitives per iteration as interleaved independent chains.
# The Boys function of a primitive quartet comes from the class's compact table
# (boys_eval_c), which the quartet kernels stage in LDS.
KET_UNROLL: Dict[Tuple[int, int, int, int], int] = {}


def _suffix_block(lines, suf):
    """Rename every local declared in a block of generated lines (and K)."""
    names = set()
    for ln in lines:
        if re.match(r"\s*(const double|const PrimPair|double)\s", ln):
            names.update(re.findall(r"(?:const double|const PrimPair|double|,)\s+(\w+)\s*(?:=(?!=)|\[)", ln))
    names.add("K")
    pat = re.compile(r"(?<![.\w])(" + "|".join(sorted(names, key=len, reverse=True)) + r")\b")
    return [pat.sub(lambda m: m.group(1) + suf, ln) for ln in lines]


def gen_class(la, lb, lc, ld, uniform_bra=False):
    """One class's contracted-quartet routine.  By default the ket primitive pairs
    are wave-uniform (scalar loads) and the bra's are per lane; uniform_bra swaps
    the two (the transposed J/K kernel: one bra pair per wavefront, ket pairs per
    lane).  The arithmetic is the same."""
    L = la + lb + lc + ld
    em = Emitter()
    name = f"eri_{la}{lb}{lc}{ld}" + ("_ub" if uniform_bra else "")
    em.emit(f"// class ({'spd'[la]}{'spd'[lb]}|{'spd'[lc]}{'spd'[ld]}), L = {L}")
    em.emit(f"__device__ __forceinline__ void {name}(const PrimPair* __restrict__ bp, int nbp,")
    em.emit(f"        const PrimPair* __restrict__ kp, int nkp, double ABx, double ABy, double ABz,")
    em.emit(f"        double CDx, double CDy, double CDz, const double* __restrict__ btab, double* __restrict__ out) {{")
    body = Emitter()
    accmap, nvrr, targets = gen_vrr(la, lb, lc, ld, body)
    acc = {t: f"acc{n}" for n, t in enumerate(targets)}
    em.emit("    " + " ".join(f"double {acc[t]} = 0.0;" for t in targets))
    if uniform_bra:
        em.emit("    const UniformPrimPairs bu = as_uniform(bp);  // the bra list is the same for the whole wavefront")
    else:
        em.emit("    const UniformPrimPairs kc = as_uniform(kp);  // the ket list is the same for the whole wavefront")
    em.emit("    for (int ib = 0; ib < nbp; ++ib) {")
    em.emit("      const PrimPair B = " + ("load_uniform(bu + ib);" if uniform_bra else "bp[ib];"))
    if la + lb > 0:
        em.emit("      const double PAx = B.PAx, PAy = B.PAy, PAz = B.PAz, oo2p = B.oo2p;")
    blk = Emitter()  # one primitive quartet; K is the ket primitive pair
    blk.emit("        const PrimPair K = " + ("kp[IK];" if uniform_bra else "load_uniform(kc + IK);"))
    blk.emit("        const double s = B.p + K.p;")
    blk.emit("        const double r = rsqrt_fp64(s), rs = r * r;")
    blk.emit("        const double PQx = B.Px - K.Px, PQy = B.Py - K.Py, PQz = B.Pz - K.Pz;")
    blk.emit("        const double rho = B.p * K.p * rs;")
    blk.emit("        const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);")
    blk.emit("        const double pref = B.K * K.K * r;")
    blk.emit(f"        double Fb[{L + 1}];")
    blk.emit(f"        boys_eval_c<{L}>(T, pref, btab, Fb);  // pref * F_m(T)")
    for m in range(L + 1):
        blk.emit(f"        const double F{m} = Fb[{m}];")
    if la + lb > 0:
        blk.emit("        const double WPx = -K.p * rs * PQx, WPy = -K.p * rs * PQy, WPz = -K.p * rs * PQz;")
        blk.emit("        const double rop = K.p * rs;")
    if lc + ld > 0:
        blk.emit("        const double QCx = K.PAx, QCy = K.PAy, QCz = K.PAz, oo2q = K.oo2p;")
        blk.emit("        const double WQx = B.p * rs * PQx, WQy = B.p * rs * PQy, WQz = B.p * rs * PQz;")
        blk.emit("        const double roq = B.p * rs;")
        if la + lb > 0:
            blk.emit("        const double oo2s = 0.5 * rs;")
    for ln in body.lines:
        blk.emit("    " + ln)
    for t in targets:
        blk.emit(f"        {acc[t]} += {accmap[t]};")
    unroll = KET_UNROLL.get((la, lb, lc, ld), 1)
    if unroll == 2:
        b0 = _suffix_block(blk.lines, "_u0")
        b1 = _suffix_block(blk.lines, "_u1")
        em.emit("      int ik = 0;")
        em.emit("      for (; ik + 1 < nkp; ik += 2) {")
        for x, y in zip(b0, b1):  # interleaved: two independent chains
            em.emit(x.replace("IK", "ik"))
            em.emit(y.replace("IK", "ik + 1"))
        em.emit("      }")
        em.emit("      if (ik < nkp) {")
        for x in b0:
            em.emit(x.replace("IK", "ik"))
        em.emit("      }")
    else:
        em.emit("      for (int ik = 0; ik < nkp; ++ik) {")
        for x in blk.lines:
            em.emit(x.replace("IK", "ik"))
        em.emit("      }")
    em.flops += blk.flops
    em.emit("    }")
    hr = Emitter()
    gen_hrr(la, lb, lc, ld, acc, hr)
    em.lines += hr.lines
    em.emit("}")
    # per primitive quartet: setup ~ 25 flops + boys (approx 2*8 + 3*L + exp~20) + VRR + accumulate
    per_prim = body.flops + len(targets) + 25 + 16 + (3 * L + 20 if L > 0 else 0)
    info = dict(cls=f"{la}{lb}{lc}{ld}", L=L, vrr_terms=nvrr, targets=len(targets),
                flops_per_primitive_quartet=per_prim, flops_hrr=hr.flops)
    return "\n".join(em.lines), info


# Classes evaluated by the cooperative kernel (one wavefront per contracted
# quartet, intermediates in LDS): the (dd| classes with p or d kets, whose
# one-lane-per-quartet code spills thousands of registers and whose quartets are
# few.  Classes with many smaller quartets ((dp|pp), (dp|dp), (dd|ds)) measured
# faster on the straight-line kernels despite their spills.
COOP_DEFAULT = [(2, 2, 1, 1), (2, 2, 2, 1), (2, 2, 2, 2)]
# Tables exist for every class whose straight-line code spills; the batch
# tuner may move the others onto the cooperative kernel.
COOP_CLASSES = [(2, 1, 1, 1), (2, 1, 2, 0), (2, 1, 2, 1), (2, 2, 1, 0), (2, 2, 1, 1), (2, 2, 2, 0),
                (2, 2, 2, 1), (2, 2, 2, 2)]


def coop_plan(la, lb, lc, ld):
    """Step tables of one class for the cooperative kernel.

    LDS slots: [0, nT) the contracted targets [e0|f0] (accumulated over primitive
    quartets); [nT, nT + nV) the VRR intermediates of one primitive quartet;
    after the primitive loop the HRR intermediates reuse [nT, nT + nH).
    VRR step (8 ints): out, a, b, c, d, e, code, 0 with
      code = kind | dir << 2 | n1 << 4 | n2 << 8, kind 1 = bra, 2 = ket:
      bra: out = PA_dir S[a] + WP_dir S[b] + n1 oo2p (S[c] - rop S[d])
      ket: out = QC_dir S[a] + WQ_dir S[b] + n1 oo2q (S[c] - roq S[d]) + n2 oo2s S[e]
    (absent terms: n1 or n2 = 0).  Base (2 ints): slot, m: S[slot] = pref F_m.
    HRR step (4 ints): out, a, b, code = dir | 4 * (ket): out = S[a] + (AB or CD)_dir S[b].
    Steps are grouped in levels; a level only reads earlier levels."""
    order, recipe, targets = vrr_plan(la, lb, lc, ld)
    nT = len(targets)
    acc = {t: i for i, t in enumerate(targets)}
    vslot = {k: nT + i for i, k in enumerate(order)}
    base, lev, vsteps = [], {}, []
    for k in order:
        r = recipe[k]
        if r[0] == "base":
            base.append((vslot[k], r[1]))
            lev[k] = 0
            continue
        a = b = c = d = e = -1
        n1 = n2 = 0
        deps = []
        for part in r[1]:
            if part[0] in ("PA", "QC"):
                a = vslot[part[2]]; deps.append(part[2]); dr = part[1]
            elif part[0] in ("WP", "WQ"):
                b = vslot[part[2]]; deps.append(part[2])
            elif part[0] in ("E2P", "F2Q"):
                n1 = part[2]; c = vslot[part[3]]; d = vslot[part[4]]; deps += [part[3], part[4]]
            elif part[0] == "E2S":
                n2 = part[2]; e = vslot[part[3]]; deps.append(part[3])
        kind = 1 if r[0] == "bra" else 2
        lev[k] = 1 + max(lev[x] for x in deps)
        vsteps.append((lev[k], (vslot[k], a, b, c, d, e, kind | dr << 2 | n1 << 4 | n2 << 8, 0)))
    vsteps.sort(key=lambda x: x[0])
    nvl = max([l for l, _ in vsteps] + [0])
    vlev = [0] * (nvl + 1)
    for l, _ in vsteps:
        vlev[l] += 1
    voff = [0]
    for l in range(1, nvl + 1):
        voff.append(voff[-1] + vlev[l])
    targ = [(acc[t], vslot[(t[0], t[1], 0)]) for t in targets]
    # HRR
    fs = [f for lf in range(lc, lc + ld + 1) for f in cart(lf)]
    nxt = [nT]
    hlev_of = {}
    hsteps = []

    def new_slot():
        v = nxt[0]
        nxt[0] += 1
        return v

    bseen, border = {}, []

    def bra(a_, b_):
        key = (a_, b_)
        if key in bseen:
            return key
        if sum(b_) == 0:
            bseen[key] = ("base",)
        else:
            i = next(i for i in range(3) if b_[i] > 0)
            b1 = sub(b_, i)
            bseen[key] = ("rec", i, bra(add(a_, i), b1), bra(a_, b1))
        border.append(key)
        return key

    for a_ in cart(la):
        for b_ in cart(lb):
            bra(a_, b_)
    bslot = {}

    def bvar(key, f):
        if sum(key[1]) == 0:
            return acc[(key[0], f)]
        return bslot[(key, f)]

    for key in border:
        r = bseen[key]
        if r[0] == "base":
            continue
        for f in fs:
            o = new_slot()
            bslot[(key, f)] = o
            sa, sb = bvar(r[2], f), bvar(r[3], f)
            hlev_of[o] = 1 + max(hlev_of.get(sa, 0), hlev_of.get(sb, 0))
            hsteps.append((hlev_of[o], (o, sa, sb, r[1])))
    kseen, korder = {}, []

    def ket(c_, d_):
        key = (c_, d_)
        if key in kseen:
            return key
        if sum(d_) == 0:
            kseen[key] = ("base",)
        else:
            i = next(i for i in range(3) if d_[i] > 0)
            d1 = sub(d_, i)
            kseen[key] = ("rec", i, ket(add(c_, i), d1), ket(c_, d1))
        korder.append(key)
        return key

    for c_ in cart(lc):
        for d_ in cart(ld):
            ket(c_, d_)
    kslot = {}

    def kvar(ab, key):
        if sum(key[1]) == 0:
            return bvar(ab, key[0])
        return kslot[(ab, key)]

    out_slot = []
    for a_ in cart(la):
        for b_ in cart(lb):
            ab = (a_, b_)
            for key in korder:
                r = kseen[key]
                if r[0] == "base":
                    continue
                o = new_slot()
                kslot[(ab, key)] = o
                sa, sb = kvar(ab, r[2]), kvar(ab, r[3])
                hlev_of[o] = 1 + max(hlev_of.get(sa, 0), hlev_of.get(sb, 0))
                hsteps.append((hlev_of[o], (o, sa, sb, r[1] | 4)))
            for c_ in cart(lc):
                for d_ in cart(ld):
                    out_slot.append(kvar(ab, (c_, d_)))
    hsteps.sort(key=lambda x: x[0])
    nhl = max([l for l, _ in hsteps] + [0])
    hcount = [0] * (nhl + 1)
    for l, _ in hsteps:
        hcount[l] += 1
    hoff = [0]
    for l in range(1, nhl + 1):
        hoff.append(hoff[-1] + hcount[l])
    nV, nH = len(order), nxt[0] - nT
    return dict(nT=nT, nV=nV, nH=nH, lds=nT + max(nV, nH), base=base, voff=voff,
                vsteps=[st for _, st in vsteps], targ=targ, hoff=hoff, hsteps=[st for _, st in hsteps],
                out=out_slot)


def write_coop(path: str) -> None:
    """coop_tables.inc: one flat int array and a descriptor per cooperative class."""
    data: List[int] = []
    descs = []

    def align4():  # every section starts on a 16-byte boundary (the kernel reads int4 records)
        while len(data) % 4:
            data.append(0)

    for cls in COOP_CLASSES:
        t = coop_plan(*cls)
        d = dict(cls=cls, nT=t["nT"], nV=t["nV"], nH=t["nH"], lds=t["lds"])
        align4()
        d["base"] = len(data); d["nbase"] = len(t["base"])
        for sl, m in t["base"]:
            data += [sl, m]
        align4()
        d["voff"] = len(data); d["nvl"] = len(t["voff"]) - 1
        data += t["voff"]
        align4()
        d["vsteps"] = len(data)
        for st in t["vsteps"]:
            data += list(st)
        align4()
        d["targ"] = len(data)
        for a_, v_ in t["targ"]:
            data += [a_, v_]
        align4()
        d["hoff"] = len(data); d["nhl"] = len(t["hoff"]) - 1
        data += t["hoff"]
        align4()
        d["hsteps"] = len(data)
        for st in t["hsteps"]:
            data += list(st)
        align4()
        d["out"] = len(data)
        data += t["out"]
        descs.append(d)
    lines = ["// Generated by gen_eri.py -- do not edit.",
             "// Step tables of the cooperative (one wavefront per quartet) ERI kernel.",
             f"constexpr int kNCoop = {len(descs)};",
             f"constexpr int kCoopDataLen = {len(data)};",
             "// device-resident from module load (no allocation or upload at run time)",
             "static __device__ __attribute__((aligned(16))) int kCoopData[kCoopDataLen] = {"]
    for i in range(0, len(data), 24):
        lines.append("  " + ", ".join(str(x) for x in data[i:i + 24]) + ",")
    lines.append("};")
    lines.append("static const CoopDesc kCoopDesc[kNCoop] = {")
    for d in descs:
        la, lb, lc, ld = d["cls"]
        dflt = 1 if d["cls"] in COOP_DEFAULT else 0
        lines.append(f"  {{{la}, {lb}, {lc}, {ld}, {d['nT']}, {d['nV']}, {d['nH']}, {d['lds']}, {d['nbase']}, {d['base']}, "
                     f"{d['nvl']}, {d['voff']}, {d['vsteps']}, {d['targ']}, {d['nhl']}, {d['hoff']}, {d['hsteps']}, {d['out']}, "
                     f"{dflt}}},")
    lines.append("};")
    text = "\n".join(lines) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)


def main(out_path: str) -> Dict:
    chunks = ["// Generated by gen_eri.py -- do not edit.", "",
              "template <int LA, int LB, int LC, int LD>",
              "__device__ void eri_block(const PrimPair* __restrict__ bp, int nbp, const PrimPair* __restrict__ kp,",
              "    int nkp, double ABx, double ABy, double ABz, double CDx, double CDy, double CDz,",
              "    const double* __restrict__ btab, double* __restrict__ out);",
              "template <int LA, int LB, int LC, int LD>",
              "__device__ void eri_block_ub(const PrimPair* __restrict__ bp, int nbp, const PrimPair* __restrict__ kp,",
              "    int nkp, double ABx, double ABy, double ABz, double CDx, double CDy, double CDz,",
              "    const double* __restrict__ btab, double* __restrict__ out);", ""]
    infos = []
    for (la, lb, lc, ld) in quartet_classes():
        code, info = gen_class(la, lb, lc, ld)
        chunks.append(code)
        if (la, lb) != (lc, ld):
            code_ub, _ = gen_class(la, lb, lc, ld, uniform_bra=True)
            chunks.append(code_ub)
            chunks.append(f"template <> __device__ __forceinline__ void eri_block_ub<{la}, {lb}, {lc}, {ld}>(")
            chunks.append("    const PrimPair* __restrict__ bp, int nbp, const PrimPair* __restrict__ kp, int nkp,")
            chunks.append("    double ABx, double ABy, double ABz, double CDx, double CDy, double CDz,")
            chunks.append("    const double* __restrict__ btab, double* __restrict__ out) {")
            chunks.append(f"  eri_{la}{lb}{lc}{ld}_ub(bp, nbp, kp, nkp, ABx, ABy, ABz, CDx, CDy, CDz, btab, out);")
            chunks.append("}")
        chunks.append(f"template <> __device__ __forceinline__ void eri_block<{la}, {lb}, {lc}, {ld}>(")
        chunks.append("    const PrimPair* __restrict__ bp, int nbp, const PrimPair* __restrict__ kp, int nkp,")
        chunks.append("    double ABx, double ABy, double ABz, double CDx, double CDy, double CDz,")
        chunks.append("    const double* __restrict__ btab, double* __restrict__ out) {")
        chunks.append(f"  eri_{la}{lb}{lc}{ld}(bp, nbp, kp, nkp, ABx, ABy, ABz, CDx, CDy, CDz, btab, out);")
        chunks.append("}")
        chunks.append("")
        infos.append(info)
    text = "\n".join(chunks) + "\n"
    # algorithmic FP64 operation counts per class, in launch order (bra class major)
    flops = ["// Generated by gen_eri.py -- do not edit.",
             "// FP64 operations per primitive quartet (Boys, VRR, accumulate) and per",
             "// contracted quartet (HRR), in quartet-class order.",
             "constexpr double kFlopsPerPrimQuartet[21] = {" +
             ", ".join(f"{i['flops_per_primitive_quartet']}.0" for i in infos) + "};",
             "constexpr double kFlopsHrr[21] = {" + ", ".join(f"{i['flops_hrr']}.0" for i in infos) + "};", ""]
    fpath = os.path.join(os.path.dirname(out_path), "eri_flops.inc")
    ftext = "\n".join(flops)
    if not os.path.exists(fpath) or open(fpath).read() != ftext:
        with open(fpath, "w") as fh:
            fh.write(ftext)
    if not os.path.exists(out_path) or open(out_path).read() != text:
        with open(out_path, "w") as fh:
            fh.write(text)
    return {i["cls"]: i for i in infos}


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    infos = main(os.path.join(here, "eri_gen.inc"))
    write_coop(os.path.join(here, "coop_tables.inc"))
    json.dump(infos, sys.stdout, indent=1)
