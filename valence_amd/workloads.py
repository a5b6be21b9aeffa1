"""Seeded synthetic wave functions for the benchmark and parity configurations.

Each generator returns a VsvbSystem (bohr coordinates) built from the recipe
of one workload; the seed is 1000 + workload index.  Choices the recipes leave
open, fixed here and used everywhere:

* Primitive counts are read literally, shell by shell: s shells take 6, 3, 1
  primitives in order and p shells 3, 1 (so an H with 2s1p gets s(6) s(3) p(3),
  and 3s1p gets s(6) s(3) s(1) p(3)); d shells take 1.
* Exponents of a shell are drawn log-uniform in the element's range and sorted
  descending; contraction weights U(0.1, 1).  One basis per element type.
* Orbital weights are N(0,1) over every AO of the orbital's OBS atoms and then
  scaled to unit Euclidean norm (the energy path renormalizes by S_pp anyway).
* ntol_d = 24 so the Givens rotation skip (src/givens.F90:245) never alters a
  determinant; ntol_i = 12 (Schwarz 1e-12) for every workload; max_iter = 0.
* Alkanes are all-trans, staggered, C-C 1.54 A, C-H 1.09 A, tetrahedral angles.
* A Givens trial rotation mixes two orbitals with angle U(-0.05, 0.05); the
  rotated orbitals are expanded over the union of the two OBS.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np

from .system import ANGSTROM_TO_BOHR, VsvbSystem

SEED_BASE = 1000
S_PRIMS = [6, 3, 1]
P_PRIMS = [3, 1]
D_PRIMS = [1]

# element -> (Z, shells "3s2p1d", exponent range)
ELEMENTS: Dict[str, Tuple[float, Tuple[int, int, int], Tuple[float, float]]] = {
    "O": (8.0, (3, 2, 1), (0.1, 5000.0)),
    "H": (1.0, (2, 1, 0), (0.1, 20.0)),
    "C": (6.0, (3, 2, 1), (0.1, 3000.0)),
    "Hc": (1.0, (3, 1, 0), (0.05, 30.0)),   # H of the H12 chain: 3s1p, [0.05, 30]
    "Li": (3.0, (3, 2, 0), (0.05, 700.0)),
    "F": (9.0, (3, 2, 1), (0.1, 10000.0)),
}


def _basis_for(rng: np.random.Generator, elem: str):
    z, (ns, np_, nd), (lo, hi) = ELEMENTS[elem]
    shells = []
    for l, count, prims in ((0, ns, S_PRIMS), (1, np_, P_PRIMS), (2, nd, D_PRIMS)):
        for k in range(count):
            nprim = prims[k] if k < len(prims) else 1
            ex = np.sort(np.exp(rng.uniform(math.log(lo), math.log(hi), nprim)))[::-1]
            co = rng.uniform(0.1, 1.0, nprim) if nprim > 1 else np.ones(1)
            shells.append((l, ex, co))
    return z, shells


def _ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def assemble(elems: Sequence[str], coords_ang: np.ndarray, basis: Dict[str, tuple],
             orbitals: Sequence[Tuple[Sequence[int], np.ndarray]], *, npair: int = 0, nunpd: int = 0,
             pair_sc=None, ntol_i: int = 12, ntol_d: int = 24, title: str = "") -> VsvbSystem:
    """Build a VsvbSystem: orbitals are (OBS atom list, weights over the OBS AOs)."""
    types = []
    for e in elems:
        if e not in types:
            types.append(e)
    atom_type = np.array([types.index(e) for e in elems], np.int32)
    nuc = np.array([basis[t][0] for t in types])
    tss, sl, sps, pe, pc = [0], [], [0], [], []
    for t in types:
        for (l, ex, co) in basis[t][1]:
            sl.append(l)
            pe.extend(ex.tolist())
            pc.extend(co.tolist())
            sps.append(len(pe))
        tss.append(len(sl))
    norbs = len(orbitals)
    mxctr = max(len(o[0]) for o in orbitals)
    orb_natom = np.array([len(o[0]) for o in orbitals], np.int32)
    orb_atoms = -np.ones((norbs, mxctr), np.int32)
    xp, cf, starts = [], [], [0]
    for p, (atoms, w) in enumerate(orbitals):
        orb_atoms[p, :len(atoms)] = atoms
        xp.extend(range(1, len(w) + 1))
        cf.extend(np.asarray(w, float).tolist())
        starts.append(len(xp))
    ndocc = norbs - npair * 2 - nunpd
    ps = np.zeros((1, npair, 2), np.int32) if pair_sc is None else np.asarray(pair_sc, np.int32).reshape(1, npair, 2)
    s = VsvbSystem(
        natom=len(elems), npair=npair, nunpd=nunpd, ndocc=ndocc, ndf=0, nspinc=1 if npair else 0,
        nang=max(sl), mxctr=mxctr, ntol_c=0, ntol_d=ntol_d, ntol_i=ntol_i, ntol_e_min=0, ntol_e_max=0,
        max_iter=0, ptbnmax=0.0, feather=0.0, orbset=np.zeros((0, 2), np.int32),
        atom_type=atom_type, coords=np.asarray(coords_ang, float) * ANGSTROM_TO_BOHR, nuc_charge=nuc,
        type_shell_start=np.array(tss, np.int32), shell_l=np.array(sl, np.int32),
        shell_prim_start=np.array(sps, np.int32), prim_exp=np.array(pe), prim_coef=np.array(pc),
        pair_sc=ps, coeff_sc=np.ones(1), xorb=np.zeros(0, np.int32), xroot=np.zeros(0, np.int32),
        orb_natom=orb_natom, orb_atoms=orb_atoms, orb_start=np.array(starts, np.int32),
        xpset=np.array(xp, np.int32), coeff=np.array(cf), title=title)
    s.validate()
    return s


def _nao(basis, elem) -> int:
    return sum(_ncart(l) for (l, _, _) in basis[elem][1])


def _rand_orbital(rng, basis, elems, atoms) -> np.ndarray:
    n = sum(_nao(basis, elems[a]) for a in atoms)
    w = rng.standard_normal(n)
    return w / np.linalg.norm(w)


# ---------------------------------------------------------------------------
# geometries
# ---------------------------------------------------------------------------
TET = math.acos(-1.0 / 3.0)  # 109.47 degrees


def alkane_geometry(n: int, r_cc: float = 1.54, r_ch: float = 1.09):
    """All-trans staggered CnH2n+2 (Angstrom).  Returns elements, coords, C-H bond list."""
    half = TET / 2
    dx, dy = r_cc * math.sin(half), r_cc * math.cos(half)
    C = np.array([[k * dx, (k % 2) * dy, 0.0] for k in range(n)])
    elems = ["C"] * n
    coords = [c for c in C]
    bonds_ch: List[Tuple[int, int]] = []

    def add_h(ci, direction):
        d = direction / np.linalg.norm(direction)
        coords.append(C[ci] + r_ch * d)
        elems.append("H")
        bonds_ch.append((ci, len(elems) - 1))

    ez = np.array([0.0, 0.0, 1.0])
    for k in range(n):
        nbrs = [j for j in (k - 1, k + 1) if 0 <= j < n]
        if n == 1:
            # methane: tetrahedron
            for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]):
                add_h(k, np.array(v, float))
            continue
        if len(nbrs) == 2:
            b1 = (C[nbrs[0]] - C[k]); b1 /= np.linalg.norm(b1)
            b2 = (C[nbrs[1]] - C[k]); b2 /= np.linalg.norm(b2)
            u = -(b1 + b2); u /= np.linalg.norm(u)
            w = np.cross(b1, b2); w /= np.linalg.norm(w)
            for sgn in (1.0, -1.0):
                add_h(k, u * math.cos(half) + sgn * w * math.sin(half))
        else:
            j = nbrs[0]
            b = (C[j] - C[k]); b /= np.linalg.norm(b)
            # virtual continuation of the zigzag: anti to the next carbon, in plane
            virt = np.array([(k - (j - k)) * dx, ((k - (j - k)) % 2) * dy, 0.0])
            h0 = virt - C[k]; h0 /= np.linalg.norm(h0)
            for ang in (0.0, 2 * math.pi / 3, 4 * math.pi / 3):
                add_h(k, _rotate(h0, b, ang))
    return elems, np.array(coords), bonds_ch


def _rotate(v, axis, ang):
    axis = axis / np.linalg.norm(axis)
    return (v * math.cos(ang) + np.cross(axis, v) * math.sin(ang) + axis * np.dot(axis, v) * (1 - math.cos(ang)))


def alkane_system(n: int, basis, rng, coords_override=None, title: str = "") -> VsvbSystem:
    elems, xyz, bonds_ch = alkane_geometry(n)
    if coords_override is not None:
        xyz = coords_override
    orbs = []
    for k in range(n):                       # C 1s cores
        orbs.append(([k], _rand_orbital(rng, basis, elems, [k])))
    for k in range(n - 1):                   # C-C bonds
        orbs.append(([k, k + 1], _rand_orbital(rng, basis, elems, [k, k + 1])))
    for (c, h) in bonds_ch:                  # C-H bonds
        orbs.append(([c, h], _rand_orbital(rng, basis, elems, [c, h])))
    return assemble(elems, xyz, basis, orbs, title=title)


# ---------------------------------------------------------------------------
# the five workloads
# ---------------------------------------------------------------------------
def config0() -> VsvbSystem:
    """Water VSHF: O 3s2p1d, H 2s1p, five docc (core, two lone pairs, two O-H)."""
    rng = np.random.default_rng(SEED_BASE + 0)
    basis = {e: _basis_for(rng, e) for e in ("O", "H")}
    r, th = 0.9572, math.radians(104.5)
    xyz = np.array([[0, 0, 0], [r * math.sin(th / 2), r * math.cos(th / 2), 0],
                    [-r * math.sin(th / 2), r * math.cos(th / 2), 0]])
    elems = ["O", "H", "H"]
    orbs = [([0], _rand_orbital(rng, basis, elems, [0])) for _ in range(3)]
    orbs += [([0, 1], _rand_orbital(rng, basis, elems, [0, 1])), ([0, 2], _rand_orbital(rng, basis, elems, [0, 2]))]
    return assemble(elems, xyz, basis, orbs, title="water VSHF (seed 1000)")


def _alkane_basis(rng):
    return {"C": _basis_for(rng, "C"), "H": _basis_for(rng, "H")}


def config1() -> VsvbSystem:
    """C8H18 VSHF: 33 docc (8 C cores, 7 C-C, 18 C-H)."""
    rng = np.random.default_rng(SEED_BASE + 1)
    basis = _alkane_basis(rng)
    return alkane_system(8, basis, rng, title="C8H18 VSHF (seed 1001)")


def config2() -> VsvbSystem:
    """H12 chain, six Rumer-coupled pairs (1,2),(3,4),...; orbital i over atoms i-1, i, i+1."""
    rng = np.random.default_rng(SEED_BASE + 2)
    basis = {"Hc": _basis_for(rng, "Hc")}
    spacing = 1.4 / ANGSTROM_TO_BOHR
    xyz = np.array([[0.0, 0.0, k * spacing] for k in range(12)])
    elems = ["Hc"] * 12
    orbs = []
    for i in range(12):
        atoms = [j for j in (i - 1, i, i + 1) if 0 <= j < 12]
        orbs.append((atoms, _rand_orbital(rng, basis, elems, atoms)))
    pairs = [(2 * k + 1, 2 * k + 2) for k in range(6)]
    return assemble(elems, xyz, basis, orbs, npair=6, pair_sc=pairs, title="H12 spin-coupled (seed 1002)")


def rotate_pair(s: VsvbSystem, p: int, q: int, theta: float) -> VsvbSystem:
    """Givens rotation of orbitals p, q; both expanded over the union of their OBS."""
    out = s.copy()
    atoms_p = list(s.orb_atoms[p, :s.orb_natom[p]])
    atoms_q = list(s.orb_atoms[q, :s.orb_natom[q]])
    union = atoms_p + [a for a in atoms_q if a not in atoms_p]

    def dense(o):
        # weights of orbital o over the union OBS
        atoms = list(s.orb_atoms[o, :s.orb_natom[o]])
        sizes = {a: _atom_nao(s, a) for a in union}
        off = {}
        k = 0
        for a in union:
            off[a] = k
            k += sizes[a]
        w = np.zeros(k)
        local_off = {}
        k2 = 0
        for a in atoms:
            local_off[a] = k2
            k2 += sizes[a]
        for idx in range(s.orb_start[o], s.orb_start[o + 1]):
            x = s.xpset[idx] - 1
            for a in atoms:
                if local_off[a] <= x < local_off[a] + sizes[a]:
                    w[off[a] + x - local_off[a]] += s.coeff[idx]
                    break
        return w

    wp, wq = dense(p), dense(q)
    c, sn = math.cos(theta), math.sin(theta)
    new = {p: c * wp + sn * wq, q: -sn * wp + c * wq}
    orbs = []
    for o in range(s.norbs):
        if o in new:
            orbs.append((union, new[o]))
        else:
            atoms = list(s.orb_atoms[o, :s.orb_natom[o]])
            orbs.append((atoms, s.coeff[s.orb_start[o]:s.orb_start[o + 1]].copy()))
    return _rebuild_orbitals(out, orbs)


def _atom_nao(s: VsvbSystem, a: int) -> int:
    t = s.atom_type[a]
    return sum(_ncart(int(s.shell_l[k])) for k in range(s.type_shell_start[t], s.type_shell_start[t + 1]))


def _rebuild_orbitals(s: VsvbSystem, orbs) -> VsvbSystem:
    mxctr = max(len(o[0]) for o in orbs)
    s.mxctr = mxctr
    s.orb_natom = np.array([len(o[0]) for o in orbs], np.int32)
    s.orb_atoms = -np.ones((len(orbs), mxctr), np.int32)
    xp, cf, st = [], [], [0]
    for p, (atoms, w) in enumerate(orbs):
        s.orb_atoms[p, :len(atoms)] = atoms
        xp.extend(range(1, len(w) + 1))
        cf.extend(np.asarray(w).tolist())
        st.append(len(xp))
    s.orb_start = np.array(st, np.int32)
    s.xpset = np.array(xp, np.int32)
    s.coeff = np.array(cf)
    s.validate()
    return s


def config3(n_displaced: int = 48, n_rotated: int = 32, n_alkanes: int = 16) -> List[VsvbSystem]:
    """96 wave functions: displaced C8H18 geometries, trial orbital rotations of one
    C8H18 wave function, and CnH2n+2 with n ~ U{2..12}; all with the C8H18 basis."""
    rng = np.random.default_rng(SEED_BASE + 3)
    base = config1()
    basis = {"C": (6.0, _type_shells(base, 0)), "H": (1.0, _type_shells(base, 1))}
    out: List[VsvbSystem] = []
    elems, xyz0, _ = alkane_geometry(8)
    for i in range(n_displaced):
        s = base.copy()
        s.coords = (xyz0 + rng.normal(0.0, 0.02, xyz0.shape)) * ANGSTROM_TO_BOHR
        s.title = f"C8H18 displaced #{i}"
        out.append(s)
    for i in range(n_rotated):
        s = base
        used = set()
        for _ in range(3):
            while True:
                p, q = rng.choice(base.norbs, 2, replace=False)
                if p not in used and q not in used:
                    break
            used |= {p, q}
            s = rotate_pair(s, int(p), int(q), float(rng.uniform(-0.05, 0.05)))
        s.title = f"C8H18 rotated #{i}"
        out.append(s)
    for i in range(n_alkanes):
        n = int(rng.integers(2, 13))
        out.append(alkane_system(n, basis, rng, title=f"C{n}H{2 * n + 2} #{i}"))
    return out


def _type_shells(s: VsvbSystem, t: int):
    shells = []
    for k in range(s.type_shell_start[t], s.type_shell_start[t + 1]):
        p0, p1 = s.shell_prim_start[k], s.shell_prim_start[k + 1]
        shells.append((int(s.shell_l[k]), s.prim_exp[p0:p1].copy(), s.prim_coef[p0:p1].copy()))
    return shells


def lif_slab(nx: int, ny: int, nz: int, seed: int = SEED_BASE + 4) -> VsvbSystem:
    """LiF rock-salt slab nx x ny x nz (alternating Li/F, 2.01 A); atom-centred
    docc: Li 1s cores, F 1s cores and four F valence pairs."""
    rng = np.random.default_rng(seed)
    basis = {"Li": _basis_for(rng, "Li"), "F": _basis_for(rng, "F")}
    a = 2.01
    elems, xyz = [], []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                elems.append("Li" if (i + j + k) % 2 == 0 else "F")
                xyz.append([i * a, j * a, k * a])
    xyz = np.array(xyz)
    orbs = []
    for at, e in enumerate(elems):
        nv = 1 if e == "Li" else 5
        for _ in range(nv):
            orbs.append(([at], _rand_orbital(rng, basis, elems, [at])))
    return assemble(elems, xyz, basis, orbs, title=f"LiF {nx}x{ny}x{nz} slab (seed {seed})")


def config4() -> VsvbSystem:
    """LiF rock-salt slab 4x4x8 (128 atoms, 2.01 A)."""
    return lif_slab(4, 4, 8)


def lif8() -> VsvbSystem:
    """The 2x2x2 corner of the LiF recipe (8 atoms, same seed and basis recipe):
    small enough for the oracle's full energy."""
    return lif_slab(2, 2, 2)


CONFIGS = {0: config0, 1: config1, 2: config2, 3: config3, 4: config4}
EXTRA = {"lif8": lif8}
