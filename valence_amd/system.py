"""VALENCE wave-function input: parser, writer and packed (SoA) form.

The text format is the one the reference reads with Fortran list-directed
input in ``xm_getdims`` / ``xm_input`` (src/xm_module.F90:25-69, 76-335):

* line A: 15 integers ``natom natom_t npair nunpd ndocc totlen xpmax nspinc
  num_sh num_pr nang ndf nset nxorb mxctr`` (xm_module.F90:40-41);
* line B: ``ntol_c ntol_d ntol_i ntol_e_min ntol_e_max max_iter ptbnmax
  feather (orbset(1:2,i), i=1..nset)`` (xm_module.F90:106-108);
* one record per atom: ``type x y z`` in Angstrom (xm_module.F90:113-119);
* per atom type: ``charge nshell`` then per shell ``l nprim`` followed by
  ``exponent`` (nprim == 1) or ``nprim`` records ``exponent coeff``
  (xm_module.F90:124-167);
* spin couplings (xm_module.F90:178-198), excited-orbital roots (:202-206);
* per orbital (spin-coupled, unpaired, doubly occupied, NDF in that order):
  ``natoms atom_1..atom_n nterm`` then ``(ao_index coeff)*nterm``
  (xm_module.F90:213-296).

List-directed semantics that matter here: every READ starts on a fresh
record, values may continue over following records, and whatever is left of
the last record a READ touched is skipped.  Blank records carry no values.
A ``/`` ends the READ early (the remaining items keep their prior values).
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

ANGSTROM_TO_BOHR = 1.889725987722  # src/tools_module.F90:11


class InputError(ValueError):
    """Malformed VALENCE input (the reference would call xm_abort)."""


class _ListDirectedReader:
    """Record-oriented reader with Fortran list-directed READ semantics."""

    def __init__(self, text: str):
        self._lines = text.splitlines()
        self._pos = 0
        self._pending: List[str] = []   # unread values of the current statement's record
        self._stopped = False

    def begin(self) -> "_ListDirectedReader":
        """Start a READ statement: it begins on a fresh record."""
        self._pending = []
        self._stopped = False
        return self

    def take(self, count: int) -> List[str]:
        """Pull ``count`` more values for the current statement."""
        values: List[str] = []
        while len(values) < count:
            if self._stopped:
                raise InputError("'/' ended a READ before all its values were given")
            if not self._pending:
                if self._pos >= len(self._lines):
                    raise InputError(f"end of input while reading {count} values (got {len(values)})")
                line = self._lines[self._pos]
                self._pos += 1
                toks: List[str] = []
                for tok in line.replace(",", " ").split():
                    if tok.startswith("/"):
                        self._stopped = True
                        break
                    if "*" in tok:  # r*value repeat form
                        rep, _, val = tok.partition("*")
                        toks.extend([val] * int(rep))
                    else:
                        toks.append(tok)
                self._pending = toks
                continue
            values.append(self._pending.pop(0))
        return values

    def read(self, count: int) -> List[str]:
        """One complete READ of ``count`` values; the rest of its last record is skipped."""
        return self.begin().take(count)


def _f(tok: str) -> float:
    try:
        return float(tok.replace("d", "e").replace("D", "e"))
    except ValueError:
        raise InputError(f"expected a real number, found {tok!r}") from None


def _i(tok: str) -> int:
    try:
        return int(tok)
    except ValueError:
        raise InputError(f"expected an integer, found {tok!r}") from None


@dataclasses.dataclass
class VsvbSystem:
    """One wave function plus its molecule and basis, as the reference holds it.

    Index conventions: atom types, atoms and orbitals are 0-based here;
    ``xpset`` keeps the reference's 1-based OBS-local AO numbering (values < 1
    name NDFs, xm_module.F90 / README).  ``coords`` are in bohr (the reference
    converts with angs2bohr right after reading, valence_initialize_module.F90:22).
    """

    natom: int
    npair: int
    nunpd: int
    ndocc: int
    ndf: int
    nspinc: int
    nang: int
    mxctr: int
    ntol_c: int
    ntol_d: int
    ntol_i: int
    ntol_e_min: int
    ntol_e_max: int
    max_iter: int
    ptbnmax: float
    feather: float
    orbset: np.ndarray          # (nset, 2) int
    atom_type: np.ndarray       # (natom,) int, 0-based type index
    coords: np.ndarray          # (natom, 3) float, bohr
    nuc_charge: np.ndarray      # (ntype,) float
    type_shell_start: np.ndarray  # (ntype+1,) int into shell arrays
    shell_l: np.ndarray         # (nshell_type,) int
    shell_prim_start: np.ndarray  # (nshell_type+1,) int into prim arrays
    prim_exp: np.ndarray        # (nprim,) float
    prim_coef: np.ndarray       # (nprim,) float, unnormalized
    pair_sc: np.ndarray         # (nspinc_eff, npair, 2) int, 1-based electron slots
    coeff_sc: np.ndarray        # (nspinc_eff,) float
    xorb: np.ndarray            # (nxorb,) int
    xroot: np.ndarray           # (nxorb,) int
    orb_natom: np.ndarray       # (norbs,) int
    orb_atoms: np.ndarray       # (norbs, mxctr) int, 0-based atom indices (-1 pad)
    orb_start: np.ndarray       # (norbs+1,) int into xpset/coeff
    xpset: np.ndarray           # (totlen,) int, 1-based OBS AO index or <1 for NDF
    coeff: np.ndarray           # (totlen,) float
    title: str = ""

    # ---- derived sizes -------------------------------------------------
    @property
    def ntype(self) -> int:
        return len(self.nuc_charge)

    @property
    def norbs(self) -> int:
        return len(self.orb_natom)

    @property
    def totlen(self) -> int:
        return len(self.xpset)

    @property
    def nelec(self) -> int:
        return 2 * self.npair + 2 * self.ndocc + self.nunpd

    @property
    def nalpha(self) -> int:
        return self.npair + self.nunpd + self.ndocc

    @property
    def nbeta(self) -> int:
        return self.npair + self.ndocc

    @property
    def nspatial(self) -> int:
        """Orbitals that carry electrons (everything except NDFs)."""
        return 2 * self.npair + self.nunpd + self.ndocc

    def shell_size(self, l: int) -> int:
        return (l + 1) * (l + 2) // 2

    def nbf(self) -> int:
        """Cartesian AO count of the whole molecule."""
        n = 0
        for a in range(self.natom):
            t = self.atom_type[a]
            for s in range(self.type_shell_start[t], self.type_shell_start[t + 1]):
                n += self.shell_size(int(self.shell_l[s]))
        return n

    def copy(self) -> "VsvbSystem":
        return dataclasses.replace(
            self, **{f.name: (getattr(self, f.name).copy() if isinstance(getattr(self, f.name), np.ndarray)
                              else getattr(self, f.name)) for f in dataclasses.fields(self)})

    def validate(self) -> None:
        """Host-side checks the reference leaves to luck (ignored stat= codes)."""
        if self.npair > 0 and self.nspinc < 1:
            raise InputError("no spin couplings;")  # valence_initialize_module.F90:58
        if np.any(self.atom_type < 0) or np.any(self.atom_type >= self.ntype):
            raise InputError("atom type out of range")
        if self.orb_start[-1] != len(self.xpset):
            raise InputError("orbital expansion lengths do not add up to totlen")
        for p in range(self.norbs):
            na = int(self.orb_natom[p])
            if na < 1 or na > self.mxctr:
                raise InputError(f"orbital {p + 1}: OBS atom count {na} outside 1..mxctr")
            if np.any(self.orb_atoms[p, :na] < 0) or np.any(self.orb_atoms[p, :na] >= self.natom):
                raise InputError(f"orbital {p + 1}: OBS atom index out of range")
            nobs = self.obs_size(p)
            xp = self.xpset[self.orb_start[p]:self.orb_start[p + 1]]
            if np.any(xp > nobs):
                raise InputError(f"orbital {p + 1}: AO index beyond its OBS ({nobs})")
            if np.any(xp < 1) and np.any(xp + self.norbs <= self.nspatial):
                raise InputError(f"orbital {p + 1}: NDF reference does not name an NDF")
        if self.npair > 0:
            slots = self.pair_sc.reshape(-1)
            if np.any(slots < 1) or np.any(slots > 2 * self.npair):
                raise InputError("spin-coupling pair names a slot outside 1..2*npair")

    def obs_size(self, p: int) -> int:
        n = 0
        for k in range(int(self.orb_natom[p])):
            t = self.atom_type[self.orb_atoms[p, k]]
            for s in range(self.type_shell_start[t], self.type_shell_start[t + 1]):
                n += self.shell_size(int(self.shell_l[s]))
        return n


def parse_input(text: str) -> VsvbSystem:
    """Parse a VALENCE input deck (see module docstring for the grammar)."""
    r = _ListDirectedReader(text)
    dims = [_i(t) for t in r.read(15)]
    (natom, natom_t, npair, nunpd, ndocc, totlen, xpmax, nspinc,
     num_sh, num_pr, nang, ndf, nset, nxorb, mxctr) = dims
    if natom < 1 or natom_t < 1:
        raise InputError("need at least one atom and one atom type")
    ctl = r.read(8 + 2 * nset)
    ntol_c, ntol_d, ntol_i, ntol_e_min, ntol_e_max, max_iter = (_i(t) for t in ctl[:6])
    ptbnmax, feather = _f(ctl[6]), _f(ctl[7])
    orbset = np.array([_i(t) for t in ctl[8:]], dtype=np.int32).reshape(nset, 2) if nset else np.zeros((0, 2), np.int32)

    atom_type = np.zeros(natom, np.int32)
    coords = np.zeros((natom, 3))
    for a in range(natom):
        v = r.read(4)
        atom_type[a] = _i(v[0]) - 1
        coords[a] = [_f(x) for x in v[1:4]]
    coords *= ANGSTROM_TO_BOHR

    nuc_charge = np.zeros(natom_t)
    type_shell_start = [0]
    shell_l: List[int] = []
    shell_prim_start = [0]
    prim_exp: List[float] = []
    prim_coef: List[float] = []
    for t in range(natom_t):
        v = r.read(2)
        nuc_charge[t] = _f(v[0])
        nshell = _i(v[1])
        for _ in range(nshell):
            lv = r.read(2)
            l, clen = _i(lv[0]), _i(lv[1])
            if l < 0 or l > 2:
                raise InputError(f"angular momentum {l} not supported (s, p, d only)")
            shell_l.append(l)
            if clen == 1:
                prim_exp.append(_f(r.read(1)[0]))
                prim_coef.append(1.0)
            else:
                for _k in range(clen):
                    ev = r.read(2)
                    prim_exp.append(_f(ev[0]))
                    prim_coef.append(_f(ev[1]))
            shell_prim_start.append(len(prim_exp))
        type_shell_start.append(len(shell_l))

    nsc = max(nspinc, 1)
    pair_sc = np.zeros((nsc, npair, 2), np.int32)
    coeff_sc = np.ones(nsc)
    if npair > 0:
        if nspinc == 1:
            v = [_i(x) for x in r.read(2 * npair)]
            pair_sc[0] = np.array(v, np.int32).reshape(npair, 2)
        elif nspinc > 1:
            v = r.read(nspinc * (1 + 2 * npair))
            k = 0
            for j in range(nspinc):
                coeff_sc[j] = _f(v[k]); k += 1
                for i in range(npair):
                    pair_sc[j, i, 0] = _i(v[k]); pair_sc[j, i, 1] = _i(v[k + 1]); k += 2
    xorb = np.zeros(nxorb, np.int32)
    xroot = np.zeros(nxorb, np.int32)
    if nxorb > 0:
        v = [_i(x) for x in r.read(2 * nxorb)]
        xorb[:] = v[0::2]
        xroot[:] = v[1::2]

    norbs = 2 * npair + nunpd + ndocc + ndf
    orb_natom = np.zeros(norbs, np.int32)
    orb_atoms = -np.ones((norbs, max(mxctr, 1)), np.int32)
    orb_start = [0]
    xpset: List[int] = []
    coeff: List[float] = []
    for p in range(norbs):
        st = r.begin()
        na = _i(st.take(1)[0])
        if na < 1 or na > mxctr:
            raise InputError(f"orbital {p + 1}: {na} OBS atoms, mxctr is {mxctr}")
        full = st.take(na + 1)
        orb_natom[p] = na
        orb_atoms[p, :na] = [_i(x) - 1 for x in full[:na]]
        nterm = _i(full[na])
        v = r.read(2 * nterm)
        xpset.extend(_i(x) for x in v[0::2])
        coeff.extend(_f(x) for x in v[1::2])
        orb_start.append(len(xpset))

    sysobj = VsvbSystem(
        natom=natom, npair=npair, nunpd=nunpd, ndocc=ndocc, ndf=ndf, nspinc=nspinc,
        nang=nang, mxctr=max(mxctr, 1), ntol_c=ntol_c, ntol_d=ntol_d, ntol_i=ntol_i,
        ntol_e_min=ntol_e_min, ntol_e_max=ntol_e_max, max_iter=max_iter,
        ptbnmax=ptbnmax, feather=feather, orbset=orbset,
        atom_type=atom_type, coords=coords, nuc_charge=nuc_charge,
        type_shell_start=np.array(type_shell_start, np.int32),
        shell_l=np.array(shell_l, np.int32),
        shell_prim_start=np.array(shell_prim_start, np.int32),
        prim_exp=np.array(prim_exp), prim_coef=np.array(prim_coef),
        pair_sc=pair_sc, coeff_sc=coeff_sc, xorb=xorb, xroot=xroot,
        orb_natom=orb_natom, orb_atoms=orb_atoms,
        orb_start=np.array(orb_start, np.int32),
        xpset=np.array(xpset, np.int32), coeff=np.array(coeff))
    # The header counts size the reference's allocations (read_allocate_input,
    # valence_initialize_module.F90:63-101); inputs routinely over-allocate, and
    # exceeding them would overrun the reference's arrays.
    if len(sysobj.shell_l) > num_sh or len(sysobj.prim_exp) > num_pr:
        raise InputError(f"basis larger than the header allows: shells {len(sysobj.shell_l)} > {num_sh} "
                         f"or primitives {len(sysobj.prim_exp)} > {num_pr}")
    if sysobj.totlen > totlen + xpmax:
        raise InputError(f"orbital expansion length {sysobj.totlen} exceeds totlen+xpmax {totlen + xpmax}")
    sysobj.validate()
    return sysobj


def read_input(path: str) -> VsvbSystem:
    with open(path) as fh:
        return parse_input(fh.read())


def write_input(s: VsvbSystem, orbitals: bool = True) -> str:
    """Write ``s`` in the reference's input format (inverse of parse_input).

    Coordinates go out in Angstrom with 17 significant digits so a
    parse/write/parse round trip is exact.  orbitals=False stops before the
    orbital records (and the title), for a deck completed by an 'orbitals'
    file (format_orbitals), as the reference's `cat input orbitals` does.
    """
    out: List[str] = []
    nxorb = len(s.xorb)
    out.append(" ".join(str(x) for x in (
        s.natom, s.ntype, s.npair, s.nunpd, s.ndocc, s.totlen, _xpmax(s), s.nspinc,
        len(s.shell_l), len(s.prim_exp), s.nang, s.ndf, len(s.orbset), nxorb, s.mxctr)))
    ctl = [s.ntol_c, s.ntol_d, s.ntol_i, s.ntol_e_min, s.ntol_e_max, s.max_iter]
    out.append(" ".join(str(x) for x in ctl) + f" {float(s.ptbnmax)!r} {float(s.feather)!r} "
               + " ".join(f"{a} {b}" for a, b in s.orbset))
    for a in range(s.natom):
        x, y, z = (s.coords[a] / ANGSTROM_TO_BOHR)
        out.append(f"{s.atom_type[a] + 1} {x:.17g} {y:.17g} {z:.17g}")
    for t in range(s.ntype):
        sh0, sh1 = s.type_shell_start[t], s.type_shell_start[t + 1]
        out.append(f"{float(s.nuc_charge[t])!r} {sh1 - sh0}")
        for sh in range(sh0, sh1):
            p0, p1 = s.shell_prim_start[sh], s.shell_prim_start[sh + 1]
            out.append(f"{s.shell_l[sh]} {p1 - p0}")
            if p1 - p0 == 1:
                out.append(f"{float(s.prim_exp[p0])!r}")
                if s.prim_coef[p0] != 1.0:
                    raise InputError("single-primitive shells carry unit weight in this format")
            else:
                for k in range(p0, p1):
                    out.append(f"{float(s.prim_exp[k])!r} {float(s.prim_coef[k])!r}")
    if s.npair > 0:
        if s.nspinc == 1:
            out.append(" ".join(f"{a} {b}" for a, b in s.pair_sc[0]))
        else:
            for j in range(s.nspinc):
                out.append(f"{float(s.coeff_sc[j])!r} " + " ".join(f"{a} {b}" for a, b in s.pair_sc[j]))
    if nxorb:
        out.append(" ".join(f"{a} {b}" for a, b in zip(s.xorb, s.xroot)))
    if not orbitals:
        return "\n".join(out) + "\n"
    for p in range(s.norbs):
        na = s.orb_natom[p]
        o0, o1 = s.orb_start[p], s.orb_start[p + 1]
        out.append(f"{na} " + " ".join(str(a + 1) for a in s.orb_atoms[p, :na]) + f" {o1 - o0}")
        out.append(" ".join(f"{s.xpset[k]} {float(s.coeff[k])!r}" for k in range(o0, o1)))
    out.append("")
    if s.title:
        out.append(s.title)
    return "\n".join(out) + "\n"


# ---------------------------------------------------------------------------
# The 'orbitals' and 'nelecwfn' files (xm_output, src/xm_module.F90:464-577).
# The reference writes them with Fortran formats; their text is restated here
# field for field so that a file from either side reads the same:
#   1  (5x,1i2,5x,10i4)          orbital header: natoms, atoms, nterm
#   2  (4(1i4,1x,1f13.8))        orbital terms: AO index, weight (4 a record)
#   3  (1x,7(2i3,2x))            orbital pairs (one coupling) / excited roots
#   4  (1x,1f13.8,7(2i3,2x))     coupling weight and orbital pairs
#   5  (/,1x,'total energy in atomic units ',1f32.16)
#   7  (1x,'converged to ',1e10.2,' kCal/mol')
# Format reversion (more items than one pass of a format takes) starts a new
# record at the last top-level group, or at the beginning when there is none;
# position edits after the last item write nothing (gfortran).
# ---------------------------------------------------------------------------

def _fi(v: int, w: int) -> str:
    t = str(int(v))
    return t.rjust(w) if len(t) <= w else "*" * w


def _ff(v: float, w: int, d: int) -> str:
    t = f"{float(v):{w}.{d}f}"
    return t if len(t) <= w else "*" * w


def _fe(v: float, w: int, d: int) -> str:
    """Fortran Ew.d: sign, 0.ddd, E, signed two-digit exponent, right-justified."""
    x = float(v)
    if x == 0.0:
        mant, ex = 0, 0
    else:
        ex = int(np.floor(np.log10(abs(x)))) + 1
        mant = int(round(abs(x) / 10.0 ** ex * 10 ** d))
        if mant >= 10 ** d:  # rounding carried into a new digit
            mant //= 10
            ex += 1
    t = ("-" if x < 0 else "") + "0." + str(mant).rjust(d, "0") + "E" + ("-" if ex < 0 else "+") + f"{abs(ex):02d}"
    return t.rjust(w) if len(t) <= w else "*" * w


def _fmt_header(items: Sequence[int]) -> List[str]:
    """Format 1: each record takes one item as i2 and then up to ten as i4."""
    lines, k = [], 0
    while k < len(items):
        line = "     " + _fi(items[k], 2)
        k += 1
        if k < len(items):
            line += "     " + "".join(_fi(x, 4) for x in items[k:k + 10])
            k += 10
        lines.append(line)
    return lines


def _fmt_terms(idx: Sequence[int], w: Sequence[float]) -> List[str]:
    """Format 2: (i4, 1x, f13.8) pairs, four to a record."""
    pairs = [_fi(a, 4) + " " + _ff(b, 13, 8) for a, b in zip(idx, w)]
    return ["".join(pairs[k:k + 4]) for k in range(0, len(pairs), 4)]


def _fmt_pairs(lead: str, pairs: Sequence[Sequence[int]]) -> List[str]:
    """Formats 3 and 4: ``lead`` (1x, or 1x and a weight), then (2i3, 2x)
    groups, seven to a record; continuation records repeat only the group."""
    groups = [_fi(a, 3) + _fi(b, 3) for a, b in pairs]
    if not groups:
        return [lead.rstrip()]
    lines = []
    for k in range(0, len(groups), 7):
        lines.append((lead if k == 0 else "") + "  ".join(groups[k:k + 7]))
    return lines


def format_nelecwfn(s: VsvbSystem) -> str:
    """The 'nelecwfn' file the reference writes when nspinc > 1 (xm_module.F90:480-503):
    every coupling's weight and orbital pairs, then the excited-orbital roots,
    then a blank record.  Empty for nspinc <= 1 (the reference writes none)."""
    if s.nspinc <= 1:
        return ""
    out: List[str] = []
    for j in range(s.nspinc):
        out += _fmt_pairs(" " + _ff(s.coeff_sc[j], 13, 8), [tuple(p) for p in s.pair_sc[j]])
    out += _fmt_pairs(" ", list(zip(s.xorb, s.xroot)))
    out.append("")
    return "\n".join(out) + "\n"


def format_orbitals(s: VsvbSystem, energy: float, mode: str = "save", etol: float = 0.0) -> str:
    """The 'orbitals' file (xm_module.F90:506-570): every orbital (spin-coupled,
    unpaired, doubly occupied), a blank record, the NDFs, then the total energy,
    with mode 'done' the convergence tolerance etol (kcal/mol), and a blank
    record.  The orbital records are the orbital part of an input deck, which
    is how the reference continues a run from them (valence_api.F90:70-80)."""
    out: List[str] = []
    nreg = 2 * s.npair + s.nunpd + s.ndocc

    def orbital(p: int):
        na = int(s.orb_natom[p])
        o0, o1 = int(s.orb_start[p]), int(s.orb_start[p + 1])
        out.extend(_fmt_header([na] + [int(a) + 1 for a in s.orb_atoms[p, :na]] + [o1 - o0]))
        out.extend(_fmt_terms(s.xpset[o0:o1], s.coeff[o0:o1]))

    for p in range(nreg):
        orbital(p)
    out.append("")
    for p in range(nreg, nreg + s.ndf):
        orbital(p)
    out.append("")
    out.append(" total energy in atomic units " + _ff(energy, 32, 16))
    if mode == "done":
        out.append(" converged to " + _fe(etol, 10, 2) + " kCal/mol")
    out.append("")
    return "\n".join(out) + "\n"


def write_orbitals(s: VsvbSystem, energy: float, directory: str = ".", mode: str = "save", etol: float = 0.0,
                   append: bool = False) -> None:
    """xm_output: 'orbitals' (and 'nelecwfn' when nspinc > 1) in ``directory``,
    replaced or (append) extended."""
    import os
    how = "a" if append else "w"
    if s.nspinc > 1:
        with open(os.path.join(directory, "nelecwfn"), how) as fh:
            fh.write(format_nelecwfn(s))
    with open(os.path.join(directory, "orbitals"), how) as fh:
        fh.write(format_orbitals(s, energy, mode, etol))


def _xpmax(s: VsvbSystem) -> int:
    return max([s.obs_size(p) for p in range(s.norbs)] + [1])


# ---------------------------------------------------------------------------
# Packed, C-ABI-ready view
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Packed:
    """Contiguous int32/float64 arrays in the order the C structs expect."""
    i32: dict
    f64: dict
    scalars: dict


def pack(s: VsvbSystem) -> Packed:
    i32 = dict(
        atom_type=np.ascontiguousarray(s.atom_type, np.int32),
        type_shell_start=np.ascontiguousarray(s.type_shell_start, np.int32),
        shell_l=np.ascontiguousarray(s.shell_l, np.int32),
        shell_prim_start=np.ascontiguousarray(s.shell_prim_start, np.int32),
        orb_natom=np.ascontiguousarray(s.orb_natom, np.int32),
        orb_atoms=np.ascontiguousarray(s.orb_atoms, np.int32),
        orb_start=np.ascontiguousarray(s.orb_start, np.int32),
        xpset=np.ascontiguousarray(s.xpset, np.int32),
        pair_sc=np.ascontiguousarray(s.pair_sc, np.int32),
    )
    f64 = dict(
        coords=np.ascontiguousarray(s.coords, np.float64),
        nuc_charge=np.ascontiguousarray(s.nuc_charge, np.float64),
        prim_exp=np.ascontiguousarray(s.prim_exp, np.float64),
        prim_coef=np.ascontiguousarray(s.prim_coef, np.float64),
        coeff=np.ascontiguousarray(s.coeff, np.float64),
        coeff_sc=np.ascontiguousarray(s.coeff_sc, np.float64),
    )
    scalars = dict(natom=s.natom, ntype=s.ntype, nshell_type=len(s.shell_l), nprim=len(s.prim_exp),
                   norbs=s.norbs, mxctr=s.mxctr, totlen=s.totlen, npair=s.npair, nunpd=s.nunpd,
                   ndocc=s.ndocc, ndf=s.ndf, nspinc=s.nspinc, ntol_d=s.ntol_d, ntol_i=s.ntol_i)
    return Packed(i32, f64, scalars)


def nuclear_repulsion(s: VsvbSystem) -> float:
    """E_nuc exactly as get_nuclear_repulsion_energy (src/tools_module.F90:19-40)."""
    e = 0.0
    for i in range(1, s.natom):
        for j in range(i):
            zij = s.nuc_charge[s.atom_type[i]] * s.nuc_charge[s.atom_type[j]]
            if abs(zij) > 1e-12:
                d = s.coords[i] - s.coords[j]
                rsq = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
                if rsq > 1e-5:
                    e += zij * rsq ** -0.5
    return e


def spin_orbital_slots(s: VsvbSystem) -> Sequence[int]:
    """Electron slot -> spatial orbital (0-based), as guess_energy builds bra/ket
    (src/valence.F90:324-336): non-docc slots map 1:1, each docc orbital takes
    two consecutive slots (alpha then beta)."""
    nnd = 2 * s.npair + s.nunpd
    slots = list(range(nnd))
    for d in range(nnd, nnd + s.ndocc):
        slots += [d, d]
    return slots


def ensure(cond: bool, msg: str) -> None:
    if not cond:
        raise InputError(msg)


def maybe(x: Optional[np.ndarray]) -> np.ndarray:
    return np.zeros(0) if x is None else x
